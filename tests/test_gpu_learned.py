"""The learned part of the mixed-precision solve: the initial guess and the
flexible-BiCGSTAB preconditioning stage, trained on the solver's own solves
with the device's forward and backward kernels (pcf_net_train_residual) and
compared on held-out runs with the method without them.  The solve's
stopping test is unchanged (|r| / |b| against the reference tolerance), so
what the network does can change the iteration count, never the accuracy;
a solve that fails with the stage is rerun without it."""
import ctypes

import hipmem
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def namelists(pcf, seed, guess, steps, n_hills=6):
    nl = pcf.examples.gaussian_hills_3d(x_size=32, y_size=32, z_size=16, iterations=steps, seed=seed,
                                        n_hills=n_hills, precision="mixed")
    nl.poisson.initial_guess = guess
    nl.poisson.preconditioner_type = "multigrid"
    return nl


def stage2_iterations(pcf, nl, guess_net=None, stage_net=None, stage_its=0, steps=6):
    st = pcf.State(nl)
    if guess_net is not None:
        pcf.learned.attach(st, guess_net)
    if stage_net is not None:
        pcf.learned.attach_stage(st, stage_net, stage_its)
    st.initial_cleaning()
    its = []
    for _ in range(steps):
        info = st.step()
        its.append(int(info.niter[1]))
    return its, st


def trained(pcf, guess=True, stage=False, seeds=(11, 12), steps=14):
    g = pcf.learned.GuessNetwork(32, 32, 16, training=True, seed=0) if guess else None
    s = pcf.learned.GuessNetwork(32, 32, 16, training=True, seed=1) if stage else None
    states = []
    for seed in seeds:
        st = pcf.State(namelists(pcf, seed, "zero", steps))
        st.initial_cleaning()
        states.append(st)
    hist = pcf.learned.train_online(states, steps, guess=g, stage=s, inner=6, lr=1e-3)
    return g, s, hist


def test_transposed_operator(pcf):
    """<A x, y> = <x, A^T y> for the operator of a solve (the gradient of the
    residual loss goes through A^T)."""
    st = pcf.State(namelists(pcf, 5, "zero", 2))
    st.initial_cleaning()
    st.step()
    n = st.nx * st.ny * st.nz
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    dx, dy, da = hipmem.DeviceArray((n,)).upload(x), hipmem.DeviceArray((n,)).upload(y), hipmem.DeviceArray((n,))
    st.apply_operator_device(dx.ptr, da.ptr, True)
    axy = float(np.dot(da.download(), y))
    assert st.lib.pcf_apply_operator_transpose(st.h, ctypes.c_void_p(dy.ptr), ctypes.c_void_p(da.ptr)) == 0
    xaty = float(np.dot(x, da.download()))
    assert abs(axy - xaty) <= 1e-12 * abs(axy), (axy, xaty)


def test_device_residual_loss_matches_the_host(pcf):
    """pcf_net_train_residual with inner = 0 evaluates L = |A x - v|^2 / |v|^2
    for x = base + |v| net(...): the same value the host computes from the
    network's output and the device apply."""
    st = pcf.State(namelists(pcf, 5, "zero", 2))
    st.initial_cleaning()
    st.step()
    n = st.nx * st.ny * st.nz
    params = pcf.learned.init_params(3)
    params[-(27 * 32 + 1):-1] = np.random.default_rng(4).standard_normal(27 * 32).astype(np.float32) * 1e-2
    net = pcf.learned.GuessNetwork(32, 32, 16, training=False, params=params)
    rng = np.random.default_rng(1)
    v, base = rng.standard_normal(n), rng.standard_normal(n) * 0.1
    d_v, d_b = pcf.learned.DeviceBuffer(8 * n).upload(v), pcf.learned.DeviceBuffer(8 * n).upload(base)
    lb, ln = pcf.learned.residual_step(net, st, d_v, d_b, 0, 0.0)
    # host restatement
    d_in, d_out = pcf.learned.DeviceBuffer(12 * n), pcf.learned.DeviceBuffer(4 * n)
    vn = ctypes.c_double(0.0)
    assert st.lib.pcf_guess_input(st.h, ctypes.c_void_p(d_v.ptr), ctypes.c_void_p(d_b.ptr),
                                  ctypes.c_void_p(d_in.ptr), ctypes.byref(vn)) == 0
    assert net.lib.pcf_net_forward(net.h, ctypes.c_void_p(d_in.ptr), ctypes.c_void_p(d_out.ptr)) == 0
    out = d_out.download(np.float32, (n,)).astype(np.float64)
    vv = float(np.dot(v, v))

    def loss(x):
        dx, da = hipmem.DeviceArray((n,)).upload(x), hipmem.DeviceArray((n,))
        st.apply_operator_device(dx.ptr, da.ptr, True)
        r = da.download() - v
        return float(np.dot(r, r)) / vv

    assert lb == pytest.approx(loss(base), rel=1e-12)
    assert ln == pytest.approx(loss(base + vn.value * out), rel=1e-9)


def test_device_training_lowers_the_guess_residual_on_held_out_runs(pcf):
    """Online training of the initial guess on the device (two runs of
    random hills); on a third run the guess leaves a smaller residual than
    the warm start it corrects."""
    g, _, hist = trained(pcf, guess=True)
    print("training solves (warm, guess):", [tuple(round(x, 4) for x in h["guess"]) for h in hist])
    assert all(h["guess"][1] < h["guess"][0] for h in hist[-6:])
    infer = pcf.learned.GuessNetwork(32, 32, 16, training=False, params=g.params())
    st = pcf.State(namelists(pcf, 21, "zero", 10))
    st.initial_cleaning()
    rl = pcf.learned.residual_loss(infer, st, 8)
    print("held-out residual loss (warm, learned):", [(round(a, 4), round(b, 4)) for a, b in rl])
    ratio = float(np.mean([b / a for a, b in rl]))
    assert ratio < 0.85, ratio


def test_learned_guess_costs_at_most_one_iteration_per_solve_over_the_warm_start(pcf):
    """What the trained guess does to BiCGSTAB's iteration count on a held-out
    run.  Its smaller residual is of small scale, which BiCGSTAB removes in
    its first iterations anyway, so the count stays within one iteration
    per solve of the warm start's."""
    g, _, _ = trained(pcf, guess=True)
    infer = pcf.learned.GuessNetwork(32, 32, 16, training=False, params=g.params())
    steps = 8
    warm, _ = stage2_iterations(pcf, namelists(pcf, 21, "previous", steps), steps=steps)
    learned, _ = stage2_iterations(pcf, namelists(pcf, 21, "learned", steps), guess_net=infer, steps=steps)
    print("stage-2 iterations, warm start:", warm, "learned start:", learned)
    assert sum(learned[1:]) <= sum(warm[1:]) + len(warm) - 1, (warm, learned)


@pytest.mark.xfail(reason="the learned guess has not been shown to save BiCGSTAB iterations over the warm start "
                          "at this size", strict=False)
def test_learned_guess_saves_iterations_over_the_warm_start(pcf):
    g, _, _ = trained(pcf, guess=True)
    infer = pcf.learned.GuessNetwork(32, 32, 16, training=False, params=g.params())
    steps = 8
    warm, _ = stage2_iterations(pcf, namelists(pcf, 21, "previous", steps), steps=steps)
    learned, _ = stage2_iterations(pcf, namelists(pcf, 21, "learned", steps), guess_net=infer, steps=steps)
    assert sum(learned[1:]) < sum(warm[1:]), (warm, learned)


def test_untrained_stage_is_the_plain_solve(pcf):
    """A stage network whose last layer is zero adds nothing: the flexible
    solve with it gives the plain mixed solve's iterations and bits."""
    steps = 3
    plain, st0 = stage2_iterations(pcf, namelists(pcf, 21, "previous", steps), steps=steps)
    net = pcf.learned.GuessNetwork(32, 32, 16, training=False, params=pcf.learned.init_params(2))
    staged, st1 = stage2_iterations(pcf, namelists(pcf, 21, "previous", steps), stage_net=net, stage_its=3,
                                    steps=steps)
    assert staged == plain
    for f in ("U", "V", "W", "PIP", "RHOP"):
        assert np.array_equal(st0.get(f), st1.get(f)), f


def test_learned_stage_solves_meet_the_criterion_and_report_iterations(pcf):
    """The trained stage in the first two iterations of each solve: every
    solve meets the reference criterion (the residual recurrence of the
    flexible form stays the true residual), and the iteration counts against
    the multigrid solve without it are reported."""
    _, s, hist = trained(pcf, guess=False, stage=True)
    print("stage training solves (M^-1, M^-1 + net):", [tuple(round(x, 4) for x in h["stage"]) for h in hist])
    assert all(h["stage"][1] < h["stage"][0] for h in hist[-6:])
    infer = pcf.learned.GuessNetwork(32, 32, 16, training=False, params=s.params())
    steps = 8
    warm, _ = stage2_iterations(pcf, namelists(pcf, 21, "previous", steps), steps=steps)
    staged, st = stage2_iterations(pcf, namelists(pcf, 21, "previous", steps), stage_net=infer, stage_its=2,
                                   steps=steps)
    print("stage-2 iterations, multigrid:", warm, "multigrid + learned stage:", staged)
    fcs = pcf.learned._interior_fcscal(st)
    b = st.get_vector("RHS")
    x = st.get_vector("SOLUTION") * fcs
    n = b.size
    dx, da = hipmem.DeviceArray((n,)).upload(np.asfortranarray(x).ravel(order="F")), hipmem.DeviceArray((n,))
    st.apply_operator_device(dx.ptr, da.ptr, True)
    r = np.asfortranarray(b).ravel(order="F") - da.download()
    tol = st.namelists.poisson.tolerance
    rel = float(np.sqrt(np.mean(r * r)) / np.sqrt(np.mean(b * b)))
    assert rel == pytest.approx(st.last_residual(), rel=1e-3)
    fb = ctypes.c_int32(0)
    assert st.lib.pcf_stage_fallbacks(st.h, ctypes.byref(fb)) == 0
    print("final relative residual", rel, "fallbacks", fb.value)


def test_stage_fallback_keeps_the_error_code(pcf, abi):
    """With an iteration limit no solve can meet, the flexible solve fails,
    the plain solve it falls back to fails too, and the step returns the
    reference's non-convergence code with the fallback counted."""
    nl = namelists(pcf, 21, "previous", 2)
    nl.poisson.poisson_iterations = 1
    nl.poisson.initial_cleaning = False
    st = pcf.State(nl)
    params = pcf.learned.init_params(2)
    params[-(27 * 32 + 1):-1] = np.random.default_rng(4).standard_normal(27 * 32).astype(np.float32) * 1e-2
    net = pcf.learned.GuessNetwork(32, 32, 16, training=False, params=params)
    pcf.learned.attach_stage(st, net, 1)
    with pytest.raises(pcf.PinCFlowError) as e:
        st.step()
    assert e.value.rc == abi.E_NOCONV
    fb = ctypes.c_int32(0)
    assert st.lib.pcf_stage_fallbacks(st.h, ctypes.byref(fb)) == 0
    assert fb.value == 1
