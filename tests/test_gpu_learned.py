"""The learned initial guess: trained on the solver's own solves with the
device's forward and backward kernels, used as BiCGSTAB's starting point in
the mixed-precision solve (multigrid preconditioner), and compared on held-out
runs with the zero start and with the warm start it corrects.  The solve's
stopping test is unchanged (|r| / |b| against the reference tolerance), so
the guess affects the iteration count only."""
import ctypes

import hipmem
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def namelists(pcf, seed, guess, steps):
    nl = pcf.examples.gaussian_hills_3d(x_size=32, y_size=32, z_size=16, iterations=steps, seed=seed, n_hills=6,
                                        precision="mixed")
    nl.poisson.initial_guess = guess
    nl.poisson.preconditioner_type = "multigrid"
    return nl


def stage_iterations(pcf, nl, net=None, steps=6):
    st = pcf.State(nl)
    if net is not None:
        pcf.learned.attach(st, net)
    st.initial_cleaning()
    its = []
    for _ in range(steps):
        info = st.step()
        its.append(int(info.niter[1]))
    return its


def test_learned_guess_gains_over_zero_start(pcf):
    train_runs = []
    for seed in (11, 12):
        st = pcf.State(namelists(pcf, seed, "zero", 14))
        st.initial_cleaning()
        train_runs += pcf.learned.collect_samples(st, 14)
    net = pcf.learned.GuessNetwork(32, 32, 16, training=True, seed=0)
    losses = pcf.learned.train(net, train_runs, epochs=25)
    print("training loss (normalised)", losses[0], "->", losses[-1])
    assert losses[-1] < 0.9 * losses[0]
    infer = pcf.learned.GuessNetwork(32, 32, 16, training=False, params=net.params())
    held_out = 21
    st = pcf.State(namelists(pcf, held_out, "zero", 10))
    st.initial_cleaning()
    held = pcf.learned.collect_samples(st, 10)
    rel = pcf.learned.evaluate(infer, held)
    print("held-out error of the learned guess relative to the warm start's", rel)
    # the trained network at least halves the squared error of the warm start
    # it corrects, on runs it has not seen
    assert rel < 0.5
    steps = 8
    zero = stage_iterations(pcf, namelists(pcf, held_out, "zero", steps), steps=steps)
    warm = stage_iterations(pcf, namelists(pcf, held_out, "previous", steps), steps=steps)
    learned = stage_iterations(pcf, namelists(pcf, held_out, "learned", steps), net=infer, steps=steps)
    # the first step has no previous solution; compare the steps after it
    z, w, lrn = sum(zero[1:]), sum(warm[1:]), sum(learned[1:])
    print("stage-2 iterations, zero start:", zero, "warm start:", warm, "learned start:", learned)
    # BiCGSTAB's count depends on the initial residual, which the smaller
    # error of the guess lowers little (its remaining error is of small
    # scale, which the operator amplifies): the learned start may not cost
    # more than an iteration per solve over either start
    assert lrn <= min(z, w) + len(zero) - 1, (zero, warm, learned)


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


def test_residual_trained_guess(pcf):
    """Online training on the residual |A x0 - b|^2 / |b|^2 of the guess
    (learned.train_residual) on two runs; on a held-out run the guess leaves
    a smaller residual than the warm start it corrects.  The iteration
    counts stay those of the warm start: the guess removes the fast residual
    components, which BiCGSTAB removes in its first iterations anyway."""
    net = pcf.learned.GuessNetwork(32, 32, 16, training=True, seed=0)
    states = []
    for seed in (11, 12):
        st = pcf.State(namelists(pcf, seed, "zero", 14))
        st.initial_cleaning()
        states.append(st)
    hist = pcf.learned.train_residual(net, states, 14, inner=6, lr=1e-3)
    print("training solves (warm, guess):", [(round(a, 4), round(b, 4)) for a, b in hist])
    assert all(b < a for a, b in hist[-6:])
    infer = pcf.learned.GuessNetwork(32, 32, 16, training=False, params=net.params())
    held_out = 21
    st = pcf.State(namelists(pcf, held_out, "zero", 10))
    st.initial_cleaning()
    rl = pcf.learned.residual_loss(infer, st, 8)
    print("held-out residual loss (warm, learned):", [(round(a, 4), round(b, 4)) for a, b in rl])
    ratio = float(np.mean([b / a for a, b in rl]))
    assert ratio < 0.85, ratio
    steps = 8
    warm = stage_iterations(pcf, namelists(pcf, held_out, "previous", steps), steps=steps)
    learned = stage_iterations(pcf, namelists(pcf, held_out, "learned", steps), net=infer, steps=steps)
    print("stage-2 iterations, warm start:", warm, "learned start:", learned)
    assert sum(learned[1:]) <= sum(warm[1:]) + len(warm) - 1, (warm, learned)
