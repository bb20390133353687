"""ensemble(simulation, parameters, keywords) (src/ensemble.jl:20-74): argument
checks and error reporting on the CPU, the seeded bubble members on the
oracle, and on the GPU members that share the device reproduce their
standalone runs bit for bit."""
import numpy as np
import pytest


def test_output_files_are_required_and_distinct(pcf, tmp_path):
    sim = lambda **kw: pcf.examples.bubble_member(**kw)
    with pytest.raises(ValueError, match="output_file"):
        pcf.ensemble(sim, {"seed": (1, 2)}, {})
    with pytest.raises(ValueError, match="duplicate"):
        pcf.ensemble(sim, {"seed": (1, 2), "output_file": ("a.npz", "a.npz")}, {})
    with pytest.raises(ValueError):
        pcf.ensemble(sim, {"seed": (1, 2, 3), "output_file": ("a.npz", "b.npz")}, {})


def test_member_errors_name_the_member(pcf, tmp_path):
    def sim(seed, **kw):
        raise RuntimeError(f"member with seed {seed} fails")
    with pytest.raises(pcf.EnsembleError) as e:
        pcf.ensemble(sim, {"seed": (7, 8), "output_file": (str(tmp_path / "a.npz"), str(tmp_path / "b.npz"))}, {})
    assert e.value.member == 1 and "seed 7" in str(e.value)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_bubble_members_run_on_the_oracle(pcf, abi, oracle_mod, seed):
    """The members are valid configurations that stay stable: the time step
    stays at dtmax and each solve converges well inside poisson_iterations."""
    nl = pcf.examples.bubble_member(seed=seed, x_size=24, z_size=24, iterations=3)
    _, recs, steps = oracle_mod.integrate(nl, abi, pcf.output)
    assert len(steps) == 3 and len(recs) == 4
    assert all(dt == pytest.approx(60.0) for _, dt, _, _ in steps), steps
    assert max(max(a, b) for _, _, a, b in steps) < 200, steps
    assert np.all(np.isfinite(recs[-1]["thetap"]))


def test_member_seeds_of_the_oracle_comparison(pcf):
    """the seeds test_gpu_configs.py compares with the oracle give members of
    the real ensemble sizes and both models"""
    shapes = [(m.domain.x_size, m.domain.y_size, m.domain.z_size, m.atmosphere.model)
              for m in (pcf.examples.bubble_member(seed=s) for s in (11, 34, 23))]
    assert shapes == [(64, 64, 64, "Boussinesq"), (64, 64, 64, "PseudoIncompressible"),
                      (64, 64, 96, "PseudoIncompressible")]


@pytest.mark.gpu
def test_members_sharing_the_gpu_match_their_standalone_runs(pcf, tmp_path):
    seeds = (0, 1, 2, 3, 4)
    sizes = (16, 24, 16, 32, 24)
    files = tuple(str(tmp_path / f"member{s}.npz") for s in seeds)
    sim = lambda seed, x_size, **kw: pcf.examples.bubble_member(seed=seed, x_size=x_size, z_size=12, iterations=3)
    res = pcf.ensemble(sim, {"seed": seeds, "x_size": sizes, "output_file": files}, {}, batched=False)
    for c, (seed, nx) in enumerate(zip(seeds, sizes)):
        alone, recs, _ = pcf.integrate(sim(seed, nx))
        st, recs_e, _ = res[c]
        for f in ("RHOP", "U", "V", "W", "PIP", "RHO"):
            assert np.array_equal(st.get(f), alone.get(f)), (seed, f)
        with np.load(files[c], allow_pickle=False) as data:
            assert data["thetap"].shape[-1] == len(recs)


def _members(pcf, precision, seeds, sizes, iterations, pc=None, poisson_iterations=None):
    def sim(seed, x_size, iterations, pc=None, poisson_iterations=None):
        # "both": FP64 and mixed members by seed parity, two solve kinds in one batch
        prec = ("fp64", "mixed")[seed % 2] if precision == "both" else precision
        nl = pcf.examples.bubble_member(seed=seed, x_size=x_size, z_size=12, iterations=iterations,
                                        precision=prec)
        if pc:
            nl.poisson.preconditioner_type = pc
        if poisson_iterations:
            nl.poisson.poisson_iterations = poisson_iterations
        return nl
    params = {"seed": seeds, "x_size": sizes, "iterations": iterations}
    if pc:
        params["pc"] = pc
    if poisson_iterations:
        params["poisson_iterations"] = poisson_iterations
    return sim, params


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed", "both"])
def test_batched_solves_give_each_member_its_standalone_bits(pcf, tmp_path, precision):
    """Members of uneven grids (16-32 cells square, 12 levels) and uneven run
    lengths whose solves run as one launch per kernel over all of them
    (SolveBatch), joining and leaving the batch as they start and converge:
    every member's iteration counts and fields equal its own standalone
    run's bit for bit.  One member uses the multigrid preconditioner, which
    the batch does not cover: its solves run on their own; the shorter runs
    detach (and the tables are rebuilt) while the others go on.  "both":
    FP64 and mixed members in one ensemble, two solve kinds batched side by
    side."""
    seeds = (0, 1, 2, 3, 4, 5)
    sizes = (16, 24, 16, 32, 24, 32)
    iterations = (3, 3, 1, 3, 2, 3)
    pcs = ("line_relaxation",) * 5 + ("multigrid" if precision in ("mixed", "both") else "line_relaxation",)
    sim, params = _members(pcf, precision, seeds, sizes, iterations, pc=pcs)
    params["output_file"] = tuple(str(tmp_path / f"m{s}.npz") for s in seeds)
    batch = pcf.SolveBatch()
    res = pcf.ensemble(sim, params, {}, write=False, batch=batch)
    stats = batch.stats()
    batch.close()
    print(precision, "batch statistics:", stats)
    assert stats["member_solves"] > 0 and stats["batched_iterations"] > 0
    if precision != "fp64":
        assert stats["solves_alone"] > 0
    for c, seed in enumerate(seeds):
        alone, _, steps = pcf.integrate(sim(seed, sizes[c], iterations[c], pcs[c]))
        st, _, steps_e = res[c]
        assert [s[2:] for s in steps_e] == [s[2:] for s in steps], (seed, steps_e, steps)
        for f in ("RHOP", "U", "V", "W", "PIP", "RHO"):
            assert np.array_equal(st.get(f), alone.get(f)), (seed, f)
        alone.close()


@pytest.mark.gpu
def test_batched_solve_that_fails_stops_at_its_own_limit(pcf, abi, tmp_path):
    """A member whose poisson_iterations no solve can meet fails with the
    reference's non-convergence error after exactly that many iterations,
    as on its own, while the batch goes on with the others; the ensemble
    then raises the failure with the member named (reduce_exceptions)."""
    seeds = (0, 1, 2)
    sim, params = _members(pcf, "fp64", seeds, (16, 24, 16), (2, 2, 2), poisson_iterations=(1000, 3, 1000))
    params["output_file"] = tuple(str(tmp_path / f"m{s}.npz") for s in seeds)
    nl = sim(1, 24, 2, poisson_iterations=3)
    nl.poisson.initial_cleaning = False
    with pytest.raises(pcf.PoissonNotConverged) as alone:
        pcf.integrate(nl)
    keywords = {}
    params_nc = dict(params)
    batch = pcf.SolveBatch()
    with pytest.raises(pcf.EnsembleError) as e:
        pcf.ensemble(lambda **kw: _no_cleaning(sim(**kw)), params_nc, keywords, write=False, batch=batch)
    stats = batch.stats()
    batch.close()
    assert e.value.member == 2
    assert isinstance(e.value.error, pcf.PoissonNotConverged)
    assert e.value.error.rc == alone.value.rc == abi.E_NOCONV
    assert str(e.value.error) == str(alone.value)
    assert stats["member_solves"] > 0


def _no_cleaning(nl):
    nl.poisson.initial_cleaning = False
    return nl


@pytest.mark.gpu
def test_ensemble_attach_errors(pcf, abi):
    """The ensemble's argument checks: a state attaches to one ensemble at a
    time, an ensemble with members attached is not destroyed, and detach is
    idempotent (pcf_destroy detaches too)."""
    import ctypes
    nl = pcf.examples.bubble_member(seed=0, x_size=16, z_size=12, iterations=1)
    st = pcf.State(nl)
    a, b = pcf.SolveBatch(), pcf.SolveBatch()
    a.attach(st)
    with pytest.raises(pcf.PinCFlowError) as e:
        b.attach(st)
    assert e.value.rc == abi.E_ARG
    assert st.lib.pcf_ensemble_destroy(a.h) == abi.E_ARG
    a.detach(st)
    a.detach(st)
    b.attach(st)
    st.step()
    st.close()  # detaches from b
    a.close()
    b.close()
    assert st.lib.pcf_ensemble_attach(None, None) == abi.E_ARG
    assert st.lib.pcf_ensemble_stats(None, (ctypes.c_int64 * 4)()) == abi.E_ARG
