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
    res = pcf.ensemble(sim, {"seed": seeds, "x_size": sizes, "output_file": files}, {})
    for c, (seed, nx) in enumerate(zip(seeds, sizes)):
        alone, recs, _ = pcf.integrate(sim(seed, nx))
        st, recs_e, _ = res[c]
        for f in ("RHOP", "U", "V", "W", "PIP", "RHO"):
            assert np.array_equal(st.get(f), alone.get(f)), (seed, f)
        with np.load(files[c], allow_pickle=False) as data:
            assert data["thetap"].shape[-1] == len(recs)
