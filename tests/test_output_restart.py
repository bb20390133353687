"""Output records, the output file and restart (write_output.jl, create_output.jl,
read_input!.jl, integrate.jl:168-196)."""
import numpy as np
import pytest


def _records(pcf, oracle_mod, steps=3):
    nl = pcf.examples.periodic_hill(x_size=12, z_size=10, output_steps=True, iterations=steps)
    nl.output.prepare_restart = True
    _, recs, _ = oracle_mod.integrate(nl, pcf._abi, pcf.output)
    return nl, recs


def test_output_file_layout_and_read_input(pcf, oracle_mod, tmp_path):
    nl, recs = _records(pcf, oracle_mod)
    path = tmp_path / "out.npz"
    pcf.output.write_file(str(path), recs)
    with np.load(path, allow_pickle=False) as f:
        assert f["t"].shape == (len(recs),)
        assert f["us"].shape == recs[0]["us"].shape + (len(recs),)
        assert f["z"].shape == recs[0]["z"].shape  # grid datasets once
        assert f["us"].dtype == np.float32
    for iin, rec in ((2, recs[1]), (-1, recs[-1])):
        t, fields = pcf.output.read_input(str(path), iin)
        assert t == pytest.approx(float(rec["t"][0]))
        for k in ("rhop", "us", "vs", "wts", "pip"):
            np.testing.assert_array_equal(fields[k], rec[k])


def test_compressible_records_carry_p(pcf, oracle_mod, tmp_path):
    """write_output.jl:159-165: the compressible model writes P with every
    record; read_input! restores it (read_input!.jl:73-77)."""
    nl = pcf.examples.hot_bubble(x_size=8, z_size=8)
    nl.output.output_steps = True
    nl.output.iterations = 2
    _, recs, _ = oracle_mod.integrate(nl, pcf._abi, pcf.output)
    assert all("p" in r for r in recs)
    path = tmp_path / "hb.npz"
    pcf.output.write_file(str(path), recs)
    with np.load(path, allow_pickle=False) as f:
        assert f["p"].shape == recs[0]["p"].shape + (len(recs),)
    _, fields = pcf.output.read_input(str(path), 2)
    np.testing.assert_array_equal(fields["p"], recs[1]["p"])


def test_read_input_needs_the_restart_datasets(pcf, oracle_mod, tmp_path):
    nl = pcf.examples.periodic_hill(x_size=12, z_size=10, output_steps=True, iterations=1)
    nl.output.prepare_restart = False
    nl.output.output_variables = ("w",)
    _, recs, _ = oracle_mod.integrate(nl, pcf._abi, pcf.output)
    path = tmp_path / "out.npz"
    pcf.output.write_file(str(path), recs)
    with pytest.raises(ValueError):
        pcf.output.read_input(str(path))


@pytest.mark.gpu
@pytest.mark.parametrize("example", ["periodic_hill", "mountain_wave", "vortex", "hot_bubble"])
def test_restart_continues_the_run(pcf, tmp_path, example):
    """A run restarted from its stored record tracks the uninterrupted run.
    The record holds FP32 values (as the reference's file), so the two agree
    to RESTART_RTOL, not bitwise."""
    RESTART_RTOL = 1e-4
    make = {"periodic_hill": lambda: pcf.examples.periodic_hill(x_size=24, z_size=20, output_steps=True),
            "mountain_wave": lambda: pcf.examples.mountain_wave(x_size=12, y_size=10, z_size=9),
            "vortex": lambda: pcf.examples.vortex(x_size=12, y_size=10),
            "hot_bubble": lambda: pcf.examples.hot_bubble(x_size=12, z_size=10)}[example]

    def namelists(iterations):
        nl = make()
        nl.output.output_steps = True
        nl.output.iterations = iterations
        nl.output.prepare_restart = True
        nl.output.output_file = str(tmp_path / "out.npz")
        nl.output.input_file = str(tmp_path / "out.npz")
        return nl

    full, _, steps_full = pcf.integrate(namelists(5))
    first, _, _ = pcf.integrate(namelists(3), write=True)
    t3, n3 = first.time()
    nl = namelists(2)
    nl.output.restart = True
    nl.poisson.initial_cleaning = False
    rest, recs, steps_rest = pcf.integrate(nl)
    assert recs[0]["t"][0] == pytest.approx(t3, rel=1e-6)
    assert rest.time()[0] == pytest.approx(full.time()[0], rel=1e-6)
    tracer = ("CHI",) if example == "vortex" else ()
    comp = ("P", "PBAR") if example == "hot_bubble" else ()  # P restored, pbar synchronized
    for f in ("RHOP", "U", "V", "W", "PIP") + tracer + comp:
        a, b = rest.get(f), full.get(f)
        scale = max(float(np.max(np.abs(b))), 1e-30)
        assert float(np.max(np.abs(a - b))) / scale <= RESTART_RTOL, f


@pytest.mark.gpu
def test_restart_past_tmax_is_an_error(pcf):
    nl = pcf.examples.periodic_hill(x_size=12, z_size=10)
    st = pcf.State(nl)
    with pytest.raises(pcf.PinCFlowError):
        st.restart(nl.output.tmax * 2, {k: st.get("U")[3:-3, 3:-3, 3:-3] * 0 for k in
                                         ("rhop", "us", "vs", "wts", "pip")})


@pytest.mark.gpu
@pytest.mark.parametrize("writer", ["oracle", "gpu"])
@pytest.mark.parametrize("example", ["periodic_hill", "mountain_wave", "vortex", "hot_bubble"])
def test_restart_matches_the_oracle(pcf, oracle_mod, tmp_path, example, writer):
    """One side (the oracle or the device) runs three steps and writes its
    records; the device and the oracle both restart from the last record
    (output.read_input, then read_input! and integrate.jl:168-186 on each
    side) and continue for two steps.  With the Poisson tolerance tightened
    to 1e-12 (relative) the two continued runs agree to STEP_RTOL, and their
    iteration counts within the spread of reordered sums (iter_slack)."""
    from test_oracle_reduction_order import iter_slack
    STEP_RTOL = 1e-9
    make = {"periodic_hill": lambda: pcf.examples.periodic_hill(x_size=24, z_size=20),
            "mountain_wave": lambda: pcf.examples.mountain_wave(x_size=12, y_size=10, z_size=9),
            "vortex": lambda: pcf.examples.vortex(x_size=12, y_size=10),
            "hot_bubble": lambda: pcf.examples.hot_bubble(x_size=12, z_size=10)}[example]

    def namelists(iterations):
        nl = make()
        nl.output.output_steps = True
        nl.output.iterations = iterations
        nl.output.prepare_restart = True
        nl.poisson.tolerance = 1e-12
        nl.poisson.tolerance_is_relative = True
        return nl

    path = str(tmp_path / "first.npz")
    nl = namelists(3)
    if writer == "gpu":
        _, recs, _ = pcf.integrate(nl)
    else:
        _, recs, _ = oracle_mod.integrate(nl, pcf._abi, pcf.output)
    pcf.output.write_file(path, recs)
    time_s, fields = pcf.output.read_input(path, -1)
    nl = namelists(2)
    nl.poisson.initial_cleaning = False
    gpu = pcf.State(nl)
    orc = oracle_mod.Oracle(pcf.to_config(nl), pcf._abi)
    gpu.restart(time_s, fields)
    orc.restart(nl, pcf.output, time_s, fields)
    for _ in range(2):
        a, b = gpu.step(), orc.step()
        assert a.time == pytest.approx(b.time, rel=1e-12)
        assert a.dt == pytest.approx(b.dt, rel=1e-10)
        for q in range(2):
            assert abs(a.niter[q] - b.niter[q]) <= iter_slack(b.niter[q]), (a.niter[q], b.niter[q])
    names = ["RHOP", "U", "V", "W", "PIP"] + (["RHO"] if nl.atmosphere.model != "Boussinesq" else [])
    names += (["CHI"] if example == "vortex" else []) + (["P", "PBAR"] if example == "hot_bubble" else [])
    for f in names:
        x, y = gpu.get(f), orc.get(f)
        scale = max(float(np.max(np.abs(y))), 1e-30)
        assert float(np.max(np.abs(x - y))) / scale <= STEP_RTOL, (f, float(np.max(np.abs(x - y))) / scale)
