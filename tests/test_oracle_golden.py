"""The oracle against the reference's own end-to-end fixtures.

test/test_{periodic_hill,mountain_wave,cold_bubble,hot_bubble,vortex}.jl run each example
at reduced size to t = 3600 s and compare the L2 and L-infinity norms of every
output dataset (t = 0 after the initial cleaning, and t = 3600 s) with Float32
literals at rtol = sqrt(eps(Float32)).  The literals are in
tests/golden/reference_norms.json.  They pin the whole FP64 path of the oracle
in aggregate: grid, background, initial cleaning, the 25-point operator, the
preconditioner, BiCGSTAB, the corrector, MUSCL fluxes, RK3 and boundaries; hot_bubble
adds the compressible model (P fluxes, heat conduction, the pi' update, the
J P wind weighting and the dP/dpi operator term).
"""
import json
from pathlib import Path

import pytest

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_norms.json").read_text())


@pytest.mark.parametrize("example", ["periodic_hill", "mountain_wave", "cold_bubble", "hot_bubble", "vortex"])
def test_reference_norms(pcf, abi, oracle_mod, example):
    case = GOLD[example]
    nl = getattr(pcf.examples, example)(**case["keywords"])
    _, recs, steps = oracle_mod.integrate(nl, abi, pcf.output)
    assert len(recs) == 2
    l2, linf = pcf.output.norms(recs)
    rtol = GOLD["rtol"]
    for norms, ref in ((l2, case["l2"]), (linf, case["linf"])):
        for key, want in ref.items():
            assert norms[key] == pytest.approx(want, rel=rtol, abs=0.0 if want else 1e-30), (example, key)
