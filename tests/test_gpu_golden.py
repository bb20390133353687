"""The HIP library against the reference's own end-to-end fixtures.

Same cases and tolerance as tests/test_oracle_golden.py (the L2 and L-infinity
norms of every output dataset at t = 0 and t = 3600 s, rtol =
sqrt(eps(Float32))), run through integrate(namelists) on the GPU."""
import json
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_norms.json").read_text())


@pytest.mark.parametrize("example", ["periodic_hill", "mountain_wave", "cold_bubble", "hot_bubble", "vortex"])
def test_reference_norms_gpu(pcf, example):
    case = GOLD[example]
    nl = getattr(pcf.examples, example)(**case["keywords"])
    _, recs, steps = pcf.integrate(nl)
    assert len(recs) == 2
    l2, linf = pcf.output.norms(recs)
    rtol = GOLD["rtol"]
    for norms, ref in ((l2, case["l2"]), (linf, case["linf"])):
        for key, want in ref.items():
            assert norms[key] == pytest.approx(want, rel=rtol, abs=0.0 if want else 1e-30), (example, key)
