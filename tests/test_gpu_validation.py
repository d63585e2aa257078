"""The reference's validation suite (src/cases.ts:61-427, src/cases3d.ts:11-81) on the gfx950 solver,
with the reference's acceptance ranges (cells_amd/validate.py restates each case unchanged)."""
import pytest

from cells_amd import validate

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", list(validate.CASES))
def test_validation_case(case):
    print(f"validation case {case}", flush=True)
    r = validate.run_case(case)
    for m in r["metrics"]:
        print(f"  {m['name']}: {m['value']:.6g} accept {m['accept']} {'ok' if m['pass_'] else 'MISS'}")
    assert r["pass"], [m for m in r["metrics"] if not m["pass_"]]
