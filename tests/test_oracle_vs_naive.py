"""The FP64 oracle against an independently written naive FP64 implementation (tests/naive_lbm.py)
on tiny grids that exercise every boundary rule: the two agree to round-off after 1, 10 and 100
steps, populations and momentum-exchange forces alike."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests import naive_lbm
from tests.parity_cases import CASES2, case3, taylor_green_macro


@pytest.mark.parametrize("name", list(CASES2))
def test_d2q9_oracle_matches_naive(name):
    cfg, sdf_fn = CASES2[name]
    W, H = cfg["width"], cfg["height"]
    sdf = sdf_fn(W, H) if sdf_fn else None
    macro = taylor_green_macro(W) if name == "box" else None
    s = O.OracleSolver(dict(cfg), 64)
    if sdf is not None:
        s.set_sdf(sdf)
    s.init_field(macro)
    done = 0
    for n in (1, 10, 100):
        s.steps(n - done)
        done = n
        f_naive, forces = naive_lbm.run2d(dict(cfg), sdf, n, macro)
        f_orc = s.read_populations()
        err = np.abs(f_orc - f_naive).max()
        assert err < 1e-13, (name, n, err)
    steps, vals = s.read_forces_raw()
    assert len(steps) == 100
    scale = max(np.abs(forces).max(), 1e-30)
    assert np.abs(vals - forces).max() <= 1e-12 * max(scale, 1.0), (name, np.abs(vals - forces).max())


@pytest.mark.parametrize("smag", [0.0, 0.12])
def test_d3q19_oracle_matches_naive(smag):
    cfg, sdf = case3("fp32", smag=smag)
    s = O.OracleSolver3D(dict(cfg), 64)
    s.set_sdf(sdf)
    s.init_field()
    done = 0
    for n in (1, 10, 60):
        s.steps(n - done)
        done = n
        f_naive, forces = naive_lbm.run3d(dict(cfg), sdf, n)
        f_orc = s.read_populations()
        err = np.abs(f_orc - f_naive).max()
        assert err < 1e-13, (n, err)
    steps, vals = s.read_forces_raw()
    assert np.abs(vals - forces).max() < 1e-12 * max(1.0, np.abs(forces).max())
    assert np.abs(forces).max() > 1e-4  # the bodies really are in the flow
