"""Runs the Node.js wrapper's own tests (tests/js) under the system node."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ADDON = os.path.join(os.path.dirname(HERE), "cells_amd", "node", "lbm_napi.node")


def _node(script: str, timeout: int):
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed")
    if not os.path.exists(ADDON):
        pytest.skip("the N-API addon has not been built")
    r = subprocess.run([node, os.path.join(HERE, "js", script)], capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr


def test_node_wrapper_host_logic():
    _node("cpu.test.js", 60)


@pytest.mark.gpu
def test_node_wrapper_on_gpu():
    _node("gpu.test.js", 120)


@pytest.mark.gpu
def test_benchmark_harness_rows(tmp_path):
    """The benchmark page's harness (src/bench.ts:97-186) in its quick form: the row fields and
    log lines of the reference, every scenario present, and the 3D medium grid over the reference's
    fp32 / fp16 floors (src/bench.ts:180-184)."""
    import json
    import re

    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed")
    out = tmp_path / "bench.json"
    script = os.path.join(os.path.dirname(HERE), "cells_amd", "node", "bench.js")
    r = subprocess.run([node, script, "--quick", "--json", str(out)], capture_output=True, text=True, timeout=110)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    res = json.loads(out.read_text())
    assert "gfx950" in res["adapter"]
    rows = res["rows"]
    keys = {"scenario", "width", "height", "workgroup", "forceEvery", "steps", "ms", "mlups", "gbps", "timer"}
    for row in rows:
        assert keys <= set(row) and row["timer"] == "gpu-timestamps" and row["mlups"] > 0
        assert abs(row["mlups"] - row["width"] * row["height"] * row.get("depth", 1) * row["steps"] / (row["ms"] * 1e3)) \
            <= 1e-9 * row["mlups"]
    by = {(x["scenario"], x["workgroup"], x["forceEvery"]): x for x in rows}
    for key in (("copy roofline", 128, 0), ("periodic", 128, 0), ("tunnel+cylinder", 128, 0),
                ("tunnel+cylinder", 128, 1), ("d3q19 sphere fp32", "auto", 4), ("d3q19 empty fp16", 128, 0)):
        assert key in by, key
    assert re.search(r"^copy roofline\s+512x256\s+wg 128\s+\d+ MLUPS\s+\d+ GB/s", r.stdout, re.M)
    assert re.search(r"^d3q19 empty fp16\s+192x96x96\s+wg 128\s+force off\s+\d+ MLUPS\s+\d+ GB/s\s+\(\d+ steps, median of 3",
                     r.stdout, re.M)
    assert "d3q19 floor fp32 (empty, wg 128):" in r.stdout and "MLUPS: met" in r.stdout
    assert "missed" not in r.stdout


@pytest.mark.gpu
def test_validation_suite_from_node():
    """The validation suite (src/cases.ts:61-427, src/cases3d.ts:11-81) driven from Node through the
    Solver / Solver3D wrappers, every case inside the reference's acceptance ranges."""
    import json

    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed")
    if not os.path.exists(ADDON):
        pytest.skip("the N-API addon has not been built")
    script = os.path.join(HERE, "validation", "validate.js")
    r = subprocess.run([node, script], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    results = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    assert [x["name"] for x in results] == [
        "Plane Poiseuille", "Periodic seam invariance", "Taylor-Green vortex", "Schäfer-Turek 2D-1 (Re 20)",
        "Schäfer-Turek 2D-2 (Re 100)", "Unconfined cylinder (Re 100)", "Ghia lid-driven cavity (Re 1000)",
        "NACA0012 (Re 500, alpha 0)", "Sphere C_D, FP16 vs FP32 (Re 100)"]
    for x in results:
        assert x["pass"], [m for m in x["metrics"] if not m["pass"]]
    assert r.returncode == 0


def test_node_stream_function_and_vortex_closed_form():
    """streamFunction / primaryVortex (src/analysis.ts:28-55) in the Node port: for the linear field
    ux = -a (y - yc), uy = b (x - xc) the trapezoid integration is exact, psi is the paraboloid
    -a ((y - yc)^2 - yc^2) / 2 - b ((x - xc)^2 - xc^2) / 2 and the quadratic refinement recovers the
    centre (xc, yc) exactly; the Python helpers agree to rounding."""
    import json

    import numpy as np

    from cells_amd.analysis import primary_vortex, stream_function

    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed")
    W, H, a, b, xc, yc = 37, 29, 0.013, 0.021, 17.3, 12.6
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    ux = (-a * (y - yc)).reshape(-1)
    uy = (b * (x - xc)).reshape(-1)
    geo = os.path.join(os.path.dirname(HERE), "cells_amd", "node", "geometry.js")
    src = (f"const g = require({json.dumps(geo)}); const W = {W}, H = {H};"
           f"const ux = Float64Array.from({json.dumps(ux.tolist())}); const uy = Float64Array.from({json.dumps(uy.tolist())});"
           "const psi = g.streamFunction(ux, uy, W, H);"
           "console.log(JSON.stringify({psi: Array.from(psi), v: g.primaryVortex(psi, W, H)}));")
    r = subprocess.run([node, "-e", src], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    exact = -a * ((y - yc) ** 2 - yc ** 2) / 2 - b * ((x - xc) ** 2 - xc ** 2) / 2
    np.testing.assert_allclose(out["psi"], exact.reshape(-1), rtol=0, atol=1e-12)
    assert abs(out["v"][0] - xc) < 1e-9 and abs(out["v"][1] - yc) < 1e-9
    psi_py = stream_function(ux, uy, W, H)
    np.testing.assert_allclose(out["psi"], psi_py, rtol=0, atol=1e-13)
    np.testing.assert_allclose(out["v"], primary_vortex(psi_py, W, H), rtol=0, atol=1e-12)
