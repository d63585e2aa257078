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
