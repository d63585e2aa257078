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
