"""ctypes binding of the oracle (oracle/liblbm_oracle.so). Test infrastructure only.

The oracle is the CPU restatement of the reference's kernels (oracle/solver.hpp) that the HIP
library is checked against. Only tests, the smoke check and the CPU-baseline leg of the benchmark
use it. The classes mirror cells_amd.lbm.Solver / Solver3D so that a test drives both the same way.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from cells_amd.lbm import (DEFAULTS2, DEFAULTS3, HISTORY_LEN, KEYS2, KEYS3, PRECISIONS, Config2D, Config3D,
                           ForceSample, ForceSample3, make_config2d, make_config3d)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "orc_create_2d": ([C.POINTER(Config2D), C.c_int], P),
        "orc_create_3d": ([C.POINTER(Config3D), C.c_int], P),
        "orc_destroy": ([P], None),
        "orc_set_threads": ([C.c_int], None),
        "orc_max_threads": ([], C.c_int),
        "orc_update_2d": ([P, C.POINTER(Config2D)], None),
        "orc_update_3d": ([P, C.POINTER(Config3D)], None),
        "orc_set_sdf": ([P, C.c_void_p], None),
        "orc_brush": ([P, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int], None),
        "orc_rebuild_flags": ([P], None),
        "orc_read_sdf": ([P, C.c_void_p], None),
        "orc_init_field": ([P, C.c_void_p], None),
        "orc_reset_forces": ([P], None),
        "orc_set_ramp": ([P, C.c_int, C.c_double, C.c_double], None),
        "orc_set_open_faces": ([P, C.c_int], None),
        "orc_steps": ([P, C.c_int64], None),
        "orc_macro": ([P, C.c_void_p], None),
        "orc_read_forces": ([P, C.c_void_p, C.c_void_p, C.c_int64], C.c_int64),
        "orc_read_populations": ([P, C.c_void_p], None),
        "orc_raw_bytes": ([P], C.c_size_t),
        "orc_read_raw": ([P, C.c_void_p], None),
        "orc_read_flags": ([P, C.c_void_p], None),
        "orc_read_slots": ([P, C.c_void_p], None),
        "orc_info": ([P, C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32)], None),
        "orc_params": ([P, C.c_void_p], None),
        "orc_tables2": ([C.c_void_p] * 5, None),
        "orc_tables3": ([C.c_void_p] * 7, None),
        "orc_tables27": ([C.c_void_p] * 7, None),
        "orc_flag_bits": ([C.c_int, C.c_int], C.c_uint32),
        "orc_trt_lambda3": ([C.c_double], C.c_double),
        "orc_feq2": ([C.c_int, C.c_double, C.c_double, C.c_double], C.c_double),
        "orc_feq3": ([C.c_int, C.c_double, C.c_double, C.c_double, C.c_double], C.c_double),
        "orc_f32_to_f16": ([C.c_float], C.c_uint16),
        "orc_f32_to_f16_sr": ([C.c_float, C.c_uint32], C.c_uint16),
        "orc_step_random": ([C.c_uint32, C.c_uint32, C.c_int], C.c_uint32),
        "orc_f16_to_f32": ([C.c_uint16], C.c_float),
        "orc_f32_to_bf16": ([C.c_float], C.c_uint16),
        "orc_bf16_to_f32": ([C.c_uint16], C.c_float),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def set_threads(n: int):
    load().orc_set_threads(int(n))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class _Base:
    dim = 2
    Q = 9

    def __init__(self, h, cfg):
        if not h:
            raise ValueError("oracle rejected the configuration")
        self._h = h
        self.cfg = cfg
        dims = np.zeros(3, dtype=np.int32)
        n, st, q = C.c_int64(), C.c_int64(), C.c_int32()
        _lib.orc_info(self._h, _ptr(dims), C.byref(n), C.byref(st), C.byref(q))
        self.W, self.H, self.D = (int(v) for v in dims)
        self.N = int(n.value)
        self.Q = int(q.value)

    @property
    def step(self):
        dims = np.zeros(3, dtype=np.int32)
        n, st, q = C.c_int64(), C.c_int64(), C.c_int32()
        _lib.orc_info(self._h, _ptr(dims), C.byref(n), C.byref(st), C.byref(q))
        return int(st.value)

    def set_sdf(self, sdf):
        a = np.ascontiguousarray(sdf, dtype=np.float32)
        assert a.size == self.N
        _lib.orc_set_sdf(self._h, _ptr(a))

    def read_sdf(self) -> np.ndarray:
        out = np.empty(self.N, dtype=np.float32)
        _lib.orc_read_sdf(self._h, _ptr(out))
        return out

    def reset_forces(self):
        _lib.orc_reset_forces(self._h)

    def set_inlet_ramp(self, kind, u_target, steps):
        _lib.orc_set_ramp(self._h, {"off": 0, "linear": 1, "smoothstep": 2}[kind], float(u_target), float(steps))

    def steps(self, n):
        _lib.orc_steps(self._h, int(n))

    run = steps

    def macro(self) -> np.ndarray:
        out = np.empty(self.N * 4, dtype=np.float64)
        _lib.orc_macro(self._h, _ptr(out))
        return out

    read_macro = macro

    def read_forces_raw(self):
        steps = np.empty(HISTORY_LEN, dtype=np.int64)
        vals = np.empty(HISTORY_LEN * self.dim, dtype=np.float64)
        k = int(_lib.orc_read_forces(self._h, _ptr(steps), _ptr(vals), HISTORY_LEN))
        return steps[:k].copy(), vals[: k * self.dim].reshape(k, self.dim).copy()

    def read_forces(self):
        steps, vals = self.read_forces_raw()
        T = ForceSample if self.dim == 2 else ForceSample3
        return [T(int(s), *map(float, v)) for s, v in zip(steps, vals)]

    def read_populations(self) -> np.ndarray:
        out = np.empty(self.Q * self.N, dtype=np.float64)
        _lib.orc_read_populations(self._h, _ptr(out))
        return out.reshape(self.Q, self.N)

    def read_raw(self) -> np.ndarray:
        nb = int(_lib.orc_raw_bytes(self._h))
        out = np.empty(nb, dtype=np.uint8)
        _lib.orc_read_raw(self._h, _ptr(out))
        return out

    def read_flags(self) -> np.ndarray:
        out = np.empty(self.N, dtype=np.uint32)
        _lib.orc_read_flags(self._h, _ptr(out))
        return out

    def read_slots(self) -> np.ndarray:
        out = np.empty(self.N, dtype=np.uint32)
        _lib.orc_read_slots(self._h, _ptr(out))
        return out

    def params(self) -> np.ndarray:
        out = np.zeros(13, dtype=np.float64)
        _lib.orc_params(self._h, _ptr(out))
        return out[: 13 if self.dim == 2 else 7]

    def destroy(self):
        if self._h:
            _lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class OracleSolver(_Base):
    """D2Q9 oracle; real_bits 32 (float arithmetic and storage) or 64 (double)."""

    dim = 2
    Q = 9

    def __init__(self, config: dict, real_bits: int = 32):
        load()
        full = dict(DEFAULTS2)
        full.update({KEYS2.get(k, k): v for k, v in config.items()})
        c = make_config2d(full)
        super().__init__(_lib.orc_create_2d(C.byref(c), int(real_bits)), full)

    def update(self, changes=None, **kw):
        ch = dict(changes or {}, **kw)
        self.cfg.update({KEYS2.get(k, k): v for k, v in ch.items()})
        _lib.orc_update_2d(self._h, C.byref(make_config2d(self.cfg)))

    def init_field(self, macro=None):
        if macro is None:
            _lib.orc_init_field(self._h, None)
        else:
            m = np.ascontiguousarray(macro, dtype=np.float32)
            assert m.size == self.N * 4
            _lib.orc_init_field(self._h, _ptr(m))

    def rebuild_flags(self):
        _lib.orc_rebuild_flags(self._h)

    def brush(self, cx, cy, r, erase=False, rebuild=True):
        _lib.orc_brush(self._h, float(cx), float(cy), float(r), int(bool(erase)), int(bool(rebuild)))

    def brush_stroke(self, points, r, erase=False):
        """src/solver.ts:244-247: every disc stamped, then one rebuild."""
        for x, y in points:
            self.brush(x, y, r, erase, False)
        self.rebuild_flags()


class OracleSolver3D(_Base):
    """D3Q19 oracle; real_bits 32 uses the configured storage codec, 64 stores doubles."""

    dim = 3
    Q = 19

    def __init__(self, config: dict, real_bits: int = 32):
        load()
        full = dict(DEFAULTS3)
        full.update({KEYS3.get(k, k): v for k, v in config.items()})
        c = make_config3d(full)
        super().__init__(_lib.orc_create_3d(C.byref(c), int(real_bits)), full)

    def _push(self):
        _lib.orc_update_3d(self._h, C.byref(make_config3d(self.cfg)))

    def set_tau(self, tau):
        self.cfg["tau"] = tau
        self._push()

    def set_smagorinsky(self, cs):
        self.cfg["smagorinsky"] = cs
        self._push()

    def set_inlet(self, u):
        self.cfg["u_in"] = u
        self._push()

    def init_field(self):
        _lib.orc_init_field(self._h, None)

    def set_open_faces(self, kind):
        """'equilibrium' (the reference's inlet and outlet planes) or 'zou-he' (Hecht-Harting faces)."""
        _lib.orc_set_open_faces(self._h, {"equilibrium": 0, "zou-he": 1}[kind])


def tables2():
    lib = load()
    cx, cy, opp, my = (np.zeros(9, np.int32) for _ in range(4))
    w = np.zeros(9)
    lib.orc_tables2(_ptr(cx), _ptr(cy), _ptr(w), _ptr(opp), _ptr(my))
    return dict(cx=cx, cy=cy, w=w, opp=opp, mirror_y=my)


def tables3():
    lib = load()
    cx, cy, cz, opp, my, mz = (np.zeros(19, np.int32) for _ in range(6))
    w = np.zeros(19)
    lib.orc_tables3(_ptr(cx), _ptr(cy), _ptr(cz), _ptr(w), _ptr(opp), _ptr(my), _ptr(mz))
    return dict(cx=cx, cy=cy, cz=cz, w=w, opp=opp, mirror_y=my, mirror_z=mz)


def tables27():
    lib = load()
    cx, cy, cz, opp, my, mz = (np.zeros(27, np.int32) for _ in range(6))
    w = np.zeros(27)
    lib.orc_tables27(_ptr(cx), _ptr(cy), _ptr(cz), _ptr(w), _ptr(opp), _ptr(my), _ptr(mz))
    return dict(cx=cx, cy=cy, cz=cz, w=w, opp=opp, mirror_y=my, mirror_z=mz)


def flag_bits(lattice: int):
    lib = load()
    return tuple(int(lib.orc_flag_bits(lattice, k)) for k in range(3))


def trt_lambda3(tau: float) -> float:
    return float(load().orc_trt_lambda3(tau))


def feq2(i, rho, ux, uy):
    return float(load().orc_feq2(i, rho, ux, uy))


def feq3(i, rho, ux, uy, uz):
    return float(load().orc_feq3(i, rho, ux, uy, uz))


def f32_to_f16(x: float) -> int:
    return int(load().orc_f32_to_f16(x))


def f32_to_f16_sr(x: float, r: int) -> int:
    return int(load().orc_f32_to_f16_sr(x, r))


def step_random(node: int, step: int, i: int) -> int:
    return int(load().orc_step_random(node, step, i))


def f32_to_bf16(x: float) -> int:
    return int(load().orc_f32_to_bf16(x))


__all__ = ["OracleSolver", "OracleSolver3D", "set_threads", "f32_to_f16", "f32_to_bf16", "PRECISIONS"]
