"""ctypes binding of the C ABI (include/lbm.h) with the reference's Solver / Solver3D surface.

Method names and argument meaning follow src/solver.ts:50-340 and src/solver3d.ts:42-319 (both the
camelCase names of the reference and snake_case aliases exist). All compute runs in the gfx950
library ``cells_amd/liblbm_hip.so``; there is no CPU path: if the library is missing or no GPU is
present, creating a solver raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from collections import namedtuple
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_hip.so")

LBM_OK = 0
X_PERIODIC, X_WALL, X_OPEN = 0, 1, 2  # XMode (src/lattice.ts:17-21)
Y_PERIODIC, Y_NOSLIP, Y_SLIP, Y_MOVING = 0, 1, 2, 3  # YMode (src/lattice.ts:23-28)
PRECISIONS = {"fp32": 0, "fp16": 1, "bf16": 2, "fp16sr": 3}
HISTORY_LEN = 8192
MAGIC_LAMBDA = 3 / 16

ForceSample = namedtuple("ForceSample", "step fx fy")
ForceSample3 = namedtuple("ForceSample3", "step fx fy fz")


class Config2D(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("left", C.c_int32), ("right", C.c_int32), ("bottom", C.c_int32), ("top", C.c_int32),
        ("parabolic", C.c_int32), ("workgroup_size", C.c_int32), ("force_every", C.c_int32), ("_pad0", C.c_int32),
        ("tau", C.c_double), ("lambda_", C.c_double), ("smagorinsky", C.c_double), ("u_in", C.c_double),
        ("u_lid", C.c_double), ("sponge_fraction", C.c_double), ("sponge_tau", C.c_double),
        ("body_force", C.c_double * 2), ("absorb", C.c_double), ("inlet_layer_fraction", C.c_double),
    ]


class Config3D(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("depth", C.c_int32), ("precision", C.c_int32),
        ("workgroup_size", C.c_int32), ("force_every", C.c_int32),
        ("tau", C.c_double), ("lambda_", C.c_double), ("smagorinsky", C.c_double), ("u_in", C.c_double),
        ("sponge_fraction", C.c_double), ("sponge_tau", C.c_double), ("absorb", C.c_double),
        ("velocity_set", C.c_int32), ("_pad1", C.c_int32),
    ]


class AdjointConfig(C.Structure):
    """lbm_adjoint_config (include/lbm.h)."""
    _fields_ = [
        ("steps", C.c_int64), ("window_start", C.c_int64), ("direction", C.c_double * 3), ("scale", C.c_double),
        ("checkpoint_every", C.c_int32), ("ramp_kind", C.c_int32), ("ramp_u", C.c_double),
        ("ramp_steps", C.c_double), ("record_budget_bytes", C.c_int64),
    ]


# Reference config keys (camelCase, src/solver.ts:15-40, src/solver3d.ts:11-31) -> struct fields.
KEYS2 = {
    "width": "width", "height": "height", "left": "left", "right": "right", "bottom": "bottom", "top": "top",
    "tau": "tau", "lambda": "lambda_", "smagorinsky": "smagorinsky", "uIn": "u_in", "parabolic": "parabolic",
    "uLid": "u_lid", "spongeFraction": "sponge_fraction", "spongeTau": "sponge_tau", "bodyForce": "body_force",
    "absorb": "absorb", "inletLayerFraction": "inlet_layer_fraction", "workgroupSize": "workgroup_size",
    "forceEvery": "force_every",
}
KEYS3 = {
    "width": "width", "height": "height", "depth": "depth", "precision": "precision", "tau": "tau",
    "lambda": "lambda_", "smagorinsky": "smagorinsky", "uIn": "u_in", "spongeFraction": "sponge_fraction",
    "spongeTau": "sponge_tau", "absorb": "absorb", "workgroupSize": "workgroup_size", "forceEvery": "force_every",
    "velocitySet": "velocity_set",
}

DEFAULTS2 = dict(lambda_=MAGIC_LAMBDA, smagorinsky=0.0, u_in=0.0, parabolic=0, u_lid=0.0, sponge_fraction=0.0,
                 sponge_tau=1.0, body_force=(0.0, 0.0), absorb=0.0, inlet_layer_fraction=0.0, workgroup_size=128,
                 force_every=1)
DEFAULTS3 = dict(lambda_=float("nan"), smagorinsky=0.0, u_in=0.0, sponge_fraction=0.0, sponge_tau=1.0, absorb=0.0,
                 workgroup_size=128, force_every=1, velocity_set=19)


def make_config2d(cfg: dict) -> Config2D:
    """SolverConfig with the reference defaults (src/solver.ts:98-112)."""
    c = Config2D()
    vals = dict(DEFAULTS2)
    for k, v in cfg.items():
        vals[KEYS2.get(k, k)] = v
    for k, v in vals.items():
        if k == "body_force":
            c.body_force[0], c.body_force[1] = float(v[0]), float(v[1])
        elif k == "parabolic":
            c.parabolic = 1 if v else 0
        else:
            setattr(c, k, v)
    return c


def make_config3d(cfg: dict) -> Config3D:
    """Solver3DConfig with the reference defaults (src/solver3d.ts:110-119); lambda None or NaN follows
    trtLambda3(tau)."""
    c = Config3D()
    vals = dict(DEFAULTS3)
    for k, v in cfg.items():
        vals[KEYS3.get(k, k)] = v
    for k, v in vals.items():
        if k == "precision":
            v = PRECISIONS[v] if isinstance(v, str) else int(v)
        if k == "lambda_" and v is None:
            v = float("nan")
        setattr(c, k, v)
    return c


def trt_lambda3(tau: float) -> float:
    """trtLambda3 (src/lattice3d.ts:29-31)."""
    return min(MAGIC_LAMBDA, 50 * (tau - 0.5) ** 2)


_lib = None


ABI_VERSION = 2  # LBM_ABI_VERSION of include/lbm.h that the structures below follow


def load_library(path: str = LIB_PATH):
    """Loads the gfx950 library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    P = C.c_void_p
    sig = {
        "lbm_abi_version": ([], C.c_int),
        "lbm_last_error": ([], C.c_char_p),
        "lbm_device_name": ([C.c_int, C.c_char_p, C.c_size_t], C.c_int),
        "lbm_config2d_defaults": ([C.POINTER(Config2D)], None),
        "lbm_config3d_defaults": ([C.POINTER(Config3D)], None),
        "lbm_create_2d": ([C.POINTER(Config2D), C.c_int, C.POINTER(P)], C.c_int),
        "lbm_create_3d": ([C.POINTER(Config3D), C.c_int, C.POINTER(P)], C.c_int),
        "lbm_destroy": ([P], None),
        "lbm_update_2d": ([P, C.POINTER(Config2D)], C.c_int),
        "lbm_update_3d": ([P, C.POINTER(Config3D)], C.c_int),
        "lbm_set_inlet_ramp": ([P, C.c_int, C.c_double, C.c_double], C.c_int),
        "lbm_set_open_faces": ([P, C.c_int], C.c_int),
        "lbm_debug_fail_rebuilds": ([P, C.c_int], C.c_int),
        "lbm_set_sdf": ([P, C.c_void_p, C.c_size_t], C.c_int),
        "lbm_rebuild_flags": ([P], C.c_int),
        "lbm_brush": ([P, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int], C.c_int),
        "lbm_brush_stroke": ([P, C.c_void_p, C.c_size_t, C.c_double, C.c_int], C.c_int),
        "lbm_init_field": ([P, C.c_void_p, C.c_size_t], C.c_int),
        "lbm_reset_forces": ([P], C.c_int),
        "lbm_enqueue_steps": ([P, C.c_int64], C.c_int),
        "lbm_enqueue_steps_noalloc": ([P, C.c_int64], C.c_int),
        "lbm_enqueue_macro": ([P], C.c_int),
        "lbm_run": ([P, C.c_int64, C.c_int64], C.c_int),
        "lbm_sync": ([P], C.c_int),
        "lbm_read_macro": ([P, C.c_void_p, C.c_size_t], C.c_int),
        "lbm_read_forces_begin": ([P, C.POINTER(C.c_int64)], C.c_int),
        "lbm_read_forces_wait": ([P, C.c_int64], C.c_int),
        "lbm_read_forces_end": ([P, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)], C.c_int),
        "lbm_read_forces": ([P, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)], C.c_int),
        "lbm_diverged": ([P, C.POINTER(C.c_int64)], C.c_int),
        "lbm_batch_diverged": ([P, C.c_void_p, C.c_size_t], C.c_int),
        "lbm_info": ([P, C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                      C.POINTER(C.c_int32)], C.c_int),
        "lbm_read_populations": ([P, C.c_void_p, C.c_size_t], C.c_int),
        "lbm_read_flags": ([P, C.c_void_p, C.c_size_t], C.c_int),
        "lbm_time_steps": ([P, C.c_int64, C.POINTER(C.c_double)], C.c_int),
        "lbm_profile_steps": ([P, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)],
                              C.c_int),
        "lbm_node_counts": ([P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
        "lbm_autotune": ([P, C.c_int32, C.POINTER(C.c_int32)], C.c_int),
        "lbm_launch_info": ([P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "lbm_set_launch": ([P, C.c_int32, C.c_int32], C.c_int),
        "lbm_population_stride": ([P, C.POINTER(C.c_int64)], C.c_int),
        "lbm_set_population_stride": ([P, C.c_int64], C.c_int),
        "lbm_population_layout": ([P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)], C.c_int),
        "lbm_set_population_layout": ([P, C.c_int32, C.c_int64], C.c_int),
        "lbm_copy_roofline": ([C.c_int, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double)],
                              C.c_int),
        "lbm_copy_roofline_filled": ([C.c_int, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_double)], C.c_int),
        "lbm_pull_copy_roofline": ([C.c_int, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    C.POINTER(C.c_double), C.POINTER(C.c_int64)], C.c_int),
        "lbm_adjoint_forward": ([P, C.POINTER(AdjointConfig), C.POINTER(C.c_double)], C.c_int),
        "lbm_adjoint_backward": ([P, C.c_double, C.c_void_p, C.c_size_t], C.c_int),
        "lbm_adjoint_info": ([P] + [C.POINTER(C.c_int64)] * 5, C.c_int),
        "lbm_adjoint_release": ([P], C.c_int),
        "lbm_batch_create_2d": ([C.POINTER(Config2D), C.c_int32, C.c_int, C.POINTER(P)], C.c_int),
        "lbm_batch_destroy": ([P], None),
        "lbm_batch_size": ([P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)], C.c_int),
        "lbm_batch_tunnel": ([P, C.c_int32], P),
        "lbm_batch_set_inlet_ramp": ([P, C.c_int, C.c_double], C.c_int),
        "lbm_batch_reset_clock": ([P], C.c_int),
        "lbm_batch_enqueue_steps": ([P, C.c_int64], C.c_int),
        "lbm_batch_sync": ([P], C.c_int),
        "lbm_batch_time_steps": ([P, C.c_int64, C.POINTER(C.c_double)], C.c_int),
        "lbm_batch_autotune": ([P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "lbm_batch_launch_info": ([P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "lbm_batch_set_launch": ([P, C.c_int32, C.c_int32], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.lbm_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{path} has ABI {lib.lbm_abi_version()}, this module expects {ABI_VERSION}: rebuild it")
    _lib = lib
    return lib


class LbmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def _check(rc: int):
    if rc != LBM_OK:
        raise LbmError(rc, _lib.lbm_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class _Base:
    dim = 2
    Q = 9

    def __init__(self, handle, cfg: dict, owned: bool = True):
        self._h = handle
        self._owned = owned
        self.cfg = cfg
        dims = np.zeros(3, dtype=np.int32)
        n = C.c_int64()
        step = C.c_int64()
        q = C.c_int32()
        prec = C.c_int32()
        _check(_lib.lbm_info(self._h, _ptr(dims), C.byref(n), C.byref(step), C.byref(q), C.byref(prec)))
        self.W, self.H, self.D = (int(v) for v in dims)
        self.N = int(n.value)
        self.Q = int(q.value)
        self.precision = int(prec.value)

    # -- readonly step counter (src/solver.ts:57)
    @property
    def step(self) -> int:
        step = C.c_int64()
        _check(_lib.lbm_info(self._h, None, None, C.byref(step), None, None))
        return int(step.value)

    def set_sdf(self, sdf):
        a = np.ascontiguousarray(sdf, dtype=np.float32)
        _check(_lib.lbm_set_sdf(self._h, _ptr(a), a.size))

    def rebuild_flags(self):
        """rebuildFlags (src/solver.ts:205-217): flags, boundary tables and refill from the current SDF."""
        _check(_lib.lbm_rebuild_flags(self._h))

    def debug_fail_rebuilds(self, n: int):
        """Test hook: the next n boundary-table rebuilds fail as a device allocation failure would."""
        _check(_lib.lbm_debug_fail_rebuilds(self._h, int(n)))

    rebuildFlags = rebuild_flags

    def reset_forces(self):
        _check(_lib.lbm_reset_forces(self._h))

    def encode_steps(self, n: int, _pass=None):
        """encodeSteps: the pass argument of the reference has no equivalent and is ignored."""
        _check(_lib.lbm_enqueue_steps(self._h, int(n)))

    def encode_steps_noalloc(self, n: int):
        """encodeSteps without any allocation: the force reduction's scratch is the library's static
        per-GPU array (calls on one GPU must not overlap)."""
        _check(_lib.lbm_enqueue_steps_noalloc(self._h, int(n)))

    def encode_macro(self, _pass=None):
        _check(_lib.lbm_enqueue_macro(self._h))

    def run(self, n: int, chunk: int = 500):
        _check(_lib.lbm_run(self._h, int(n), int(chunk)))

    def sync(self):
        _check(_lib.lbm_sync(self._h))

    def set_inlet_ramp(self, kind: str, u_target: float, steps: float):
        k = {"off": 0, "linear": 1, "smoothstep": 2}[kind]
        _check(_lib.lbm_set_inlet_ramp(self._h, k, float(u_target), float(steps)))

    def read_macro(self) -> np.ndarray:
        out = np.empty(self.N * 4, dtype=np.float32)
        _check(_lib.lbm_read_macro(self._h, _ptr(out), out.size))
        return out

    def read_forces_raw(self):
        """(steps int64[n], values float64[n, dim]) since the last read."""
        steps = np.empty(HISTORY_LEN, dtype=np.int64)
        vals = np.empty(HISTORY_LEN * self.dim, dtype=np.float64)
        n = C.c_int64()
        _check(_lib.lbm_read_forces(self._h, _ptr(steps), _ptr(vals), HISTORY_LEN, C.byref(n)))
        k = int(n.value)
        return steps[:k].copy(), vals[: k * self.dim].reshape(k, self.dim).copy()

    def read_forces(self):
        steps, vals = self.read_forces_raw()
        T = ForceSample if self.dim == 2 else ForceSample3
        return [T(int(s), *map(float, v)) for s, v in zip(steps, vals)]

    def diverged_at(self) -> int:
        """Step of the first non-finite force sample since initField / resetForces, 0 if none
        (hasDiverged, src/tunnel3d.ts:97-99, kept on the device)."""
        v = C.c_int64()
        _check(_lib.lbm_diverged(self._h, C.byref(v)))
        return int(v.value)

    def has_diverged(self) -> bool:
        return self.diverged_at() != 0

    hasDiverged = has_diverged

    def read_populations(self) -> np.ndarray:
        """Current distribution buffer, decoded to float32 (Q, N) of (f - w)."""
        if self.precision == 0:
            out = np.empty(self.Q * self.N, dtype=np.float32)
            _check(_lib.lbm_read_populations(self._h, _ptr(out), out.nbytes))
            return out.reshape(self.Q, self.N)
        raw = self.read_populations_raw()
        return decode(raw, self.precision).reshape(self.Q, self.N)

    def read_populations_raw(self) -> np.ndarray:
        dt = np.float32 if self.precision == 0 else np.uint16
        out = np.empty(self.Q * self.N, dtype=dt)
        _check(_lib.lbm_read_populations(self._h, _ptr(out), out.nbytes))
        return out

    def read_flags(self) -> np.ndarray:
        out = np.empty(self.N, dtype=np.uint32)
        _check(_lib.lbm_read_flags(self._h, _ptr(out), out.size))
        return out

    def time_steps(self, n: int) -> float:
        """Milliseconds of n steps between two HIP events on the solver's stream."""
        ms = C.c_double()
        _check(_lib.lbm_time_steps(self._h, int(n), C.byref(ms)))
        return float(ms.value)

    def profile_steps(self, n: int):
        """(total, summed per-step, slowest step) milliseconds of n steps, an event after every step."""
        t, b, l = C.c_double(), C.c_double(), C.c_double()
        _check(_lib.lbm_profile_steps(self._h, int(n), C.byref(t), C.byref(b), C.byref(l)))
        return float(t.value), float(b.value), float(l.value)

    def autotune(self, trial_steps: int = 50) -> int:
        """Picks the fastest bulk block width for this grid (state is restored); returns it."""
        c = C.c_int32()
        _check(_lib.lbm_autotune(self._h, int(trial_steps), C.byref(c)))
        return int(c.value)

    def launch_info(self):
        """(threads per bulk block, variant: bits 0-1 row order 0 fixed, 1 / 2 alternating; in 3D bit
        2 inlet/outlet nodes share their wave's stores, bit 3 loads issued before the node kind is
        known) in use."""
        b, w = C.c_int32(), C.c_int32()
        _check(_lib.lbm_launch_info(self._h, C.byref(b), C.byref(w)))
        return int(b.value), int(w.value)

    def set_launch(self, block: int, variant: int = 0):
        """Pins the launch shape instead of autotune (same meaning as launch_info)."""
        _check(_lib.lbm_set_launch(self._h, int(block), int(variant)))

    def population_stride(self) -> int:
        """Elements between consecutive populations in the device buffers: N, or for D3Q19 a padded
        stride the autotuner or set_population_stride chose (host-visible layouts stay Q x N)."""
        v = C.c_int64()
        _check(_lib.lbm_population_stride(self._h, C.byref(v)))
        return int(v.value)

    def set_population_stride(self, stride: int = 0):
        """Moves the populations to another stride (0 or N: unpadded; else a multiple of 64 >= N)."""
        _check(_lib.lbm_set_population_stride(self._h, int(stride)))

    def population_layout(self):
        """(tiled, stride): tiled 1 for tiles of 64 cells (D3Q19), else one array per population."""
        t, v = C.c_int32(), C.c_int64()
        _check(_lib.lbm_population_layout(self._h, C.byref(t), C.byref(v)))
        return int(t.value), int(v.value)

    def set_population_layout(self, tiled: bool, stride: int = 0):
        _check(_lib.lbm_set_population_layout(self._h, int(bool(tiled)), int(stride)))

    def node_counts(self):
        """(bulk, list, solid) node counts."""
        b, l, s = C.c_int64(), C.c_int64(), C.c_int64()
        _check(_lib.lbm_node_counts(self._h, C.byref(b), C.byref(l), C.byref(s)))
        return int(b.value), int(l.value), int(s.value)

    def destroy(self):
        if self._h:
            if self._owned:
                _lib.lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # camelCase names of the reference
    setSdf = set_sdf
    resetForces = reset_forces
    encodeSteps = encode_steps
    encodeMacro = encode_macro
    readMacro = read_macro
    readForces = read_forces


def decode(raw: np.ndarray, precision: int) -> np.ndarray:
    """Storage codec decode (src/shaders/common3d.ts:95): FP16 x 2^-15, BF16 as the top half of a float32."""
    if precision in (1, 3):
        return raw.view(np.float16).astype(np.float32) * np.float32(1.0 / 32768.0)
    if precision == 2:
        return (raw.astype(np.uint32) << 16).view(np.float32)
    return raw


class Solver(_Base):
    """D2Q9 solver (src/solver.ts:50-340)."""

    dim = 2
    Q = 9

    @staticmethod
    def create(config: dict, device: int = 0) -> "Solver":
        lib = load_library()
        c = make_config2d(config)
        h = C.c_void_p()
        _check(lib.lbm_create_2d(C.byref(c), int(device), C.byref(h)))
        full = dict(DEFAULTS2)
        full.update({KEYS2.get(k, k): v for k, v in config.items()})
        return Solver(h, full)

    def update(self, changes: Optional[dict] = None, **kw):
        """update (src/solver.ts:175-178): merge into the config and rewrite the parameters."""
        ch = dict(changes or {}, **kw)
        self.cfg.update({KEYS2.get(k, k): v for k, v in ch.items()})
        _check(_lib.lbm_update_2d(self._h, C.byref(make_config2d(self.cfg))))

    def init_field(self, macro=None):
        """initField (src/solver.ts:220-234): equilibrium at the (rho, ux, uy, -) field, or rest."""
        if macro is None:
            _check(_lib.lbm_init_field(self._h, None, 0))
        else:
            m = np.ascontiguousarray(macro, dtype=np.float32)
            _check(_lib.lbm_init_field(self._h, _ptr(m), m.size))

    initField = init_field

    def brush(self, cx: float, cy: float, r: float, erase: bool = False, rebuild: bool = True):
        """brush (src/solver.ts:249-268): union (or, erasing, subtract) a disc into the SDF."""
        _check(_lib.lbm_brush(self._h, float(cx), float(cy), float(r), int(bool(erase)), int(bool(rebuild))))

    def brush_stroke(self, points, r: float, erase: bool = False):
        """brushStroke (src/solver.ts:243-247): a disc at every (x, y) point, then one rebuild."""
        xy = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 2))
        _check(_lib.lbm_brush_stroke(self._h, _ptr(xy), xy.shape[0], float(r), int(bool(erase))))

    brushStroke = brush_stroke


class Solver3D(_Base):
    """D3Q19 solver (src/solver3d.ts:42-319)."""

    dim = 3
    Q = 19

    @staticmethod
    def create(config: dict, device: int = 0) -> "Solver3D":
        lib = load_library()
        c = make_config3d(config)
        h = C.c_void_p()
        _check(lib.lbm_create_3d(C.byref(c), int(device), C.byref(h)))
        full = dict(DEFAULTS3)
        full.update({KEYS3.get(k, k): v for k, v in config.items()})
        return Solver3D(h, full)

    def _push(self):
        _check(_lib.lbm_update_3d(self._h, C.byref(make_config3d(self.cfg))))

    def set_tau(self, tau: float):
        """setTau (src/solver3d.ts:182-185)."""
        self.cfg["tau"] = tau
        self._push()

    def set_smagorinsky(self, cs: float):
        """setSmagorinsky (src/solver3d.ts:187-190)."""
        self.cfg["smagorinsky"] = cs
        self._push()

    def set_inlet(self, u: float):
        """setInlet (src/solver3d.ts:192-195)."""
        self.cfg["u_in"] = u
        self._push()

    def set_open_faces(self, kind: str):
        """Open x planes: 'equilibrium' (the reference's, src/shaders/step3d.ts:190-209) or 'zou-he'
        (Hecht-Harting velocity inlet and rho = 1 pressure outlet, the nodes then collide)."""
        _check(_lib.lbm_set_open_faces(self._h, {"equilibrium": 0, "zou-he": 1}[kind]))

    def init_field(self):
        """initField (src/solver3d.ts:231-242): inflow equilibrium in fluid, zero in solids."""
        _check(_lib.lbm_init_field(self._h, None, 0))

    setTau = set_tau
    setSmagorinsky = set_smagorinsky
    setInlet = set_inlet
    initField = init_field


class Batch:
    """Batched D2Q9 tunnels advanced by one launch per step (include/lbm.h, lbm_batch_*)."""

    @staticmethod
    def create(configs, device: int = 0) -> "Batch":
        lib = load_library()
        arr = (Config2D * len(configs))(*[make_config2d(c) for c in configs])
        h = C.c_void_p()
        _check(lib.lbm_batch_create_2d(arr, len(configs), int(device), C.byref(h)))
        b = Batch()
        b._h = h
        b.tunnels = []
        for i, c in enumerate(configs):
            th = lib.lbm_batch_tunnel(h, i)
            full = dict(DEFAULTS2)
            full.update({KEYS2.get(k, k): v for k, v in c.items()})
            b.tunnels.append(Solver(th, full, owned=False))
        return b

    def size(self):
        n = C.c_int32()
        cells = C.c_int64()
        _check(_lib.lbm_batch_size(self._h, C.byref(n), C.byref(cells)))
        return int(n.value), int(cells.value)

    def tunnel(self, i: int) -> "Solver":
        _lib.lbm_batch_tunnel(self._h, int(i))  # marks the batch tables for refresh
        return self.tunnels[i]

    def set_inlet_ramp(self, kind: str, steps: float):
        _check(_lib.lbm_batch_set_inlet_ramp(self._h, {"off": 0, "linear": 1, "smoothstep": 2}[kind], float(steps)))

    def reset_clock(self):
        _check(_lib.lbm_batch_reset_clock(self._h))

    def enqueue_steps(self, n: int):
        _check(_lib.lbm_batch_enqueue_steps(self._h, int(n)))

    def run(self, n: int):
        self.enqueue_steps(n)
        self.sync()

    def sync(self):
        _check(_lib.lbm_batch_sync(self._h))

    def diverged_at(self) -> np.ndarray:
        """Per tunnel, the step of its first non-finite force sample (0: none so far)."""
        out = np.zeros(len(self.tunnels), dtype=np.int64)
        _check(_lib.lbm_batch_diverged(self._h, _ptr(out), out.size))
        return out

    def time_steps(self, n: int) -> float:
        ms = C.c_double()
        _check(_lib.lbm_batch_time_steps(self._h, int(n), C.byref(ms)))
        return float(ms.value)

    def autotune(self, trial_steps: int = 20):
        """Times the tile widths, tunnel orders and bulk-tile directions on the current state
        (restored); returns the chosen (tile, order): order bit 0 largest tunnel first, bit 1 bulk
        tiles walked from the end on every other step."""
        t, o = C.c_int32(), C.c_int32()
        _check(_lib.lbm_batch_autotune(self._h, int(trial_steps), C.byref(t), C.byref(o)))
        return int(t.value), int(o.value)

    def launch_info(self):
        t, o = C.c_int32(), C.c_int32()
        _check(_lib.lbm_batch_launch_info(self._h, C.byref(t), C.byref(o)))
        return int(t.value), int(o.value)

    def set_launch(self, tile: int, order: int = 0):
        _check(_lib.lbm_batch_set_launch(self._h, int(tile), int(order)))

    def destroy(self):
        if getattr(self, "_h", None):
            for t in self.tunnels:
                t._h = None
            _lib.lbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def device_name(device: int = 0) -> str:
    buf = C.create_string_buffer(256)
    _check(load_library().lbm_device_name(int(device), buf, len(buf)))
    return buf.value.decode()


def copy_roofline(n_cells: int, q: int = 19, elem_bytes: int = 4, iters: int = 200, device: int = 0,
                  workgroup: int = 256, filled: bool = False) -> float:
    """GB/s of the copy-roofline kernel: q populations + one flag word read, q populations written; over
    zeroed buffers as the reference page does, or (filled) over flow-like values."""
    lib = load_library()
    ms = C.c_double()
    _check(lib.lbm_copy_roofline_filled(int(device), int(n_cells), int(q), int(elem_bytes), int(workgroup),
                                        int(bool(filled)), int(iters), C.byref(ms)))
    return n_cells * (q * elem_bytes * 2 + 4) * iters / (ms.value * 1e-3) / 1e9


def pull_copy_roofline(W: int, H: int, D: int, elem_bytes: int = 4, iters: int = 200, device: int = 0,
                       filled: bool = True) -> float:
    """GB/s of the pull-copy kernel (a D3Q19 step's bytes and shifted streams, no arithmetic), over
    flow-like values (filled, the step's ceiling) or zeroed buffers."""
    lib = load_library()
    ms, cells = C.c_double(), C.c_int64()
    _check(lib.lbm_pull_copy_roofline(int(device), int(W), int(H), int(D), int(elem_bytes), int(bool(filled)),
                                      int(iters), C.byref(ms), C.byref(cells)))
    return cells.value * (19 * elem_bytes * 2 + 4) * iters / (ms.value * 1e-3) / 1e9


def bytes_per_cell(q: int, precision: int) -> int:
    """Algorithmic bytes per cell update (src/bench.ts:24-28): Q loads + Q stores + one flag word."""
    b = 4 if precision == 0 else 2
    return q * b * 2 + 4


def lambda_for(tau: float, lam: Optional[float]) -> float:
    return trt_lambda3(tau) if lam is None or (isinstance(lam, float) and math.isnan(lam)) else lam
