"""Python binding of the gfx950 C ABI (include/airbox_hip.h) via ctypes.

The host-side mirror of the reference's backend API for this path
(src/engine.js: prepare -> setupSim/GpuSim -> steps/probe/read -> evaluate).
It loads ``libairbox_hip.so`` built in-tree; when the library or a gfx950 GPU
is missing every call raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libairbox_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "airbox_hip.h")

F32, F32_F64, F64 = 0, 1, 2
PRECISIONS = {"f32": F32, "f32_f64": F32_F64, "f64": F64}

# D3Q19 velocity set in the reference's numbering (engine.js:4-9)
CX = (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0)
CY = (0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1)
CZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1)


class AbxError(RuntimeError):
    """A failed library call; the message is the library's (the reference's
    own text where the reference raises the same error)."""


class PrepParams(C.Structure):
    _fields_ = [("dx", C.c_double), ("sealAxis", C.c_int), ("sealSign", C.c_int), ("seal", C.c_double),
                ("inletAxis", C.c_int), ("inletSign", C.c_int), ("inletYaw", C.c_double), ("cy", C.c_double),
                ("sy", C.c_double), ("has_click", C.c_int), ("click", C.c_double * 3), ("pleat", C.c_double),
                ("has_outlet", C.c_int), ("out_center", C.c_double * 3), ("out_diameter", C.c_double),
                ("out_height", C.c_double), ("max_cells", C.c_double)]


class GeomInfo(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("N", C.c_int64), ("nf", C.c_int64),
                ("top", C.c_int), ("jm", C.c_int), ("kw0", C.c_int), ("h", C.c_double),
                ("origin", C.c_double * 3), ("mn", C.c_double * 3), ("mx", C.c_double * 3),
                ("nHoles", C.c_int), ("nOpen", C.c_int), ("nOut", C.c_int), ("Af", C.c_int), ("leak", C.c_int),
                ("pleatOk", C.c_int), ("has_outlet", C.c_int),
                ("volume_l", C.c_double), ("inlet_cm2", C.c_double), ("filter_cm2", C.c_double),
                ("outlet_cm2", C.c_double), ("outlet_cm2_real", C.c_double),
                ("nOpenPts", C.c_int64), ("nFilterPts", C.c_int64), ("nOutletPts", C.c_int64),
                ("dAx", C.c_int), ("iAx", C.c_int), ("oAx", C.c_int), ("sd", C.c_int), ("si", C.c_int),
                ("seal", C.c_double), ("yaw", C.c_double), ("cy", C.c_double), ("sy", C.c_double),
                ("pleat", C.c_double)]


class PackDesc(C.Structure):
    """abx_pack_desc: one layer of abx_conv3d_pack_all."""
    _fields_ = [("w", C.c_void_p), ("fwd", C.c_void_p), ("bwd", C.c_void_p), ("co", C.c_int), ("taps", C.c_int),
                ("ci", C.c_int)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("fluid", C.c_void_p), ("opening", C.c_void_p),
                ("outlet", C.c_void_p), ("dc", C.c_void_p), ("top", C.c_int), ("jm", C.c_int),
                ("pleat", C.c_double)]


class RunParams(C.Structure):
    _fields_ = [("K", C.c_double), ("aniso", C.c_double), ("u", C.c_double), ("tau", C.c_double)]


class SimInfo(C.Structure):
    _fields_ = [("NF", C.c_int64), ("Ain", C.c_int), ("Af", C.c_int), ("Lp", C.c_int), ("nM", C.c_int),
                ("nT", C.c_int), ("u0", C.c_double), ("tau0", C.c_double), ("sig", C.c_double),
                ("aniso", C.c_double), ("cs2s", C.c_double), ("precision", C.c_int), ("B", C.c_int),
                ("device_bytes", C.c_int64), ("plane", C.c_int64), ("nrows", C.c_int), ("nseg", C.c_int)]


class EvalGeom(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("top", C.c_int), ("jm", C.c_int),
                ("h", C.c_double), ("pleat", C.c_double), ("mn_z", C.c_double), ("opening", C.c_void_p),
                ("xc", C.c_void_p), ("dc", C.c_void_p), ("zc", C.c_void_p), ("oAx", C.c_int), ("dAx", C.c_int),
                ("iAx", C.c_int), ("sd", C.c_int), ("si", C.c_int), ("seal", C.c_double), ("cy", C.c_double),
                ("sy", C.c_double)]


class Metrics(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("gamma", "gammaCore", "lowCore", "highCore", "reverse", "maxToMean",
                                          "zeta", "inlet_cm2", "filter_cm2")] + \
               [(k, C.c_int) for k in ("ic", "jc", "jlow", "jb", "kB")] + \
               [("rhoRef", C.c_double), ("pin", C.c_double), ("n", C.c_int), ("cn", C.c_int)]


METRIC_KEYS = ("gamma", "gammaCore", "lowCore", "highCore", "reverse", "maxToMean", "zeta", "inlet_cm2",
               "filter_cm2")


class SteadyParams(C.Structure):
    _fields_ = [("window", C.c_int), ("max_steps", C.c_int), ("step0", C.c_int), ("tol_gamma", C.c_double),
                ("tol_zeta", C.c_double)]


class SteadyWindow(C.Structure):
    _fields_ = [("step", C.c_int), ("gamma", C.c_double), ("massRatio", C.c_double), ("metrics", Metrics)]


class SteadyResult(C.Structure):
    _fields_ = [("converged", C.c_int), ("steps", C.c_int), ("windows", C.c_int), ("metrics", Metrics)]


STEADY_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(SteadyWindow))


def metrics_dict(M: Metrics) -> dict:
    return {k: getattr(M, k) for k in METRIC_KEYS}

G_SOLID, G_FLUID, G_OPENING, G_OUTLET, G_XC, G_DC, G_ZC, G_OPENPTS, G_FILTERPTS, G_OUTLETPTS, G_FINE = range(11)

_lib = None


def lib_path() -> str:
    return LIB_PATH


def load() -> C.CDLL:
    """Load libairbox_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AbxError(f"{LIB_PATH} is missing: build it with `make -C airboxcalc_amd/csrc` "
                       "or __graft_entry__.build()")
    # One HIP runtime per process: torch's ROCm wheel carries its own
    # libamdhip64 (soname libamdhip64.so.7, but requested as libamdhip64.so).
    # Loaded after ours it becomes a second runtime that finds no GPU; loaded
    # first it also satisfies this library's libamdhip64.so.7.  So torch, when
    # installed, is imported before the library is opened.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, d, f32p = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_void_p
    sig = {
        "abx_last_error": (C.c_char_p, []),
        "abx_version": (C.c_char_p, []),
        "abx_device_open": (i32, [i32, C.POINTER(vp), C.c_char_p, C.c_size_t]),
        "abx_device_close": (None, [vp]),
        "abx_copy_bandwidth": (i32, [vp, i64, i32, C.POINTER(d), C.POINTER(d)]),
        "abx_aa_pattern_bandwidth": (i32, [vp, i64, i32, C.POINTER(d), C.POINTER(d)]),
        "abx_parse_stl": (i32, [C.c_char_p, i64, vp, i64, C.POINTER(i64)]),
        "abx_voxelize": (i32, [vp, vp, i64, vp, i32, i32, i32, d, vp]),
        "abx_prepare": (i32, [vp, vp, i64, C.POINTER(PrepParams), C.POINTER(vp)]),
        "abx_geom_info_get": (i32, [vp, C.POINTER(GeomInfo)]),
        "abx_geom_read": (i32, [vp, i32, vp, i64]),
        "abx_geom_free": (None, [vp]),
        "abx_geom_grid": (i32, [vp, C.POINTER(Grid)]),
        "abx_sim_create": (i32, [vp, C.POINTER(Grid), C.POINTER(RunParams), i32, i32, vp, C.POINTER(vp)]),
        "abx_sim_info_get": (i32, [vp, C.POINTER(SimInfo)]),
        "abx_sim_steps": (i32, [vp, i32, i32]),
        "abx_sim_probe": (i32, [vp, i32, vp]),
        "abx_sim_read_acc": (i32, [vp, i32, d, vp]),
        "abx_sim_read_acc_f64": (i32, [vp, i32, d, vp]),
        "abx_sim_read_f": (i32, [vp, i32, vp]),
        "abx_sim_write_f": (i32, [vp, i32, vp]),
        "abx_sim_init_eq": (i32, [vp, vp]),
        "abx_sim_init_eq_resample": (i32, [vp, vp, i32, i32, i32, i32, i32, i32, i32, vp, C.POINTER(d), d,
                                           C.POINTER(d), d, C.POINTER(C.c_float)]),
        "abx_sim_clear_acc": (i32, [vp]),
        "abx_sim_set_ns": (i32, [vp, vp]),
        "abx_sim_sync": (i32, [vp]),
        "abx_sim_set_launch": (i32, [vp, i32, i32, i32]),
        "abx_sim_pattern_time": (i32, [vp, i32, C.POINTER(d), C.POINTER(d)]),
        "abx_sim_time_steps": (i32, [vp, i32, i32, C.POINTER(d)]),
        "abx_sim_set_occupancy": (i32, [vp, i32, i32]),
        "abx_sim_destroy": (None, [vp]),
        "abx_sim_reset": (i32, [vp]),
        "abx_sim_set_graph": (i32, [vp, i32]),
        "abx_batch_create": (i32, [vp, C.POINTER(vp), i32, C.POINTER(vp)]),
        "abx_batch_steps": (i32, [vp, i32, i32]),
        "abx_batch_probe": (i32, [vp, i32, vp]),
        "abx_batch_sync": (i32, [vp]),
        "abx_batch_set_launch": (i32, [vp, i32, i32, i32]),
        "abx_batch_tune": (i32, [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(d)]),
        "abx_batch_set_xcd_group": (i32, [vp, i32]),
        "abx_batch_get_xcd_group": (i32, [vp]),
        "abx_batch_set_occupancy": (i32, [vp, i32]),
        "abx_batch_get_occupancy": (i32, [vp]),
        "abx_batch_time_steps": (i32, [vp, i32, i32, C.POINTER(d)]),
        "abx_batch_destroy": (None, [vp]),
        "abx_geom_eval_view": (i32, [vp, C.POINTER(EvalGeom)]),
        "abx_evaluate": (i32, [vp, C.POINTER(EvalGeom), i32, d, C.POINTER(Metrics), vp, vp, vp, vp, vp, vp, vp]),
        "abx_evaluate_metrics_noalloc_bytes": (i64, [i32, i32, i32]),
        "abx_evaluate_metrics_noalloc": (i32, [vp, C.POINTER(EvalGeom), i32, d, C.POINTER(Metrics), vp]),
        "abx_streamlines": (i32, [vp, C.POINTER(EvalGeom), i32, d, i32, vp, i64, vp, i64, vp, i64,
                                  C.POINTER(i64), C.POINTER(i64)]),
        "abx_sim_solve_to_steady": (i32, [vp, C.POINTER(EvalGeom), C.POINTER(SteadyParams), STEADY_CB, vp,
                                          C.POINTER(SteadyResult)]),
        "abx_surrogate_load": (i32, [vp, C.c_char_p, C.POINTER(vp)]),
        "abx_surrogate_info": (i32, [vp, C.POINTER(i32), C.POINTER(d)]),
        "abx_surrogate_free": (None, [vp]),
        "abx_surrogate_warm_start": (i32, [vp, vp, vp, vp, i64, C.POINTER(PrepParams), d, d]),
        # flow surrogate kernels (stream first; tensors as device addresses)
        "abx_conv3d_pack": (i32, [vp, vp, i32, i32, i32, i32, vp]),
        "abx_conv3d_pack_all": (i32, [vp, i32, vp]),
        "abx_conv3d_set_kernel": (i32, [i32]),
        "abx_surrogate_features_workspace": (i64, [i32, i32, i32]),
        "abx_surrogate_features": (i32, [vp, vp, vp, vp, vp, i32, i32, i32, i32, d, d, d, i32, i32, i32, vp, i64,
                                         vp, i32, i32]),
        "abx_conv3d_forward": (i32, [vp, vp, i32, i32, i32, i32, i32, i32, vp, vp, i32, i32, i32, i32, vp, i32]),
        "abx_conv3d_forward_gated": (i32, [vp, vp, i32, i32, i32, i32, i32, i32, vp, vp, i32, i32, i32, i32, vp, i32,
                                           vp, i32]),
        "abx_conv3d_wgrad_workspace": (i64, [i32, i32, i32, i32, i32, i32, i32]),
        "abx_conv3d_wgrad": (i32, [vp, vp, i32, vp, i32, i32, i32, i32, i32, i32, i32, i32, vp, i64, vp, vp, i32]),
        "abx_conv3d_forward_parts": (i32, [vp, vp, i32, i32, vp, i32, i32, i32, i32, i32, i32, vp, vp, i32, i32, i32,
                                           vp, i32, i32, vp, i32]),
        "abx_conv3d_wgrad_parts": (i32, [vp, vp, i32, i32, vp, i32, vp, i32, i32, i32, i32, i32, i32, i32, vp, i64,
                                         vp, vp, i32]),
        "abx_maxpool2": (i32, [vp, vp, i32, i32, i32, i32, i32, i32, vp, i32]),
        "abx_maxpool2_backward": (i32, [vp, vp, i32, vp, i32, i32, i32, i32, i32, i32, vp, i32, i32]),
        "abx_maxpool2_backward_gated": (i32, [vp, vp, i32, vp, i32, i32, i32, i32, i32, i32, vp, i32, i32]),
        "abx_upsample2": (i32, [vp, vp, i32, i32, i32, i32, i32, i32, vp, i32]),
        "abx_upsample2_backward": (i32, [vp, vp, i32, i32, i32, i32, i32, i32, vp, i32]),
        "abx_upsample2_backward_gated": (i32, [vp, vp, i32, i32, i32, i32, i32, i32, vp, i32, vp, i32]),
        "abx_relu_backward": (i32, [vp, vp, i32, vp, i32, i64, i32]),
        "abx_mse_partials": (i64, [i64]),
        "abx_head_loss_partials": (i64, [i64]),
        "abx_head_loss_workspace": (i64, [i64]),
        "abx_head_loss_backward": (i32, [vp, vp, i32, i32, vp, vp, i32, vp, i32, vp, i64, vp, C.c_float, vp, vp, i32,
                                         vp, i64, vp, vp]),
        "abx_mse_loss": (i32, [vp, vp, i32, vp, i32, vp, i64, i32, vp, C.c_float, vp, vp, i32]),
        "abx_adam": (i32, [vp, vp, vp, vp, vp, i64, C.c_float, C.c_float, C.c_float, C.c_float, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def declared_functions() -> list:
    """Names of the functions include/airbox_hip.h declares."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(abx_[a-z0-9_]+)\s*\(", txt)))


def _check(rc: int) -> None:
    if rc != 0:
        raise AbxError(load().abx_last_error().decode("utf-8"))


def _p(a: Optional[np.ndarray]):
    return None if a is None else C.c_void_p(a.ctypes.data)


class Device:
    """A gfx950 GPU (replaces gpuDevice, engine.js:479-491)."""

    def __init__(self, ordinal: int = 0):
        L = load()
        self.h = C.c_void_p()
        name = C.create_string_buffer(256)
        _check(L.abx_device_open(ordinal, C.byref(self.h), name, 256))
        self.name = name.value.decode()
        self.ordinal = int(ordinal)

    def close(self):
        if self.h:
            load().abx_device_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def aa_pattern_bandwidth(dev: "Device", positions: int, iters: int = 20):
    """(GB/s at 152 B per position, ms per pass) of the AA step's memory
    pattern over `positions` dense positions (abx_aa_pattern_bandwidth)."""
    g, ms = C.c_double(), C.c_double()
    _check(load().abx_aa_pattern_bandwidth(dev.h, int(positions), iters, C.byref(g), C.byref(ms)))
    return g.value, ms.value


def copy_bandwidth(dev: "Device", nbytes: int, iters: int = 20):
    """(GB/s, ms per copy) of a plain float4 device copy of nbytes."""
    g, ms = C.c_double(), C.c_double()
    _check(load().abx_copy_bandwidth(dev.h, int(nbytes), iters, C.byref(g), C.byref(ms)))
    return g.value, ms.value


_DEV = {}


def device(ordinal: int = 0) -> Device:
    if ordinal not in _DEV:
        _DEV[ordinal] = Device(ordinal)
    return _DEV[ordinal]


def parse_stl(data: bytes) -> np.ndarray:
    """parseSTL (engine.js:15-28)."""
    L = load()
    n = C.c_int64()
    _check(L.abx_parse_stl(data, len(data), None, 0, C.byref(n)))
    out = np.empty(n.value, np.float32)
    _check(L.abx_parse_stl(data, len(data), _p(out), out.size, C.byref(n)))
    return out


def voxelize(dev: Device, T: np.ndarray, origin: Sequence[float], nx: int, ny: int, nz: int, h: float) -> np.ndarray:
    """voxelize (engine.js:45-93) on the GPU."""
    T = np.ascontiguousarray(T, np.float32).ravel()
    o = np.ascontiguousarray(origin, np.float64)
    vox = np.empty(nx * ny * nz, np.uint8)
    _check(load().abx_voxelize(dev.h, _p(T), T.size // 9, _p(o), nx, ny, nz, h, _p(vox)))
    return vox


def prep_params(p: dict, max_cells: float = 4.5e6, cy=None, sy=None) -> PrepParams:
    P = PrepParams()
    P.dx = float(p["dx"])
    P.sealAxis, P.sealSign, P.seal = int(p["sealAxis"]), int(p["sealSign"]), float(p["seal"])
    P.inletAxis, P.inletSign = int(p["inletAxis"]), int(p["inletSign"])
    P.inletYaw = float(p.get("inletYaw") or 0.0)
    P.cy = math.nan if cy is None else cy
    P.sy = math.nan if sy is None else sy
    if p.get("click") is not None:
        P.has_click = 1
        for a in range(3):
            P.click[a] = float(p["click"][a])
    P.pleat = float(p["pleat"])
    o = p.get("outlet")
    if o and o.get("enabled"):
        P.has_outlet = 1
        for a in range(3):
            P.out_center[a] = float(o["center"][a])
        P.out_diameter, P.out_height = float(o["diameter"]), float(o["height"])
    P.max_cells = max_cells
    return P


class Geometry:
    """Prepared geometry (prepare, engine.js:110-234)."""

    def __init__(self, dev: Device, tris: np.ndarray, p: dict, max_cells: float = 4.5e6, cy=None, sy=None):
        L = load()
        tris = np.ascontiguousarray(tris, np.float32).ravel()
        P = prep_params(p, max_cells, cy, sy)
        self.ptr = C.c_void_p()
        _check(L.abx_prepare(dev.h, _p(tris), tris.size, C.byref(P), C.byref(self.ptr)))
        self.tris, self.prep = tris, P   # what a warm start needs to prepare the box at another cell size
        I = GeomInfo()
        _check(L.abx_geom_info_get(self.ptr, C.byref(I)))
        self.info = I
        for k, _ in GeomInfo._fields_:
            v = getattr(I, k)
            setattr(self, k, list(v) if isinstance(v, C.Array) else v)
        self._cache = {}

    def _read(self, what: int, n: int, dtype) -> np.ndarray:
        out = np.empty(n, dtype)
        _check(load().abx_geom_read(self.ptr, what, _p(out), n))
        return out

    def array(self, name: str) -> np.ndarray:
        if name in self._cache:
            return self._cache[name]
        N, I = self.N, self.info
        spec = {"solid": (G_SOLID, N, np.uint8), "fluid": (G_FLUID, N, np.uint8),
                "opening": (G_OPENING, I.nx * I.ny, np.uint8), "outlet": (G_OUTLET, N, np.uint8),
                "xc": (G_XC, I.nx, np.float32), "dc": (G_DC, I.ny, np.float32), "zc": (G_ZC, I.nz, np.float32),
                "openPts": (G_OPENPTS, 3 * I.nOpenPts, np.float64),
                "filterPts": (G_FILTERPTS, 3 * I.nFilterPts, np.float64),
                "outletPts": (G_OUTLETPTS, 3 * I.nOutletPts, np.float64),
                "fine": (G_FINE, 27 * N, np.uint8)}[name]
        a = self._read(*spec)
        if name != "fine":
            self._cache[name] = a
        return a

    def eval_view(self) -> "EvalGeom":
        e = EvalGeom()
        _check(load().abx_geom_eval_view(self.ptr, C.byref(e)))
        e._owner = self
        return e

    def grid(self) -> "GridArrays":
        g = Grid()
        _check(load().abx_geom_grid(self.ptr, C.byref(g)))
        return GridArrays.from_struct(g, self)

    def free(self):
        if self.ptr:
            load().abx_geom_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GridArrays:
    """The lattice a simulation runs on (abx_grid) from plain arrays."""

    def __init__(self, nx, ny, nz, fluid, opening, outlet, dc, top, jm, pleat, owner=None):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.fluid = np.ascontiguousarray(fluid, np.uint8)
        self.opening = np.ascontiguousarray(opening, np.uint8)
        self.outlet = np.ascontiguousarray(outlet, np.uint8)
        self.dc = np.ascontiguousarray(dc, np.float32)
        self.top, self.jm, self.pleat = int(top), int(jm), float(pleat)
        self.owner = owner

    @classmethod
    def from_struct(cls, g: Grid, owner):
        N = g.nx * g.ny * g.nz
        arr = lambda p, n, t: np.ctypeslib.as_array(C.cast(p, C.POINTER(t)), shape=(n,))
        return cls(g.nx, g.ny, g.nz, arr(g.fluid, N, C.c_uint8), arr(g.opening, g.nx * g.ny, C.c_uint8),
                   arr(g.outlet, N, C.c_uint8), arr(g.dc, g.ny, C.c_float), g.top, g.jm, g.pleat, owner)

    def struct(self) -> Grid:
        g = Grid()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.fluid, g.opening, g.outlet, g.dc = (self.fluid.ctypes.data, self.opening.ctypes.data,
                                              self.outlet.ctypes.data, self.dc.ctypes.data)
        g.top, g.jm, g.pleat = self.top, self.jm, self.pleat
        return g


class Sim:
    """setupSim + GpuSim (engine.js:239-264, 493-594) on the GPU."""

    def __init__(self, dev: Device, grid: GridArrays, K=3.0, aniso=10.0, u=0.06, tau=0.51, precision="f32",
                 B: int = 1, ns: Optional[np.ndarray] = None):
        L = load()
        self.prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.B = B
        rp = RunParams(K, aniso, u, tau)
        gs = grid.struct()
        self.grid = grid
        self.h = C.c_void_p()
        nsa = None if ns is None else np.ascontiguousarray(ns, np.float32)
        _check(L.abx_sim_create(dev.h, C.byref(gs), C.byref(rp), self.prec, B, _p(nsa), C.byref(self.h)))
        I = SimInfo()
        _check(L.abx_sim_info_get(self.h, C.byref(I)))
        self.info = I
        self.NF = I.NF
        self.sdtype = np.float64 if self.prec == F64 else np.float32

    def steps(self, n: int, doacc: bool = False) -> None:
        _check(load().abx_sim_steps(self.h, n, 1 if doacc else 0))

    def time_steps(self, n: int, doacc: bool = False) -> float:
        ms = C.c_double()
        _check(load().abx_sim_time_steps(self.h, n, 1 if doacc else 0, C.byref(ms)))
        return ms.value

    def probe(self, step: int = 0):
        out = np.zeros(3)
        _check(load().abx_sim_probe(self.h, step, _p(out)))
        return float(out[0]), float(out[1]), float(out[2])

    def read_acc(self, divisor: float = 1.0, b: int = 0, f64: bool = False) -> np.ndarray:
        if f64:
            out = np.empty(4 * self.NF, np.float64)
            _check(load().abx_sim_read_acc_f64(self.h, b, divisor, _p(out)))
        else:
            out = np.empty(4 * self.NF, np.float32)
            _check(load().abx_sim_read_acc(self.h, b, divisor, _p(out)))
        return out

    def read_rho_u(self, b: int = 0) -> np.ndarray:
        """(rho, ux, uy, uz) of the current populations in cid order (the
        moments of read_f; host side)."""
        f = self.read_f(b).astype(np.float64).reshape(19, self.NF)
        rho = f.sum(0)
        u = [(np.asarray(c, np.float64)[:, None] * f).sum(0) / rho for c in (CX, CY, CZ)]
        return np.concatenate([rho] + u).astype(np.float32)

    def read_f(self, b: int = 0) -> np.ndarray:
        out = np.empty(19 * self.NF, self.sdtype)
        _check(load().abx_sim_read_f(self.h, b, _p(out)))
        return out

    def write_f(self, f: np.ndarray, b: int = -1) -> None:
        f = np.ascontiguousarray(f, self.sdtype)
        assert f.size == 19 * self.NF
        _check(load().abx_sim_write_f(self.h, b, _p(f)))

    def init_eq(self, rho_u: np.ndarray) -> None:
        r = np.ascontiguousarray(rho_u, np.float32)
        assert r.size == 4 * self.NF
        _check(load().abx_sim_init_eq(self.h, _p(r)))

    def init_eq_resample(self, vol_ptr: int, cstride: int, vol_dims, coarse_dims, coarse_fluid_ptr: int,
                         origin_c, h_c: float, origin_f, h_f: float, affine) -> None:
        """abx_sim_init_eq_resample: equilibrium from a coarse predicted volume
        already on the device (vol_ptr: bf16, coarse_fluid_ptr: uint8)."""
        oc = (C.c_double * 3)(*[float(v) for v in origin_c])
        of = (C.c_double * 3)(*[float(v) for v in origin_f])
        af = (C.c_float * 8)(*[float(v) for v in affine])
        _check(load().abx_sim_init_eq_resample(self.h, C.c_void_p(vol_ptr), int(cstride), *[int(v) for v in vol_dims],
                                               *[int(v) for v in coarse_dims], C.c_void_p(coarse_fluid_ptr), oc,
                                               float(h_c), of, float(h_f), af))

    def evaluate(self, eg: "EvalGeom", divisor: float = 1.0, b: int = 0, field3d: bool = True) -> dict:
        """evaluate + field3d (engine.js:632-719) on the GPU."""
        nx, ny, nz = eg.nx, eg.ny, eg.nz
        N = nx * ny * nz
        M = Metrics()
        fmap = np.empty(nx * nz, np.float32)
        pI = np.empty(3 * nx * nz, np.float32)
        pL = np.empty(3 * nx * nz, np.float32)
        lon = np.empty(3 * nz * ny, np.float32)
        vol = [np.empty(N, np.float32) for _ in range(3)] if field3d else [None, None, None]
        _check(load().abx_evaluate(self.h, C.byref(eg), b, divisor, C.byref(M), _p(fmap), _p(pI), _p(pL), _p(lon),
                                   *[_p(v) for v in vol]))
        out = dict(metrics={k: getattr(M, k) for k in METRIC_KEYS}, fmap=fmap, planInlet=pI.reshape(3, -1),
                   planLow=pL.reshape(3, -1), lon=lon.reshape(3, -1), ic=M.ic, jc=M.jc, jlow=M.jlow, jb=M.jb,
                   kB=M.kB, rhoRef=M.rhoRef, pin=M.pin)
        if field3d:
            out["vol"] = dict(spd=vol[0], pst=vol[1], loss=vol[2], kB=M.kB)
        return out

    def solve_to_steady(self, eg: "EvalGeom", window: int = 1000, tol_gamma: float = 0.005, tol_zeta: float = 0.02,
                        max_steps: int = 60000, step0: int = 0, on_window=None) -> dict:
        """abx_sim_solve_to_steady: windows of `window` averaged steps until
        gamma-core and zeta agree between two consecutive windows.  Returns
        {converged, steps, metrics (mean of the two agreeing windows),
        windows: [(step, gamma, massRatio, gammaCore, zeta)]}; on_window(dict)
        is called after every window (returning True stops the solve)."""
        wins, raised = [], []

        def cb(_user, wp):
            w = wp.contents
            m = metrics_dict(w.metrics)
            wins.append((w.step, w.gamma, w.massRatio, m["gammaCore"], m["zeta"]))
            if on_window is not None:
                try:
                    return 1 if on_window(dict(step=w.step, gamma=w.gamma, massRatio=w.massRatio, metrics=m)) else 0
                except BaseException as e:  # stop the solve, re-raise below
                    raised.append(e)
                    return 1
            return 0

        P = SteadyParams(int(window), int(max_steps), int(step0), float(tol_gamma), float(tol_zeta))
        R = SteadyResult()
        fn = STEADY_CB(cb)
        _check(load().abx_sim_solve_to_steady(self.h, C.byref(eg), C.byref(P), fn, None, C.byref(R)))
        if raised:
            raise raised[0]
        return dict(converged=bool(R.converged), steps=R.steps, metrics=metrics_dict(R.metrics), windows=wins,
                    ic=R.metrics.ic, jc=R.metrics.jc, jlow=R.metrics.jlow, jb=R.metrics.jb, kB=R.metrics.kB)

    def streamlines(self, eg: "EvalGeom", divisor: float = 1.0, b: int = 0, max_seeds: int = 220) -> dict:
        """streamlines (engine.js:721-758) on the GPU."""
        cap = 2002 * max_seeds + 16
        pts = np.empty(3 * cap, np.float32)
        spd = np.empty(cap, np.float32)
        off = np.zeros(max_seeds + 2, np.uint32)
        npts, nl = C.c_int64(), C.c_int64()
        _check(load().abx_streamlines(self.h, C.byref(eg), b, divisor, max_seeds, _p(pts), pts.size, _p(spd),
                                      spd.size, _p(off), off.size, C.byref(npts), C.byref(nl)))
        n = npts.value
        return dict(pts=pts[:3 * n].copy(), spd=spd[:n].copy(), off=off[:nl.value + 1].copy())

    def clear_acc(self) -> None:
        _check(load().abx_sim_clear_acc(self.h))

    def set_ns(self, ns: Optional[np.ndarray]) -> None:
        nsa = None if ns is None else np.ascontiguousarray(ns, np.float32)
        _check(load().abx_sim_set_ns(self.h, _p(nsa)))

    def set_launch(self, kernel: int = 4, xcd_map: int = -1, nt: int = 19) -> None:
        """kernel 4 = AA pattern in one lattice (default), 2 = two lattices
        with pull streaming.  xcd_map: block order across XCDs (0 round
        robin, 1 contiguous, G > 1 groups of G blocks per XCD; negative: the
        library's default, groups of 4).  nt: non-temporal access flags of
        the AA steps, even step in bits 0-7, odd step in bits 8-15
        (abx_sim_set_launch; default 19)."""
        _check(load().abx_sim_set_launch(self.h, kernel, xcd_map, nt))

    def set_occupancy(self, even_waves_per_cu: int = 16, odd_waves_per_cu: int = 0) -> None:
        """Occupancy caps of the AA steps in waves per CU, 0 = none
        (abx_sim_set_occupancy; the defaults are the library's)."""
        _check(load().abx_sim_set_occupancy(self.h, even_waves_per_cu, odd_waves_per_cu))

    def pattern_time(self, iters: int = 20) -> tuple:
        """(ms even, ms odd): the AA steps' memory pattern alone on this
        handle's lattice (abx_sim_pattern_time).  Moves the populations."""
        me, mo = C.c_double(), C.c_double()
        _check(load().abx_sim_pattern_time(self.h, iters, C.byref(me), C.byref(mo)))
        return me.value, mo.value

    def set_graph(self, enable: bool = True) -> None:
        _check(load().abx_sim_set_graph(self.h, 1 if enable else 0))

    def reset(self) -> None:
        """Back to the cold start f0 with empty averages."""
        _check(load().abx_sim_reset(self.h))

    def sync(self) -> None:
        _check(load().abx_sim_sync(self.h))

    def destroy(self):
        if self.h:
            load().abx_sim_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """Independent simulations (different geometries, grid sizes and
    parameters) stepped together, one launch per step (abx_batch_*)."""

    def __init__(self, dev: Device, sims):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h for s in self.sims])
        self.h = C.c_void_p()
        _check(load().abx_batch_create(dev.h, arr, len(self.sims), C.byref(self.h)))

    def steps(self, n: int, doacc: bool = False) -> None:
        _check(load().abx_batch_steps(self.h, n, 1 if doacc else 0))

    def probe(self, step: int = 0) -> np.ndarray:
        out = np.zeros(3 * len(self.sims))
        _check(load().abx_batch_probe(self.h, step, _p(out)))
        return out.reshape(-1, 3)

    def sync(self) -> None:
        _check(load().abx_batch_sync(self.h))

    def set_launch(self, spb: int = 4, persistent: bool = False, graph: bool = False) -> None:
        _check(load().abx_batch_set_launch(self.h, spb, 1 if persistent else 0, 1 if graph else 0))

    def tune(self, trial_steps: int = 20):
        """Autotune the launch shape; members restart from f0 afterwards.
        The XCD group it picks stays set (set_launch does not change it)."""
        spb, pers, ms = C.c_int(), C.c_int(), C.c_double()
        _check(load().abx_batch_tune(self.h, trial_steps, C.byref(spb), C.byref(pers), C.byref(ms)))
        return dict(spb=spb.value, persistent=bool(pers.value), xcd_group=self.xcd_group,
                    even_waves_per_cu=load().abx_batch_get_occupancy(self.h), ms_per_step=ms.value)

    @property
    def xcd_group(self) -> int:
        return load().abx_batch_get_xcd_group(self.h)

    def set_xcd_group(self, group: int) -> None:
        _check(load().abx_batch_set_xcd_group(self.h, group))

    def set_occupancy(self, even_waves_per_cu: int) -> None:
        _check(load().abx_batch_set_occupancy(self.h, even_waves_per_cu))

    def time_steps(self, n: int, doacc: bool = False) -> float:
        ms = C.c_double()
        _check(load().abx_batch_time_steps(self.h, n, 1 if doacc else 0, C.byref(ms)))
        return ms.value

    def destroy(self):
        if self.h:
            load().abx_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Surrogate:
    """The trained flow surrogate on the device (abx_surrogate_load): the
    warm start of the C ABI, without torch."""

    def __init__(self, dev: Device, path: str):
        self.h = C.c_void_p()
        self.dev = dev
        _check(load().abx_surrogate_load(dev.h, path.encode(), C.byref(self.h)))
        dims = (C.c_int * 3)()
        fdx = C.c_double()
        _check(load().abx_surrogate_info(self.h, dims, C.byref(fdx)))
        self.dims, self.family_dx = tuple(dims), fdx.value

    def warm_start(self, sim: "Sim", g: "Geometry", K: float, aniso: float) -> None:
        """Seed sim (on g's lattice) at the equilibrium of the prediction
        (abx_surrogate_warm_start)."""
        _check(load().abx_surrogate_warm_start(self.h, sim.h, g.ptr, _p(g.tris), g.tris.size, C.byref(g.prep),
                                               float(K), float(aniso)))

    def free(self):
        if self.h:
            load().abx_surrogate_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gpu_available() -> bool:
    """True when libairbox_hip.so loads and a gfx950 GPU opens."""
    try:
        device(0)
        return True
    except Exception:
        return False
