"""ORACLE - test infrastructure only.

ctypes wrapper of ``libabx_oracle.so``, the plain-C restatement of the
reference compute core (src/engine.js) in ``engine_oracle.c``.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use it, always as
the checker or the CPU baseline, never as the thing measured or shipped.

Each wrapper names the engine.js lines its C function restates.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libabx_oracle.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class PrepParams(C.Structure):
    _fields_ = [("dx", C.c_double), ("sealAxis", C.c_int), ("sealSign", C.c_int), ("seal", C.c_double),
                ("inletAxis", C.c_int), ("inletSign", C.c_int), ("yaw_deg", C.c_double), ("cy", C.c_double),
                ("sy", C.c_double), ("has_click", C.c_int), ("click", C.c_double * 3), ("pleat", C.c_double),
                ("has_outlet", C.c_int), ("out_center", C.c_double * 3), ("out_diameter", C.c_double),
                ("out_height", C.c_double), ("max_cells", C.c_double)]


class GeomS(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nxf", C.c_int), ("nyf", C.c_int),
                ("nzf", C.c_int), ("N", C.c_int64), ("h", C.c_double), ("origin", C.c_double * 3),
                ("mn", C.c_double * 3), ("mx", C.c_double * 3), ("top", C.c_int), ("jm", C.c_int),
                ("kw0", C.c_int), ("nHoles", C.c_int), ("nOpen", C.c_int), ("nOut", C.c_int), ("Af", C.c_int),
                ("leak", C.c_int), ("nf", C.c_int64),
                ("fine", C.POINTER(C.c_uint8)), ("solid", C.POINTER(C.c_uint8)), ("fluid", C.POINTER(C.c_uint8)),
                ("opening", C.POINTER(C.c_uint8)), ("outlet", C.POINTER(C.c_uint8)),
                ("xc", C.POINTER(C.c_float)), ("dc", C.POINTER(C.c_float)), ("zc", C.POINTER(C.c_float)),
                ("T", C.POINTER(C.c_float)),
                ("openPts", C.POINTER(C.c_double)), ("filterPts", C.POINTER(C.c_double)),
                ("outletPts", C.POINTER(C.c_double)),
                ("nOpenPts", C.c_int64), ("nFilterPts", C.c_int64), ("nOutletPts", C.c_int64),
                ("oAx", C.c_int), ("dAx", C.c_int), ("iAx", C.c_int), ("sd", C.c_int), ("si", C.c_int),
                ("seal", C.c_double), ("cy", C.c_double), ("sy", C.c_double), ("yaw_deg", C.c_double),
                ("pleat", C.c_double), ("has_outlet", C.c_int)]


class SimS(C.Structure):
    _fields_ = [("NF", C.c_int64), ("cid", C.POINTER(C.c_int32)), ("cells", C.POINTER(C.c_int32)),
                ("nb", C.POINTER(C.c_int32)), ("jsOf", C.POINTER(C.c_int32)), ("flag", C.POINTER(C.c_uint32)),
                ("layerM", C.POINTER(C.c_int32)), ("layerT", C.POINTER(C.c_int32)), ("nM", C.c_int64),
                ("nT", C.c_int64), ("Ain", C.c_int), ("Af", C.c_int), ("Lp", C.c_int), ("u0", C.c_double),
                ("tau0", C.c_double), ("sig", C.c_double), ("aniso", C.c_double), ("cs2s", C.c_double)]


class MetricsS(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("gamma", "gammaCore", "lowCore", "highCore", "reverse", "maxToMean",
                                          "zeta", "inlet_cm2", "filter_cm2")] + \
               [(k, C.c_int) for k in ("ic", "jc", "jlow", "jb", "kB")] + \
               [("rhoRef", C.c_double), ("pin_field", C.c_double)]


class OracleError(RuntimeError):
    pass


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.orc_last_error.restype = C.c_char_p
        L.orc_prepare.argtypes = [_f32p, C.c_int64, C.POINTER(PrepParams), C.POINTER(GeomS)]
        L.orc_geom_free.argtypes = [C.POINTER(GeomS)]
        L.orc_setup_sim.argtypes = [C.POINTER(GeomS), C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.POINTER(SimS)]
        L.orc_sim_free.argtypes = [C.POINTER(SimS)]
        L.orc_cpu_step.argtypes = [C.POINTER(SimS), _f32p, _f32p, C.c_void_p, C.c_void_p]
        L.orc_cpu_step_f64.argtypes = [C.POINTER(SimS), _f64p, _f64p, C.c_void_p, C.c_void_p]
        L.orc_cpu_step_threads.argtypes = [C.POINTER(SimS), _f32p, _f32p, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_probe.argtypes = [C.POINTER(SimS), _f32p, C.c_int, _f64p]
        L.orc_simulate.argtypes = [C.POINTER(SimS), C.c_int, C.c_int, C.c_int, C.c_void_p, _f32p, _f64p,
                                   C.POINTER(C.c_int), C.c_int64]
        L.orc_evaluate.argtypes = [C.POINTER(GeomS), C.POINTER(SimS), C.c_double, _f32p, C.c_void_p,
                                   C.POINTER(MetricsS), _f32p, _f32p, _f32p, _f32p, _u8p]
        L.orc_field3d.argtypes = [C.POINTER(GeomS), C.POINTER(SimS), C.c_double, _f32p, _u8p, C.c_double,
                                  _f32p, _f32p, _f32p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_streamlines.restype = C.c_int64
        L.orc_streamlines.argtypes = [C.POINTER(GeomS), C.POINTER(SimS), C.c_double, _f32p, _u8p, C.c_int,
                                      _f32p, C.c_int64, _f32p, C.c_int64,
                                      np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS"), C.c_int64,
                                      C.POINTER(C.c_int64)]
        L.orc_voxelize.argtypes = [_f32p, C.c_int64, _f64p, C.c_int, C.c_int, C.c_int, C.c_double, _u8p]
        L.orc_voxel_counts.argtypes = [_f32p, C.c_int64, _f64p, C.c_int, C.c_int, C.c_double,
                                       np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")]
        L.orc_parse_stl.argtypes = [C.c_char_p, C.c_int64, C.POINTER(C.POINTER(C.c_float)),
                                    C.POINTER(C.c_int64)]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_js_hypot3.restype = C.c_double
        L.orc_js_hypot3.argtypes = [C.c_double] * 3
        _lib = L
    return _lib


def _err() -> str:
    return lib().orc_last_error().decode("utf-8")


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).astype(dtype, copy=True)


def parse_stl(data: bytes) -> np.ndarray:
    """parseSTL (engine.js:15-28)."""
    out = C.POINTER(C.c_float)()
    n = C.c_int64()
    if lib().orc_parse_stl(data, len(data), C.byref(out), C.byref(n)):
        raise OracleError(_err())
    a = np.ctypeslib.as_array(out, shape=(n.value,)).copy()
    lib().orc_free(out)
    return a


def frame_trig(yaw_deg: float):
    """cos/sin of the yaw as engine.js:36 forms them (deg * PI / 180)."""
    r = (yaw_deg or 0.0) * math.pi / 180
    return math.cos(r), math.sin(r)


def prep_params(p: dict, cy=None, sy=None, max_cells: float = 4.5e6) -> PrepParams:
    P = PrepParams()
    P.dx = p["dx"]
    P.sealAxis, P.sealSign, P.seal = int(p["sealAxis"]), int(p["sealSign"]), float(p["seal"])
    P.inletAxis, P.inletSign = int(p["inletAxis"]), int(p["inletSign"])
    P.yaw_deg = float(p.get("inletYaw") or 0.0)
    c0, s0 = frame_trig(P.yaw_deg)
    P.cy = c0 if cy is None else cy
    P.sy = s0 if sy is None else sy
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


@dataclass
class Geom:
    s: GeomS
    nx: int
    ny: int
    nz: int
    N: int
    nf: int
    h: float
    origin: np.ndarray
    mn: np.ndarray
    mx: np.ndarray
    top: int
    jm: int
    kw0: int
    fine: np.ndarray
    solid: np.ndarray
    fluid: np.ndarray
    opening: np.ndarray
    outlet: np.ndarray
    xc: np.ndarray
    dc: np.ndarray
    zc: np.ndarray
    T: np.ndarray
    openPts: np.ndarray
    filterPts: np.ndarray
    outletPts: np.ndarray
    summary: dict

    def free(self):
        lib().orc_geom_free(C.byref(self.s))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def prepare(tris: np.ndarray, p: dict, cy=None, sy=None, max_cells: float = 4.5e6) -> Geom:
    """prepare (engine.js:110-234)."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).ravel()
    P = prep_params(p, cy, sy, max_cells)
    g = GeomS()
    if lib().orc_prepare(tris, tris.size, C.byref(P), C.byref(g)):
        msg = _err()
        lib().orc_geom_free(C.byref(g))
        raise OracleError(msg)
    N, nf = g.N, g.nf
    nfine = g.nxf * g.nyf * g.nzf
    h = g.h
    summ = dict(grid=[g.nx, g.ny, g.nz], cells=nf, volume_l=nf * h ** 3 / 1e6, inlet_cm2=g.nOpen * h * h / 100,
                filter_cm2=g.Af * h * h / 100, outlet_cm2=g.leak * h * h / 100, nHoles=g.nHoles,
                outlet_cm2_real=(g.nOut * h * h / 100) if g.has_outlet else 0, pleatOk=g.jm > g.top)
    G = Geom(s=g, nx=g.nx, ny=g.ny, nz=g.nz, N=N, nf=nf, h=h, origin=np.array(g.origin[:]),
             mn=np.array(g.mn[:]), mx=np.array(g.mx[:]), top=g.top, jm=g.jm, kw0=g.kw0,
             fine=_arr(g.fine, nfine, np.uint8), solid=_arr(g.solid, N, np.uint8),
             fluid=_arr(g.fluid, N, np.uint8), opening=_arr(g.opening, g.nx * g.ny, np.uint8),
             outlet=_arr(g.outlet, N, np.uint8), xc=_arr(g.xc, g.nx, np.float32),
             dc=_arr(g.dc, g.ny, np.float32), zc=_arr(g.zc, g.nz, np.float32), T=_arr(g.T, tris.size, np.float32),
             openPts=_arr(g.openPts, 3 * g.nOpenPts, np.float64),
             filterPts=_arr(g.filterPts, 3 * g.nFilterPts, np.float64),
             outletPts=_arr(g.outletPts, 3 * g.nOutletPts, np.float64), summary=summ)
    return G


@dataclass
class Sim:
    s: SimS
    NF: int
    cid: np.ndarray
    cells: np.ndarray
    nb: np.ndarray
    flag: np.ndarray
    layerM: np.ndarray
    layerT: np.ndarray
    Ain: int
    Af: int
    Lp: int
    sig: float
    cs2s: float
    u0: float

    def f0(self, dtype=np.float32) -> np.ndarray:
        """Cold start f = WT (engine.js:260-261)."""
        w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float64)
        return np.repeat(w, self.NF).astype(dtype)

    def free(self):
        lib().orc_sim_free(C.byref(self.s))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def setup_sim(g: Geom, K=3.0, aniso=10.0, u=0.06, tau=0.51) -> Sim:
    """setupSim (engine.js:239-264)."""
    S = SimS()
    if lib().orc_setup_sim(C.byref(g.s), K, aniso, u, tau, C.byref(S)):
        raise OracleError(_err())
    NF = S.NF
    return Sim(s=S, NF=NF, cid=_arr(S.cid, g.N, np.int32), cells=_arr(S.cells, NF, np.int32),
               nb=_arr(S.nb, 19 * NF, np.int32), flag=_arr(S.flag, NF, np.uint32),
               layerM=_arr(S.layerM, S.nM, np.int32), layerT=_arr(S.layerT, S.nT, np.int32), Ain=S.Ain, Af=S.Af,
               Lp=S.Lp, sig=S.sig, cs2s=S.cs2s, u0=S.u0)


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def cpu_step(S: Sim, f: np.ndarray, f2: np.ndarray, acc: Optional[np.ndarray] = None,
             ns: Optional[np.ndarray] = None, threads: int = 1) -> None:
    """cpuStep (engine.js:266-321): f32 storage, f64 arithmetic.  With f64
    arrays the same step runs with f64 storage (the FP64 reference).  threads
    > 1 (f32 only) splits the cells over OpenMP threads; same result."""
    if ns is not None:
        assert ns.dtype == np.float32 and ns.size == S.NF
    if f.dtype == np.float64:
        assert acc is None or acc.dtype == np.float64
        lib().orc_cpu_step_f64(C.byref(S.s), f, f2, _ptr(acc), _ptr(ns))
    else:
        assert acc is None or acc.dtype == np.float32
        if threads > 1:
            lib().orc_cpu_step_threads(C.byref(S.s), f, f2, _ptr(acc), _ptr(ns), int(threads))
        else:
            lib().orc_cpu_step(C.byref(S.s), f, f2, _ptr(acc), _ptr(ns))


def run_steps(S: Sim, n: int, f: Optional[np.ndarray] = None, ns=None, acc=None, dtype=np.float32):
    """n cpuSteps from f (default f0), ping-pong like simulate (engine.js:343-349)."""
    f = S.f0(dtype) if f is None else f.copy()
    f2 = np.empty_like(f)
    for _ in range(n):
        cpu_step(S, f, f2, acc, ns)
        f, f2 = f2, f
    return f


def probe(S: Sim, f: np.ndarray, step: int = 0):
    """probeCPU + progressInfo (engine.js:324-339) -> (gamma, massRatio, qo)."""
    out = np.zeros(3)
    if lib().orc_probe(C.byref(S.s), f, step, out):
        raise OracleError(_err())
    return float(out[0]), float(out[1]), float(out[2])


def simulate(S: Sim, steps: int, avg: int, every: int, ns=None, step_limit: int = 0):
    """simulate (engine.js:341-358) up to the averaged acc: (acc, hist)."""
    acc = np.zeros(4 * S.NF, dtype=np.float32)
    hist = np.zeros(3 * (steps // max(every, 1) + 2))
    nh = C.c_int()
    if lib().orc_simulate(C.byref(S.s), steps, avg, every, _ptr(ns), acc, hist, C.byref(nh), step_limit):
        raise OracleError(_err())
    return acc, hist[: 3 * nh.value].reshape(-1, 3)


def evaluate(g: Geom, S: Sim, u0: float, acc: np.ndarray, ns=None) -> dict:
    """evaluate (engine.js:632-691) without field3d/streamlines."""
    M = MetricsS()
    nx, ny, nz = g.nx, g.ny, g.nz
    fmap = np.zeros(nx * nz, np.float32)
    pI = np.zeros(3 * nx * nz, np.float32)
    pL = np.zeros(3 * nx * nz, np.float32)
    lon = np.zeros(3 * nz * ny, np.float32)
    fl = np.zeros(g.N, np.uint8)
    if lib().orc_evaluate(C.byref(g.s), C.byref(S.s), u0, acc, _ptr(ns), C.byref(M), fmap, pI, pL, lon, fl):
        raise OracleError(_err())
    met = {k: getattr(M, k) for k in ("gamma", "gammaCore", "lowCore", "highCore", "reverse", "maxToMean",
                                      "zeta", "inlet_cm2", "filter_cm2")}
    return dict(metrics=met, fmap=fmap, planInlet=pI.reshape(3, -1), planLow=pL.reshape(3, -1),
                lon=lon.reshape(3, -1), fluid=fl, ic=M.ic, jc=M.jc, jlow=M.jlow, jb=M.jb, rhoRef=M.rhoRef)


def field3d(g: Geom, S: Sim, u0: float, acc: np.ndarray, fluid: np.ndarray, rhoRef: float) -> dict:
    """field3d (engine.js:698-719)."""
    spd = np.zeros(g.N, np.float32)
    pst = np.zeros(g.N, np.float32)
    loss = np.zeros(g.N, np.float32)
    kB = C.c_int()
    pin = C.c_double()
    lib().orc_field3d(C.byref(g.s), C.byref(S.s), u0, acc, fluid, rhoRef, spd, pst, loss, C.byref(kB),
                      C.byref(pin))
    return dict(spd=spd, pst=pst, loss=loss, kB=kB.value, pin=pin.value)


def streamlines(g: Geom, S: Sim, u0: float, acc: np.ndarray, fluid: np.ndarray, max_seeds: int = 220) -> dict:
    """streamlines (engine.js:721-758)."""
    cap = 2200 * max_seeds + 16
    pts = np.zeros(3 * cap, np.float32)
    spd = np.zeros(cap, np.float32)
    off = np.zeros(4 * max_seeds + 8, np.uint32)
    npts = C.c_int64()
    nl = lib().orc_streamlines(C.byref(g.s), C.byref(S.s), u0, acc, fluid, max_seeds, pts, pts.size, spd,
                               spd.size, off, off.size, C.byref(npts))
    n = npts.value
    return dict(pts=pts[: 3 * n].copy(), spd=spd[:n].copy(), off=off[: nl + 1].copy())


def voxelize(T: np.ndarray, origin, nx: int, ny: int, nz: int, h: float) -> np.ndarray:
    """voxelize (engine.js:45-93): parity fill along Z, Uint8 (nx*ny*nz)."""
    T = np.ascontiguousarray(T, dtype=np.float32).ravel()
    vox = np.zeros(nx * ny * nz, np.uint8)
    lib().orc_voxelize(T, T.size // 9, np.asarray(origin, np.float64), nx, ny, nz, h, vox)
    return vox


def voxel_counts(T: np.ndarray, origin, nx: int, ny: int, h: float) -> np.ndarray:
    """Per-column hit counts of voxelize's count pass (engine.js:50-75)."""
    T = np.ascontiguousarray(T, dtype=np.float32).ravel()
    cnt = np.zeros(nx * ny, np.int32)
    lib().orc_voxel_counts(T, T.size // 9, np.asarray(origin, np.float64), nx, ny, h, cnt)
    return cnt


def js_hypot3(a, b, c) -> float:
    return lib().orc_js_hypot3(a, b, c)
