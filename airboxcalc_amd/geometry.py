"""Procedural air-filter boxes, written as triangulated STL.

The reference ships no sample geometry (its CLI names ``Koerper1.stl``,
airbox_cfd.py:45, which is not in the repository), so every workload here is
generated.  A box is the lower half of an air-filter housing:

* model axes: X free, Y = depth from the seal plane (box in +Y, seal at Y=0,
  floor at Y=LY), Z = from the inlet side into the box;
* walls of thickness ``t`` on X-min, X-max, Z-max and the floor (Y=LY); the
  seal side (Y=0) is open, that is where the pleated filter sits;
* the inlet wall carries a circular hole.  With ``yaw_deg`` != 0 the inlet wall
  is tilted about Y so that its outward normal is (sin yaw, 0, -cos yaw): the
  situation the reference's inlet picking turns into ``inletYaw``
  (ui.html:425-431).  ``prepare`` then rotates its lattice to the inlet axis;
* optional internal baffles hang from the floor towards the seal plane.

The solid is emitted as a union of closed, pairwise disjoint pieces.  The
reference voxelizer fills by ray parity along the local Z axis
(engine.js:45-93); parity of a union of disjoint closed pieces is the union as
long as coincident interface faces are emitted as *identical* triangles (the
two hits then coincide exactly and cancel).  Every face that two pieces share
goes through :func:`_quad`, which canonicalises vertex order and diagonal.

Randomised variants draw from ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import math
import struct
from dataclasses import asdict, dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

Point = Tuple[float, float, float]


@dataclass
class Baffle:
    """Plate hanging from the floor.  ``normal`` is "x" or "z" (model axis the
    plate is perpendicular to); ``pos`` its lower coordinate along that axis;
    ``lo``/``hi`` its extent along the other in-plane axis; ``top`` the depth
    Y of its upper edge (it reaches down to the floor)."""
    normal: str
    pos: float
    thick: float
    lo: float
    hi: float
    top: float


@dataclass
class AirboxSpec:
    LX: float = 300.0
    LY: float = 150.0
    LZ: float = 200.0
    wall: float = 6.0
    inlet_d: float = 70.0
    yaw_deg: float = 0.0
    pleat: float = 40.0
    outlet_d: Optional[float] = 60.0
    outlet_h: float = 50.0
    baffles: List[Baffle] = field(default_factory=list)
    segments: int = 64          # polygon segments of the inlet hole
    subdiv: int = 1             # split every planar quad into subdiv x subdiv cells

    # -- derived -----------------------------------------------------------
    @property
    def cx(self) -> float:
        return 0.5 * self.LX

    def zo(self, x: float) -> float:
        """Z of the outer inlet-wall face at model X (face passes Z=0 at cx)."""
        return math.tan(math.radians(self.yaw_deg)) * (x - self.cx)

    def zi(self, x: float) -> float:
        """Z of the inner inlet-wall face at model X."""
        return self.zo(x) + self.wall / math.cos(math.radians(self.yaw_deg))

    def inlet_centre(self) -> Point:
        return (self.cx, 0.5 * self.LY, 0.0)

    def outlet_centre(self) -> Point:
        return (self.cx, 0.0, 0.5 * (self.zi(self.cx) + self.LZ - self.wall))


class _Mesh:
    def __init__(self, subdiv: int):
        self.chunks: List[np.ndarray] = []
        self.subdiv = max(1, int(subdiv))

    def tri(self, a: Point, b: Point, c: Point) -> None:
        self.chunks.append(np.array([(*a, *b, *c)], dtype=np.float64))

    def quad(self, p00: Point, p10: Point, p11: Point, p01: Point, n: int = None) -> None:
        """Planar quad p00-p10-p11-p01, subdivided n x n, split the same way
        whatever corner it is given from (canonical order).  Vertex (iu, iv) is
        the bilinear point at u = iu/n, v = iv/n (corners exact); cells in
        (iu, iv) order, two triangles each."""
        n = self.subdiv if n is None else n
        corners = [p00, p10, p11, p01]
        # canonical start: the lexicographically smallest corner, then the
        # smaller of its two neighbours as the "u" direction
        s = min(range(4), key=lambda q: corners[q])
        nxt, prv = corners[(s + 1) % 4], corners[(s + 3) % 4]
        if nxt <= prv:
            order = [corners[(s + q) % 4] for q in range(4)]
        else:
            order = [corners[(s - q) % 4] for q in range(4)]
        a, b, c, d = (np.array(v, dtype=np.float64) for v in order)
        t = np.arange(n + 1) / n
        u = t[:, None, None]
        v = t[None, :, None]
        Q = (1 - u) * (1 - v) * a + u * (1 - v) * b + u * v * c + (1 - u) * v * d
        Q[0, 0], Q[n, 0], Q[n, n], Q[0, n] = a, b, c, d
        q00, q10, q11, q01 = Q[:-1, :-1], Q[1:, :-1], Q[1:, 1:], Q[:-1, 1:]
        t1 = np.concatenate([q00, q10, q11], -1)
        t2 = np.concatenate([q00, q11, q01], -1)
        self.chunks.append(np.stack([t1, t2], 2).reshape(-1, 9))

    def array(self) -> np.ndarray:
        if not self.chunks:
            return np.zeros((0, 9), np.float32)
        return np.concatenate(self.chunks).astype(np.float32)


def _prism(m: _Mesh, poly_xz: Sequence[Tuple[float, float]], y0: float, y1: float) -> None:
    """Closed prism: convex polygon in model X-Z extruded over Y in [y0, y1]."""
    P = list(poly_xz)
    for q in range(len(P)):
        a, b = P[q], P[(q + 1) % len(P)]
        if a == b:
            continue
        m.quad((a[0], y0, a[1]), (b[0], y0, b[1]), (b[0], y1, b[1]), (a[0], y1, a[1]))
    # caps (parallel to every Z ray: they never produce a hit, but keep it closed)
    for y in (y0, y1):
        for q in range(1, len(P) - 1):
            m.tri((P[0][0], y, P[0][1]), (P[q][0], y, P[q][1]), (P[q + 1][0], y, P[q + 1][1]))


def _ring_face(u_lo, u_hi, d_lo, d_hi, dh, r, angles):
    """Rectangle [u_lo,u_hi]x[d_lo,d_hi] minus a disc (centre (0,dh), radius r),
    as ring quads between the disc polygon and the rectangle boundary, one per
    angle interval.  Returns (circle_pts, rect_pts, rect_side) per angle; side
    codes: 0 u_lo, 1 u_hi, 2 d_lo, 3 d_hi."""
    circ, rect, side = [], [], []
    for th in angles:
        c, s = math.cos(th), math.sin(th)
        best, code = math.inf, -1
        if c > 1e-15 and (u_hi / c) < best:
            best, code = u_hi / c, 1
        if c < -1e-15 and (u_lo / c) < best:
            best, code = u_lo / c, 0
        if s > 1e-15 and ((d_hi - dh) / s) < best:
            best, code = (d_hi - dh) / s, 3
        if s < -1e-15 and ((d_lo - dh) / s) < best:
            best, code = (d_lo - dh) / s, 2
        u, d = best * c, dh + best * s
        if code == 0:
            u = u_lo
        elif code == 1:
            u = u_hi
        elif code == 2:
            d = d_lo
        else:
            d = d_hi
        circ.append((r * c, dh + r * s))
        rect.append((u, d))
        side.append(code)
    return circ, rect, side


def airbox_triangles(spec: AirboxSpec) -> np.ndarray:
    """Triangle soup (float32, shape (ntri, 9)) of the closed airbox shell."""
    S = spec
    t, LX, LY, LZ = S.wall, S.LX, S.LY, S.LZ
    D1 = LY - t
    phi = math.radians(S.yaw_deg)
    cp, sp = math.cos(phi), math.sin(phi)
    m = _Mesh(S.subdiv)

    # canonical corners of the tilted inlet wall (model X-Z)
    A0 = (t, S.zo(t))            # outer face at the left end
    A1 = (LX - t, S.zo(LX - t))  # outer face at the right end
    B0 = (t, S.zi(t))            # inner face, left end
    B1 = (LX - t, S.zi(LX - t))  # inner face, right end
    L0 = (0.0, S.zo(0.0))
    R1 = (LX, S.zo(LX))

    # side walls run from the outer inlet line to the back; the back wall sits
    # between them.  Vertices on shared edges are listed so that shared faces
    # are exactly the same quads.
    _prism(m, [L0, A0, B0, (t, LZ - t), (t, LZ), (0.0, LZ)], 0.0, D1)
    _prism(m, [A1, R1, (LX, LZ), (LX - t, LZ), (LX - t, LZ - t), B1], 0.0, D1)
    _prism(m, [(t, LZ - t), (LX - t, LZ - t), (LX - t, LZ), (t, LZ)], 0.0, D1)
    # floor under everything
    _prism(m, [L0, A0, A1, R1, (LX, LZ), (LX - t, LZ), (t, LZ), (0.0, LZ)], D1, LY)

    # inlet wall with its hole, built in the wall frame (u along the wall,
    # d = model Y, w = depth along the inward normal (-sin, cos))
    def wall_to_model(u, d, w):
        return (S.cx + u * cp - w * sp, d, u * sp + w * cp)

    dh, r = 0.5 * LY, 0.5 * S.inlet_d
    ua0, ua1 = (t - S.cx) / cp, (LX - t - S.cx) / cp
    ub0, ub1 = (t - S.cx + t * sp) / cp, (LX - t - S.cx + t * sp) / cp
    if not (r < dh - 1.0 and dh + r < D1 - 1.0 and -r > max(ua0, ub0) + 1.0 and r < min(ua1, ub1) - 1.0):
        raise ValueError("inlet hole does not fit the inlet wall")
    n = max(8, int(S.segments))
    angles = [2 * math.pi * q / n for q in range(n)]
    for (u_lo, u_hi) in ((ua0, ua1), (ub0, ub1)):
        for (uu, dd) in ((u_lo, 0.0), (u_hi, 0.0), (u_hi, D1), (u_lo, D1)):
            angles.append(math.atan2(dd - dh, uu) % (2 * math.pi))
    angles = sorted(set(angles))

    faces = []
    for w, (u_lo, u_hi), corner_lo, corner_hi in ((0.0, (ua0, ua1), A0, A1), (t, (ub0, ub1), B0, B1)):
        circ, rect, side = _ring_face(u_lo, u_hi, 0.0, D1, dh, r, angles)

        def to_model_rect(ud, code, clo=corner_lo, chi=corner_hi, w=w):
            u, d = ud
            if code == 0:     # on the left end: exact corner X/Z, only Y varies
                return (clo[0], d, clo[1])
            if code == 1:
                return (chi[0], d, chi[1])
            return wall_to_model(u, d, w)
        cm = [wall_to_model(u, d, w) for (u, d) in circ]
        rm = [to_model_rect(ud, code) for ud, code in zip(rect, side)]
        faces.append((cm, rm))
        na = len(angles)
        for q in range(na):
            q1 = (q + 1) % na
            # ring quad; subdivided like any other quad
            m.quad(cm[q], cm[q1], rm[q1], rm[q], n=1)
    # hole wall (parallel to the inlet axis)
    (cf, _), (cb, _) = faces
    for q in range(len(angles)):
        q1 = (q + 1) % len(angles)
        m.quad(cf[q], cf[q1], cb[q1], cb[q], n=1)
    # ends of the inlet wall: exactly the quads the side walls carry
    m.quad((A0[0], 0.0, A0[1]), (B0[0], 0.0, B0[1]), (B0[0], D1, B0[1]), (A0[0], D1, A0[1]))
    m.quad((A1[0], 0.0, A1[1]), (B1[0], 0.0, B1[1]), (B1[0], D1, B1[1]), (A1[0], D1, A1[1]))
    # caps of the inlet wall at Y=0 and Y=D1 (parallel to the rays)
    for y in (0.0, D1):
        m.quad((A0[0], y, A0[1]), (A1[0], y, A1[1]), (B1[0], y, B1[1]), (B0[0], y, B0[1]), n=1)

    for b in S.baffles:
        if b.normal == "x":
            poly = [(b.pos, b.lo), (b.pos + b.thick, b.lo), (b.pos + b.thick, b.hi), (b.pos, b.hi)]
        else:
            poly = [(b.lo, b.pos), (b.hi, b.pos), (b.hi, b.pos + b.thick), (b.lo, b.pos + b.thick)]
        _prism(m, poly, b.top, D1)
    return m.array()


def random_baffles(rng: np.random.Generator, spec: AirboxSpec, nmax: int = 3,
                   thick=(2.0, 4.0), margin: float = 8.0) -> List[Baffle]:
    """0..nmax floor-hung baffles, uniformly placed inside the cavity, pairwise
    disjoint and clear of the walls, the inlet hole and the pleat layer."""
    S = spec
    nb = int(rng.integers(0, nmax + 1))
    out: List[Baffle] = []
    t = S.wall
    tries = 0
    while len(out) < nb and tries < 200:
        tries += 1
        th = float(rng.uniform(*thick))
        top = float(rng.uniform(S.pleat + 20.0, S.LY - t - 30.0))
        if rng.random() < 0.5:
            pos = float(rng.uniform(t + margin, S.LX - t - margin - th))
            zmin = max(S.zi(pos), S.zi(pos + th)) + margin + 20.0
            zmax = S.LZ - t - margin
            if zmax - zmin < 30.0:
                continue
            lo = float(rng.uniform(zmin, zmax - 20.0))
            hi = float(rng.uniform(lo + 20.0, zmax))
            cand = Baffle("x", pos, th, lo, hi, top)
            box = (pos, pos + th, lo, hi)
        else:
            lo = float(rng.uniform(t + margin, S.LX - t - margin - 20.0))
            hi = float(rng.uniform(lo + 20.0, S.LX - t - margin))
            zmin = max(S.zi(lo), S.zi(hi)) + margin + 20.0
            zmax = S.LZ - t - margin - th
            if zmax - zmin < 5.0:
                continue
            pos = float(rng.uniform(zmin, zmax))
            cand = Baffle("z", pos, th, lo, hi, top)
            box = (lo, hi, pos, pos + th)
        ok = True
        for o in out:
            ob = (o.pos, o.pos + o.thick, o.lo, o.hi) if o.normal == "x" else (o.lo, o.hi, o.pos, o.pos + o.thick)
            if not (box[1] + 2.0 < ob[0] or ob[1] + 2.0 < box[0] or box[3] + 2.0 < ob[2] or ob[3] + 2.0 < box[2]):
                ok = False
                break
        if ok:
            out.append(cand)
    return out


def seeded_variant(seed: int, kind: str = "roofline") -> Tuple[AirboxSpec, dict]:
    """Seeded box variants.

    kind="roofline": 300x150x200 mm, 0-3 baffles (2-4 mm), yaw uniform 0-30 deg.
    kind="batch":    outer dims uniform 200-360 mm per axis, dx in {2,2.5,3,4},
                     K uniform 1.5-6, aniso uniform 5-20.
    kind="surrogate": the "roofline" box of the same seed (same draws, same
                     geometry) plus K uniform 1.5-6 and aniso uniform 5-20.
    Returns (spec, extra) where extra holds per-variant run parameters."""
    rng = np.random.default_rng(seed)
    extra = {}
    if kind in ("roofline", "surrogate"):
        yaw = round(float(rng.uniform(0.0, 30.0)), 1)
        spec = AirboxSpec(yaw_deg=yaw)
        spec.baffles = random_baffles(rng, spec)
        if kind == "surrogate":
            extra["K"] = float(rng.uniform(1.5, 6.0))
            extra["aniso"] = float(rng.uniform(5.0, 20.0))
    elif kind == "batch":
        LX, LY, LZ = (float(v) for v in rng.uniform(200.0, 360.0, size=3))
        spec = AirboxSpec(LX=LX, LY=LY, LZ=LZ, inlet_d=min(70.0, 0.4 * min(LX, LY)))
        extra["dx"] = float(rng.choice([2.0, 2.5, 3.0, 4.0]))
        extra["K"] = float(rng.uniform(1.5, 6.0))
        extra["aniso"] = float(rng.uniform(5.0, 20.0))
    else:
        raise ValueError(kind)
    return spec, extra


def prepare_params(spec: AirboxSpec, dx: float, outlet: bool = True, click: bool = True) -> dict:
    """Parameter object of the reference's ``prepare`` (ui.html:477-479) for
    this box: seal plane Y=0 with the box in +Y, inlet on the Z-min side."""
    p = {
        "dx": float(dx), "sealAxis": 1, "sealSign": -1, "seal": 0.0,
        "inletAxis": 2, "inletSign": -1, "inletYaw": float(spec.yaw_deg),
        "click": list(spec.inlet_centre()) if click else None,
        "pleat": float(spec.pleat),
        "outlet": None,
    }
    if outlet and spec.outlet_d:
        p["outlet"] = {"enabled": True, "center": list(spec.outlet_centre()),
                       "diameter": float(spec.outlet_d), "height": float(spec.outlet_h)}
    return p


def run_params(steps: int, avg: int, every: int, K: float = 3.0, aniso: float = 10.0,
               u: float = 0.06, tau: float = 0.51, dx: float = 4.0) -> dict:
    """Run parameter object (ui.html:502-504)."""
    return {"device": "auto", "K": float(K), "aniso": float(aniso), "u": float(u), "tau": float(tau),
            "steps": int(steps), "avg": int(avg), "every": int(every), "dx": float(dx), "outlet": ""}


def stl_binary_bytes(tris: np.ndarray) -> bytes:
    """Binary STL (80-byte header, count, 50-byte records) of a triangle soup."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    rec = np.zeros(len(tris), dtype=[("n", "<f4", 3), ("v", "<f4", 9), ("a", "<u2")])
    rec["v"] = tris
    return b"airboxcalc_amd procedural airbox".ljust(80, b" ") + struct.pack("<I", len(tris)) + rec.tobytes()


def write_stl_binary(path: str, tris: np.ndarray) -> None:
    with open(path, "wb") as fh:
        fh.write(stl_binary_bytes(tris))


def write_stl_ascii(path: str, tris: np.ndarray) -> None:
    tris = np.asarray(tris, dtype=np.float32).reshape(-1, 9)
    with open(path, "w") as fh:
        fh.write("solid airbox\n")
        for t in tris:
            fh.write(" facet normal 0 0 0\n  outer loop\n")
            for v in range(3):
                fh.write("   vertex %.9g %.9g %.9g\n" % (t[3 * v], t[3 * v + 1], t[3 * v + 2]))
            fh.write("  endloop\n endfacet\n")
        fh.write("endsolid airbox\n")


def spec_dict(spec: AirboxSpec) -> dict:
    """Plain-dict form of a spec (JSON-serialisable)."""
    return asdict(spec)


def spec_from_dict(d: dict) -> AirboxSpec:
    d = dict(d)
    d["baffles"] = [Baffle(**b) for b in d.get("baffles", [])]
    return AirboxSpec(**d)


# -- variant sets on disk: STL files plus one manifest -----------------------

# cell size each family is solved at when its variants carry none
FAMILY_DX = {"roofline": 1.0, "surrogate": 2.0}


def _variant_name(kind: str, seed: int) -> str:
    return f"{kind}_{int(seed):07d}.stl"


def write_variant_set(out_dir: str, kind: str, seeds: Sequence[int], subdiv: int = 1, ascii: bool = False) -> dict:
    """Write the seeded variants of one family as STL files into out_dir and
    describe them in out_dir/manifest.json: per variant its seed, file name,
    triangle count, the SHA-256 of the file, the full spec (enough to
    regenerate the mesh), the per-variant run parameters (dx, K, aniso), and
    the parameter objects of the reference's prepare and run for it.  Returns
    the manifest."""
    import hashlib
    import json
    import os
    os.makedirs(out_dir, exist_ok=True)
    variants = []
    for seed in seeds:
        spec, extra = seeded_variant(int(seed), kind)
        spec.subdiv = int(subdiv)
        tris = airbox_triangles(spec)
        name = _variant_name(kind, seed)
        path = os.path.join(out_dir, name)
        (write_stl_ascii if ascii else write_stl_binary)(path, tris)
        with open(path, "rb") as fh:
            sha = hashlib.sha256(fh.read()).hexdigest()
        dx = float(extra.get("dx", FAMILY_DX.get(kind, 2.0)))
        variants.append({"seed": int(seed), "file": name, "format": "ascii" if ascii else "binary",
                         "triangles": int(len(tris)), "sha256": sha, "spec": spec_dict(spec),
                         "run": {"dx": dx, "K": float(extra.get("K", 3.0)), "aniso": float(extra.get("aniso", 10.0))},
                         "prepare": prepare_params(spec, dx)})
    manifest = {"generator": "airboxcalc_amd.geometry.seeded_variant", "kind": kind, "subdiv": int(subdiv),
                "count": len(variants), "variants": variants}
    with open(os.path.join(out_dir, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)
    return manifest


def verify_variant_set(out_dir: str) -> List[str]:
    """Check a written set against its manifest: every file's SHA-256, its
    triangle count, and that regenerating the mesh from the recorded spec
    gives the file's triangles bit for bit.  Returns the problems found."""
    import hashlib
    import json
    import os
    with open(os.path.join(out_dir, "manifest.json")) as fh:
        manifest = json.load(fh)
    problems = []
    for v in manifest["variants"]:
        path = os.path.join(out_dir, v["file"])
        with open(path, "rb") as fh:
            data = fh.read()
        if hashlib.sha256(data).hexdigest() != v["sha256"]:
            problems.append(f"{v['file']}: SHA-256 differs")
        tris = airbox_triangles(spec_from_dict(v["spec"]))
        if len(tris) != v["triangles"]:
            problems.append(f"{v['file']}: {len(tris)} triangles regenerated, {v['triangles']} recorded")
        elif v["format"] == "binary" and data != stl_binary_bytes(tris):
            problems.append(f"{v['file']}: the recorded spec does not regenerate the file")
    return problems


def _parse_seeds(text: str) -> List[int]:
    """'0-63' or '1,5,9' or a mix ('0-3,10')."""
    out = []
    for part in text.split(","):
        a, _, b = part.partition("-")
        out.extend(range(int(a), int(b) + 1) if b else [int(a)])
    return out


def main(argv: Optional[Sequence[str]] = None) -> int:
    import argparse
    ap = argparse.ArgumentParser(prog="python -m airboxcalc_amd.geometry",
                                 description="write seeded airbox variants as STL with a manifest, or verify a set")
    ap.add_argument("--kind", default="batch", choices=["roofline", "batch", "surrogate"])
    ap.add_argument("--seeds", default="0-63", help="'0-63', '1,5,9' or a mix")
    ap.add_argument("--subdiv", type=int, default=1, help="split every planar quad into subdiv x subdiv cells")
    ap.add_argument("--ascii", action="store_true")
    ap.add_argument("--out", required=True, help="output folder (STL files and manifest.json)")
    ap.add_argument("--verify", action="store_true", help="check an existing folder against its manifest")
    a = ap.parse_args(argv)
    if a.verify:
        problems = verify_variant_set(a.out)
        for p in problems:
            print(p)
        return 1 if problems else 0
    m = write_variant_set(a.out, a.kind, _parse_seeds(a.seeds), a.subdiv, a.ascii)
    print(f"{m['count']} {a.kind} variants written to {a.out} ({sum(v['triangles'] for v in m['variants'])} triangles)")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
