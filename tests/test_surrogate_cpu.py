"""CPU checks of the surrogate's data preparation (oracle geometry, no GPU)."""
from __future__ import annotations

import types

import numpy as np
import pytest

from airboxcalc_amd import abx
from airboxcalc_amd import geometry as G
from airboxcalc_amd import surrogate as S
from oracle import oracle as O
from tests.gpu_util import grid_of_oracle


def _box_geom(seed=3, dx=S.FAMILY_DX):
    box = S.family([seed])[0]
    g = O.prepare(G.airbox_triangles(box.spec), G.prepare_params(box.spec, dx), max_cells=0)
    return box, g


def test_family_is_the_roofline_geometry_plus_k_and_aniso():
    for seed in (0, 7, 42):
        a, _ = G.seeded_variant(seed, "roofline")
        b = S.family([seed])[0]
        assert G.spec_dict(a) == G.spec_dict(b.spec)
        assert 1.5 <= b.K <= 6.0 and 5.0 <= b.aniso <= 20.0


@pytest.mark.parametrize("seed", [0, 11, 133])
def test_family_fits_the_volume(seed):
    _, g = _box_geom(seed)
    assert g.nx <= S.GRID[0] and g.ny <= S.GRID[1] and g.nz <= S.GRID[2]


def test_features_and_targets_follow_the_cell_flags():
    box, g = _box_geom()
    grid = grid_of_oracle(g)
    x = S.features(grid, box.K, box.aniso)
    assert x.shape == S.GRID + (16,)
    # cell flags of setupSim (engine.js:250-255): 1 inlet, 2 outlet, 3 porous
    sim = O.setup_sim(g, K=box.K, aniso=box.aniso)
    flag = np.asarray(sim.flag)
    fl = g.fluid.reshape(g.nx, g.ny, g.nz).astype(bool)
    cid_cells = np.nonzero(fl.ravel())[0]
    vol = lambda ch: x[:g.nx, :g.ny, :g.nz, ch].ravel()[cid_cells]
    np.testing.assert_array_equal(vol(2) > 0, flag == 1)
    np.testing.assert_array_equal(vol(5) > 0, flag == 2)
    np.testing.assert_array_equal(vol(1) > 0, flag == 3)
    np.testing.assert_allclose(vol(3)[flag == 3], box.K / 6, rtol=1e-6)
    # solid outside the box's lattice and on its solid cells; distance sign
    assert (x[g.nx:, ..., 0] == 1).all() and (x[:, g.ny:, :, 0] == 1).all() and (x[:, :, g.nz:, 0] == 1).all()
    assert ((x[..., 4] > 0) == (x[..., 0] == 0)).all()
    # targets: the averaged fields scattered to the fluid cells
    NF = cid_cells.size
    rng = np.random.default_rng(0)
    acc = rng.standard_normal(4 * NF).astype(np.float32)
    acc[3 * NF:] = 1 + 0.01 * acc[3 * NF:]
    y, m = S.targets(grid, acc)
    assert m.sum() == NF and y.shape == S.GRID + (4,)
    got = y[:g.nx, :g.ny, :g.nz].reshape(-1, 4)[cid_cells]
    np.testing.assert_allclose(got[:, 0], acc[:NF] / S.U0, rtol=1e-6)
    np.testing.assert_allclose(got[:, 3], (acc[3 * NF:] - 1) / S.RHO_SCALE, rtol=1e-4, atol=1e-4)
    # and back: the prediction of the targets gives the fields again
    ru = S.rho_u_of_prediction(y, grid)
    np.testing.assert_allclose(ru[:NF], acc[3 * NF:], rtol=1e-6)
    np.testing.assert_allclose(ru[NF:2 * NF], acc[:NF], rtol=1e-6, atol=1e-7)


def _geom_view(g):
    """What resample reads of a geometry: grid(), origin, h."""
    grid = grid_of_oracle(g)
    return types.SimpleNamespace(grid=lambda: grid, origin=list(g.origin), h=g.h)


def test_resample_same_lattice_is_identity_and_linear_fields_survive():
    box, g = _box_geom(5)
    gv = _geom_view(g)
    grid = gv.grid()
    fl = g.fluid.reshape(g.nx, g.ny, g.nz).astype(bool)
    rng = np.random.default_rng(1)
    pred = rng.standard_normal(S.GRID + (4,)).astype(np.float32)
    ru = S.resample(pred, gv, gv)
    np.testing.assert_allclose(ru, S.rho_u_of_prediction(pred, grid), rtol=1e-5, atol=1e-6)
    # a field linear in position is reproduced on a finer lattice away from walls
    _, gf = _box_geom(5, dx=3.0)
    gfv = _geom_view(gf)
    I, J, K = np.meshgrid(np.arange(S.GRID[0]), np.arange(S.GRID[1]), np.arange(S.GRID[2]), indexing="ij")
    pos = [g.origin[a] + (idx + 0.5) * g.h for a, idx in enumerate((I, J, K))]
    lin = np.stack([0.01 * pos[0], 0.02 * pos[1], -0.01 * pos[2], 0.005 * pos[0]], -1).astype(np.float32)
    ru = S.resample(lin, gv, gfv)
    ffl = gf.fluid.reshape(gf.nx, gf.ny, gf.nz).astype(bool)
    idx = np.nonzero(ffl)
    fpos = [gf.origin[a] + (idx[a] + 0.5) * gf.h for a in range(3)]
    NFf = idx[0].size
    ux = ru[NFf:2 * NFf]
    # cells whose whole coarse interpolation stencil is fluid
    ci = [np.clip((fpos[a] - g.origin[a]) / g.h - 0.5, 0, n - 1) for a, n in enumerate((g.nx, g.ny, g.nz))]
    i0 = [np.minimum(np.floor(c).astype(int), n - 2) for c, n in zip(ci, (g.nx, g.ny, g.nz))]
    full = np.ones(NFf, bool)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                full &= fl[i0[0] + dx, i0[1] + dy, i0[2] + dz]
    inside = full & np.all([(fpos[a] - g.origin[a]) / g.h - 0.5 >= 0 for a in range(3)], 0)
    assert inside.sum() > 0.5 * NFf
    np.testing.assert_allclose(ux[inside], S.U0 * 0.01 * fpos[0][inside], rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("path", [S.WEIGHTS, S.WEIGHTS_DX2])
def test_trained_weights_file_matches_the_network(path):
    from airboxcalc_amd import unet as U
    z = S.read_weights(path)
    assert z["w"].shape == (U.n_params(),) and np.isfinite(z["w"]).all()
    assert z["norm"]["mean"].shape == (4,) and (z["norm"]["std"] > 0).all()
    assert z["meta"]["grid"] == list(S.GRID) and z["meta"]["dx"] == S.FAMILY_DX
    assert z["grid"] == S.GRID and z["family_dx"] == S.FAMILY_DX
    assert (z["u0_scale"], z["rho_scale"]) == (S.U0, S.RHO_SCALE)
    # the unused head channels stay zero
    head = U.layers()[-1]
    W = z["w"][head.w_off:head.w_off + head.nw].reshape(head.co, head.taps, head.ci)
    assert not W[U.OUT_CH:].any()


def test_weight_file_round_trip(tmp_path):
    from airboxcalc_amd import unet as U
    rng = np.random.default_rng(5)
    w = rng.standard_normal(U.n_params()).astype(np.float32)
    norm = dict(mean=rng.standard_normal(4).astype(np.float32), std=rng.uniform(0.1, 1, 4).astype(np.float32))
    p = str(tmp_path / "w.abxw")
    S.write_weights(p, w, norm, {"boxes": 3, "note": "x" * 17})
    r = S.read_weights(p)
    assert np.array_equal(r["w"], w) and r["meta"] == {"boxes": 3, "note": "x" * 17}
    assert np.array_equal(r["norm"]["mean"], norm["mean"]) and np.array_equal(r["norm"]["std"], norm["std"])
    with pytest.raises(ValueError):
        S.write_weights(p, w[:-1], norm, {})


def test_restricted_targets_of_a_same_lattice_solve_are_the_targets():
    """Restriction from a lattice onto itself returns targets() of the field."""
    box, g = _box_geom(7)
    gv = _geom_view(g)
    grid = gv.grid()
    NF = int(g.fluid.sum())
    rng = np.random.default_rng(3)
    acc = rng.standard_normal(4 * NF).astype(np.float32) * 0.01
    acc[3 * NF:] += 1.0
    y0, m0 = S.targets(grid, acc)
    y1, m1 = S.restrict_targets(acc, gv, gv)
    assert np.array_equal(m0, m1)
    np.testing.assert_allclose(y1, y0, rtol=1e-5, atol=1e-5)


def test_cli_npz_layout_matches_the_reference_cli():
    """The keys, shapes and types airbox_cfd.py's --from-npz reads
    (airbox_cfd.py:160-166, 287-288): masks and fields (nx, ny, nz), axes 1-D,
    hist rows (step, gamma, ratio)."""
    import io
    from airboxcalc_amd import engine as E
    box, g = _box_geom(9)
    grid = grid_of_oracle(g)
    NF = int(g.fluid.sum())
    view = types.SimpleNamespace(grid=lambda: grid, array=lambda k: getattr(g, k))
    acc = np.random.default_rng(0).standard_normal(4 * NF).astype(np.float32)
    buf = io.BytesIO()
    np.savez_compressed(buf, **E.cli_arrays(view, acc), u=0.06, pleat=40.0, hist=np.array([[100, 0.9, 0.5]]))
    buf.seek(0)
    d = np.load(buf)
    for k in ("fluid", "porous", "solid", "ux", "uy", "uz", "rho"):
        assert d[k].shape == (g.nx, g.ny, g.nz), k
    for k, n in (("xc", g.nx), ("dc", g.ny), ("zc", g.nz)):
        assert d[k].shape == (n,)
    assert d["fluid"].dtype == bool and d["fluid"].sum() == NF
    np.testing.assert_array_equal(d["ux"][d["fluid"]], acc[:NF].astype(np.float64))
    assert (d["rho"][~d["fluid"]] == 0).all()
    assert d["fluid"][:, :, 0].shape == (g.nx, g.ny)       # the opening argument of report()
    assert d["hist"].shape[1] == 3
    rep = E.cli_report(dict(gamma=0.9, gammaCore=0.95, lowCore=0.01, highCore=0.02, reverse=0.0, maxToMean=1.5,
                            zeta=1.234), 3.0, 4.0, 2000, [(100, 0.9, 0.5)])
    assert rep["dp_rohluft_mbar_bei_30ms"] == round(1.234 * 0.5 * 1.2 * 900 / 100, 2)
    assert set(rep) == {"gamma", "gamma_ohne_randzellen", "anteil_unter_50pct_kern", "anteil_ueber_150pct_kern",
                        "anteil_rueckstroemung", "max_zu_mittel", "zeta_rohluftseite", "dp_rohluft_mbar_bei_30ms",
                        "K", "dx", "steps", "verlauf"}
