"""Results of a HIP solve written in the layout of the reference's command-line
tool (airbox_cfd.py:287-288 for the NPZ, 360-366 for the JSON report), so its
--from-npz report (airbox_cfd.py:160-166) reads them: keys, dtypes and shapes,
and fields equal to the solve's averaged acc."""
import json

import numpy as np
import pytest

from airboxcalc_amd import engine
from tests import golden_util as GU
from tests.gpu_util import dev, gpu_geometry, say

pytestmark = pytest.mark.gpu


def test_cli_npz_and_report_from_a_hip_solve(tmp_path):
    fx = GU.load("small_yaw")
    say("CLI outputs from a HIP solve")
    g = gpu_geometry(fx)
    r = fx["desc"]["run"]
    res = engine.simulate_hip(dev(), g, r, precision="f32_f64", keep_sim=True)
    ev = res.sim.evaluate(g.eval_view(), divisor=max(res.nacc, 1), field3d=False)
    res.sim.destroy()
    path = str(tmp_path / "airbox_result.npz")
    engine.save_cli_npz(path, g, res.acc, r["u"], g.eval_view().pleat, res.hist)
    d = np.load(path, allow_pickle=False)
    assert set(d.files) == {"fluid", "porous", "solid", "xc", "dc", "zc", "ux", "uy", "uz", "rho", "u", "pleat",
                            "hist"}
    nx, ny, nz = g.nx, g.ny, g.nz
    for k in ("fluid", "porous", "solid"):
        assert d[k].shape == (nx, ny, nz) and d[k].dtype == bool, k
    for k in ("ux", "uy", "uz", "rho"):
        assert d[k].shape == (nx, ny, nz) and d[k].dtype == np.float64, k
    for k, n in (("xc", nx), ("dc", ny), ("zc", nz)):
        assert d[k].shape == (n,), k
    np.testing.assert_array_equal(d["xc"], g.array("xc").astype(np.float64))
    fluid = d["fluid"]
    assert np.array_equal(fluid.ravel(), g.array("fluid").astype(bool))
    assert np.array_equal(d["solid"].ravel(), g.array("solid").astype(bool))
    NF = res.NF
    assert fluid.sum() == NF
    # fields in the reference's flat order: fluid cells ascending = cid order
    for c, k in enumerate(("ux", "uy", "uz", "rho")):
        np.testing.assert_array_equal(d[k][fluid], res.acc[c * NF:(c + 1) * NF].astype(np.float64))
        assert (d[k][~fluid] == 0).all()
    assert float(d["u"]) == r["u"]
    np.testing.assert_array_equal(d["hist"], np.array(res.hist, dtype=np.float64))
    # the f32_f64 kernel's averaged field is engine.js's: its metrics, hence the report
    m = ev["metrics"]
    for k, e in fx["sim"]["sim"]["metrics"].items():
        assert m[k] == pytest.approx(GU.x64(e), rel=1e-12, abs=1e-14), k
    rep = engine.cli_report(m, r["K"], r["dx"], r["steps"], res.hist)
    text = json.dumps(rep, indent=1)
    back = json.loads(text)
    assert back["gamma"] == round(m["gamma"], 4) and back["zeta_rohluftseite"] == round(m["zeta"], 3)
    assert back["gamma_ohne_randzellen"] == round(m["gammaCore"], 4)
    assert back["anteil_unter_50pct_kern"] == round(m["lowCore"], 4)
    assert back["dp_rohluft_mbar_bei_30ms"] == round(m["zeta"] * 0.5 * 1.2 * 30 ** 2 / 100, 2)
    assert back["verlauf"] == [list(map(float, row)) for row in res.hist]
    assert (back["K"], back["dx"], back["steps"]) == (r["K"], r["dx"], r["steps"])
