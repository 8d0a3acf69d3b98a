"""Time to steady (the convergence criterion) and the FP64 error report."""
import numpy as np
import pytest

from airboxcalc_amd import steady
from tests import golden_util as GU
from tests.gpu_util import dev, gpu_geometry, say

pytestmark = pytest.mark.gpu


def test_cold_then_warm_start_reach_the_same_steady_state():
    fx = GU.load("small")
    g = gpu_geometry(fx)
    p = dict(fx["desc"]["run"])
    say("steady cold")
    cold = steady.solve_to_steady(dev(), g, p, window=200, max_steps=30000, keep_acc=True)
    assert cold.converged, cold.windows
    NF = cold.acc.size // 4
    say("steady warm")
    warm = steady.solve_to_steady(dev(), g, p, window=200, max_steps=30000,
                                  init_rho_u=steady.rho_u_from_acc(cold.acc, NF))
    assert warm.converged
    assert warm.steps < cold.steps
    assert abs(warm.metrics["gammaCore"] - cold.metrics["gammaCore"]) < 0.005
    assert abs(warm.metrics["zeta"] - cold.metrics["zeta"]) <= 0.02 * abs(cold.metrics["zeta"])


def test_rho_u_round_trip():
    acc = np.arange(40, dtype=np.float32)
    ru = steady.rho_u_from_acc(acc, 10)
    assert np.array_equal(ru[:10], acc[30:]) and np.array_equal(ru[10:], acc[:30])


@pytest.mark.parametrize("precision", ["f32", "f32_f64"])
def test_fp64_error_report(precision):
    fx = GU.load("small_yaw")
    g = gpu_geometry(fx)
    p = dict(fx["desc"]["run"])
    say("fp64 report", precision)
    e = steady.fp64_errors(dev(), g, p, precision=precision)
    assert e["field_u_rel_l2"] < 1e-4
    assert e["metric_gammaCore"]["abs_err"] < 1e-4
    assert e["metric_zeta"]["abs_err"] < 1e-3 * max(1.0, abs(e["metric_zeta"]["fp64"]))
    assert e["field3d_spd_rel_l2"] < 1e-4
