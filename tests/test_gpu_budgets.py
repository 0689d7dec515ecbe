"""Mass and energy budgets and wave-gauge elevation histories of the GPU
run against the FP64 oracle over the run (a piston flume, 100 steps,
compared every 10), and the force balance of fluid at hydrostatic rest."""
import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G
from sph2d_amd import hip as H
from sph2d_amd import params as P

pytestmark = pytest.mark.gpu

XS = [0.25, 0.6, 1.0, 1.4]
SEARCH_N = 2.0


def budgets_np(prm, st):
    """The budget of rrsph_diagnostics from a state in numpy: fluid mass,
    kinetic energy, potential energy m g y, Tait elastic energy
    m B / rho0 ((x^6 - 1) / 6 + 1 / x - 1) with x = rho / rho0 (the integral
    of p / rho^2 from rho0), max speed, density range."""
    fl = st["itype"] > 0
    m, g = prm.mass, prm.g
    v2 = np.sum(st["v"][fl] ** 2, axis=1)
    x = st["rho"][fl] / prm.rho0
    B = prm.eos_c_sqr * prm.rho0 / 7.0
    return {
        "mass_fluid": m * fl.sum(),
        "kinetic_energy": float(np.sum(0.5 * m * v2)),
        "potential_energy": float(np.sum(m * g * st["r"][fl, 1])),
        "internal_energy": float(np.sum(m * B / prm.rho0 * ((x ** 6 - 1) / 6 + 1 / x - 1))),
        "max_speed": float(np.sqrt(v2.max())),
        "rho_min": float(st["rho"][fl].min()),
        "rho_max": float(st["rho"][fl].max()),
    }


def gauges_np(prm, st, xs, search_n, ptype=P.TYPE_WATER):
    """HeightTesting::maxInLayer (tools/WaterProfile/include/HeightTesting.h:25-54):
    max y of particles of type ptype with |x - x_g| < search_n h, from 0."""
    sel = st["itype"] == ptype
    x, y = st["r"][sel, 0], st["r"][sel, 1]
    out = []
    for xg in xs:
        w = np.abs(x - xg) < search_n * prm.hsml
        out.append(max(0.0, float(y[w].max())) if w.any() else 0.0)
    return np.array(out)


@pytest.mark.parametrize("precision", [H.FP64, H.FP32], ids=["fp64", "fp32"])
def test_budgets_and_gauge_histories_match_the_oracle(precision):
    lay = G.piston_tank(delta=0.02, length=1.6, depth=0.3, wall_height=0.6)
    prm = lay.params
    print("budgets n", prm.ntotal, "precision", precision, flush=True)
    gpu = H.HipStep(prm, precision=precision)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=prm.start_simulation_time, dt=prm.dt)
    # FP64: summation order only; FP32: the FP32 step's drift from FP64
    tol = {"budget": 1e-10, "gauge": 1e-12} if precision == H.FP64 else {"budget": 2e-4, "gauge": 1e-6}
    e_scale = None
    for block in range(10):
        gpu.step(10)
        assert ora.step(10) == 0
        st = ora.state()
        want = budgets_np(prm, st)
        got = gpu.diagnostics()
        if e_scale is None:
            e_scale = want["potential_energy"] + want["internal_energy"]
        assert got["mass_fluid"] == want["mass_fluid"]
        assert got["n_fluid_alive"] == int((st["itype"] > 0).sum())
        for k in ("kinetic_energy", "potential_energy", "internal_energy"):
            assert abs(got[k] - want[k]) <= tol["budget"] * e_scale, (block, k, got[k], want[k])
        assert abs(got["max_speed"] - want["max_speed"]) <= tol["budget"] * 10 * max(want["max_speed"], 1e-3)
        for k in ("rho_min", "rho_max"):
            assert abs(got[k] - want[k]) <= tol["budget"] * prm.rho0, (block, k, got[k], want[k])
        gg = gpu.gauges(XS, search_n=SEARCH_N)
        gw = gauges_np(prm, st, XS, SEARCH_N)
        assert np.all(gw > 0)
        assert np.max(np.abs(gg - gw)) <= tol["gauge"], (block, gg, gw)
    gpu.close()


def test_hydrostatic_rest_force_balance():
    """Fluid at hydrostatic density, at rest, away from walls and the free
    surface: the horizontal acceleration vanishes to rounding (the lattice
    is left-right symmetric about each particle) and the vertical one is a
    small residual of -g + pressure gradient, as in the oracle."""
    lay = G.piston_tank(delta=0.02, length=1.6, depth=0.4, wall_height=0.7)
    prm = lay.params
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    gpu.set_export_fields(True)
    gpu.step(1)
    a = gpu.fields()["a"]
    gpu.close()
    r, it = lay.r, lay.itype
    h2 = 2 * prm.hsml
    x0, x1 = r[it > 0, 0].min(), r[it > 0, 0].max()
    y0, y1 = r[it > 0, 1].min(), r[it > 0, 1].max()
    interior = (it > 0) & (r[:, 0] > x0 + 2 * h2) & (r[:, 0] < x1 - 2 * h2) & (r[:, 1] > y0 + 2 * h2) & \
               (r[:, 1] < y1 - 2 * h2)
    assert interior.sum() > 100
    assert np.max(np.abs(a[interior, 0])) < 1e-9 * prm.g
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(1) == 0
    ref = ora.fields()["a"]
    assert np.max(np.abs(a[interior] - ref[interior])) < 1e-9 * prm.g
    assert np.max(np.abs(a[interior, 1])) < 0.05 * prm.g
