"""The renormalised summation density (the boundary-deficiency correction of
docs/latex/ModelParams.tex:124-133, density_normalization = 1) on the GPU
against the FP64 oracle: rho_j = sum m W / sum (m / rho_i) W on every n-th
step, plain summation in between; 2D and 3D, list and sorted steps."""
import copy

import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G
from sph2d_amd import hip as H
from sph2d_amd import params as P

pytestmark = pytest.mark.gpu


def _compare(lay, prm, nsteps, precision):
    gpu = H.HipStep(prm, precision=precision)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    gpu.step(nsteps)
    g = gpu.download()
    gpu.close()
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(nsteps) == 0
    return g, ora.state()


@pytest.mark.parametrize("lists", ["1", "0"])
@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_renormalised_summation_2d(monkeypatch, lists, precision):
    """C0 dam break, summation density renormalised every 3rd step, 9 steps."""
    monkeypatch.setenv("RRSPH_HIP_LIST", lists)
    lay = G.dam_break(density_treatment=P.DENSITY_SUMMATION)
    prm = copy.copy(lay.params)
    prm.density_normalization = 1
    prm.density_normalization_every = 3
    g, r = _compare(lay, prm, 9, precision)
    plain_prm = copy.copy(lay.params)
    gp, _ = _compare(lay, plain_prm, 9, H.FP64) if precision == H.FP64 else (None, None)
    rho_err = float(np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]))
    if precision == H.FP64:
        assert rho_err < 1e-12 and np.max(np.abs(g["r"] - r["r"])) < 1e-13
        assert np.max(np.abs(gp["rho"] - r["rho"])) > 1.0  # the correction changed the run
    else:
        assert rho_err < 1e-5 and np.max(np.abs(g["r"] - r["r"])) < 1e-6 * prm.delta


@pytest.mark.parametrize("lists", ["1", "0"])
@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_renormalised_summation_3d(monkeypatch, lists, precision):
    """The small 3D dam break (delta 0.04 m), renormalised every 2nd step, 6 steps."""
    monkeypatch.setenv("RRSPH_HIP_LIST", lists)
    lay = G.dam_break_3d(delta=0.04)
    prm = copy.copy(lay.params)
    prm.density_normalization = 1
    prm.density_normalization_every = 2
    g, r = _compare(lay, prm, 6, precision)
    rho_err = float(np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]))
    if precision == H.FP64:
        assert rho_err < 1e-12 and np.max(np.abs(g["r"] - r["r"])) < 1e-13
    else:
        assert rho_err < 1e-5


@pytest.mark.parametrize("lists", ["1", "0"])
def test_renormalised_summation_3d_after_a_wall_disappears(monkeypatch, lists):
    """The small 3D dam break renormalised every 2nd step, the wall behind
    the column (x <= 0, moved to the end of the arrays as the wave maker's
    range) disappearing (NWM_METHOD_WALL_DISAPPEAR) at step 31, when builds
    have become rare (the fluid barely moves yet): the lists built before
    keep the vanished particles, whose records then hold density 0, until
    the next build; the renormalisation must skip them.  40 FP64 steps
    against the oracle."""
    monkeypatch.setenv("RRSPH_HIP_LIST", lists)
    lay = G.dam_break_3d(delta=0.04)
    prm = copy.copy(lay.params)
    back = (lay.itype == P.TYPE_BOUNDARY) & (lay.r[:, 0] <= 1e-9)
    order = np.concatenate([np.nonzero(~back)[0], np.nonzero(back)[0]])
    lay = G.Layout(prm, lay.r[order], lay.v[order], lay.rho[order], lay.itype[order])
    prm.density_normalization = 1
    prm.density_normalization_every = 2
    prm.nwm = P.NWM_METHOD_WALL_DISAPPEAR
    prm.boundary_motion = True
    prm.nwm_particles_start = prm.ntotal - int(back.sum())
    prm.nwm_particles_end = prm.ntotal
    prm.nwm_time_start = 30.5 * prm.dt
    P.post_fill_in_model_params(prm)
    g, r = _compare(lay, prm, 40, H.FP64)
    assert np.all(g["itype"][prm.nwm_particles_start:] == 0)
    assert np.array_equal(g["itype"], r["itype"])
    alive = r["itype"] != 0  # the densities of non-existent particles are not compared (as for FIX)
    assert np.all(np.isfinite(g["rho"][alive]))
    assert np.max(np.abs(g["rho"][alive] - r["rho"][alive]) / r["rho"][alive]) < 1e-10
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-12
