"""FuncAtPoint on the GPU (rrsph_hip_probe) against the oracle's
(tools/FuncAtPoint/lib/src/fatp.cpp:14-58, :121-144) on the same state: the
2D dam break after some steps (FP64 and FP32 contexts), dead particles among
the probed cells (the reference's grid holds them, rho = m W(0)), the 3D dam
break, and the input errors of check_input (fatp.cpp:92-119)."""
import copy

import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G
from sph2d_amd import hip as H

pytestmark = pytest.mark.gpu


def _points(prm, lay, n, seed=5):
    """Points within two cells of random fluid and wall particles, and a
    tenth anywhere in the box (mostly empty space), all inside the box."""
    rng = np.random.default_rng(seed)
    dim = lay.r.shape[1]
    cell = prm.cell_scale_k * prm.hsml
    box_lo = np.array([prm.x_mingeom, prm.y_mingeom, prm.z_mingeom][:dim])
    box_hi = np.array([prm.x_maxgeom, prm.y_maxgeom, prm.z_maxgeom][:dim])
    near = lay.r[rng.integers(0, len(lay.r), n - n // 10)] + rng.uniform(-2, 2, (n - n // 10, dim)) * cell
    anywhere = box_lo + rng.uniform(0, 1, (n // 10, dim)) * (box_hi - box_lo)
    return np.clip(np.concatenate([near, anywhere]), box_lo, box_hi)


def _compare(prm, gpu, pts, fields, tol):
    st = gpu.download()
    for f in fields:
        got = gpu.probe(pts, f)
        ost, want = oracle.probe(prm, st, pts, f)
        assert ost == 0
        scale = float(np.max(np.abs(want))) or 1.0
        print(f, "max |got - want| / scale", float(np.max(np.abs(got - want))) / scale, flush=True)
        assert np.count_nonzero(want) > len(pts) // 5  # not vacuous (p is 0 around walls at rest)
        assert np.max(np.abs(got - want)) <= tol * scale, (f, got, want)


@pytest.mark.parametrize("precision,tol", [(H.FP64, 1e-12), (H.FP32, 1e-5)], ids=["fp64", "fp32"])
def test_probe_matches_the_oracle_2d(precision, tol):
    lay = G.dam_break()
    prm = lay.params
    gpu = H.HipStep(prm, precision=precision)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    gpu.step(200)  # the column has started to fall: every field is nonzero
    _compare(prm, gpu, _points(prm, lay, 300), ("vx", "vy", "p", "rho"), tol)
    gpu.close()


def test_probe_counts_dead_particles_as_the_reference_grid_does():
    lay = G.dam_break()
    prm = lay.params
    itype = lay.itype.copy()
    dead = np.flatnonzero(itype > 0)[::7]  # every 7th fluid particle, left where it is
    itype[dead] = 0
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(lay.r, lay.v + 0.1, lay.rho, itype)
    pts = lay.r[dead[:200]] + 0.3 * prm.delta
    _compare(prm, gpu, pts, ("vx", "rho"), 1e-12)
    gpu.close()


def test_probe_matches_the_oracle_3d():
    lay = G.dam_break_3d()
    prm = lay.params
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(lay.r, lay.v + np.array([0.05, -0.02, 0.03]), lay.rho, lay.itype)  # nonzero velocities
    gpu.step(5)
    _compare(prm, gpu, _points(prm, lay, 200), ("vx", "vz", "p", "rho"), 1e-12)
    gpu.close()


def test_probe_input_errors():
    lay = G.dam_break()
    prm = lay.params
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    with pytest.raises(H.RrsphError, match="point is outside bounds"):
        gpu.probe([[prm.x_maxgeom + 1e-9, 0.5]], "p")
    with pytest.raises(H.RrsphError, match="wrong target value"):
        gpu.probe([[1.0, 0.5]], "vz")
    # the box is [min, min + (max - min)] with that rounding, inclusive (RRSPH_FIO.h:15-38)
    top = [prm.x_mingeom + (prm.x_maxgeom - prm.x_mingeom), prm.y_mingeom + (prm.y_maxgeom - prm.y_mingeom)]
    corner = gpu.probe([[prm.x_mingeom, prm.y_mingeom], top], "p")
    assert np.all(np.isfinite(corner))
    with pytest.raises(H.RrsphError, match="point is outside bounds"):
        gpu.probe([[np.nextafter(top[0], np.inf), top[1]]], "p")
    gpu.close()
