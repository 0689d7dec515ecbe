"""The 3D step on the GPU against the FP64 oracle's 3D path: cell keys and
neighbour sets bit for bit, fields and trajectories within the stated
tolerances, and the incremental re-sort against the full sort."""
import copy

import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G
from sph2d_amd import hip as H

pytestmark = pytest.mark.gpu


def _scale_err(a, b):
    s = float(np.max(np.abs(b))) or 1.0
    return float(np.max(np.abs(a - b))) / s


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_3d_keys_and_neighbour_sets(precision):
    """3D Morton keys (GridUtils.h:145-151) and neighbour sets bit-exact."""
    lay = G.dam_break_3d()
    prm = lay.params
    print("3D keys/neighbours n", prm.ntotal, "precision", precision, flush=True)
    gpu = H.HipStep(prm, precision=precision)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    keys = gpu.cell_keys()
    L = oracle.lib()
    op = oracle.from_params(prm)
    want = np.array([L.oracle_cell_idx(op, oracle.ptr(np.ascontiguousarray(lay.r[i]))) for i in range(prm.ntotal)],
                    dtype=np.uint64)
    assert np.array_equal(keys.astype(np.uint64), want)
    nb_gpu = gpu.neighbours()
    st, nb_ref = oracle.grid_find(prm, lay.r, lay.itype)
    assert st == 0
    sg, sr = oracle.neighbour_sets(nb_gpu, prm.ntotal), oracle.neighbour_sets(nb_ref, prm.ntotal)
    assert all(np.array_equal(a, b) for a, b in zip(sg, sr))
    gpu.close()


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_3d_neighbour_sets_across_cell_corners_and_edges(precision):
    """Centres at the cut-off (to within ulps .. 1e-3) from a corner or an
    edge of their cell and satellites in the cells across it, at the cut-off
    to a few ulps: the GPU walk finds exactly the oracle's neighbours there."""
    from helpers import corner_cutoff_case, corner_pairs_at_cutoff

    prm, r, itype = corner_cutoff_case(dim=3)
    print("3D corner case n", prm.ntotal, "precision", precision, flush=True)
    gpu = H.HipStep(prm, precision=precision)
    gpu.upload(r, np.zeros_like(r), np.full(prm.ntotal, prm.rho0), itype)
    nb_gpu = gpu.neighbours()
    gpu.close()
    st, nb_ref = oracle.grid_find(prm, r, itype)
    assert st == 0
    sg, sr = oracle.neighbour_sets(nb_gpu, prm.ntotal), oracle.neighbour_sets(nb_ref, prm.ntotal)
    assert corner_pairs_at_cutoff(prm, r, sr) > 500
    assert all(np.array_equal(a, b) for a, b in zip(sg, sr))


@pytest.mark.parametrize("density_treatment", [0, 1])
def test_3d_step_fp64_matches_oracle(density_treatment):
    """3D dam break (summation or continuity density), FP64: fields after one
    step within 1e-12 of their scale, state after five steps within 1e-12."""
    lay = G.dam_break_3d(density_treatment=density_treatment)
    prm = lay.params
    print("3D step n", prm.ntotal, "density", density_treatment, flush=True)
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.set_export_fields(True)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    gpu.step(1)
    assert ora.step(1) == 0
    f, fr = gpu.fields(), ora.fields()
    assert _scale_err(f["a"], fr["a"]) < 1e-12
    if density_treatment:
        assert _scale_err(f["drho"], fr["drho"]) < 1e-12
    assert _scale_err(f["av"], fr["av"]) < 1e-12
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-14
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-13
    gpu.step(4)
    assert ora.step(4) == 0
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-12
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-12
    assert np.array_equal(g["itype"], r["itype"])
    t, dt, step = gpu.time()
    assert step == 5 and abs(t - r["time"]) < 1e-15
    d = gpu.diagnostics()
    assert d["n_fluid_alive"] == prm.nfluid
    assert abs(d["mass_fluid"] - prm.nfluid * prm.mass) < 1e-12 * prm.nfluid * prm.mass
    gpu.close()


def test_3d_fp32_error_vs_fp64_oracle():
    """FP32 pair arithmetic in 3D: density within 1e-6 relative after one
    step and 1e-4 after 20 steps."""
    lay = G.dam_break_3d(density_treatment=1)
    prm = lay.params
    gpu = H.HipStep(prm, precision=H.FP32)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    gpu.step(1)
    assert ora.step(1) == 0
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-6
    gpu.step(19)
    assert ora.step(19) == 0
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-4
    gpu.close()


def test_3d_incremental_resort_equals_full_sort(monkeypatch):
    lay = G.dam_break_3d()
    prm = lay.params
    monkeypatch.setenv("RRSPH_HIP_LIST", "0")  # the sorted step of every step
    out = {}
    for mode in ("full", "incremental"):
        if mode == "full":
            monkeypatch.setenv("RRSPH_HIP_SORT", "full")
        else:
            monkeypatch.delenv("RRSPH_HIP_SORT", raising=False)
        gpu = H.HipStep(prm, precision=H.FP32)
        gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
        gpu.step(60)
        out[mode] = (gpu.download(), gpu.sort_stats())
        gpu.close()
    assert out["incremental"][1]["incremental"] == 59
    for k in ("r", "v", "rho", "p", "itype"):
        assert np.array_equal(out["full"][0][k], out["incremental"][0][k]), k


@pytest.mark.parametrize("lists", ["1", "0"])
def test_3d_neighbour_overflow_is_the_reference_error(monkeypatch, lists):
    """3D, list and sorted steps: more than max_neighbours - 1 neighbours
    with the consistency check stops with the reference's error
    (GridFind.cpp:186-188) before whole_step, as the oracle does; the state
    is the uploaded one."""
    monkeypatch.setenv("RRSPH_HIP_LIST", lists)
    lay = G.dam_break_3d(delta=0.04)
    prm = copy.copy(lay.params)
    prm.max_neighbours = 16  # a 3D particle has ~75
    for precision in (H.FP64, H.FP32):
        gpu = H.HipStep(prm, precision=precision)
        gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
        with pytest.raises(H.RrsphError) as e:
            gpu.step(2)
        assert e.value.status == 3 and "making neighbours grid: error occured" in str(e.value)
        t, _dt, step = gpu.time()
        st = gpu.download()
        gpu.close()
        assert step == 0 and t == prm.start_simulation_time
        assert np.array_equal(st["r"], lay.r) and np.array_equal(st["itype"], lay.itype)
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(1) == 1


@pytest.mark.parametrize("lists", ["1", "0"])
def test_3d_fix_turns_escaping_fluid_into_non_existent_particles(monkeypatch, lists):
    """3D FIX (consistency_treatment = 2), list and sorted steps: four
    isolated fluid particles above the column fly out of the domain and
    become itype 0; the rest of the state follows the oracle."""
    from sph2d_amd import params as P

    monkeypatch.setenv("RRSPH_HIP_LIST", lists)
    lay = G.dam_break_3d()
    prm = copy.copy(lay.params)
    prm.consistency_treatment = P.CONSISTENCY_FIX
    nf = prm.nfluid
    extra_r = np.array([[0.2, 0.9, 0.2], [0.5, 0.95, 0.1], [0.7, 0.9, 0.3], [0.3, 0.92, 0.35]])
    extra_v = np.array([[0.0, 900.0, 0.0], [0.0, 1200.0, 0.0], [3000.0, 0.0, 0.0], [0.0, 500.0, -3000.0]])
    r = np.concatenate([lay.r[:nf], extra_r, lay.r[nf:]])
    v = np.concatenate([lay.v[:nf], extra_v, lay.v[nf:]])
    rho = np.concatenate([lay.rho[:nf], np.full(4, prm.rho0), lay.rho[nf:]])
    itype = np.concatenate([lay.itype[:nf], np.full(4, P.TYPE_WATER, dtype=np.int32), lay.itype[nf:]])
    prm.nfluid, prm.ntotal = nf + 4, prm.ntotal + 4
    prm.nwm_particles_start = prm.nwm_particles_end = prm.ntotal
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(r, v, rho, itype)
    gpu.step(6)
    got = gpu.download()
    gpu.close()
    ora = oracle.Oracle(prm)
    ora.set_state(r, v, rho, itype, time=0.0, dt=prm.dt)
    assert ora.step(6) == 0
    ref = ora.state()
    assert np.all(got["itype"][nf:nf + 4] == 0), got["itype"][nf:nf + 4]
    assert np.array_equal(got["itype"], ref["itype"])
    assert np.array_equal(got["r"][nf:nf + 4], ref["r"][nf:nf + 4])  # dead particles stay where they died
    alive = ref["itype"] != 0
    assert np.max(np.abs(got["r"][alive] - ref["r"][alive])) < 1e-12
    assert np.max(np.abs(got["rho"][alive] - ref["rho"][alive]) / ref["rho"][alive]) < 1e-11
