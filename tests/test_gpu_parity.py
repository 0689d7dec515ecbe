"""GPU parity of the RRSPH_HIP step against the FP64 oracle and the
reference's golden vectors, through the C-ABI (ctypes over librrsph_hip.so).

Bars: cell keys and neighbour sets bit-exact; FP64 fields within 1e-12
relative to the field's scale (summation order differs inside a cell);
FP32 fields within the tolerances written in each test.
"""
import copy

import numpy as np
import pytest

import oracle
from helpers import load_dambreak_case
from sph2d_amd import generators as G
from sph2d_amd import hip as H

pytestmark = pytest.mark.gpu


def _sets(nb, n):
    return oracle.neighbour_sets(nb, n)


def _scale_err(a, b):
    """max |a - b| / max |b| (error relative to the field's scale)."""
    s = float(np.max(np.abs(b))) or 1.0
    return float(np.max(np.abs(a - b))) / s


@pytest.fixture(scope="module")
def golden():
    prm, st, tg = load_dambreak_case()
    return prm, st, tg


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_keys_and_neighbour_sets_golden_case(golden, precision):
    """Morton keys (GridUtils.h:139-144) and neighbour sets (GridFind.cpp:112-189)
    of the reference's dam-break test case, bit-exact."""
    prm, st, _tg = golden
    gpu = H.HipStep(prm, precision=precision)
    gpu.upload(st["r"], st["v"], st["rho"], st["itype"])
    keys = gpu.cell_keys()
    L = oracle.lib()
    op = oracle.from_params(prm)
    want = np.array([L.oracle_cell_idx(op, oracle.ptr(np.ascontiguousarray(st["r"][i]))) for i in range(prm.ntotal)],
                    dtype=np.uint64)
    assert np.array_equal(keys.astype(np.uint64), want)
    nb_gpu = gpu.neighbours()
    status, nb_ref = oracle.grid_find(prm, st["r"], st["itype"])
    assert status == 0
    sg, sr = _sets(nb_gpu, prm.ntotal), _sets(nb_ref, prm.ntotal)
    assert all(np.array_equal(a, b) for a, b in zip(sg, sr))
    gpu.close()


def test_first_step_fields_match_golden_targets(golden):
    """Step 1 of the golden case: drho = density_con target; a = gravity +
    pressure (approx. 2) + artificial viscosity targets summed; mu = AV target.
    (dt = 0 on the first predict, so v^ = v and rho^ = rho, exactly the inputs
    of the reference's module tests.)"""
    prm, st, tg = golden
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.set_export_fields(True)
    gpu.upload(st["r"], st["v"], st["rho"], st["itype"])
    gpu.step(1)
    f = gpu.fields()
    drho_t = tg["OMP_con_density_target__drho"]
    assert np.all(np.abs(f["drho"] - drho_t) <= 1e-8 + 1e-12 * np.abs(drho_t) + 1e-9 * np.max(np.abs(drho_t)))
    ax_t = tg["OMP_exdvdt_dynamic__exdvdt_x"] + tg["OMP_intf2_no_artp_target__ax"] + tg["OMP_art_visc_target__ax"]
    ay_t = tg["OMP_exdvdt_dynamic__exdvdt_y"] + tg["OMP_intf2_no_artp_target__ay"] + tg["OMP_art_visc_target__ay"]
    # three printed targets (6 or 8 decimals) summed: 3 units of the 6th decimal
    tol = 3e-6 + 1e-11 * np.abs(ay_t)
    assert np.all(np.abs(f["a"][:, 0] - ax_t) <= tol + 1e-11 * np.abs(ax_t))
    assert np.all(np.abs(f["a"][:, 1] - ay_t) <= tol)
    assert np.all(np.abs(f["arvmu"] - tg["OMP_art_visc_target__mu"]) <= 1e-8 + 1e-12 * tg["OMP_art_visc_target__mu"])
    gpu.close()


def _run_pair(prm, st, nsteps, precision, export=True):
    gpu = H.HipStep(prm, precision=precision)
    gpu.set_export_fields(export)
    gpu.upload(st["r"], st["v"], st["rho"], st["itype"])
    ora = oracle.Oracle(prm)
    ora.set_state(st["r"], st["v"], st["rho"], st["itype"], time=prm.start_simulation_time, dt=prm.dt)
    gpu.step(nsteps)
    assert ora.step(nsteps) == 0
    return gpu, ora


@pytest.mark.parametrize("density_treatment", [1, 0])
def test_dam_break_fp64_matches_oracle(density_treatment):
    """C0 dam break (6,923 particles, fixed dt): one step and ten steps, FP64."""
    lay = G.dam_break(density_treatment=density_treatment)
    prm = lay.params
    gpu, ora = _run_pair(prm, lay.state, 1, H.FP64)
    f, fr = gpu.fields(), ora.fields()
    assert _scale_err(f["a"], fr["a"]) < 1e-12
    if density_treatment:
        assert _scale_err(f["drho"], fr["drho"]) < 1e-12
    assert _scale_err(f["av"], fr["av"]) < 1e-12
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-14
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-13
    assert _scale_err(g["p"], r["p"]) < 1e-11
    assert np.array_equal(g["itype"], r["itype"])
    gpu.step(9)
    assert ora.step(9) == 0
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-12
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-11
    t, dt, step = gpu.time()
    assert step == 10 and abs(t - r["time"]) < 1e-15 and dt == r["dt"]
    gpu.close()


def test_dam_break_fp32_error_vs_fp64_oracle():
    """FP32 pair arithmetic (FP64 positions): one-step density within 1e-6
    relative and accelerations within 1e-3 of the field scale; 100 steps:
    density within 1e-4 relative (the FP32+compensated accuracy target)."""
    lay = G.dam_break()
    prm = lay.params
    gpu, ora = _run_pair(prm, lay.state, 1, H.FP32)
    f, fr = gpu.fields(), ora.fields()
    assert _scale_err(f["a"], fr["a"]) < 1e-3
    assert _scale_err(f["drho"], fr["drho"]) < 1e-3
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-6
    gpu.step(99)
    assert ora.step(99) == 0
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-4
    gpu.close()


def test_piston_tank_dynamic_dt_matches_oracle():
    """Piston flume at a coarse spacing (boundary motion, Fourier piston,
    dynamic CFL dt from the max reductions): 20 steps FP64."""
    lay = G.piston_tank(delta=0.02, length=4.0, depth=0.4)
    prm = lay.params
    gpu, ora = _run_pair(prm, lay.state, 20, H.FP64)
    g, r = gpu.download(), ora.state()
    t, dt, step = gpu.time()
    assert step == 20
    assert abs(dt - r["dt"]) <= 1e-12 * r["dt"]
    assert abs(t - r["time"]) <= 1e-12 * r["time"]
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-11
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-11
    gpu.close()


@pytest.mark.parametrize("variant", ["approx1_artp", "delta_wendland", "gauss_summation"])
def test_general_pair_path_matches_oracle(variant):
    """Configurations outside the cubic fast path run the general pair kernel
    (kernel-type switch per term, artificial pressure, delta-SPH): FP64 one
    step and five steps against the oracle."""
    lay = G.dam_break()
    prm = copy.copy(lay.params)
    if variant == "approx1_artp":
        prm.intf_sph_approximation = 1
        prm.artificial_pressure = True
        prm.artificial_pressure_coef = 0.2
        prm.artificial_pressure_index = 4.0
    elif variant == "delta_wendland":
        prm.density_treatment = 2
        prm.density_delta_sph_coef = 0.1
        prm.density_skf = prm.intf_skf = prm.artificial_viscosity_skf = prm.average_velocity_skf = 3
    else:
        prm.density_treatment = 0
        prm.density_skf = prm.intf_skf = prm.artificial_viscosity_skf = prm.average_velocity_skf = 2
        prm.artificial_pressure_skf = 2
        from sph2d_amd import params as PP
        PP.fill_in_computing_params(prm)
        PP.post_fill_in_model_params(prm)
    gpu, ora = _run_pair(prm, lay.state, 1, H.FP64)
    f, fr = gpu.fields(), ora.fields()
    assert _scale_err(f["a"], fr["a"]) < 1e-12
    assert _scale_err(f["drho"], fr["drho"]) < 1e-12
    assert _scale_err(f["av"], fr["av"]) < 1e-12
    gpu.step(4)
    assert ora.step(4) == 0
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-12
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-12
    gpu.close()


@pytest.mark.parametrize("impl", ["tile", "general", "walk"])
def test_alternative_pair_kernels_match_oracle(impl, monkeypatch):
    """The LDS-tiled pair kernel, the single-phase walk and the general kernel (selected with
    RRSPH_HIP_PAIR_KERNEL at context creation) give the oracle's result on
    the C0 dam break, FP64 and FP32."""
    monkeypatch.setenv("RRSPH_HIP_PAIR_KERNEL", impl)
    lay = G.dam_break()
    prm = lay.params
    gpu, ora = _run_pair(prm, lay.state, 3, H.FP64)
    f, fr = gpu.fields(), ora.fields()
    assert _scale_err(f["a"], fr["a"]) < 1e-12
    assert _scale_err(f["drho"], fr["drho"]) < 1e-12
    g, r = gpu.download(), ora.state()
    assert np.max(np.abs(g["r"] - r["r"])) < 1e-13
    gpu.close()
    gpu32, ora = _run_pair(prm, lay.state, 3, H.FP32)
    g, r = gpu32.download(), ora.state()
    assert np.max(np.abs(g["rho"] - r["rho"]) / r["rho"]) < 1e-6
    gpu32.close()


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_neighbour_sets_near_cutoff(precision):
    """Thousands of pairs within a few ulps of the cut-off (a fused
    multiply-add in the distance test would flip ~1% of them): the GPU
    neighbour sets equal the oracle's bit for bit."""
    from helpers import near_cutoff_case

    prm, r, itype = near_cutoff_case()
    print("near-cutoff case n", prm.ntotal, "precision", precision, flush=True)
    gpu = H.HipStep(prm, precision=precision)
    v = np.zeros_like(r)
    rho = np.full(prm.ntotal, prm.rho0)
    gpu.upload(r, v, rho, itype)
    nb_gpu = gpu.neighbours()
    st, nb_ref = oracle.grid_find(prm, r, itype)
    assert st == 0
    sg, sr = _sets(nb_gpu, prm.ntotal), _sets(nb_ref, prm.ntotal)
    assert all(np.array_equal(a, b) for a, b in zip(sg, sr))
    gpu.close()


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_neighbour_sets_across_cell_corners(precision):
    """Centres at the cut-off (to within ulps .. 1e-3) from a corner of their
    cell and satellites in the cell diagonally across it, at the cut-off to a
    few ulps: the GPU walk finds exactly the oracle's neighbours there."""
    from helpers import corner_cutoff_case, corner_pairs_at_cutoff

    prm, r, itype = corner_cutoff_case(dim=2)
    print("corner case n", prm.ntotal, "precision", precision, flush=True)
    gpu = H.HipStep(prm, precision=precision)
    gpu.upload(r, np.zeros_like(r), np.full(prm.ntotal, prm.rho0), itype)
    nb_gpu = gpu.neighbours()
    gpu.close()
    st, nb_ref = oracle.grid_find(prm, r, itype)
    assert st == 0
    sg, sr = _sets(nb_gpu, prm.ntotal), _sets(nb_ref, prm.ntotal)
    assert corner_pairs_at_cutoff(prm, r, sr) > 1000  # the case reaches the trimmed cells
    assert all(np.array_equal(a, b) for a, b in zip(sg, sr))


def _run_with_env(monkeypatch, env, prm, state, nsteps, precision):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    gpu = H.HipStep(prm, precision=precision)
    gpu.upload(state["r"], state["v"], state["rho"], state["itype"])
    gpu.step(nsteps)
    out, stats = gpu.download(), gpu.sort_stats()
    gpu.close()
    for k in env:
        monkeypatch.delenv(k)
    return out, stats


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_incremental_resort_equals_full_sort(monkeypatch, precision):
    """The incremental re-sort (merge of the slots that kept their cell with
    the few that changed) yields the permutation of the full stable radix
    sort, so 200 dam-break steps are bit-identical either way; a re-sort
    capacity of 4 forces the fallback (re-run with the full sort) on most
    steps and changes nothing either."""
    lay = G.dam_break()
    prm = lay.params
    print("incremental re-sort, dam break n", prm.ntotal, "precision", precision, flush=True)
    monkeypatch.setenv("RRSPH_HIP_LIST", "0")  # the sorted step of every step
    full, sf = _run_with_env(monkeypatch, {"RRSPH_HIP_SORT": "full"}, prm, lay.state, 200, precision)
    inc, si = _run_with_env(monkeypatch, {}, prm, lay.state, 200, precision)
    assert sf["full"] == 200 and sf["incremental"] == 0
    assert si["full"] == 1 and si["incremental"] == 199 and si["fallbacks"] == 0
    for k in ("r", "v", "rho", "p", "itype"):
        assert np.array_equal(full[k], inc[k]), k
    tiny, st = _run_with_env(monkeypatch, {"RRSPH_HIP_RESORT_CAP": "4"}, prm, lay.state, 200, precision)
    assert st["fallbacks"] > 0
    for k in ("r", "v", "rho", "p", "itype"):
        assert np.array_equal(full[k], tiny[k]), k


def test_incremental_resort_above_16384_wavefronts_equals_full_sort(monkeypatch):
    """Above 16,384 wavefronts (n > 1,048,576) the incremental re-sort counts
    the changed slots of its ballot masks in several passes: 30 sorted steps
    of a 1.1M-particle flume (jittered start, so particles change cell every
    step) are bit-identical to the full radix sort's."""
    lay = G.piston_tank(delta=0.003, jitter=1e-4)
    prm = lay.params
    print("incremental re-sort, flume n", prm.ntotal, flush=True)
    assert prm.ntotal > 16384 * 64
    monkeypatch.setenv("RRSPH_HIP_LIST", "0")
    full, sf = _run_with_env(monkeypatch, {"RRSPH_HIP_SORT": "full"}, prm, lay.state, 30, H.FP32)
    inc, si = _run_with_env(monkeypatch, {}, prm, lay.state, 30, H.FP32)
    print("sort stats", sf, si, flush=True)
    assert sf["incremental"] == 0 and si["incremental"] >= 20 and si["last_changed"] > 0
    for k in ("r", "v", "rho", "p", "itype"):
        assert np.array_equal(full[k], inc[k]), k


def test_unstaged_resort_placement_equals_full_sort(monkeypatch):
    """resort_scatter stages the changed composites that fall inside a
    block's key range in LDS when there are at most 256 of them, and places
    the block's slots by binary search over the global list otherwise.
    RRSPH_HIP_RESORT_STAGE=0 forces the binary search on every block: 200
    dam-break steps stay bit-identical to the full radix sort."""
    lay = G.dam_break()
    prm = lay.params
    monkeypatch.setenv("RRSPH_HIP_LIST", "0")
    full, _ = _run_with_env(monkeypatch, {"RRSPH_HIP_SORT": "full"}, prm, lay.state, 200, H.FP32)
    inc, si = _run_with_env(monkeypatch, {"RRSPH_HIP_RESORT_STAGE": "0"}, prm, lay.state, 200, H.FP32)
    assert si["incremental"] == 199 and si["fallbacks"] == 0
    for k in ("r", "v", "rho", "p", "itype"):
        assert np.array_equal(full[k], inc[k]), k
