"""GPU parity at the measured size: the 2D piston flume at ~1M particles
(20 m x 0.5 m of water at delta = 0.0031623 m, 1,020,263 particles), the
workload bench.py times, against the FP64 oracle on the same input; and the
incremental re-sort above 2^20 particles, where its changed-slot scan runs
in more than one pass."""
import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G
from sph2d_amd import hip as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_1m():
    lay = G.piston_tank(delta=G.c1_1m_delta())
    print("C1-1M flume n", lay.params.ntotal, flush=True)
    return lay


def _canonical_sets(nb, n):
    """Neighbour lists (reference layout, sentinel n) with each row's entries
    sorted and everything after the sentinel blanked: equal arrays <=> equal
    sets."""
    nb = nb.copy()
    first = np.argmax(nb == n, axis=1)
    first[~(nb == n).any(axis=1)] = nb.shape[1]
    nb[np.arange(nb.shape[1])[None, :] >= first[:, None]] = n
    return np.sort(nb, axis=1)


def test_c1_1m_keys_and_neighbour_sets(c1_1m):
    prm = c1_1m.params
    gpu = H.HipStep(prm, precision=H.FP32)
    gpu.upload(c1_1m.r, c1_1m.v, c1_1m.rho, c1_1m.itype)
    nb_gpu = gpu.neighbours()
    gpu.close()
    st, nb_ref = oracle.grid_find(prm, c1_1m.r, c1_1m.itype)
    assert st == 0
    assert np.array_equal(_canonical_sets(nb_gpu, prm.ntotal), _canonical_sets(nb_ref, prm.ntotal))


def test_c1_1m_ten_steps_match_the_oracle(c1_1m):
    """Ten steps of the ~1M-particle flume (continuity, dynamic CFL dt,
    piston): FP64 within 1e-11 of the oracle; the FP32 step's density
    within 1e-4 relative of the oracle's FP64 density (the accuracy target)."""
    prm = c1_1m.params
    ora = oracle.Oracle(prm)
    ora.set_state(c1_1m.r, c1_1m.v, c1_1m.rho, c1_1m.itype, time=0.0, dt=prm.dt)
    assert ora.step(10) == 0
    ref = ora.state()
    for precision in (H.FP64, H.FP32):
        gpu = H.HipStep(prm, precision=precision)
        gpu.upload(c1_1m.r, c1_1m.v, c1_1m.rho, c1_1m.itype)
        gpu.step(10)
        g = gpu.download()
        t, dt, step = gpu.time()
        gpu.close()
        assert step == 10 and np.array_equal(g["itype"], ref["itype"])
        rho_err = float(np.max(np.abs(g["rho"] - ref["rho"]) / ref["rho"]))
        r_err = float(np.max(np.abs(g["r"] - ref["r"])))
        print("precision", precision, "rho rel", rho_err, "r", r_err, "dt", dt, ref["dt"], flush=True)
        if precision == H.FP64:
            assert abs(dt - ref["dt"]) <= 1e-12 * ref["dt"] and abs(t - ref["time"]) <= 1e-12 * ref["time"]
            assert r_err < 1e-11 and rho_err < 1e-11
            assert float(np.max(np.abs(g["v"] - ref["v"]))) < 1e-9 * max(float(np.max(np.abs(ref["v"]))), 1e-3)
        else:
            assert rho_err < 1e-4
            assert r_err < 1e-5 * prm.delta
            assert abs(dt - ref["dt"]) <= 1e-4 * ref["dt"]


def test_incremental_resort_above_2_to_the_20_equals_full_sort(monkeypatch):
    """1.13M particles (more than 16,384 wavefronts of ballots: the
    changed-slot scan of resort_sort_small runs two passes): from a state
    past the start-up transient (40 steps in, when a step moves a few hundred
    particles to another cell, under the 8,192 the incremental path takes),
    20 steps with the incremental re-sort are bit-identical to 20 with the
    full sort."""
    lay = G.piston_tank(delta=0.003)
    prm = lay.params
    assert prm.ntotal > 1 << 20
    print("re-sort above 2^20: n", prm.ntotal, flush=True)
    monkeypatch.setenv("RRSPH_HIP_LIST", "0")  # the sorted step of every step
    monkeypatch.setenv("RRSPH_HIP_SORT", "full")
    warm = H.HipStep(prm, precision=H.FP32)
    monkeypatch.delenv("RRSPH_HIP_SORT")
    warm.upload(lay.r, lay.v, lay.rho, lay.itype)
    warm.step(40)
    st0 = warm.download()
    warm.close()
    out = {}
    for mode in ("full", "incremental"):
        if mode == "full":
            monkeypatch.setenv("RRSPH_HIP_SORT", "full")
        gpu = H.HipStep(prm, precision=H.FP32)
        monkeypatch.delenv("RRSPH_HIP_SORT", raising=False)
        gpu.upload(st0["r"], st0["v"], st0["rho"], st0["itype"])
        gpu.step(20)
        out[mode] = (gpu.download(), gpu.sort_stats())
        gpu.close()
    si = out["incremental"][1]
    print("sort stats", si, flush=True)
    assert si["incremental"] == 19 and si["fallbacks"] == 0 and 0 < si["last_changed"] <= 8192
    for k in ("r", "v", "rho", "p", "itype"):
        assert np.array_equal(out["full"][0][k], out["incremental"][0][k]), k
