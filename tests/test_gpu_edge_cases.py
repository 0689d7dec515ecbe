"""Edge cases of the GPU step against the FP64 oracle: tiny and ragged
particle counts (1 particle, wave and block boundaries), neighbour-list
overflow (the reference's error with the consistency check on), and FIX
turning escaping fluid into non-existent particles (ConsistencyCheck.cpp,
WholeStep.cpp)."""
import copy

import numpy as np
import pytest

import oracle
from sph2d_amd import generators as G
from sph2d_amd import hip as H
from sph2d_amd import params as P

pytestmark = pytest.mark.gpu


def _cloud(n, seed=3):
    """n fluid particles on a jittered lattice in the dam-break tank, with
    small random velocities and densities (so pressure, viscosity and XSPH
    all act), no walls."""
    base = G.dam_break().params
    prm = copy.copy(base)
    d = prm.delta
    side = int(np.ceil(np.sqrt(n)))
    rng = np.random.default_rng(seed)
    k = np.arange(n)
    r = np.stack([1.0 + d * (k % side), 0.5 + d * (k // side)], axis=1) + rng.uniform(-0.2 * d, 0.2 * d, (n, 2))
    v = rng.normal(0.0, 0.05, (n, 2))
    rho = prm.rho0 * (1.0 + 0.002 * rng.normal(size=n))
    itype = np.full(n, P.TYPE_WATER, dtype=np.int32)
    prm.ntotal = prm.nfluid = n
    prm.nvirt = 0
    prm.maxn = max(prm.maxn, n)
    return prm, r, v, rho, itype


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 257, 1023, 1025, 4097])
def test_tiny_and_ragged_counts_match_the_oracle(n):
    prm, r, v, rho, itype = _cloud(n)
    print("cloud n", n, flush=True)
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(r, v, rho, itype)
    gpu.step(5)
    got = gpu.download()
    gpu.close()
    ora = oracle.Oracle(prm)
    ora.set_state(r, v, rho, itype, time=0.0, dt=prm.dt)
    assert ora.step(5) == 0
    ref = ora.state()
    assert np.max(np.abs(got["r"] - ref["r"])) < 1e-13
    assert np.max(np.abs(got["v"] - ref["v"])) < 1e-10
    assert np.max(np.abs(got["rho"] - ref["rho"]) / ref["rho"]) < 1e-12


def test_neighbour_overflow_is_the_reference_error():
    """More than max_neighbours - 1 neighbours: with the consistency check on
    the step stops with "making neighbours grid: error occured"
    (GridFind.cpp:186-188), as the oracle does; with it off the step runs and
    keeps max_neighbours - 1 of them."""
    lay = G.dam_break()
    prm = copy.copy(lay.params)
    prm.max_neighbours = 8  # a dam-break particle has ~17
    for precision in (H.FP64, H.FP32):
        gpu = H.HipStep(prm, precision=precision)
        gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
        with pytest.raises(H.RrsphError) as e:
            gpu.step(1)
        assert e.value.status == 3 and "making neighbours grid: error occured" in str(e.value)
        # the reference throws before whole_step: the step is not taken (the
        # fused pair pass had integrated it; the context puts the state back)
        t, _dt, step = gpu.time()
        assert step == 0 and t == prm.start_simulation_time
        st = gpu.download()
        gpu.close()
        assert np.array_equal(st["r"], lay.r) and np.array_equal(st["itype"], lay.itype)
        if precision == H.FP64:
            assert np.array_equal(st["v"], lay.v) and np.array_equal(st["rho"], lay.rho)
        else:
            assert np.array_equal(st["v"], lay.v.astype(np.float32).astype(np.float64))
            assert np.array_equal(st["rho"], lay.rho.astype(np.float32).astype(np.float64))
    ora = oracle.Oracle(prm)
    ora.set_state(lay.r, lay.v, lay.rho, lay.itype, time=0.0, dt=prm.dt)
    assert ora.step(1) == 1
    prm.consistency_check = False
    gpu = H.HipStep(prm, precision=H.FP64)
    gpu.upload(lay.r, lay.v, lay.rho, lay.itype)
    gpu.step(2)
    st = gpu.download()
    nb = gpu.neighbours()
    gpu.close()
    assert np.all(np.isfinite(st["r"])) and np.all(np.isfinite(st["rho"]))
    counts = (nb != prm.ntotal).sum(axis=1)
    assert counts.max() == prm.max_neighbours - 1


def test_fix_turns_escaping_fluid_into_non_existent_particles():
    """FIX (consistency_treatment = 2): four isolated fluid particles above the
    tank fly out of the domain and become itype 0 (no neighbours, no motion,
    no longer fluid); the rest of the state follows the oracle."""
    lay = G.dam_break()
    prm = copy.copy(lay.params)
    prm.consistency_treatment = P.CONSISTENCY_FIX
    nf = prm.nfluid
    extra_r = np.array([[0.5, 1.9], [1.2, 1.9], [2.9, 1.9], [2.8, 1.95]])
    extra_v = np.array([[0.0, 900.0], [0.0, 1200.0], [3000.0, 0.0], [-3000.0, 500.0]])
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
    diag = gpu.diagnostics()
    gpu.close()
    ora = oracle.Oracle(prm)
    ora.set_state(r, v, rho, itype, time=0.0, dt=prm.dt)
    assert ora.step(6) == 0
    ref = ora.state()
    assert np.all(got["itype"][nf:nf + 4] == 0), got["itype"][nf:nf + 4]
    assert np.array_equal(got["itype"], ref["itype"])
    assert diag["n_fluid_alive"] == nf
    assert np.array_equal(got["r"][nf:nf + 4], ref["r"][nf:nf + 4])  # dead particles stay where they died
    alive = ref["itype"] != 0
    assert np.max(np.abs(got["r"][alive] - ref["r"][alive])) < 1e-12
    assert np.max(np.abs(got["rho"][alive] - ref["rho"][alive]) / ref["rho"][alive]) < 1e-11
