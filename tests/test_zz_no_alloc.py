"""rrsph_hip_smooth_gauges_noalloc (the smooth gauges with no allocation)
against rrsph_hip_smooth_gauges through the C-ABI: bit-identical results on
consecutive calls, the second with a smaller input (fewer particles and
fewer gauges), in 2D and 3D; more than 64 gauges is refused before any
device work."""
import numpy as np
import pytest

from sph2d_amd import generators as G
from sph2d_amd import hip as H

pytestmark = pytest.mark.gpu


def _sim(lay, precision, steps):
    sim = H.HipStep(lay.params, precision=precision)
    sim.upload(lay.r, lay.v, lay.rho, lay.itype)
    if steps:
        sim.step(steps)
    return sim


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_noalloc_gauges_equal_the_allocating_entry_point(precision):
    big = G.piston_tank(delta=0.01, length=3.0, depth=0.4)
    small = G.piston_tank(delta=0.04, length=1.0, depth=0.2)
    print("no-alloc gauges: n", big.params.ntotal, "then", small.params.ntotal, flush=True)
    cases = [(big, list(np.linspace(0.1, 2.9, 12)), 0.4), (small, [0.2, 0.5, 0.8], 0.2)]
    for lay, xs, depth in cases:
        sim = _sim(lay, precision, 5)
        d = lay.params.delta
        spec = H.gauge_spec(xs, radius=3 * d, tau=d / 4, y_ref=depth)
        want = sim.smooth_gauges(spec)
        got = sim.smooth_gauges_noalloc(spec)
        assert np.all(np.isfinite(got)) and np.all(got > 0)
        assert np.array_equal(got, want), (got, want)
        again = sim.smooth_gauges_noalloc(spec)  # the scratch keeps the last call's contents
        assert np.array_equal(again, want)
        sim.close()


def test_noalloc_gauges_3d_and_the_gauge_limit():
    lay = G.dam_break_3d(delta=0.04)
    sim = _sim(lay, H.FP64, 3)
    xs, zs = [0.1, 0.2, 0.3, 0.5], [0.1, 0.2, 0.1, 0.3]
    spec = H.gauge_spec(xs, radius=0.12, tau=0.01, y_ref=0.3, zs=zs)
    assert np.array_equal(sim.smooth_gauges_noalloc(spec), sim.smooth_gauges(spec))
    spec.n = 65  # beyond the scratch
    with pytest.raises(H.RrsphError) as e:
        sim.smooth_gauges_noalloc(spec)
    assert e.value.status == 1 and "at most 64 gauges" in str(e.value)
    sim.close()
