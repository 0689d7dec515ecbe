"""The CPU baseline leg of bench.py on a small flume (no GPU): the FP64
oracle runs the requested steps in three consecutive runs after two warm-up
steps, and the reported rate is the median of the three."""
import bench
from sph2d_amd import generators as G


def test_cpu_leg_reports_the_median_of_three_runs():
    lay = G.piston_tank(delta=0.02, length=1.2, depth=0.3, wall_height=0.6)
    rate, threads, secs, rates = bench.cpu_leg(lay, 9)
    assert len(rates) == 3 and threads >= 1 and secs > 0
    assert rate == sorted(rates)[1]
    assert all(r > 0 for r in rates)
