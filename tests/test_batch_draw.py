"""The batched workload's instance draw (host logic, no GPU): 96 instances
from one seed, sides powers of two in [32, 512], Ra in [1e4, 1e7], viscosity
contrast in [1, 1e3]; the draw is reproducible and changes with the seed."""
import math

from mantle_convection_lab_amd.batch import draw_instances


def test_draw_ranges_and_reproducibility():
    a = draw_instances(96, 20261202)
    assert len(a) == 96
    for side, Ra, ctr, pseed in a:
        assert side in (32, 64, 128, 256, 512)
        assert 1e4 <= Ra <= 1e7
        assert 1.0 <= ctr <= 1e3
        assert pseed >= 1
    assert a == draw_instances(96, 20261202)
    assert a != draw_instances(96, 20261203)
    # log-uniform sides: every size occurs, the middle three most often
    sides = [s[0] for s in a]
    for s in (32, 64, 128, 256, 512):
        assert sides.count(s) > 0
    assert sides.count(32) + sides.count(512) < sides.count(64) + sides.count(128) + sides.count(256)


def test_side_rounding_in_log_space():
    # the nearest power of two in log space: 45 -> 32, 46 -> 64 (sqrt(2) * 32 = 45.25)
    for s, want in ((45.0, 32), (46.0, 64), (362.0, 256), (363.0, 512)):
        assert int(2 ** round(math.log2(s))) == want
