"""The adaptive budget calibration of bench.py (CPU): the chosen scale meets
the RMS target from either side and lands within 97-100 % of it, the way the
uniform rate is chosen, so the adaptive-vs-uniform comparison is at equal
error."""
import importlib.util
import pathlib

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("c", [0.006, 0.0088, 0.0125, 0.02])
def test_converges_from_both_sides(bench, c):
    # error ~ c / sqrt(scale), as for Neyman allocations without a floor
    calls = []

    def err_at(scale):
        calls.append(scale)
        return c / scale ** 0.5, scale * 100

    scale, e, extra = bench.calibrate_budget(err_at, 0.01)
    assert 0.97 * 0.01 <= e <= 0.01
    assert extra == scale * 100
    assert len(calls) <= 6


def test_floor_limited_error(bench):
    # the error stops falling at 0.011 (a sample floor): no scale meets the
    # target, so the last scale tried is returned with its error
    def err_at(scale):
        return max(0.011, 0.02 / scale ** 0.5), scale

    scale, e, _ = bench.calibrate_budget(err_at, 0.01)
    assert e == pytest.approx(0.011)


def test_smallest_scale_meeting_the_target(bench):
    # a curve that overshoots downwards: among the scales meeting the target,
    # the smallest (cheapest) is kept
    table = {1.0: 0.005, 0.25: 0.0105}

    def err_at(scale):
        for k, v in table.items():
            if abs(scale - k) < 1e-9:
                return v, scale
        return 0.0099, scale

    scale, e, _ = bench.calibrate_budget(err_at, 0.01)
    assert e <= 0.01
    assert scale < 1.0


def test_neyman_bound_closed_forms(bench):
    """Equal needs: the allocation is uniform.  Two halves needing a and b
    samples: Neyman gives mean (sqrt a + sqrt b)^2 / 4 against (a + b) / 2
    uniform.  A floor above the optimum of the small half raises the mean to
    what the large half then needs on top of the floor."""
    import numpy as np

    assert abs(bench.neyman_mean_spp(np.full(100, 37.0)) - 37.0) < 1e-6
    a, b = 4.0, 100.0
    n = np.array([a] * 50 + [b] * 50)
    assert abs(bench.neyman_mean_spp(n) - (a ** 0.5 + b ** 0.5) ** 2 / 4) < 1e-6
    # without a floor the small half gets 12 samples; with a floor of 20 it gets
    # 20 (error share 4 / 20 = 0.2), and the large half needs 100 / n_b = 1.8
    m = bench.neyman_mean_spp(n, floor=20.0)
    assert abs(m - (20.0 + 100.0 / 1.8) / 2) < 1e-6
