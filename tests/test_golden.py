"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must keep reproducing them, and the GPU path must meet its bars on
them."""
import glob
import os

import numpy as np
import pytest

import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def t_end(g):
    pts = g["y_exact"].shape[0]
    h = float(g["h"])
    return (pts - 1) * h + 1e-4 * h


def test_fixtures_present():
    assert len(FIXTURES) >= 4


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_fixture(path):
    g = load(path)
    s, h = int(g["solver"]), float(g["h"])
    ye = orc.simulate_mirror_exact(s, g["A"], g["B"], g["C"], g["D"], g["x0"], 0.0, t_end(g), h,
                                   Uhalf=g["U"])
    np.testing.assert_array_equal(ye, g["y_exact"])
    t, ym = orc.simulate_mirror(s, g["A"], g["B"], g["C"], g["D"], g["x0"], 0.0, t_end(g), h,
                                Uhalf=g["U"])
    np.testing.assert_array_equal(ym, g["y_mirror"])
    np.testing.assert_array_equal(t, g["t"])
    f = lambda a: a.astype(np.float64)
    y64 = orc.simulate_f64(s, f(g["A"]), f(g["B"]), f(g["C"]), f(g["D"]), f(g["x0"]), f(g["U"]),
                           g["y_f64"].shape[0], h)
    np.testing.assert_array_equal(y64, g["y_f64"])


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_fixture_modes_agree(path):
    # three restatements of one solve: FP32 rounding of M separates them by
    # O(1e-6) at most at these short horizons
    g = load(path)
    assert orc.max_rel_err(g["y_mirror"], g["y_exact"]) <= 1e-5
    assert orc.max_rel_err(g["y_exact"], g["y_f64"]) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_gpu_meets_fixture(path):
    from parode_amd import api
    g = load(path)
    s, h = int(g["solver"]), float(g["h"])
    ids, e = api.get_available_gpus()
    assert api.initialize_selected_gpus(ids[:1]) == 0
    try:
        api.set_precision(api.PLAIN32)
        tv, y, e = api.simulate_ode_sampled(s, g["A"], g["B"], g["C"], g["D"], g["U"], 0.0,
                                            t_end(g), h, g["x0"])
        assert e == 0
        np.testing.assert_array_equal(tv, g["t"])
        assert orc.max_rel_err(y, g["y_exact"]) <= 1e-5
        api.set_precision(api.FP64)
        tv, y, e = api.simulate_ode_sampled(s, g["A"], g["B"], g["C"], g["D"],
                                            g["U"].astype(np.float64), 0.0, t_end(g), h, g["x0"])
        assert e == 0
        assert orc.max_rel_err(y, g["y_f64"]) <= 1e-6
    finally:
        api.set_precision(api.PLAIN32)
        api.close_gpu()
