"""The roofline model restated from the reference (scripts/perf_model.py:
206-218 predict_perf, 368-381 intensities), pinned by hand-computed values
of its closed form, and the HBM bound of the multiply's algorithmic bytes
against the figures worked out independently for the workloads."""
import pytest

from gcoospdm_amd import perf_model as P


def test_predict_perf_both_sides_of_the_ridge():
    # P100: ridge 9500 / 732 = 12.98 FLOP/B
    assert P.predict_perf(40.0, "p100") == pytest.approx(0.85 * 9500e9)
    assert P.predict_perf(4.0, "p100") == pytest.approx(732e9 * 4.0)
    # GTX 980: ridge 4981 / 224 = 22.2; at r = 22 still bandwidth-bound
    assert P.predict_perf(22.0, "980") == pytest.approx(224e9 * 22.0)
    assert P.predict_perf(22.3, "980") == pytest.approx(0.85 * 4981e9)
    # MI355X: ridge 157300 / 8000 = 19.66
    assert P.predict_perf(19.0, "mi355x") == pytest.approx(8000e9 * 19.0)
    assert P.predict_perf(20.0, "mi355x") == pytest.approx(0.85 * 157300e9)


def test_gcoo_time_closed_form():
    # n = 8000, s = 0.99: r = 2 * 8000 * 0.01 / 4 = 40, flops = 1.024e10
    t, r = P.time_of_s(0.99, 8000, "p100")
    assert r == pytest.approx(40.0)
    assert t == pytest.approx(1.024e10 / (0.85 * 9500e9))
    # s = 0.999 on a GTX 980: r = 4 < ridge, rate 896 GFLOP/s
    t, r = P.time_of_s(0.999, 8000, "980")
    assert r == pytest.approx(4.0)
    assert t == pytest.approx(1.024e9 / 896e9)
    # cuSPARSE's intensity is 2 whatever N and s; cuBLAS's is N / 6
    assert P.time_of_s_cusparse(0.99, 8000, "mi355x")[1] == pytest.approx(2.0)
    assert P.time_of_cublas(600, "mi355x")[1] == pytest.approx(100.0)


@pytest.mark.parametrize("s,us", [(0.99, 65.0), (0.995, 64.5), (0.999, 64.1)])
def test_hbm_bound_matches_the_workload_figures(s, us):
    # 12 * nnz + 8 * (2000 + 1) + 4 * 8000^2 * 2 bytes over 8 TB/s
    assert P.hbm_bound_time(s, 8000) * 1e6 == pytest.approx(us, abs=0.06)
