"""The per-point gradient error statistics reported beside the norm error (known answers)."""
import numpy as np

from accuracy import gradient_errors


def test_known_answers():
    ref = np.array([[3.0, 4.0], [0.0, 0.0], [6.0, 8.0], [0.0, 1.0]])
    got = ref.copy()
    got[1] = [0.0, 1e-3]  # an error at a point whose gradient is zero
    got[2] = [6.0, 8.01]
    e = gradient_errors(got, ref)
    rms = np.sqrt((25 + 0 + 100 + 1) / 4)
    assert np.isclose(e["max_rel"], 0.01 / 10.0)
    per = np.array([0.0, 1e-3 / (1e-3 * rms), 0.01 / (10 + 1e-3 * rms), 0.0])
    assert np.isclose(e["p50"], np.percentile(per, 50))
    assert np.isclose(e["p99"], np.percentile(per, 99))
    assert np.isclose(e["norm_rel"], np.sqrt(1e-6 + 1e-4) / np.sqrt(126.0))


def test_one_dimensional_gradients():
    ref = np.array([1.0, -2.0, 4.0])
    e = gradient_errors(ref * (1 + 1e-6), ref)
    assert np.isclose(e["max_rel"], 1e-6)
    assert e["p99"] <= 1e-6 * (1 + 1e-9)
