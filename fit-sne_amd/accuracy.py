"""Gradient error statistics of a result against a reference gradient on the same Y and P.

norm_rel:  ||dC - ref||_2 / ||ref||_2 over all coordinates;
max_rel:   max_i |dC_i - ref_i| / max_i |ref_i|, with |.| the Euclidean norm of one point's gradient;
p50, p99:  percentiles over the points of |dC_i - ref_i| / (|ref_i| + eps * rms_i |ref_i|), eps = 1e-3, so
           that points whose gradient is near zero (cluster centres) are judged against a small fraction
           of the typical gradient instead of against zero.
"""
import numpy as np

EPS = 1e-3


def gradient_errors(got, ref, eps=EPS):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if ref.ndim == 1:
        got, ref = got[:, None], ref[:, None]
    d = np.linalg.norm(got - ref, axis=1)
    r = np.linalg.norm(ref, axis=1)
    rms = float(np.sqrt(np.mean(r * r)))
    per = d / (r + eps * rms)
    return {
        "norm_rel": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
        "max_rel": float(d.max() / r.max()),
        "p50": float(np.percentile(per, 50)),
        "p99": float(np.percentile(per, 99)),
        "eps": eps,
    }
