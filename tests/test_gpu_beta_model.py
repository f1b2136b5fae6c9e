"""The learned warm start of the perplexity calibration (BF16 MFMA MLP) on the GPU.

* Forward: log(beta0) against a numpy model that rounds the same operands to BF16 (features, weights,
  first hidden layer) and accumulates in FP64, and against plain FP32 torch.
* Backward: the gradient of the mean squared error with respect to every parameter tensor against
  torch autograd in FP32 on the CPU.
* Gain: trained on kNN distances and bisection betas of three C5-style mixtures (D = 100, clusters with
  sd log-uniform in [0.3, 3], perplexity 30, K = 90; N = 16,000 each instead of C5's 4M, which the exact
  kNN cannot reach in a test), the model warm-starts the unchanged FP64 bisection on a fourth mixture:
  fewer iterations, and every row still meets |H - log(perplexity)| < 1e-5.
"""
import numpy as np
import pytest
import torch

import fitsne_gpu as fg
import synth

pytestmark = pytest.mark.gpu
K = 90


def _note(msg):
    print(f"[beta model test] {msg}", flush=True)


def bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def features(dist, mu, sd):
    f = np.log(np.maximum(dist * dist, 1e-30)).astype(np.float32)
    x = np.zeros((dist.shape[0], fg.MLP_IN))
    x[:, :K] = (f - mu[:K]) / sd[:K]
    return x


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(fg.MLP_PARAMS, dtype=np.float32)
    u = fg.mlp_unpack(p)
    u["mu"][:K] = rng.normal(-2.0, 0.5, size=K)
    u["sd"][:] = 1.0
    u["sd"][:K] = rng.uniform(0.5, 1.5, size=K)
    u["W1"][:K] = rng.normal(0, np.sqrt(2.0 / K), size=(K, fg.MLP_HID))
    u["b1"][:] = rng.normal(0, 0.1, size=fg.MLP_HID)
    u["W2"][:] = rng.normal(0, np.sqrt(2.0 / fg.MLP_HID), size=(fg.MLP_HID, fg.MLP_HID))
    u["b2"][:] = rng.normal(0, 0.1, size=fg.MLP_HID)
    u["w3"][:] = rng.normal(0, np.sqrt(1.0 / fg.MLP_HID), size=fg.MLP_HID)
    u["b3"][:] = 3.0
    return p


def random_dist(n, seed):
    rng = np.random.default_rng(seed)
    return np.sort(np.exp(rng.normal(-1.0, 0.4, size=(n, K))), axis=1)


def test_forward_matches_bf16_numpy_and_fp32_torch():
    p = random_params(3)
    dist = random_dist(3000, 4)  # not a multiple of 64: the last tile is partial
    m = fg.BetaModel(K=K)
    m.set_params(p)
    y = np.log(m.predict(dist))
    u = fg.mlp_unpack(p.astype(np.float64))
    x = features(dist, u["mu"], u["sd"])
    # BF16 operands, wide accumulation
    h1 = bf16(np.maximum(bf16(x) @ bf16(u["W1"]) + u["b1"], 0.0))
    y16 = np.maximum(h1 @ bf16(u["W2"]) + u["b2"], 0.0) @ u["w3"] + u["b3"][0]
    # plain FP32
    h = np.maximum(x @ u["W1"] + u["b1"], 0.0)
    y32 = np.maximum(h @ u["W2"] + u["b2"], 0.0) @ u["w3"] + u["b3"][0]
    e16, e32 = np.abs(y - y16).max(), np.abs(y - y32).max()
    _note(f"forward max |dy| vs BF16 model {e16:.2e}, vs FP32 {e32:.2e} (|y| ~ {np.abs(y).mean():.2f})")
    assert e16 < 5e-3
    assert e32 < 1e-1


def test_gradient_matches_torch_autograd():
    p = random_params(5)
    dist = random_dist(4000, 6)
    target = np.exp(np.random.default_rng(7).normal(3.0, 0.5, size=dist.shape[0]))
    m = fg.BetaModel(K=K)
    m.set_params(p)
    g, loss = m.gradient(dist, target)
    u = fg.mlp_unpack(p)
    t = {k: torch.tensor(v.astype(np.float32), requires_grad=k not in ("mu", "sd")) for k, v in u.items()}
    x = torch.tensor(features(dist, u["mu"], u["sd"]).astype(np.float32))
    h1 = torch.relu(x @ t["W1"] + t["b1"])
    y = torch.relu(h1 @ t["W2"] + t["b2"]) @ t["w3"] + t["b3"]
    tl = ((y - torch.tensor(np.log(target), dtype=torch.float32)) ** 2).mean()
    tl.backward()
    gu = fg.mlp_unpack(np.concatenate([np.zeros(fg.MLP_TRAIN0, np.float32), g]))
    _note(f"loss gpu {loss:.6f} torch {tl.item():.6f}")
    assert abs(loss - tl.item()) / tl.item() < 2e-2
    for k in ["W1", "b1", "W2", "b2", "w3", "b3"]:
        ref = t[k].grad.numpy().astype(np.float64)
        got = gu[k].astype(np.float64)
        if k == "W1":
            ref, got = ref[:K], got[:K]
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        _note(f"grad {k}: rel err {rel:.2e}")
        assert rel < 3e-2, k
    assert np.all(gu["W1"][K:] == 0.0)


def _calibrate(X, beta0=None):
    Xc = X - X.mean(axis=0)
    Xc /= np.abs(Xc).max()
    _, dist = fg.knn_exact(Xc, K)
    P, beta, iters = fg.distances2similarities(dist, 30.0, squared=True, beta0=beta0)
    return dist, P, beta, iters


def _entropy_err(P):
    with np.errstate(divide="ignore", invalid="ignore"):
        H = -np.sum(np.where(P > 0, P * np.log(P), 0.0), axis=1)
    return np.abs(H - np.log(30.0)).max()


def test_warm_start_saves_bisection_iterations():
    dists, betas = [], []
    for seed in (1, 2, 3):
        X, _ = synth.c5_mixture(seed=seed)
        d, _, b, it = _calibrate(X)
        dists.append(d)
        betas.append(b)
    m = fg.BetaModel(K=K, seed=11)
    loss = m.fit(np.concatenate(dists), np.concatenate(betas), epochs=40, batch=1024, lr=2e-3)
    X, _ = synth.c5_mixture(seed=4)
    dist, P_cold, b_cold, it_cold = _calibrate(X)
    beta0 = m.predict(dist)
    P_warm, b_warm, it_warm = fg.distances2similarities(dist, 30.0, squared=True, beta0=beta0)
    rel0 = np.abs(np.log(beta0 / b_cold))
    _note(f"train loss {loss:.4f}; held-out |log beta0 - log beta| median {np.median(rel0):.3f}; "
          f"bisection iterations cold {it_cold.mean():.2f} warm {it_warm.mean():.2f} "
          f"(saved {100 * (1 - it_warm.mean() / it_cold.mean()):.1f}%); "
          f"entropy error cold {_entropy_err(P_cold):.1e} warm {_entropy_err(P_warm):.1e}")
    assert _entropy_err(P_cold) < 1e-5 and _entropy_err(P_warm) < 1e-5
    assert it_warm.mean() < 0.9 * it_cold.mean()
    np.testing.assert_allclose(b_warm, b_cold, rtol=1e-3)
