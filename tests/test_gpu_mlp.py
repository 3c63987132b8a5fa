"""The learned warm start's MLP on the matrix cores against the FP64 oracle MLP.

Tolerances (BF16 weights and activations between layers, FP32 accumulate):
  forward outputs       2e-2 of the output scale
  gradient (per tensor) relative L2 error < 5e-2, cosine similarity > 0.998
"""
import numpy as np
import pytest

import splishsplash_for_cuda_amd as S
from splishsplash_for_cuda_amd._native import MLP_NPARAM

pytestmark = pytest.mark.gpu

SIZES = [(64, 12), (64,), (64, 64), (64,), (64, 64), (64,), (2, 64), (2,)]


def _params(seed):
    rng = np.random.default_rng(seed)
    out = []
    for s in SIZES:
        fan_in = s[1] if len(s) == 2 else 64
        out.append(rng.normal(scale=1.0 / np.sqrt(fan_in), size=s).ravel() if len(s) == 2
                   else rng.normal(scale=0.1, size=s))
    return np.concatenate(out)


def _bf16(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def test_mlp_forward_matches_oracle(oracle_mod):
    p = _params(3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(1000, 12)).astype(np.float32)
    s = S.Solver()
    s.mlp_set_params(p.astype(np.float32))
    y = s.mlp_forward(x, normalise=False)
    # the oracle with the weights rounded to BF16 as the device uses them
    pw = p.copy()
    off = np.cumsum([0] + [int(np.prod(z)) for z in SIZES])
    for k in (0, 2, 4, 6):
        pw[off[k]:off[k + 1]] = _bf16(p[off[k]:off[k + 1]])
    ref = oracle_mod.mlp_forward(pw, _bf16(x))
    scale = np.abs(ref).max()
    assert np.abs(y - ref).max() < 2e-2 * scale, np.abs(y - ref).max() / scale


def test_mlp_gradient_matches_oracle(oracle_mod):
    p = _params(5)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(300, 12)).astype(np.float32)      # ragged: 300 = 4 full waves + 44
    t = rng.normal(size=(300, 2)).astype(np.float32)
    s = S.Solver()
    s.mlp_set_params(p.astype(np.float32))
    loss, g = s.mlp_loss_grad(x, t)
    loss_o, g_o = oracle_mod.mlp_loss_grad(p, x, t)
    assert loss == pytest.approx(loss_o, rel=2e-2)
    off = np.cumsum([0] + [int(np.prod(z)) for z in SIZES])
    for k in range(len(SIZES)):
        a, b = g[off[k]:off[k + 1]], g_o[off[k]:off[k + 1]]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = float(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)))
        assert rel < 5e-2 and cos > 0.998, (k, rel, cos)


def test_normalization_is_scale_free():
    """Columns of the training set are standardised by their own mean and spread
    (the dt-invariant totals are O(1e-9) in SI units); a constant column gets
    a spread of max(|mean|, 1)."""
    rng = np.random.default_rng(7)
    x = rng.normal(size=(2000, 12)).astype(np.float32)
    x[:, 4] *= 3e-9
    x[:, 5] = 0.25                         # constant column
    t = np.stack([rng.normal(-6e-6, 7e-6, 2000), rng.normal(-1e-9, 7e-9, 2000)], 1).astype(np.float32)
    s = S.Solver()
    s.mlp_set_params(np.zeros(MLP_NPARAM, np.float32))
    s.mlp_dataset_upload(x, t)
    s.mlp_fit_normalization()
    _, nrm = s.mlp_get_params()
    mu, sd = nrm[:12], nrm[12:24]
    tmu, tsd = nrm[24:26], nrm[26:28]
    np.testing.assert_allclose(mu, x.astype(np.float64).mean(0), rtol=1e-4, atol=1e-12)
    np.testing.assert_allclose(sd[[0, 1, 2, 3, 4]], x[:, [0, 1, 2, 3, 4]].astype(np.float64).std(0), rtol=1e-3)
    assert sd[5] == pytest.approx(1.0)
    np.testing.assert_allclose(tmu, t.astype(np.float64).mean(0), rtol=1e-4)
    np.testing.assert_allclose(tsd, t.astype(np.float64).std(0), rtol=1e-3)
