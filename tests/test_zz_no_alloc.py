"""dfsph_mlp_forward_noalloc (the entry point that allocates no device
memory) against dfsph_mlp_forward, through the C ABI: the same kernel on the
same inputs, so the outputs are equal bit for bit, on two calls in a row
(the second smaller, so stale staging contents would show), and the size
limit and the missing model are refused with their error codes."""
import numpy as np
import pytest

import splishsplash_for_cuda_amd as S
from splishsplash_for_cuda_amd import learned
from splishsplash_for_cuda_amd import _native as N
from splishsplash_for_cuda_amd._native import MLP_NOALLOC_MAX

pytestmark = pytest.mark.gpu


def test_noalloc_forward_equals_forward():
    s = S.Solver()
    with pytest.raises(S.DFSPHError) as e:
        s.mlp_forward_noalloc(np.zeros((4, 12), np.float32))
    assert e.value.code == N.DFSPH_ERR_NO_MODEL
    params, norm = learned.load()
    s.mlp_set_params(params, norm)
    rng = np.random.default_rng(4)
    for n in (50_000, 3_001):
        x = rng.normal(size=(n, 12)).astype(np.float32) * 0.1
        a = s.mlp_forward(x, normalise=True)
        b = s.mlp_forward_noalloc(x, normalise=True)
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        np.testing.assert_array_equal(a, b)
        b_raw = s.mlp_forward_noalloc(x, normalise=False)
        np.testing.assert_array_equal(s.mlp_forward(x, normalise=False), b_raw)
    with pytest.raises(S.DFSPHError) as e:
        s.mlp_forward_noalloc(np.zeros((MLP_NOALLOC_MAX + 1, 12), np.float32))
    assert e.value.code == N.DFSPH_ERR_ARG
    s.close()
