"""Helpers shared by the GPU parity tests: run the HIP path and the oracle on
the same seeded inputs and compare by particle id."""
import numpy as np


def boundary_index_map(solver, boundary_xyz):
    """Map the GPU's cell-sorted boundary slots to the input order (positions
    are unique in every test scene)."""
    import splishsplash_for_cuda_amd as S
    bpos = np.ascontiguousarray(solver.download(S.FIELD_BOUNDARY_POS), np.float32)
    src = np.ascontiguousarray(boundary_xyz, np.float32)
    key = {src[i].tobytes(): i for i in range(len(src))}
    return np.array([key[bpos[s].tobytes()] for s in range(len(bpos))], np.int64)


def gpu_neighbor_sets(solver, bmap):
    """Per particle id: (set of fluid neighbor ids, set of boundary input indices)."""
    import splishsplash_for_cuda_amd as S
    ids = solver.download(S.FIELD_ID)
    nf, nb, L = solver.neighbors()
    res = {}
    for i in range(len(ids)):
        f = frozenset(ids[L[i, :nf[i]]].tolist())
        b = frozenset(bmap[L[i, nf[i]:nf[i] + nb[i]]].tolist())
        res[int(ids[i])] = (f, b)
    return res


def oracle_neighbor_sets(nf, nb, L):
    return {i: (frozenset(L[i, :nf[i]].tolist()), frozenset(L[i, nf[i]:nf[i] + nb[i]].tolist()))
            for i in range(len(nf))}


def max_rel(a, b, scale=None):
    """max |a - b| / scale (scale: max |b| by default)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = np.abs(b).max() if scale is None else scale
    return float(np.abs(a - b).max() / max(s, 1e-300))
