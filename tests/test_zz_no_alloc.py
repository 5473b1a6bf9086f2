"""tg_apply64_noalloc against tg_apply64 through the C ABI: the FP64
residual, its FP32 copy and its sum of squares, bit for bit, on two calls in
a row (the second on a smaller grid, so the shared static scratch holds the
first call's partials when it starts), and the size limit."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from topogpu_amd.hip import lib as L

from gpu_util import DEV, cantilever_bc, h, init, stream, t

pytestmark = pytest.mark.gpu


def _both(dims, seed):
    free, F, ndof = cantilever_bc(dims, point_support=True)
    rng = np.random.default_rng(seed)
    n_nodes = ndof // 3
    mask = np.zeros(ndof, bool)
    mask[free] = True
    nm = (mask.reshape(-1, 3) * np.array([1, 2, 4])).sum(axis=1).astype(np.uint8)
    E = 1e-9 + rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3
    u = rng.standard_normal(ndof) * mask
    rhs = rng.standard_normal(ndof) * mask
    dm, dE, du, dr = t(nm), t(E), t(u), t(rhs)
    out = {}
    for name in ("tg_apply64", "tg_apply64_noalloc"):
        y = torch.full((ndof,), np.nan, dtype=torch.float64, device=DEV)
        y32 = torch.full((ndof,), np.nan, dtype=torch.float32, device=DEV)
        ss = torch.full((1,), np.nan, dtype=torch.float64, device=DEV)
        L.call(name, L.grid(dims), L.ptr(dm), L.ptr(dE), L.ptr(du), L.ptr(dr), L.ptr(y), L.ptr(y32), L.ptr(ss),
               stream())
        out[name] = (h(y), h(y32), h(ss)[0])
        assert n_nodes == dm.numel()
    return out


def test_noalloc_residual_matches_allocating_entry_point(gpu):
    init()
    for dims, seed in (((48, 24, 24), 5), ((12, 6, 6), 6)):      # second call smaller
        out = _both(dims, seed)
        a, b = out["tg_apply64"], out["tg_apply64_noalloc"]
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), dims
        assert a[2] == b[2] and np.isfinite(a[2]) and a[2] > 0, dims
        assert a[2] == pytest.approx(float(np.sum(a[0] ** 2)), rel=1e-12)


def test_noalloc_rejects_grids_beyond_the_scratch(gpu):
    init()
    x = torch.zeros(8, dtype=torch.float64, device=DEV)
    m = torch.zeros(8, dtype=torch.uint8, device=DEV)
    with pytest.raises(L.TopoHipError, match="exceed the static scratch"):
        L.call("tg_apply64_noalloc", L.grid((400, 400, 100)), L.ptr(m), L.ptr(x), L.ptr(x), None, L.ptr(x), None,
               L.ptr(x), stream())
