"""Helpers shared by the GPU tests (not collected: no test_ prefix)."""

from __future__ import annotations

from types import SimpleNamespace as NS

import numpy as np
import torch

import oracle
from topogpu_amd import fem
from topogpu_amd.hip import lib as L

DEV = torch.device("cuda", 0)


def stream():
    return torch.cuda.current_stream(DEV).cuda_stream


_KEEP: list = []


def t(x, dtype=None):
    """Host array -> device tensor. The tensor is kept alive until the next
    h() synchronises: kernels are asynchronous, and a temporary freed right
    after L.ptr() would hand its memory to the next allocation."""
    a = np.ascontiguousarray(x)
    d = torch.from_numpy(a if dtype is None else a.astype(dtype)).to(DEV)
    _KEEP.append(d)
    return d


def h(x):
    torch.cuda.synchronize(DEV)
    out = x.cpu().numpy()
    _KEEP.clear()
    return out


def init():
    with torch.cuda.device(DEV):
        L.init_device(fem.unit_stiffness(), DEV.index)


def cantilever_bc(dims, point_support=False):
    """x=0 face clamped (optionally plus one interior pinned node), tip load."""
    nelx, nely, nelz = dims
    sup = [NS(edge="left", constraint="fixed")]
    if point_support:
        sup.append(NS(x=float(nelx // 2 + 1), y=float(nely), z=float(nelz // 2), constraint="pin_y"))
    spec = NS(nelx=nelx, nely=nely, nelz=nelz, Lx=float(nelx), Ly=float(nely), Lz=float(nelz), supports=sup,
              loads=[NS(x=float(nelx), y=nely / 2.0, z=0.0, fx=0.0, fy=0.0, fz=-1.0)])
    _, free, F, ndof = oracle.generate_bc(spec)
    return free, F, ndof


def free_mask_full(free, ndof):
    m = np.zeros(ndof, bool)
    m[free] = True
    return m


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_inf(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
