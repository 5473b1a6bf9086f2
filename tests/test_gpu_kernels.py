"""Kernel-level parity of libtopohip against the FP64 oracle (GPU).

Tolerances: FP64 kernels 1e-12 relative; FP32 vector kernels 2e-6; BF16
MFMA 1.5e-2 (inputs and KE rounded to 8-bit mantissas, FP32 accumulate);
split BF16 (hi/lo, 3 MFMA) 1e-4. Inputs of the single-precision kernels are
rounded to FP32 first so that only kernel arithmetic is measured.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from oracle import gmg as ogmg
from oracle import simp as osimp
from topogpu_amd import fem
from topogpu_amd.hip import lib as L

from gpu_util import DEV, cantilever_bc, free_mask_full, h, init, rel_inf, rel_l2, stream, t

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _init(gpu):
    init()


def _bf16_bits(x: np.ndarray) -> np.ndarray:
    return (np.asarray(x, np.float32).view(np.uint32) >> 16).astype(np.uint16)


def test_mfma_operand_layout():
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    B = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    lanes = np.arange(64)
    a = np.zeros((64, 8), np.float32)
    b = np.zeros((64, 8), np.float32)
    for l in lanes:
        for v in range(8):
            a[l, v] = A[l % 16, 8 * (l // 16) + v]
            b[l, v] = B[8 * (l // 16) + v, l % 16]
    da = t(_bf16_bits(a).ravel())
    db = t(_bf16_bits(b).ravel())
    dd = torch.zeros(256, dtype=torch.float32, device=DEV)
    L.call("tg_mfma_probe", L.ptr(da), L.ptr(db), L.ptr(dd), stream())
    d = h(dd).reshape(64, 4)
    ref = A @ B
    got = np.zeros((16, 16), np.float32)
    for l in lanes:
        for v in range(4):
            got[4 * (l // 16) + v, l % 16] = d[l, v]
    assert np.array_equal(got, ref)


def _setup(dims, seed=0, point_support=False):
    free, F, ndof = cantilever_bc(dims, point_support)
    rng = np.random.default_rng(seed)
    n_el = int(np.prod(dims))
    E = rng.uniform(1e-3, 1.0, n_el) ** 3
    nn = ndof // 3
    mask = fem.node_mask_from_free(free, nn)
    return free, F, ndof, E, mask, rng


OPS = {"fp32": (L.TG_OP_HEX_F32, 2e-6), "bf16": (L.TG_OP_HEX_BF16, 1.5e-2), "bf16x3": (L.TG_OP_HEX_BF16X3, 1e-4)}


@pytest.mark.parametrize("dims", [(5, 3, 4), (9, 7, 17), (16, 9, 31), (4, 8, 15), (2, 2, 2)])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16", "bf16x3"])
def test_hex_operator_matches_oracle(dims, prec):
    free, F, ndof, E, mask, rng = _setup(dims, 1, point_support=True)
    edof = oracle.edof_table_3d(*dims)
    fm = free_mask_full(free, ndof)
    u = rng.standard_normal(ndof)
    if prec == "fp64":
        y = torch.zeros(ndof, dtype=torch.float64, device=DEV)
        L.call("tg_apply64", L.grid(dims), L.ptr(t(mask)), L.ptr(t(E)), L.ptr(t(u)), None, L.ptr(y), None, None,
               stream())
        ref = osimp.matfree_matvec(u, E, edof, fm)
        assert rel_l2(h(y), ref) < 1e-13
        return
    kind, tol = OPS[prec]
    E32, u32 = E.astype(np.float32), u.astype(np.float32)
    ref = osimp.matfree_matvec(u32.astype(np.float64), E32.astype(np.float64), edof, fm)
    y = torch.full((ndof,), np.nan, dtype=torch.float32, device=DEV)
    dE = t(E32)
    op = L.TgOp32(kind, L.grid(dims), L.ptr(t(mask)), L.ptr(dE), None)
    ep = L.ep(L.TG_EP_STORE)
    L.call("tg_apply32", op, L.ptr(t(u32)), L.ptr(y), ep, stream())
    got = h(y)
    assert np.all(got[~fm] == 0.0)
    assert rel_l2(got, ref) < tol, (prec, rel_l2(got, ref))


@pytest.mark.parametrize("prec", ["bf16", "bf16x3", "fp32"])
def test_rigid_translation_is_exact_zero(prec):
    """KE annihilates translations; the MFMA kernel removes each element's
    first-node displacement before rounding, so a translation gives 0."""
    dims = (9, 7, 17)
    ndof = 3 * 10 * 8 * 18
    mask = np.full(ndof // 3, 7, np.uint8)
    E = np.random.default_rng(3).uniform(0.1, 1.0, int(np.prod(dims))).astype(np.float32)
    u = np.tile(np.array([0.3, -1.7, 2.5], np.float32), ndof // 3)
    y = torch.zeros(ndof, dtype=torch.float32, device=DEV)
    op = L.TgOp32(OPS[prec][0], L.grid(dims), L.ptr(t(mask)), L.ptr(t(E)), None)
    ep = L.ep(L.TG_EP_STORE)
    L.call("tg_apply32", op, L.ptr(t(u)), L.ptr(y), ep, stream())
    got = h(y)
    if prec == "fp32":
        assert np.abs(got).max() < 1e-5
    else:
        assert np.abs(got).max() == 0.0


@pytest.mark.parametrize("dims", [(128, 64, 64), (141, 33, 95)])
def test_hex_mfma_grids_larger_than_one_wave_of_workgroups(dims):
    """Grids with more tiles than workgroups resident at once (several
    rounds of workgroups, tiles cut at every grid face): every tile is
    computed once, against the FP64 device operator (itself checked against
    the oracle above), for STORE, DOT and CHEB_R, one and two right-hand
    sides. BF16 single pass to 4e-3 relative (the 128x64x64 matvec of the
    same input distributions measures 1.8e-3), split BF16 to 1e-5."""
    rng = np.random.default_rng(31)
    nn = int(np.prod([d + 1 for d in dims]))
    n = 3 * nn
    mask = np.full(nn, 7, np.uint8)
    mask.reshape(dims[0] + 1, -1)[0] = 0
    mask[rng.choice(nn, 100, replace=False)] = 2
    E = rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3
    dm, dE64, dE32 = t(mask), t(E), t(E.astype(np.float32))
    u = rng.standard_normal(2 * n).astype(np.float32)
    du = t(u)
    y64 = torch.zeros(2 * n, dtype=torch.float64, device=DEV)
    for k in range(2):
        L.call("tg_apply64", L.grid(dims), L.ptr(dm), L.ptr(dE64), L.ptr(t(u[k * n:(k + 1) * n].astype(np.float64))),
               None, L.ptr(y64[k * n:(k + 1) * n]), None, None, stream())
    ref = h(y64)
    assert np.abs(ref).max() > 0
    for kind, tol in ((L.TG_OP_HEX_BF16X3, 1e-5), (L.TG_OP_HEX_BF16, 4e-3)):
        op = L.TgOp32(kind, L.grid(dims), L.ptr(dm), L.ptr(dE32), None)
        for nr in (1, 2):
            y = torch.full((nr * n,), np.nan, dtype=torch.float32, device=DEV)
            dot = torch.zeros(nr, dtype=torch.float64, device=DEV)
            L.call("tg_apply32_b", op, nr, L.ptr(du), L.ptr(y), L.ep(L.TG_EP_DOT, dot=dot), stream())
            got, gd = h(y), h(dot)
            for k in range(nr):
                sl = slice(k * n, (k + 1) * n)
                assert rel_l2(got[sl], ref[sl]) < tol, (kind, nr, k)
                assert gd[k] == pytest.approx(float(u[sl].astype(float) @ ref[sl]), rel=10 * tol)
        # CHEB_R: r -= A u ; d = a1 dinv r + a2 u ; x += a3 u + d
        dinv = rng.uniform(0.5, 2.0, n).astype(np.float32)
        r0 = rng.standard_normal(n).astype(np.float32)
        x0 = rng.standard_normal(n).astype(np.float32)
        dr, dx, dd = t(r0), t(x0), torch.full((n,), np.nan, dtype=torch.float32, device=DEV)
        L.call("tg_apply32", op, L.ptr(du), None,
               L.ep(L.TG_EP_CHEB_R, dinv=t(dinv), d=dd, x=dx, r=dr, a1=0.3, a2=0.2, a3=1.0), stream())
        fm = np.repeat(mask, 3).reshape(-1, 3) >> np.arange(3) & 1
        fm = fm.ravel().astype(bool)
        u0 = np.where(fm, u[:n], 0.0)
        rn = np.where(fm, r0 - ref[:n], 0.0)
        dn = np.where(fm, 0.3 * dinv * rn + 0.2 * u0, 0.0)
        assert rel_l2(h(dr), rn) < tol
        assert rel_l2(h(dd), dn) < tol
        assert rel_l2(h(dx), np.where(fm, x0 + u0 + dn, 0.0)) < tol


@pytest.mark.parametrize("kind", [L.TG_OP_HEX_F32, L.TG_OP_HEX_BF16])
def test_epilogues(kind):
    dims = (6, 5, 7)
    free, F, ndof, E, mask, rng = _setup(dims, 2)
    edof = oracle.edof_table_3d(*dims)
    fm = free_mask_full(free, ndof)
    E32 = E.astype(np.float32)
    u = (rng.standard_normal(ndof) * fm).astype(np.float32)
    rhs = (rng.standard_normal(ndof) * fm).astype(np.float32)
    dinv = rng.uniform(0.5, 2.0, ndof).astype(np.float32)
    d0 = (rng.standard_normal(ndof) * fm).astype(np.float32)
    Au = osimp.matfree_matvec(u.astype(float), E32.astype(float), edof, fm)
    tol = 2e-6 if kind == L.TG_OP_HEX_F32 else 1.5e-2
    op = L.TgOp32(kind, L.grid(dims), L.ptr(t(mask)), L.ptr(t(E32)), None)
    dot = torch.zeros(2, dtype=torch.float64, device=DEV)
    y = torch.zeros(ndof, dtype=torch.float32, device=DEV)
    du, drhs = t(u), t(rhs)
    # DOT
    L.call("tg_apply32", op, L.ptr(du), L.ptr(y), L.ep(L.TG_EP_DOT, dot=dot), stream())
    assert h(dot)[0] == pytest.approx(float(u.astype(float) @ Au), rel=tol)
    # RESID_NORM
    L.call("tg_apply32", op, L.ptr(du), L.ptr(y), L.ep(L.TG_EP_RESID_NORM, rhs=drhs, dot=dot), stream())
    r = np.where(fm, rhs - Au, 0.0)
    assert rel_l2(h(y), r) < tol
    assert h(dot)[0] == pytest.approx(float(r @ r), rel=tol)
    # CHEB
    dd, xo, ddinv = t(d0), torch.zeros(ndof, dtype=torch.float32, device=DEV), t(dinv)
    a1, a2 = 0.7, 0.2
    L.call("tg_apply32", op, L.ptr(du), None, L.ep(L.TG_EP_CHEB, rhs=drhs, dinv=ddinv, d=dd, x=xo, a1=a1, a2=a2),
           stream())
    dn = np.where(fm, a1 * dinv * (rhs - Au) + a2 * d0, 0.0)
    assert rel_l2(h(dd), dn) < tol
    assert rel_l2(h(xo), np.where(fm, u + dn, 0.0)) < tol
    # CHEB_R: r -= A u ; d = a1 dinv r + a2 u ; x += a3 u + d
    x0 = (rng.standard_normal(ndof) * fm).astype(np.float32)
    dr, dx, dout = t(rhs), t(x0), torch.zeros(ndof, dtype=torch.float32, device=DEV)
    L.call("tg_apply32", op, L.ptr(du), None,
           L.ep(L.TG_EP_CHEB_R, dinv=ddinv, d=dout, x=dx, r=dr, a1=a1, a2=a2, a3=1.0), stream())
    rn = np.where(fm, rhs - Au, 0.0)
    dn = np.where(fm, a1 * dinv * rn + a2 * u, 0.0)
    assert rel_l2(h(dr), rn) < tol
    assert rel_l2(h(dout), dn) < tol
    assert rel_l2(h(dx), np.where(fm, x0 + u + dn, 0.0)) < tol
    # RESID_UPDATE
    dr = t(rhs)
    L.call("tg_apply32", op, L.ptr(du), None, L.ep(L.TG_EP_RESID_UPDATE, r=dr), stream())
    assert rel_l2(h(dr), rn) < tol


def test_diag_inv_hex():
    dims = (7, 4, 5)
    free, F, ndof, E, mask, rng = _setup(dims, 4, point_support=True)
    edof = oracle.edof_table_3d(*dims)
    fm = free_mask_full(free, ndof)
    d = osimp.matfree_diag(E, edof, fm, ndof)
    d64 = torch.zeros(ndof, dtype=torch.float64, device=DEV)
    d32 = torch.zeros(ndof, dtype=torch.float32, device=DEV)
    L.call("tg_diag_inv_hex", L.grid(dims), L.ptr(t(mask)), L.ptr(t(E)), L.ptr(d64), L.ptr(d32), stream())
    assert rel_inf(h(d64), 1.0 / d) < 1e-14
    assert rel_inf(h(d32), 1.0 / d) < 1e-6


@pytest.mark.parametrize("dims_f", [(8, 6, 4), (7, 5, 3), (3, 2, 1), (6, 2, 1)])
def test_transfers_match_oracle_prolongation(dims_f):
    dims_c = fem.coarse_dims(dims_f)
    free, F, ndof, E, mask_f, rng = _setup(dims_f, 5, point_support=dims_f[2] > 1)
    mask_c = fem.injection_mask(dims_f, dims_c, mask_f)
    P = sp.kron(ogmg.scalar_prolongation(dims_f, dims_c), sp.eye(3), format="csr")
    nfd, ncd = P.shape
    fm = np.repeat(mask_f, 3).reshape(-1, 3) >> np.arange(3) & 1
    cm = np.repeat(mask_c, 3).reshape(-1, 3) >> np.arange(3) & 1
    fm, cm = fm.ravel().astype(bool), cm.ravel().astype(bool)
    ec = (rng.standard_normal(ncd) * cm).astype(np.float32)
    xf0 = (rng.standard_normal(nfd) * fm).astype(np.float32)
    xf = t(xf0)
    L.call("tg_prolong32", L.grid(dims_f), L.grid(dims_c), L.ptr(t(mask_f)), L.ptr(t(ec)), L.ptr(xf), 1, stream())
    ref = xf0 + np.where(fm, P @ ec.astype(float), 0.0)
    assert rel_l2(h(xf), ref) < 1e-6
    xf = t(xf0)
    L.call("tg_prolong32", L.grid(dims_f), L.grid(dims_c), L.ptr(t(mask_f)), L.ptr(t(ec)), L.ptr(xf), 0, stream())
    assert rel_l2(h(xf), np.where(fm, P @ ec.astype(float), 0.0)) < 1e-6
    rf = (rng.standard_normal(nfd) * fm).astype(np.float32)
    rc = torch.zeros(ncd, dtype=torch.float32, device=DEV)
    L.call("tg_restrict32", L.grid(dims_f), L.grid(dims_c), L.ptr(t(mask_c)), L.ptr(t(rf)), L.ptr(rc), stream())
    assert rel_l2(h(rc), np.where(cm, P.T @ rf.astype(float), 0.0)) < 1e-6


def test_injection_mask_matches_oracle():
    for dims in [(8, 6, 4), (7, 5, 3), (12, 6, 3)]:
        free, F, ndof, E, mask, rng = _setup(dims, 0)
        dc = fem.coarse_dims(dims)
        mine = fem.free_from_node_mask(fem.injection_mask(dims, dc, mask))
        assert np.array_equal(mine, ogmg.coarse_free_injection(dims, dc, free))


def _oracle_levels(dims, free, E, n_levels):
    edof = oracle.edof_table_3d(*dims)
    ndof = 3 * int(np.prod([d + 1 for d in dims]))
    Kff = oracle.assemble_k(edof, E, ndof)[free][:, free]
    return ogmg.galerkin_coarse_operators(Kff, ogmg.level_dims(*dims, n_levels), free)


def _apply_stored(dims, mask, mats32, u):
    y = torch.zeros(u.size, dtype=torch.float32, device=DEV)
    op = L.TgOp32(L.TG_OP_STORED_F32, L.grid(dims), L.ptr(t(mask)), None, L.ptr(mats32))
    L.call("tg_apply32", op, L.ptr(t(u.astype(np.float32))), L.ptr(y),
           L.ep(L.TG_EP_STORE), stream())
    return h(y)


def _stencil_op(dims, mask, m64):
    nodes = int(np.prod([d + 1 for d in dims]))
    coef = torch.full((243 * nodes,), np.nan, dtype=torch.float32, device=DEV)
    dm = t(mask)
    L.call("tg_assemble_stencil", L.grid(dims), L.ptr(dm), L.ptr(m64), L.ptr(coef), stream())
    # the caller holds coef and dm for as long as it uses the operator
    return L.TgOp32(L.TG_OP_STENCIL_F32, L.grid(dims), L.ptr(dm), None, None, L.ptr(coef)), coef, dm


def _stencil_apply_host(dims, coef, u):
    """y = A u in FP64 from stencil coefficients coef[(c*81 + 3*q + d) * nodes
    + n] (neighbour q = (dx+1)*9 + (dy+1)*3 + dz+1, node n = (i*NY + j)*NZ + k)."""
    NX, NY, NZ = (d + 1 for d in dims)
    nn = NX * NY * NZ
    cf = coef.reshape(3, 27, 3, nn)
    up = np.pad(u.reshape(NX, NY, NZ, 3), ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((nn, 3))
    for q in range(27):
        dx, dy, dz = q // 9 - 1, (q // 3) % 3 - 1, q % 3 - 1
        un = up[1 + dx:1 + dx + NX, 1 + dy:1 + dy + NY, 1 + dz:1 + dz + NZ].reshape(nn, 3)
        for c in range(3):
            for d in range(3):
                y[:, c] += cf[c, q, d] * un[:, d]
    return y.ravel()


def _check_stencil_epilogues(op, K, fr, nd, rng, tol=2e-6):
    """Every epilogue of a coarse-level operator against the sparse Galerkin
    matrix K (on the free DOFs fr of an nd-vector)."""
    fm = np.zeros(nd, bool)
    fm[fr] = True
    u = np.zeros(nd)
    u[fr] = rng.standard_normal(fr.size)
    Au = np.zeros(nd)
    Au[fr] = K @ u[fr]
    rhs = np.where(fm, rng.standard_normal(nd), 0.0)
    dinv = rng.uniform(0.5, 2.0, nd)
    d0 = np.where(fm, rng.standard_normal(nd), 0.0)
    x0 = np.where(fm, rng.standard_normal(nd), 0.0)
    f32 = lambda a: t(a.astype(np.float32))   # noqa: E731
    du = f32(u)
    y = torch.zeros(nd, dtype=torch.float32, device=DEV)
    dot = torch.zeros(1, dtype=torch.float64, device=DEV)
    L.call("tg_apply32", op, L.ptr(du), L.ptr(y), L.ep(L.TG_EP_STORE), stream())
    assert rel_l2(h(y), Au) < tol
    L.call("tg_apply32", op, L.ptr(du), L.ptr(y), L.ep(L.TG_EP_DOT, dot=dot), stream())
    assert h(dot)[0] == pytest.approx(float(u @ Au), rel=tol)
    L.call("tg_apply32", op, L.ptr(du), L.ptr(y), L.ep(L.TG_EP_RESID_NORM, rhs=f32(rhs), dot=dot), stream())
    r = rhs - Au
    assert rel_l2(h(y), r) < tol and h(dot)[0] == pytest.approx(float(r @ r), rel=tol)
    a1, a2 = 0.7, 0.2
    dd, xo = f32(d0), torch.zeros(nd, dtype=torch.float32, device=DEV)
    L.call("tg_apply32", op, L.ptr(du), None,
           L.ep(L.TG_EP_CHEB, rhs=f32(rhs), dinv=f32(dinv), d=dd, x=xo, a1=a1, a2=a2), stream())
    dn = np.where(fm, a1 * dinv * r + a2 * d0, 0.0)
    assert rel_l2(h(dd), dn) < tol and rel_l2(h(xo), np.where(fm, u + dn, 0.0)) < tol
    dr, dx, dout = f32(rhs), f32(x0), torch.zeros(nd, dtype=torch.float32, device=DEV)
    L.call("tg_apply32", op, L.ptr(du), None,
           L.ep(L.TG_EP_CHEB_R, dinv=f32(dinv), d=dout, x=dx, r=dr, a1=a1, a2=a2, a3=1.0), stream())
    dn = np.where(fm, a1 * dinv * r + a2 * u, 0.0)
    assert rel_l2(h(dr), r) < tol and rel_l2(h(dout), dn) < tol
    assert rel_l2(h(dx), np.where(fm, x0 + u + dn, 0.0)) < tol
    dr = f32(rhs)
    L.call("tg_apply32", op, L.ptr(du), None, L.ep(L.TG_EP_RESID_UPDATE, r=dr), stream())
    assert rel_l2(h(dr), r) < tol


@pytest.mark.parametrize("dims,point", [((8, 4, 4), False), ((8, 4, 4), True), ((7, 5, 6), False),
                                        ((12, 6, 3), False), ((26, 18, 34), False)])
def test_galerkin_levels_match_ptkp(dims, point):
    free, F, ndof, E, mask0, rng = _setup(dims, 6, point_support=point)
    levels = _oracle_levels(dims, free, E, 3)
    ldims = fem.level_dims(dims, 3)
    masks = [mask0]
    for lv in (1, 2):
        masks.append(fem.injection_mask(ldims[lv - 1], ldims[lv], masks[-1]))
    fine_m = None
    for lv in (1, 2):
        df, dc = ldims[lv - 1], ldims[lv]
        use_mask = int(not fem.masks_consistent(df, dc, masks[lv - 1], masks[lv]))
        ne = int(np.prod(dc))
        m64 = torch.zeros(576 * ne, dtype=torch.float64, device=DEV)
        m32 = torch.zeros(576 * ne, dtype=torch.float32, device=DEV)
        L.call("tg_galerkin_elems", L.grid(df), L.grid(dc), L.ptr(t(masks[lv - 1])), use_mask,
               L.ptr(t(E)) if lv == 1 else None, L.ptr(fine_m), L.ptr(m64), L.ptr(m32), stream())
        fr, Kc, _ = levels[lv]
        nd = 3 * int(np.prod([d + 1 for d in dc]))
        u = np.zeros(nd)
        u[fr] = rng.standard_normal(fr.size)
        y = _apply_stored(dc, masks[lv], m32, u)
        assert rel_l2(y[fr], Kc @ u[fr]) < 5e-6, lv
        op_s, coef, dm_s = _stencil_op(dc, masks[lv], m64)
        assert not torch.isnan(coef).any()
        _check_stencil_epilogues(op_s, Kc, fr, nd, rng)
        # BF16 storage: the bits are the FP32 coefficients rounded to nearest
        # even, and the operator applies exactly that rounded stencil
        cb = torch.zeros(243 * (nd // 3), dtype=torch.int16, device=DEV)
        L.call("tg_assemble_stencil_bf16", L.grid(dc), L.ptr(dm_s), L.ptr(m64), L.ptr(cb), stream())
        bits = h(coef).view(np.uint32).astype(np.uint64)
        rne = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16).astype(np.uint16)
        assert np.array_equal(h(cb).view(np.uint16), rne)
        op_b = L.TgOp32(L.TG_OP_STENCIL_BF16, L.grid(dc), L.ptr(dm_s), None, None, L.ptr(cb))
        yb = torch.zeros(nd, dtype=torch.float32, device=DEV)
        u32 = u.astype(np.float32)
        L.call("tg_apply32", op_b, L.ptr(t(u32)), L.ptr(yb), L.ep(L.TG_EP_STORE), stream())
        cref = (rne.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        assert rel_l2(h(yb), _stencil_apply_host(dc, cref, u32.astype(np.float64))) < 1e-6
        _check_stencil_epilogues(op_b, Kc, fr, nd, rng, tol=1e-2)
        assert np.array_equal(fem.free_from_node_mask(masks[lv]), fr)
        if lv == 1 and not point and all(d % 2 == 0 for d in dims):
            yq = torch.zeros(nd, dtype=torch.float32, device=DEV)
            op = L.TgOp32(L.TG_OP_Q8_F32, L.grid(dc), L.ptr(t(masks[1])), L.ptr(t(E.astype(np.float32))), None)
            L.call("tg_apply32", op, L.ptr(t(u.astype(np.float32))), L.ptr(yq),
                   L.ep(L.TG_EP_STORE), stream())
            assert rel_l2(h(yq)[fr], Kc @ u[fr]) < 5e-6
            # the same operator with the child products on BF16 MFMA
            dE32, dm1 = t(E.astype(np.float32)), t(masks[1])
            for kind, tol in ((L.TG_OP_Q8_BF16X3, 2e-5), (L.TG_OP_Q8_BF16, 2e-2)):
                opm = L.TgOp32(kind, L.grid(dc), L.ptr(dm1), L.ptr(dE32), None)
                _check_stencil_epilogues(opm, Kc, fr, nd, rng, tol=tol)
            d32 = torch.zeros(nd, dtype=torch.float32, device=DEV)
            L.call("tg_diag_inv_q8", L.grid(dc), L.ptr(t(masks[1])), L.ptr(t(E)), L.ptr(d32), stream())
            assert rel_inf(h(d32)[fr], 1.0 / Kc.diagonal()) < 1e-6
        d32 = torch.zeros(nd, dtype=torch.float32, device=DEV)
        L.call("tg_diag_inv_stored", L.grid(dc), L.ptr(t(masks[lv])), L.ptr(m64), L.ptr(d32), stream())
        assert rel_inf(h(d32)[fr], 1.0 / Kc.diagonal()) < 1e-6
        fine_m = m64
        if lv == 2:
            # dense coarsest: assembly, explicit inverse, application
            fidx = -np.ones(nd, np.int32)
            fidx[fr] = np.arange(fr.size, dtype=np.int32)
            A = torch.zeros(fr.size * fr.size, dtype=torch.float64, device=DEV)
            L.call("tg_assemble_dense", L.grid(dc), L.ptr(m64), L.ptr(t(fidx)), fr.size, L.ptr(A), stream())
            Kd = Kc.toarray()
            assert np.abs(h(A).reshape(Kd.shape) - Kd).max() < 1e-12 * np.abs(Kd).max()
            L.call("tg_dense_spd_inverse", L.ptr(A), fr.size, stream())
            Ainv = h(A).reshape(Kd.shape)
            assert np.abs(Ainv @ Kd - np.eye(fr.size)).max() < 1e-8
            rc = np.zeros(nd, np.float32)
            rc[fr] = rng.standard_normal(fr.size)
            ec = torch.full((nd,), np.nan, dtype=torch.float32, device=DEV)
            L.call("tg_coarsest_apply", L.ptr(A), fr.size, L.ptr(t(fr.astype(np.int32))), nd, L.ptr(t(rc)),
                   L.ptr(ec), stream())
            e = h(ec)
            assert rel_l2(e[fr], np.linalg.solve(Kd, rc[fr].astype(float))) < 1e-6
            assert np.all(e[np.setdiff1d(np.arange(nd), fr)] == 0.0)


def _bcr_solve(dims, mask, m64, rc):
    """tg_bcr_setup + tg_bcr_apply_b on nrhs = len(rc) right-hand sides:
    the FP32 output and the FP64 solution kept in the scratch vector."""
    nw = C.c_int64(0)
    L.call("tg_bcr_size", L.grid(dims), C.byref(nw))
    work = torch.zeros(int(nw.value), dtype=torch.float64, device=DEV)
    L.call("tg_bcr_setup", L.grid(dims), L.ptr(t(mask)), L.ptr(m64), L.ptr(work), stream())
    nrhs, nd = rc.shape
    ec = torch.full((nrhs * nd,), np.nan, dtype=torch.float32, device=DEV)
    vec = torch.full((2 * nrhs * nd,), np.nan, dtype=torch.float64, device=DEV)
    L.call("tg_bcr_apply_b", L.grid(dims), L.ptr(work), nrhs, L.ptr(t(rc.ravel())), L.ptr(ec), L.ptr(vec),
           stream())
    return h(ec).reshape(nrhs, nd), h(vec).reshape(nrhs, 2, nd)[:, 1]


# plane counts 2, 3, 6, 7, 21: odd and even active counts at every level,
# the last odd plane with and without a right neighbour; interior pinned
# nodes put fixed DOFs inside planes; (20, 10, 10) is the coarsest level of
# the 4-level 160x80x80 hierarchy (363-DOF planes)
@pytest.mark.parametrize("dims,point", [((1, 2, 2), False), ((2, 3, 1), True), ((5, 2, 2), False),
                                        ((6, 4, 3), True), ((20, 3, 2), True), ((20, 10, 10), False)])
def test_block_cyclic_reduction_solves_coarsest(dims, point):
    """The plane-block cyclic reduction solve equals a direct solve of the
    assembled free-DOF matrix (backward error at FP64 rounding level), for
    two right-hand sides in one launch, with fixed DOFs returned as zero."""
    import scipy.sparse.linalg as spla
    free, F, ndof, E, mask, rng = _setup(dims, 4, point_support=point)
    m64 = t((E[:, None, None] * fem.unit_stiffness()[None]).ravel())
    K = oracle.assemble_k(oracle.edof_table_3d(*dims), E, ndof)[free][:, free].tocsc()
    # the element matrices are the oracle's K, entry by entry (dense assembly)
    if free.size <= 2048:
        fidx = -np.ones(ndof, np.int32)
        fidx[free] = np.arange(free.size, dtype=np.int32)
        A = torch.zeros(free.size * free.size, dtype=torch.float64, device=DEV)
        L.call("tg_assemble_dense", L.grid(dims), L.ptr(m64), L.ptr(t(fidx)), free.size, L.ptr(A), stream())
        Kd = K.toarray()
        assert np.abs(h(A).reshape(Kd.shape) - Kd).max() < 1e-12 * np.abs(Kd).max()
    rc = np.zeros((2, ndof), np.float32)
    rc[:, free] = rng.standard_normal((2, free.size))
    e32, x64 = _bcr_solve(dims, mask, m64, rc)
    fixed = np.setdiff1d(np.arange(ndof), free)
    knorm = abs(K).sum(axis=0).max()
    for k in range(2):
        b = rc[k, free].astype(np.float64)
        x = x64[k, free]
        berr = np.linalg.norm(K @ x - b) / (knorm * np.linalg.norm(x) + np.linalg.norm(b))
        assert berr < 1e-13, berr
        assert rel_l2(e32[k, free], spla.spsolve(K, b)) < 1e-6
        assert np.all(e32[k, fixed] == 0.0) and np.all(x64[k, fixed] == 0.0)


def test_block_cyclic_reduction_reports_not_spd():
    dims = (5, 2, 2)
    free, F, ndof, E, mask, rng = _setup(dims, 4)
    m64 = t(np.zeros(576 * int(np.prod(dims))))   # free rows with a zero diagonal
    with pytest.raises(L.TopoHipError, match="not SPD"):
        _bcr_solve(dims, mask, m64, np.zeros((1, ndof), np.float32))


@pytest.mark.parametrize("dims", [(8, 4, 4), (16, 8, 12)])
def test_galerkin_from_fine_matches_two_step(dims):
    """Gather + DGEMM path (8 child / 64 grandchild matrices) equals the
    per-element Galerkin kernel applied once / twice without masks."""
    free, F, ndof, E, mask0, rng = _setup(dims, 8)
    d1, d2 = fem.coarse_dims(dims), fem.coarse_dims(fem.coarse_dims(dims))
    n1, n2 = int(np.prod(d1)), int(np.prod(d2))
    dE = t(E)
    ref1 = torch.zeros(576 * n1, dtype=torch.float64, device=DEV)
    ref2 = torch.zeros(576 * n2, dtype=torch.float64, device=DEV)
    L.call("tg_galerkin_elems", L.grid(dims), L.grid(d1), L.ptr(t(mask0)), 0, L.ptr(dE), None, L.ptr(ref1), None,
           stream())
    m1 = fem.injection_mask(dims, d1, mask0)
    L.call("tg_galerkin_elems", L.grid(d1), L.grid(d2), L.ptr(t(m1)), 0, None, L.ptr(ref1), L.ptr(ref2), None,
           stream())
    work = torch.zeros(64 * n2 + 8 * n1, dtype=torch.float64, device=DEV)
    got1 = torch.zeros_like(ref1)
    got2 = torch.zeros_like(ref2)
    g32 = torch.zeros(576 * n2, dtype=torch.float32, device=DEV)
    L.call("tg_galerkin_from_fine", L.grid(dims), 1, L.ptr(dE), L.ptr(work), L.ptr(got1), None, stream())
    L.call("tg_galerkin_from_fine", L.grid(dims), 2, L.ptr(dE), L.ptr(work), L.ptr(got2), L.ptr(g32), stream())
    a1, b1, a2, b2, c32 = h(got1), h(ref1), h(got2), h(ref2), h(g32)
    assert np.abs(a1 - b1).max() < 1e-12 * np.abs(b1).max()
    assert np.abs(a2 - b2).max() < 1e-12 * np.abs(b2).max()
    assert np.array_equal(c32, a2.astype(np.float32))


def test_not_spd_is_reported():
    A = t(np.array([1.0, 2.0, 2.0, 1.0]))
    with pytest.raises(L.TopoHipError, match="not SPD"):
        L.call("tg_dense_spd_inverse", L.ptr(A), 2, stream())
    # the register path, the LDS path (in its first, a middle and its last
    # pivot block), then the blocked path in its last, middle and first block
    for n, bad in ((50, 33), (130, 0), (135, 70), (140, 139), (400, 399), (200, 100), (170, 0)):
        B = np.eye(n)
        B[bad, bad] = -1.0
        with pytest.raises(L.TopoHipError, match=f"not SPD \\(pivot {bad + 1}\\)"):
            L.call("tg_dense_spd_inverse", L.ptr(t(B.ravel())), n, stream())
    B = np.eye(2100)   # Cholesky path
    B[1000, 1000] = -1.0
    with pytest.raises(L.TopoHipError, match="not SPD"):
        L.call("tg_dense_spd_inverse", L.ptr(t(B.ravel())), 2100, stream())


def test_dense_spd_inverse_async_matches_host_synchronous():
    """tg_dense_spd_inverse_async (the one-workgroup inverses queued without
    a host wait, status in a device word) gives the bits of
    tg_dense_spd_inverse for SPD matrices, the same first failing pivot + 1
    for non-SPD ones, and rejects sizes above the one-workgroup limit."""
    rng = np.random.default_rng(11)
    info = torch.full((1,), -7, dtype=torch.int32, device=DEV)
    for n in (1, 7, 33, 64, 65, 135, 140):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        A = (Q * np.logspace(0, 6, n)) @ Q.T
        A = 0.5 * (A + A.T)
        ref, got = t(A.ravel()), t(A.ravel())
        L.call("tg_dense_spd_inverse", L.ptr(ref), n, stream())
        L.call("tg_dense_spd_inverse_async", L.ptr(got), n, L.ptr(info), stream())
        torch.cuda.synchronize()
        assert int(h(info)[0]) == 0, n
        assert np.array_equal(h(got), h(ref)), n
    for n, bad in ((50, 33), (135, 0), (135, 70), (140, 139)):
        B = np.eye(n)
        B[bad, bad] = -1.0
        L.call("tg_dense_spd_inverse_async", L.ptr(t(B.ravel())), n, L.ptr(info), stream())
        torch.cuda.synchronize()
        assert int(h(info)[0]) == bad + 1, (n, bad)
    with pytest.raises(L.TopoHipError, match="above 140"):
        L.call("tg_dense_spd_inverse_async", L.ptr(t(np.eye(141).ravel())), 141, L.ptr(info), stream())


@pytest.mark.parametrize("n", [1, 7, 32, 33, 64, 65, 97, 135, 140, 141, 160, 161, 192, 193, 600, 1080, 2100])
def test_dense_spd_inverse_sizes(n):
    """One-workgroup Gauss-Jordan in registers (n <= 64) and in LDS (n <=
    140), blocked Gauss-Jordan (16-wide pivot blocks, n <= 2048) and the
    library Cholesky inverse (above) invert an ill-conditioned SPD matrix to
    FP64 accuracy, symmetric."""
    rng = np.random.default_rng(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, 8, n)   # condition number 1e8
    A = (Q * ev) @ Q.T
    A = 0.5 * (A + A.T)
    d = t(A.ravel())
    L.call("tg_dense_spd_inverse", L.ptr(d), n, stream())
    Ai = h(d).reshape(n, n)
    assert np.array_equal(Ai, Ai.T)
    ref = np.linalg.inv(A)
    assert np.abs(Ai - ref).max() < 1e-6 * np.abs(ref).max()
    # forward error of a coarse solve (cond 1e8: both inverses err ~1e-8)
    b = rng.standard_normal((n, 4))
    x = np.linalg.solve(A, b)
    assert np.linalg.norm(Ai @ b - x) < 1e-6 * np.linalg.norm(x)


@pytest.mark.parametrize("rmin", [1.5, 2.0, 2.7])
@pytest.mark.parametrize("dims", [(9, 5, 7), (20, 17, 35)])
def test_filter_and_transpose(rmin, dims):
    n = int(np.prod(dims))
    H = oracle.build_density_filter(*dims, rmin)
    offs, w = fem.filter_stencil(rmin)
    # Hs = W 1 from the raw stencil
    ones = t(np.ones(n))
    hsd = torch.zeros(n, dtype=torch.float64, device=DEV)
    L.call("tg_stencil64", L.grid(dims), w.size, offs.ravel().ctypes.data, w.ctypes.data, L.ptr(ones), None, None,
           L.ptr(hsd), stream())
    hs_inv = 1.0 / h(hsd)
    x = np.random.default_rng(1).uniform(0, 1, n)
    out = torch.zeros(n, dtype=torch.float64, device=DEV)
    L.call("tg_stencil64", L.grid(dims), w.size, offs.ravel().ctypes.data, w.ctypes.data, L.ptr(t(x)), None,
           L.ptr(t(hs_inv)), L.ptr(out), stream())
    assert rel_inf(h(out), H @ x) < 1e-14
    L.call("tg_stencil64", L.grid(dims), w.size, offs.ravel().ctypes.data, w.ctypes.data, L.ptr(t(x)),
           L.ptr(t(hs_inv)), None, L.ptr(out), stream())
    assert rel_inf(h(out), H.T @ x) < 1e-14


def test_density_chain_and_sensitivity():
    dims = (6, 4, 5)
    n = int(np.prod(dims))
    rng = np.random.default_rng(9)
    rf = rng.uniform(0.0, 1.0, n)
    outs = [torch.zeros(n, dtype=torch.float64, device=DEV) for _ in range(3)]
    e32 = torch.zeros(n, dtype=torch.float32, device=DEV)
    L.call("tg_density_chain", n, L.ptr(t(rf)), 1.0, 0.5, 3.0, 1e-9, 1.0, *[L.ptr(o) for o in outs], L.ptr(e32),
           stream())
    rp, dh, E = (h(o) for o in outs)
    assert rel_inf(rp, oracle.heaviside(rf, 1.0)) < 1e-14
    assert rel_inf(dh, oracle.heaviside_deriv(rf, 1.0)) < 1e-14
    assert rel_inf(E, 1e-9 + (1 - 1e-9) * oracle.heaviside(rf, 1.0) ** 3) < 1e-14
    ndof = 3 * int(np.prod([d + 1 for d in dims]))
    U = rng.standard_normal(ndof)
    ce = torch.zeros(n, dtype=torch.float64, device=DEV)
    dcf = torch.zeros(n, dtype=torch.float64, device=DEV)
    L.call("tg_sensitivity64", L.grid(dims), L.ptr(t(U)), L.ptr(t(rp)), L.ptr(t(dh)), 3.0, 1 - 1e-9, L.ptr(ce),
           L.ptr(dcf), stream())
    ce_ref = oracle.element_compliance(U, oracle.edof_table_3d(*dims))
    assert rel_inf(h(ce), ce_ref) < 1e-13
    assert rel_inf(h(dcf), -3.0 * (1 - 1e-9) * rp ** 2 * ce_ref * dh) < 1e-13


@pytest.mark.parametrize("scale,volfrac,move,spread", [
    (1e-6, 0.5, 0.2, 3.0),      # multiplier below 1: bracket [0, 1], early stop
    (1e6, 0.5, 0.2, 3.0),       # twenty doublings of hi
    (1e25, 0.5, 0.2, 3.0),      # past 2^64: a second doubling pass
    (1.0, 0.30, 0.01, 0.0),     # target out of reach: hi reaches 1e40, all 80 bisections
])
def test_oc_mean_matches_oracle(scale, volfrac, move, spread):
    """Mean-volume OC of the reference GPU loop (solver.py:94-117) against
    the oracle's restatement, bit for bit."""
    rng = np.random.default_rng(17)
    n = 12_345
    rho = rng.uniform(0.05, 1.0, n) if spread else np.full(n, 0.5)
    dc = -scale * 10.0 ** rng.uniform(-spread, 0.0, n)
    dc[::97] = 0.0                  # clamped to -1e-12
    ref = oracle.oc_update_mean(rho, dc, volfrac, move, 1e-3)
    state = torch.zeros(L.OC_MEAN_STATE_LEN, dtype=torch.float64, device=DEV)
    dr, ddc = t(rho), t(dc)
    out = torch.zeros(n, dtype=torch.float64, device=DEV)
    chg = torch.zeros(1, dtype=torch.float64, device=DEV)
    L.call("tg_oc_mean_prepare", L.ptr(state), stream())
    for _ in range(40):
        L.call("tg_oc_mean_step", n, L.ptr(dr), L.ptr(ddc), volfrac, move, 1e-3, L.ptr(state), stream())
    L.call("tg_oc_mean_finalize", n, L.ptr(dr), L.ptr(ddc), move, 1e-3, L.ptr(state), L.ptr(out), L.ptr(chg),
           stream())
    st = h(state)
    assert st[7] == 1.0                                  # stopped
    got = h(out)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    assert h(chg)[0] == np.abs(ref - rho).max()
    if move == 0.01:
        assert st[3] == 80                               # every bisection taken
    else:
        assert abs(ref.mean() - volfrac) < 1e-8


@pytest.mark.parametrize("dims,seed,volfrac,spread", [
    ((10, 6, 5), 12, 0.45, 0.0),    # crossing inside the bracket
    ((10, 6, 5), 12, 0.30, 0.0),    # volume above target even at lambda = 1e9: all steps go right
    ((24, 12, 6), 3, 0.30, 6.0),    # sensitivities over six decades, crossing near 35
    ((24, 12, 6), 3, 0.45, 6.0),
    ((30, 20, 10), 5, 0.45, 6.0),   # crossing near 5e-5
])
def test_oc_bisection_matches_oracle(dims, seed, volfrac, spread):
    """The planned passes reach the sequential bisection's last midpoint and
    step count (pub_simp_solver.py:319-340 via the oracle), and the same
    updated densities."""
    n = int(np.prod(dims))
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.05, 0.95, n)
    dc = -rng.uniform(0.01, 5.0, n) * 10.0 ** rng.uniform(-spread, 0.0, n)
    dv = rng.uniform(0.5, 1.5, n)
    H = oracle.build_density_filter(*dims, 1.5)
    ref, lm_ref, steps_ref = oracle.oc_update_filtered(rho, dc, dv, volfrac, 0.2, H, 1.0, 0.5, True,
                                                       return_state=True)
    offs, w = fem.filter_stencil(1.5)
    # Hs = W 1 from the raw stencil on the host
    Wsum = np.zeros(dims)
    for (dx, dy, dz), wk in zip(offs, w):
        c = np.zeros(dims)
        c[max(0, -dx):dims[0] - max(0, dx), max(0, -dy):dims[1] - max(0, dy), max(0, -dz):dims[2] - max(0, dz)] = 1
        Wsum += wk * c
    hs_inv = 1.0 / Wsum.ravel()
    q, lo, hi, rn = (torch.zeros(n, dtype=torch.float64, device=DEV) for _ in range(4))
    state = torch.zeros(L.OC_STATE_LEN, dtype=torch.float64, device=DEV)
    chg = torch.zeros(1, dtype=torch.float64, device=DEV)
    drho = t(rho)
    L.call("tg_oc_prepare", n, L.ptr(drho), L.ptr(t(dc)), L.ptr(t(dv)), 0.2, L.ptr(q), L.ptr(lo), L.ptr(hi),
           L.ptr(state), stream())
    dhs = t(hs_inv)
    for _ in range(50):   # calls after convergence are no-ops
        L.call("tg_oc_bisect_step", L.grid(dims), w.size, offs.ravel().ctypes.data, w.ctypes.data, L.ptr(dhs),
               L.ptr(q), L.ptr(lo), L.ptr(hi), 1.0, 0.5, volfrac, 1, L.ptr(state), stream())
    L.call("tg_oc_finalize", n, L.ptr(drho), L.ptr(q), L.ptr(lo), L.ptr(hi), L.ptr(state), L.ptr(rn), L.ptr(chg),
           stream())
    st = h(state)
    assert st[3] == 1.0
    assert int(st[4]) == steps_ref
    assert st[2] == lm_ref
    assert st[5] <= 8          # evaluation passes (the tree method needs steps/4)
    got = h(rn)
    assert np.abs(got - ref).max() < 1e-12
    assert h(chg)[0] == pytest.approx(np.abs(ref - rho).max(), abs=1e-12)
    # a second update on the same state starts its first pass around the
    # previous result (kept in entry 27): with the crossing moved a little,
    # far (outside the neighbours, inside the bisection's range) or not at
    # all, the decisions stay the sequential loop's
    assert st[27] == lm_ref
    for scale in (1.0 + 1e-3 * rng.standard_normal(n), np.full(n, 1e4), np.ones(n)):
        dc2 = dc * scale
        ref2, lm2, steps2 = oracle.oc_update_filtered(rho, dc2, dv, volfrac, 0.2, H, 1.0, 0.5, True,
                                                      return_state=True)
        L.call("tg_oc_prepare", n, L.ptr(drho), L.ptr(t(dc2)), L.ptr(t(dv)), 0.2, L.ptr(q), L.ptr(lo), L.ptr(hi),
               L.ptr(state), stream())
        for _ in range(50):
            L.call("tg_oc_bisect_step", L.grid(dims), w.size, offs.ravel().ctypes.data, w.ctypes.data, L.ptr(dhs),
                   L.ptr(q), L.ptr(lo), L.ptr(hi), 1.0, 0.5, volfrac, 1, L.ptr(state), stream())
        L.call("tg_oc_finalize", n, L.ptr(drho), L.ptr(q), L.ptr(lo), L.ptr(hi), L.ptr(state), L.ptr(rn),
               L.ptr(chg), stream())
        st = h(state)
        assert st[3] == 1.0 and int(st[4]) == steps2 and st[2] == lm2
        assert np.abs(h(rn) - ref2).max() < 1e-12


def test_krylov_vector_kernels():
    n = 100_003
    rng = np.random.default_rng(2)
    a, b, c = (rng.standard_normal(n).astype(np.float32) for _ in range(3))
    s = torch.zeros(8, dtype=torch.float64, device=DEV)
    L.call("tg_dot32", n, L.ptr(t(a)), L.ptr(t(b)), L.ptr(t(c)), L.ptr(s[0:1]), L.ptr(s[1:2]), stream())
    r = h(s)
    assert r[0] == pytest.approx(float(a.astype(float) @ b), rel=1e-12)
    assert r[1] == pytest.approx(float(a.astype(float) @ c), rel=1e-12)
    # determinism: same launch, same bits
    L.call("tg_dot32", n, L.ptr(t(a)), L.ptr(t(b)), None, L.ptr(s[2:3]), None, stream())
    assert h(s)[2] == r[0]
    x, rr_, p, q = (t(v) for v in (a, b, c, a[::-1].copy()))
    s[0], s[1] = 2.0, 4.0
    L.call("tg_pcg_xr", n, L.ptr(x), L.ptr(rr_), L.ptr(p), L.ptr(q), L.ptr(s[0:1]), L.ptr(s[1:2]), L.ptr(s[2:3]),
           stream())
    assert np.allclose(h(x), a + 0.5 * c, rtol=1e-6)
    rnew = b - 0.5 * a[::-1]
    assert np.allclose(h(rr_), rnew, rtol=1e-6, atol=1e-6)
    assert h(s)[2] == pytest.approx(float(rnew.astype(float) @ rnew), rel=1e-6)


def test_power_step_with_and_without_reduction():
    """tg_power_step32: v.Dv = sum(v*v/dinv) in FP64 and v <- dinv*kv; with
    vdv NULL (the power iteration's steps before the last) v is updated to
    the same bits and nothing is reduced."""
    n = 100_003
    rng = np.random.default_rng(5)
    v = rng.standard_normal(n).astype(np.float32)
    kv = rng.standard_normal(n).astype(np.float32)
    dinv = rng.uniform(0.5, 2.0, n).astype(np.float32)
    s = torch.full((2,), -1.0, dtype=torch.float64, device=DEV)
    v1, v2 = t(v), t(v)
    L.call("tg_power_step32", n, L.ptr(v1), L.ptr(t(kv)), L.ptr(t(dinv)), L.ptr(s[0:1]), stream())
    L.call("tg_power_step32", n, L.ptr(v2), L.ptr(t(kv)), L.ptr(t(dinv)), None, stream())
    want = v.astype(float) ** 2 / dinv.astype(float)
    got = h(s)
    assert got[0] == pytest.approx(float(want.sum()), rel=1e-12)
    assert got[1] == -1.0
    assert np.array_equal(h(v1), dinv * kv) and np.array_equal(h(v2), h(v1))


def test_pcg_update_writes_rr_into_mapped_host_memory():
    """tg_pcg_xr_b with r.r written through tg_mapped_pointer into pinned
    host memory gives the bits of the device-scalar output, readable on the
    host once an event behind the kernel has passed (the Krylov loop's
    r.r read); pageable host memory is refused by tg_mapped_pointer."""
    from topogpu_amd.hip.readback import Mapped
    n, nr = 50_001, 3
    rng = np.random.default_rng(9)
    vecs = [rng.standard_normal(nr * n).astype(np.float32) for _ in range(4)]
    s = torch.zeros(3, nr, dtype=torch.float64, device=DEV)
    s[0], s[1] = torch.tensor([2.0, 3.0, 5.0]), torch.tensor([4.0, 1.0, 7.0])
    outs = []
    for mapped in (False, True):
        x, r, p, q = (t(v) for v in vecs)
        m = Mapped(nr) if mapped else None
        rr = m.device_ptr if mapped else L.ptr(s[2])
        L.call("tg_pcg_xr_b", n, nr, L.ptr(x), L.ptr(r), L.ptr(p), L.ptr(q), L.ptr(s[0]), L.ptr(s[1]), rr,
               stream())
        outs.append((m.start().wait() if mapped else h(s[2]), h(x), h(r)))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert all(np.array_equal(a, b) for a, b in zip(outs[0][1:], outs[1][1:]))
    want = [float(np.sum(outs[0][2][k * n:(k + 1) * n].astype(float) ** 2)) for k in range(nr)]
    assert outs[1][0] == pytest.approx(want, rel=1e-12)
    with pytest.raises(L.TopoHipError, match="pinned host memory"):
        L.call("tg_mapped_pointer", np.zeros(16).ctypes.data, C.addressof(C.c_void_p()))
