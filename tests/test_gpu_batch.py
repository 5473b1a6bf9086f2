"""Batched right-hand sides (several load cases in one launch) on the GPU.

Kernel level: each batched entry point, given nrhs slices, returns for slice k
exactly the bits that the single-RHS entry point returns on slice k alone
(the batch is a launch-grid dimension; the per-slice arithmetic and the
order of every reduction are those of the single call).

Solver level: the 8-load block solve on the multi-load bracket agrees with
the FP64 direct solution of every load case; the multi-load SIMP objective
is the sum of the single-load compliances and its sensitivity the sum of the
single-load sensitivities (linearity of the objective in the load cases; the
reference has no multi-load path to compare with).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
import topogpu_amd as tg
from topogpu_amd import fem
from topogpu_amd.bc import generate_bc
from topogpu_amd.hip import lib as L
from topogpu_amd.hip.simp import HipSIMP
from topogpu_amd.hip.state import HipStateSolver

from gpu_util import DEV, h, init, rel_l2, stream, t

pytestmark = pytest.mark.gpu

NR = 3
F32 = dict(dtype=torch.float32, device=DEV)


@pytest.fixture(autouse=True)
def _init(gpu):
    init()


def _ops(dims, rng):
    """(name, op, n, free) for every operator kind: fine-grid hex operators on
    dims, Galerkin level-1 operators (implicit Q8 and stored forms) on the
    coarse grid; free is the host bool array of free DOFs. Returns the
    tensors that must outlive the launches as well."""
    nn = int(np.prod([d + 1 for d in dims]))
    mask = np.full(nn, 7, np.uint8)
    mask.reshape(dims[0] + 1, -1)[0] = 0                       # x = 0 face clamped
    mask[rng.choice(nn, 5, replace=False)] = 5                 # a few single fixed components
    E = (rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3)
    keep = [t(mask), t(E), t(E.astype(np.float32))]
    dm, dE64, dE32 = keep
    out = []
    free0, free1 = _free(mask), None
    for kind in (L.TG_OP_HEX_F32, L.TG_OP_HEX_BF16, L.TG_OP_HEX_BF16X3):
        out.append((f"hex{kind}", L.TgOp32(kind, L.grid(dims), L.ptr(dm), L.ptr(dE32), None), 3 * nn, free0))
    dc = fem.coarse_dims(dims)
    nc = int(np.prod([d + 1 for d in dc]))
    m1 = fem.injection_mask(dims, dc, mask)
    dm1 = t(m1)
    ne = int(np.prod(dc))
    m64 = torch.zeros(576 * ne, dtype=torch.float64, device=DEV)
    m32 = torch.zeros(576 * ne, **F32)
    use_mask = int(not fem.masks_consistent(dims, dc, mask, m1))
    L.call("tg_galerkin_elems", L.grid(dims), L.grid(dc), L.ptr(dm), use_mask, L.ptr(dE64), None, L.ptr(m64),
           L.ptr(m32), stream())
    coef = torch.zeros(243 * nc, **F32)
    L.call("tg_assemble_stencil", L.grid(dc), L.ptr(dm1), L.ptr(m64), L.ptr(coef), stream())
    coef16 = torch.zeros(243 * nc, dtype=torch.int16, device=DEV)
    L.call("tg_assemble_stencil_bf16", L.grid(dc), L.ptr(dm1), L.ptr(m64), L.ptr(coef16), stream())
    keep += [dm1, m64, m32, coef, coef16]
    free1 = _free(m1)
    for kind in (L.TG_OP_Q8_F32, L.TG_OP_Q8_BF16, L.TG_OP_Q8_BF16X3):
        out.append((f"q8_{kind}", L.TgOp32(kind, L.grid(dc), L.ptr(dm1), L.ptr(dE32), None), 3 * nc, free1))
    out.append(("stored", L.TgOp32(L.TG_OP_STORED_F32, L.grid(dc), L.ptr(dm1), None, L.ptr(m32)), 3 * nc, free1))
    out.append(("stencil", L.TgOp32(L.TG_OP_STENCIL_F32, L.grid(dc), L.ptr(dm1), None, None, L.ptr(coef)), 3 * nc,
                free1))
    out.append(("stencil_bf16", L.TgOp32(L.TG_OP_STENCIL_BF16, L.grid(dc), L.ptr(dm1), None, None, L.ptr(coef16)),
                3 * nc, free1))
    return out, keep


def _free(mask):
    """Per-DOF free flags of a per-node mask (bit c = component c free)."""
    m = np.asarray(mask, np.uint8)
    return ((m[:, None] >> np.arange(3)) & 1).reshape(-1).astype(bool)


# epilogue -> (buffers it reads or writes besides u; writes y)
EPS = {
    L.TG_EP_STORE: ((), True),
    L.TG_EP_DOT: ((), True),
    L.TG_EP_RESID: (("rhs",), True),
    L.TG_EP_RESID_NORM: (("rhs",), True),
    L.TG_EP_CHEB: (("rhs", "d", "x"), False),
    L.TG_EP_CHEB_R: (("d", "x", "r"), False),
    L.TG_EP_RESID_UPDATE: (("r",), False),
    L.TG_EP_CHEB_FIRST: (("rhs", "d", "x", "r"), False),
    L.TG_EP_RESID_FIRST: (("rhs", "x", "r"), False),
    L.TG_EP_CHEB_LAST: (("d", "x", "r"), False),
}


@pytest.mark.parametrize("dims", [(8, 6, 10), (6, 4, 14)])
def test_batched_operators_bit_identical_to_single(dims):
    rng = np.random.default_rng(11)
    ops, keep = _ops(dims, rng)
    for name, op, n, _ in ops:
        dinv = t(rng.uniform(0.5, 2.0, n).astype(np.float32))
        for mode, (names, writes_y) in EPS.items():
            host = {k: rng.standard_normal(NR * n).astype(np.float32) for k in ("u", "rhs", "d", "x", "r")}
            a1, a2, a3 = 0.7, 0.2, 1.0
            # batched
            bufs = {k: t(v) for k, v in host.items()}
            yb = torch.zeros(NR * n, **F32)
            dotb = torch.zeros(NR, dtype=torch.float64, device=DEV)
            epb = L.ep(mode, dinv=dinv, dot=dotb, a1=a1, a2=a2, a3=a3, **{k: bufs[k] for k in names})
            L.call("tg_apply32_b", op, NR, L.ptr(bufs["u"]), L.ptr(yb) if writes_y else None, epb, stream())
            got = {k: h(bufs[k]) for k in names}
            got_y, got_dot = h(yb), h(dotb)
            # one slice at a time through the single-RHS entry point
            for k in range(NR):
                sl = slice(k * n, (k + 1) * n)
                b1 = {kk: t(v[sl]) for kk, v in host.items()}
                y1 = torch.zeros(n, **F32)
                dot1 = torch.zeros(1, dtype=torch.float64, device=DEV)
                ep1 = L.ep(mode, dinv=dinv, dot=dot1, a1=a1, a2=a2, a3=a3, **{kk: b1[kk] for kk in names})
                L.call("tg_apply32", op, L.ptr(b1["u"]), L.ptr(y1) if writes_y else None, ep1, stream())
                for kk in names:
                    assert np.array_equal(h(b1[kk]), got[kk][sl]), (name, mode, kk, k)
                if writes_y:
                    assert np.array_equal(h(y1), got_y[sl]), (name, mode, k)
                if mode in (L.TG_EP_DOT, L.TG_EP_RESID_NORM):
                    assert h(dot1)[0] == got_dot[k], (name, mode, k)
    del keep


def test_batched_vector_and_transfer_kernels_bit_identical():
    rng = np.random.default_rng(12)
    dims = (10, 6, 8)
    dc = fem.coarse_dims(dims)
    nf, ncn = 3 * int(np.prod([d + 1 for d in dims])), 3 * int(np.prod([d + 1 for d in dc]))
    maskf = np.full(nf // 3, 7, np.uint8)
    maskf.reshape(dims[0] + 1, -1)[0] = 0
    maskc = fem.injection_mask(dims, dc, maskf)
    dmf, dmc = t(maskf), t(maskc)
    rf = rng.standard_normal(NR * nf).astype(np.float32)
    ec = rng.standard_normal(NR * ncn).astype(np.float32)

    def each(n, fn_b, fn_1, *arrays, outs):
        """Run fn_b on the batch and fn_1 per slice; compare outputs."""
        db = [t(a) for a in arrays]
        ob = [torch.zeros(NR * m, dtype=dt, device=DEV) for m, dt in outs]
        fn_b(*db, *ob)
        gb = [h(o) for o in ob]
        for k in range(NR):
            d1 = [t(a[k * (a.size // NR):(k + 1) * (a.size // NR)]) for a in arrays]
            o1 = [torch.zeros(m, dtype=dt, device=DEV) for m, dt in outs]
            fn_1(*d1, *o1)
            for j, (m, _) in enumerate(outs):
                assert np.array_equal(h(o1[j]), gb[j][k * m:(k + 1) * m]), (fn_1, j, k)

    # restriction and prolongation (overwrite)
    each(nf, lambda r, o: L.call("tg_restrict32_b", L.grid(dims), L.grid(dc), L.ptr(dmc), NR, L.ptr(r), L.ptr(o),
                                 stream()),
         lambda r, o: L.call("tg_restrict32", L.grid(dims), L.grid(dc), L.ptr(dmc), L.ptr(r), L.ptr(o), stream()),
         rf, outs=[(ncn, torch.float32)])
    each(ncn, lambda e, o: L.call("tg_prolong32_b", L.grid(dims), L.grid(dc), L.ptr(dmf), NR, L.ptr(e), L.ptr(o), 0,
                                  stream()),
         lambda e, o: L.call("tg_prolong32", L.grid(dims), L.grid(dc), L.ptr(dmf), L.ptr(e), L.ptr(o), 0, stream()),
         ec, outs=[(nf, torch.float32)])
    # dot products (two at once) and FP64 dot
    a, b, c = (rng.standard_normal(NR * nf).astype(np.float32) for _ in range(3))
    each(nf, lambda x, y, z, o1, o2: L.call("tg_dot32_b", nf, NR, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(o1), L.ptr(o2),
                                            stream()),
         lambda x, y, z, o1, o2: L.call("tg_dot32", nf, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(o1), L.ptr(o2), stream()),
         a, b, c, outs=[(1, torch.float64), (1, torch.float64)])
    a64, b64 = rng.standard_normal(NR * nf), rng.standard_normal(NR * nf)
    each(nf, lambda x, y, o: L.call("tg_dot64_b", nf, NR, L.ptr(x), L.ptr(y), L.ptr(o), stream()),
         lambda x, y, o: L.call("tg_dot64", nf, L.ptr(x), L.ptr(y), L.ptr(o), stream()),
         a64, b64, outs=[(1, torch.float64)])
    # coarsest dense application
    fr = fem.free_from_node_mask(maskc)
    Ainv = rng.standard_normal((fr.size, fr.size))
    dA, dfl = t(Ainv.ravel()), t(fr.astype(np.int32))
    each(ncn, lambda r, o: L.call("tg_coarsest_apply_b", L.ptr(dA), fr.size, L.ptr(dfl), ncn, NR, L.ptr(r), L.ptr(o),
                                  stream()),
         lambda r, o: L.call("tg_coarsest_apply", L.ptr(dA), fr.size, L.ptr(dfl), ncn, L.ptr(r), L.ptr(o), stream()),
         ec, outs=[(ncn, torch.float32)])
    # PCG updates with per-slice scalars
    rz, pq, zq = (rng.uniform(0.5, 2.0, NR) for _ in range(3))
    x, r, p, q, z = (rng.standard_normal(NR * nf).astype(np.float32) for _ in range(5))
    dx, dr, dp = t(x), t(r), t(p)
    rr = torch.zeros(NR, dtype=torch.float64, device=DEV)
    L.call("tg_pcg_xr_b", nf, NR, L.ptr(dx), L.ptr(dr), L.ptr(dp), L.ptr(t(q)), L.ptr(t(rz)), L.ptr(t(pq)), L.ptr(rr),
           stream())
    L.call("tg_pcg_dir_b", nf, NR, L.ptr(dp), L.ptr(t(z)), L.ptr(t(zq)), L.ptr(t(pq)), stream())
    gx, gr, gp, grr = h(dx), h(dr), h(dp), h(rr)
    for k in range(NR):
        sl = slice(k * nf, (k + 1) * nf)
        x1, r1, p1 = t(x[sl]), t(r[sl]), t(p[sl])
        rr1 = torch.zeros(1, dtype=torch.float64, device=DEV)
        L.call("tg_pcg_xr", nf, L.ptr(x1), L.ptr(r1), L.ptr(p1), L.ptr(t(q[sl])), L.ptr(t(rz[k:k + 1])),
               L.ptr(t(pq[k:k + 1])), L.ptr(rr1), stream())
        L.call("tg_pcg_dir", nf, L.ptr(p1), L.ptr(t(z[sl])), L.ptr(t(zq[k:k + 1])), L.ptr(t(pq[k:k + 1])), stream())
        assert np.array_equal(h(x1), gx[sl]) and np.array_equal(h(r1), gr[sl]) and np.array_equal(h(p1), gp[sl])
        assert h(rr1)[0] == grr[k]
    # Chebyshev start with a shared dinv
    dinv = t(rng.uniform(0.5, 2.0, nf).astype(np.float32))
    rhs = rng.standard_normal(NR * nf).astype(np.float32)
    each(nf, lambda s_, d, xo, ro: L.call("tg_cheb_init32_b", nf, NR, L.ptr(s_), L.ptr(dinv), 0.6, L.ptr(d),
                                          L.ptr(xo), L.ptr(ro), stream()),
         lambda s_, d, xo, ro: L.call("tg_cheb_init32", nf, L.ptr(s_), L.ptr(dinv), 0.6, L.ptr(d), L.ptr(xo),
                                      L.ptr(ro), stream()),
         rhs, outs=[(nf, torch.float32)] * 3)


def test_batched_fp64_residual_and_sensitivity():
    rng = np.random.default_rng(13)
    dims = (9, 5, 7)
    nn = int(np.prod([d + 1 for d in dims]))
    n = 3 * nn
    mask = np.full(nn, 7, np.uint8)
    mask.reshape(dims[0] + 1, -1)[0] = 0
    E = rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3
    dm, dE = t(mask), t(E)
    U = rng.standard_normal(NR * n)
    Fv = rng.standard_normal(NR * n)
    dU, dF = t(U), t(Fv)
    y = torch.zeros(NR * n, dtype=torch.float64, device=DEV)
    y32 = torch.zeros(NR * n, **F32)
    ss = torch.zeros(NR, dtype=torch.float64, device=DEV)
    L.call("tg_apply64_b", L.grid(dims), L.ptr(dm), L.ptr(dE), NR, L.ptr(dU), L.ptr(dF), L.ptr(y), L.ptr(y32),
           L.ptr(ss), stream())
    gy, gy32, gss = h(y), h(y32), h(ss)
    for k in range(NR):
        sl = slice(k * n, (k + 1) * n)
        y1 = torch.zeros(n, dtype=torch.float64, device=DEV)
        y1_32 = torch.zeros(n, **F32)
        s1 = torch.zeros(1, dtype=torch.float64, device=DEV)
        L.call("tg_apply64", L.grid(dims), L.ptr(dm), L.ptr(dE), L.ptr(t(U[sl])), L.ptr(t(Fv[sl])), L.ptr(y1),
               L.ptr(y1_32), L.ptr(s1), stream())
        assert np.array_equal(h(y1), gy[sl]) and np.array_equal(h(y1_32), gy32[sl]) and h(s1)[0] == gss[k]
    # weighted sensitivity = weighted sum of single-load sensitivities
    rp = rng.uniform(0.1, 1.0, E.size)
    dh = rng.uniform(0.0, 1.0, E.size)
    w = np.array([0.5, 1.0, 2.0])
    ce = torch.zeros(E.size, dtype=torch.float64, device=DEV)
    dcf = torch.zeros(E.size, dtype=torch.float64, device=DEV)
    L.call("tg_sensitivity64_b", L.grid(dims), NR, L.ptr(dU), L.ptr(t(w)), L.ptr(t(rp)), L.ptr(t(dh)), 3.0, 1.0,
           L.ptr(ce), L.ptr(dcf), stream())
    edof = oracle.edof_table_3d(*dims)
    ce_ref = sum(w[k] * oracle.element_compliance(U[k * n:(k + 1) * n], edof) for k in range(NR))
    assert np.abs(h(ce) - ce_ref).max() < 1e-12 * np.abs(ce_ref).max()
    dcf_ref = -3.0 * rp ** 2 * ce_ref * dh
    assert np.abs(h(dcf) - dcf_ref).max() < 1e-12 * np.abs(dcf_ref).max()
    # weights NULL means all ones
    L.call("tg_sensitivity64_b", L.grid(dims), NR, L.ptr(dU), None, L.ptr(t(rp)), L.ptr(t(dh)), 3.0, 1.0,
           L.ptr(ce), L.ptr(dcf), stream())
    ce1 = sum(oracle.element_compliance(U[k * n:(k + 1) * n], edof) for k in range(NR))
    assert np.abs(h(ce) - ce1).max() < 1e-12 * np.abs(ce1).max()


def _bracket(dims, n_loads=8):
    prob = tg.multi_load_bracket(dims, n_loads=n_loads, seed=3)
    bc = generate_bc(prob.spec)
    return prob, bc.free_dofs, prob.force_override, bc.ndof


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_block_solve_matches_direct_per_load(precision):
    dims = (16, 8, 8)
    prob, free, F, ndof = _bracket(dims)
    rng = np.random.default_rng(14)
    E = 1e-9 + rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3
    s = HipStateSolver(dims, free, F, precision=precision, n_levels=3, tol=1e-9, max_krylov=400)
    assert s.nrhs == 8
    E64 = torch.from_numpy(E).to(DEV)
    s.set_modulus(E64, E64.float(), float(E.max()))
    st = s.solve()
    assert st.converged and st.rel_residual <= 1e-9, st
    U = h(s.u).reshape(8, ndof)
    edof = oracle.edof_table_3d(*dims)
    comps = s.compliances()
    for k in range(8):
        Uk = oracle.solve_state(edof, E, free, F[k], ndof)
        assert rel_l2(U[k], Uk) < 1e-6, k
        assert comps[k] == pytest.approx(float(F[k] @ Uk), rel=1e-8)
    # the block takes the iterations of its slowest load case, not their sum
    singles = []
    for k in (0, 3):
        s1 = HipStateSolver(dims, free, F[k], precision=precision, n_levels=3, tol=1e-9, max_krylov=400)
        s1.set_modulus(E64, E64.float(), float(E.max()))
        singles.append(s1.solve().iters)
        assert rel_l2(h(s1.u), U[k]) < 1e-6
    assert st.iters <= 2 * max(singles)


def test_multi_load_simp_is_sum_of_single_loads():
    """One SIMP step of the 4-load bracket: objective and filtered
    sensitivity equal the sums over single-load models of the same design."""
    dims = (16, 8, 8)
    prob, free, F, ndof = _bracket(dims, n_loads=4)
    kw = dict(volfrac=0.3, rmin=1.5, tol=1e-10, precision="bf16", n_levels=3)
    multi = HipSIMP(dims, free, F, **kw)
    multi.physical()
    c, st = multi.state()
    assert st.converged
    multi.sensitivities()
    dc = h(multi.dc)
    c_sum, dc_sum = 0.0, np.zeros_like(dc)
    for k in range(4):
        one = HipSIMP(dims, free, F[k], **kw)
        one.physical()
        ck, stk = one.state()
        assert stk.converged
        one.sensitivities()
        c_sum += ck
        dc_sum += h(one.dc)
    assert c == pytest.approx(c_sum, rel=1e-8)
    assert np.abs(dc - dc_sum).max() < 1e-7 * np.abs(dc_sum).max()
    # weights scale the cases
    w = np.array([1.0, 0.0, 0.0, 0.0])
    first = HipSIMP(dims, free, F, weights=w, **kw)
    first.physical()
    c1, _ = first.state()
    one = HipSIMP(dims, free, F[0], **kw)
    one.physical()
    assert c1 == pytest.approx(one.state()[0], rel=1e-8)


def test_multi_load_solver_api_runs():
    """SIMPSolver accepts a (n_loads, ndof) force_override and runs the
    weighted-compliance loop; compliance decreases over the iterations."""
    prob = tg.multi_load_bracket((16, 8, 8), n_loads=8, seed=3)
    res = tg.SIMPSolver(backend="hip", tol=1e-5, max_iter=4, move=0.2).solve(prob)
    comp = np.array([r["compliance"] for r in res.history])
    assert np.all(np.isfinite(comp)) and comp[-1] < comp[0]
    assert all(r["converged"] for r in res.history)
    assert res.rho_final.shape == (16 * 8 * 8,)


@pytest.mark.parametrize("dims", [(8, 6, 10), (6, 4, 14)])
def test_fused_smoother_modes_equal_unfused_sequences(dims):
    """The fused first and last smoothing steps return the bits of the
    sequences they replace, for every operator kind and a batch of NR:
    CHEB_FIRST(rhs) = cheb_init(rhs) then CHEB_R(d0) (a3 = 0);
    RESID_FIRST(rhs) = cheb_init(rhs) then RESID_UPDATE(d0);
    CHEB_LAST = CHEB_R restricted to x.
    The right-hand side is zero at fixed DOFs and dinv is 1 there, as in the
    smoother, so whole vectors compare equal."""
    rng = np.random.default_rng(5)
    ops, keep = _ops(dims, rng)
    a0, a1, a2 = 0.61, 1.37, 0.29
    for name, op, n, free in ops:
        dinv_h = np.where(free, rng.uniform(0.5, 2.0, n), 1.0).astype(np.float32)
        dinv = t(dinv_h)
        rhs_h = np.tile(np.where(free, 1.0, 0.0), NR).astype(np.float32) * rng.standard_normal(NR * n).astype(np.float32)
        rhs = t(rhs_h)
        z = lambda: torch.zeros(NR * n, **F32)
        # unfused: start, then the first Chebyshev step / the residual update
        d0, xu, ru, du = z(), z(), z(), z()
        L.call("tg_cheb_init32_b", n, NR, L.ptr(rhs), L.ptr(dinv), a0, L.ptr(d0), L.ptr(xu), L.ptr(ru), stream())
        r1 = ru.clone()
        L.call("tg_apply32_b", op, NR, L.ptr(d0), None,
               L.ep(L.TG_EP_CHEB_R, dinv=dinv, d=du, x=xu, r=ru, a1=a1, a2=a2), stream())
        L.call("tg_apply32_b", op, NR, L.ptr(d0), None, L.ep(L.TG_EP_RESID_UPDATE, r=r1), stream())
        # fused
        xf, rf, df = z(), z(), z()
        L.call("tg_apply32_b", op, NR, L.ptr(rhs), None,
               L.ep(L.TG_EP_CHEB_FIRST, rhs=rhs, dinv=dinv, d=df, x=xf, r=rf, a1=a1, a2=a2, a3=a0), stream())
        for a, b, what in ((rf, ru, "r"), (df, du, "d"), (xf, xu, "x")):
            assert np.array_equal(h(a), h(b)), (name, "CHEB_FIRST", what)
        x1, r2 = z(), z()
        L.call("tg_apply32_b", op, NR, L.ptr(rhs), None,
               L.ep(L.TG_EP_RESID_FIRST, rhs=rhs, dinv=dinv, x=x1, r=r2, a3=a0), stream())
        assert np.array_equal(h(r2), h(r1)), (name, "RESID_FIRST", "r")
        assert np.array_equal(h(x1), h(d0)), (name, "RESID_FIRST", "x")
        # last step: CHEB_R from the state just produced vs CHEB_LAST
        xr, rr, dr = xu.clone(), ru.clone(), z()
        xl, rl, dl = xu.clone(), ru.clone(), z()
        L.call("tg_apply32_b", op, NR, L.ptr(du), None,
               L.ep(L.TG_EP_CHEB_R, dinv=dinv, d=dr, x=xr, r=rr, a1=a1, a2=a2, a3=0.5), stream())
        L.call("tg_apply32_b", op, NR, L.ptr(du), None,
               L.ep(L.TG_EP_CHEB_LAST, dinv=dinv, d=dl, x=xl, r=rl, a1=a1, a2=a2, a3=0.5), stream())
        assert np.array_equal(h(xl), h(xr)), (name, "CHEB_LAST", "x")
        assert np.array_equal(h(rl), h(ru)) and not h(dl).any(), (name, "CHEB_LAST wrote r or d")
    del keep

