"""Operators, transfers and MG-CG on cellwise-numbered box meshes.

Renumbering only relabels nodes: cells, their colours and the order in which
each node's cell contributions are summed are unchanged.  So vmult and the
diagonal on the cellwise mesh are bitwise the lexicographic results, permuted.
Transfers agree to rounding.  MG-CG converges to the oracle's solution (on the
lexicographic mesh, permuted) with the same bars as the multigrid parity tests:
the Chebyshev eigenvalue start vector is indexed by dof, so the estimates and
the iterates differ in the last bits.
"""
import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu


def pair(lo, hi, n, p, nc, fm=1):
    cm = 7 if nc == 3 else 1
    a = mfgpu.BoxMesh(lo, hi, n, p, nc, fm, cm)
    b = mfgpu.BoxMesh(lo, hi, n, p, nc, fm, cm, numbering="cellwise")
    perm = np.full(a.n_nodes, -1, dtype=np.int64)
    perm[a.cell_nodes()] = b.cell_nodes()
    dof = (perm[:, None] * nc + np.arange(nc)[None, :]).ravel()  # cellwise dof of each lexicographic dof
    return a, b, torch.from_numpy(dof).cuda()


def solid(m, precision, formulation=1):
    op = mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1, formulation=formulation, precision=precision)
    op.cache(torch.from_numpy(m.u0(amp=1e-3, noise=1e-5, seed=4)).to("cuda", op.dtype))
    return op


@pytest.mark.parametrize("p,formulation,precision", [(2, 1, mfgpu.FP64), (4, 1, mfgpu.FP64), (4, 0, mfgpu.FP64),
                                                     (3, 1, mfgpu.FP32), (5, 1, mfgpu.FP64)])
def test_solid_vmult_and_diagonal_are_permuted_bitwise(cuda, p, formulation, precision):
    print(f"cellwise solid p={p} formulation={formulation} precision={precision}", flush=True)
    a, b, dof = pair([0, 0, 0], [2.0, 1.0, 1.0], [6, 3, 3], p, 3)
    oa, ob = solid(a, precision, formulation), solid(b, precision, formulation)
    src = mfgpu.fill_uniform(11, torch.empty(a.n_dofs, dtype=oa.dtype, device="cuda"))
    src_b = torch.empty_like(src)
    src_b[dof] = src
    ya, yb = torch.empty_like(src), torch.empty_like(src)
    oa.vmult(ya, src)
    ob.vmult(yb, src_b)
    assert torch.equal(yb[dof], ya)
    oa.compute_diagonal()
    ob.compute_diagonal()
    assert torch.equal(ob.diagonal()[dof], oa.diagonal())


@pytest.mark.parametrize("p", [1, 3, 4])
def test_heat_vmult_is_permuted_bitwise(cuda, p):
    a, b, dof = pair([0, 0, 0], [1, 1, 1], [5, 4, 3], p, 1, 63)
    coef = np.linspace(0.5, 2.0, a.n_cells * (p + 1) ** 3)
    oa, ob = mfgpu.LaplaceOperator(a, 0.1, coef), mfgpu.LaplaceOperator(b, 0.1, coef)
    src = mfgpu.fill_uniform(12, torch.empty(a.n_dofs, dtype=torch.float64, device="cuda"))
    src_b = torch.empty_like(src)
    src_b[dof] = src
    ya, yb = torch.empty_like(src), torch.empty_like(src)
    oa.vmult(ya, src)
    ob.vmult(yb, src_b)
    assert torch.equal(yb[dof], ya)


def test_transfers_are_permuted(cuda):
    lo, hi, p = [0, 0, 0], [2.0, 1.0, 1.0], 2
    ca, cb, cdof = pair(lo, hi, [2, 1, 1], p, 3)
    fa, fb, fdof = pair(lo, hi, [4, 2, 2], p, 3)
    ops = [solid(m, mfgpu.FP64) for m in (ca, fa, cb, fb)]
    ta = mfgpu.Transfer(ops[0], ops[1], mfgpu.box_children(ca, fa))
    tb = mfgpu.Transfer(ops[2], ops[3], mfgpu.box_children(cb, fb))
    xc = mfgpu.fill_uniform(13, torch.empty(ca.n_dofs, dtype=torch.float64, device="cuda"))
    xf = mfgpu.fill_uniform(14, torch.empty(fa.n_dofs, dtype=torch.float64, device="cuda"))
    xcb, xfb = torch.empty_like(xc), torch.empty_like(xf)
    xcb[cdof], xfb[fdof] = xc, xf
    for name in ("prolongate_add", "restrict_add"):
        src_a, src_b, n_out, odof = (xc, xcb, fa.n_dofs, fdof) if name == "prolongate_add" else (xf, xfb, ca.n_dofs, cdof)
        ya = torch.zeros(n_out, dtype=torch.float64, device="cuda")
        yb = torch.zeros_like(ya)
        getattr(ta, name)(ya, src_a)
        getattr(tb, name)(yb, src_b)
        assert float((yb[odof] - ya).norm()) <= 1e-14 * float(ya.norm()), name
    ya = torch.empty(ca.n_dofs, dtype=torch.float64, device="cuda")
    yb = torch.empty_like(ya)
    ta.interpolate(ya, xf)
    tb.interpolate(yb, xfb)
    assert torch.equal(yb[cdof], ya)


def test_mgcg_on_cellwise_hierarchy_matches_oracle(cuda):
    lo, hi, coarse, L, p = [0, 0, 0], [2.0, 1.0, 1.0], [2, 1, 1], 3, 2
    gms = [mfgpu.BoxMesh(lo, hi, [c << l for c in coarse], p, 3, 1, 7, numbering="cellwise") for l in range(L)]
    oms = O.box_levels(lo, hi, coarse, L, p, 3, 1, 7)
    fine_l = oms[-1]
    # cellwise dof of each lexicographic dof on the fine level
    perm = np.full(fine_l.n_nodes, -1, dtype=np.int64)
    perm[fine_l.cell_nodes()] = gms[-1].cell_nodes()
    dof = torch.from_numpy((perm[:, None] * 3 + np.arange(3)[None, :]).ravel()).cuda()
    u = fine_l.u0(amp=1e-3, noise=1e-5)
    us = [u]
    for l in range(L - 1, 0, -1):
        us.insert(0, O.interpolate(oms[l - 1], oms[l], us[0]))
    oops = [O.Solid(m, caching=1) for m in oms]
    for op, uu in zip(oops, us):
        op.cache(uu)
    gops = [mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1) for m in gms]
    trs = [mfgpu.Transfer(gops[l - 1], gops[l], mfgpu.box_children(gms[l - 1], gms[l])) for l in range(1, L)]
    ug = [None] * L
    ug[-1] = torch.empty(gms[-1].n_dofs, dtype=torch.float64, device="cuda")
    ug[-1][dof] = torch.from_numpy(u).cuda()
    for l in range(L - 1, 0, -1):
        ug[l - 1] = torch.empty(gms[l - 1].n_dofs, dtype=torch.float64, device="cuda")
        trs[l - 1].interpolate(ug[l - 1], ug[l])
    for op, uu in zip(gops, ug):
        op.cache(uu)
    mg = mfgpu.Multigrid(gops, trs, cheb_degree=4)
    mg.setup()
    xt = O.fill_uniform(2, fine_l.n_dofs)
    xt[fine_l.constrained()] = 0.0
    b = oops[-1].vmult(xt)
    x_o, its_o, _ = O.mg_cg_solve(oms, oops, b, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30,
                                  coarse_kind=0, tol_rel=1e-8)
    bg = torch.empty(fine_l.n_dofs, dtype=torch.float64, device="cuda")
    bg[dof] = torch.from_numpy(b).cuda()
    xg = torch.zeros_like(bg)
    info = mfgpu.cg_solve(gops[-1], bg, xg, mfgpu.PRECOND_MG, mg, tol_rel=1e-8)
    x = xg[dof].cpu().numpy()
    print("cellwise MG-CG iterations", info.iterations, "oracle (lexicographic)", its_o)
    assert abs(info.iterations - its_o) <= 1
    assert np.linalg.norm(x - x_o) <= 1e-8 * np.linalg.norm(x_o)
    assert np.linalg.norm(x - xt) <= 1e-6 * np.linalg.norm(xt)
