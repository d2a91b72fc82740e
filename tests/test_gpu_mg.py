"""Device multigrid parity through the C ABI: transfer operators, one V-cycle,
Chebyshev eigenvalue estimates and MG-CG against the oracle on identical
hierarchies.  Bars: transfers and the V-cycle to 1e-12 / 1e-10 relative;
MG-CG solution within 1e-8 of the oracle's and iteration count within +-1."""
import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu


def dev(x, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype)


def rel(a, b):
    a = a.detach().double().cpu().numpy() if hasattr(a, "detach") else a
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def hierarchy(kind, lo, hi, coarse, n_levels, p, precisions=None, face_mask=1, amp=1e-3, noise=1e-5,
              storage=mfgpu.T2_FOLDED):
    """Matching device and oracle hierarchies, linearised at interpolated u0;
    FP64 solid levels keep the given tensor2 point storage."""
    nc = 3 if kind == "solid" else 1
    cm = 7 if nc == 3 else 1
    precisions = precisions or [mfgpu.FP64] * n_levels
    gms = [mfgpu.BoxMesh(lo, hi, [c << l for c in coarse], p, nc, face_mask, cm) for l in range(n_levels)]
    oms = O.box_levels(lo, hi, coarse, n_levels, p, nc, face_mask, cm)
    if kind == "solid":
        gops = [mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1, precision=pr) for m, pr in zip(gms, precisions)]
        for op, pr in zip(gops, precisions):
            if pr == mfgpu.FP64:
                op.set_point_storage(storage)
        oops = [O.Solid(m, caching=1) for m in oms]
    else:
        gops = [mfgpu.LaplaceOperator(m, 0.1, None, precision=pr) for m, pr in zip(gms, precisions)]
        oops = [O.Heat(m, dt=0.1) for m in oms]
    trs = [mfgpu.Transfer(gops[l - 1], gops[l], mfgpu.box_children(gms[l - 1], gms[l])) for l in range(1, n_levels)]
    if kind == "solid":
        u = oms[-1].u0(amp=amp, noise=noise)
        us = [u]
        for l in range(n_levels - 1, 0, -1):
            us.insert(0, O.interpolate(oms[l - 1], oms[l], us[0]))
        for op, uu in zip(oops, us):
            op.cache(uu)
        # interpolate_to_mg on the device, each level in its own precision
        ug = [None] * n_levels
        ug[-1] = dev(u, gops[-1].dtype)
        for l in range(n_levels - 1, 0, -1):
            ug[l - 1] = torch.empty(gms[l - 1].n_dofs, dtype=gops[l - 1].dtype, device="cuda")
            trs[l - 1].interpolate(ug[l - 1], ug[l])
        for op, uu in zip(gops, ug):
            op.cache(uu)
    return gms, oms, gops, oops, trs


@pytest.mark.parametrize("p,nc", [(1, 3), (2, 1), (2, 3), (4, 3)])
def test_transfer_parity(cuda, p, nc):
    kind = "solid" if nc == 3 else "heat"
    gms, oms, gops, oops, trs = hierarchy(kind, [0, 0, 0], [2.0, 1.0, 1.0], [2, 1, 1], 2, p, face_mask=1 | 16)
    cm, fm = oms
    c = O.fill_uniform(11, cm.n_dofs)
    f = O.fill_uniform(12, fm.n_dofs)
    fo = f.copy()
    O.prolongate_add(cm, fm, c, fo)
    fg = dev(f)
    trs[0].prolongate_add(fg, dev(c))
    assert rel(fg, fo) < 1e-13
    co = c.copy()
    O.restrict_add(cm, fm, f, co)
    cg = dev(c)
    trs[0].restrict_add(cg, dev(f))
    assert rel(cg, co) < 1e-13
    ci = torch.empty(cm.n_dofs, dtype=torch.float64, device="cuda")
    trs[0].interpolate(ci, dev(f))
    assert rel(ci, O.interpolate(cm, fm, f)) < 1e-13


def test_vcycle_and_eigenvalues_small_config(cuda):
    """8^3 Q2 cube, 3 levels, Chebyshev degree 4 (FP64 levels)."""
    gms, oms, gops, oops, trs = hierarchy("solid", [0, 0, 0], [1, 1, 1], [2, 2, 2], 3, 2)
    mg = mfgpu.Multigrid(gops, trs, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30)
    mg.setup()
    src = O.fill_uniform(5, oms[-1].n_dofs)
    ref, eig = O.mg_vmult(oms, oops, src, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30)
    for l in range(3):
        lmax, lmin, deg = mg.level_info(l)
        assert abs(lmax - eig[l][0]) < 1e-9 * eig[l][0], (l, lmax, eig[l])
        assert deg == int(eig[l][2])
    dst = torch.empty_like(dev(src))
    mg.vmult(dst, dev(src))
    assert rel(dst, ref) < 1e-9


@pytest.mark.parametrize("coarse_kind", [0, 1])
def test_mg_cg_small_config(cuda, coarse_kind):
    gms, oms, gops, oops, trs = hierarchy("solid", [0, 0, 0], [1, 1, 1], [2, 2, 2], 3, 2)
    fine, A = oms[-1], oops[-1]
    xt = O.fill_uniform(2, fine.n_dofs)
    xt[fine.constrained()] = 0.0
    b = A.vmult(xt)
    xo, it_o, rel_o = O.mg_cg_solve(oms, oops, b, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30,
                                    coarse_kind=coarse_kind, tol_rel=1e-8)
    mg = mfgpu.Multigrid(gops, trs, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30,
                         coarse_kind=coarse_kind, coarse_tolerance=1e-3)
    mg.setup()
    x = torch.zeros(fine.n_dofs, dtype=torch.float64, device="cuda")
    info = mfgpu.cg_solve(gops[-1], dev(b), x, mfgpu.PRECOND_MG, mg, tol_rel=1e-8)
    print(f"coarse_kind={coarse_kind}: iterations gpu {info.iterations} oracle {it_o}")
    assert abs(info.iterations - it_o) <= 1
    assert info.final_residual <= 1e-8 * np.linalg.norm(b)
    assert rel(x, xo) < 1e-8
    assert rel(x, xt) < 1e-6


def test_mg_cg_gradient_storage(cuda):
    """MG-CG with every level in the 10-word tensor2 storage: the oracle's
    iteration count (+-1) and solution, as with the folded storage."""
    gms, oms, gops, oops, trs = hierarchy("solid", [0, 0, 0], [1, 1, 1], [2, 2, 2], 3, 2, storage=mfgpu.T2_GRADIENT)
    assert all(op.n_fields() == 10 for op in gops)
    fine, A = oms[-1], oops[-1]
    xt = O.fill_uniform(2, fine.n_dofs)
    xt[fine.constrained()] = 0.0
    b = A.vmult(xt)
    xo, it_o, rel_o = O.mg_cg_solve(oms, oops, b, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30,
                                    coarse_kind=0, tol_rel=1e-8)
    mg = mfgpu.Multigrid(gops, trs, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=0,
                         coarse_tolerance=1e-3)
    mg.setup()
    x = torch.zeros(fine.n_dofs, dtype=torch.float64, device="cuda")
    info = mfgpu.cg_solve(gops[-1], dev(b), x, mfgpu.PRECOND_MG, mg, tol_rel=1e-8)
    assert abs(info.iterations - it_o) <= 1
    assert rel(x, xo) < 1e-8
    assert rel(x, xt) < 1e-6


def test_jacobi_and_plain_cg(cuda):
    gms, oms, gops, oops, trs = hierarchy("solid", [0, 0, 0], [1, 1, 1], [4, 4, 4], 1, 2)
    A, G = oops[0], gops[0]
    b = A.vmult(O.fill_uniform(3, oms[0].n_dofs))
    G.compute_diagonal()
    for pre in (0, 1):
        xo, it_o, _ = O.cg_solve(A, b, precond=pre, omega=0.65, tol_rel=1e-8)
        x = torch.zeros_like(dev(b))
        info = mfgpu.cg_solve(G, dev(b), x, pre, None, omega=0.65, tol_rel=1e-8)
        assert abs(info.iterations - it_o) <= 1, (pre, info.iterations, it_o)
        assert rel(x, xo) < 1e-7
    with pytest.raises(mfgpu.MfgError) as e:
        mfgpu.cg_solve(G, dev(b), torch.zeros_like(dev(b)), 0, None, tol_rel=1e-12, max_iter=3)
    assert e.value.code == 4


def test_heat_mg_cg(cuda):
    gms, oms, gops, oops, trs = hierarchy("heat", [0, 0, 0], [1, 1, 1], [2, 2, 2], 3, 3, face_mask=63)
    fine, A = oms[-1], oops[-1]
    xt = O.fill_uniform(2, fine.n_dofs)
    xt[fine.constrained()] = 0.0
    b = A.vmult(xt)
    xo, it_o, _ = O.mg_cg_solve(oms, oops, b, cheb_degree=5, smoothing_range=15.0, eig_cg_iterations=10,
                                tol_rel=1e-10)
    mg = mfgpu.Multigrid(gops, trs, cheb_degree=5, smoothing_range=15.0, eig_cg_iterations=10)
    mg.setup()
    x = torch.zeros(fine.n_dofs, dtype=torch.float64, device="cuda")
    info = mfgpu.cg_solve(gops[-1], dev(b), x, mfgpu.PRECOND_MG, mg, tol_rel=1e-10)
    assert abs(info.iterations - it_o) <= 1
    assert rel(x, xt) < 1e-8


def test_mixed_precision_mg_cg(cuda):
    """FP32 V-cycle (levels 1.., caches from FP32 u), FP64 coarse level and
    outer CG: reaches the FP64 residual 1e-8; solution within the solver
    tolerance of the FP64-preconditioned solve."""
    prec = [mfgpu.FP64, mfgpu.FP32, mfgpu.FP32]
    gms, oms, gops, oops, trs = hierarchy("solid", [0, 0, 0], [1, 1, 1], [2, 2, 2], 3, 2, precisions=prec)
    fine_op64 = mfgpu.NeoHookOperator(gms[-1], O.MU, O.NU, O.RHO, O.ALPHA1)
    u = oms[-1].u0(amp=1e-3, noise=1e-5)
    fine_op64.cache(dev(u))
    xt = O.fill_uniform(2, oms[-1].n_dofs)
    xt[oms[-1].constrained()] = 0.0
    b = oops[-1].vmult(xt)
    mg = mfgpu.Multigrid(gops, trs, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=1)
    mg.setup()
    x = torch.zeros(oms[-1].n_dofs, dtype=torch.float64, device="cuda")
    info = mfgpu.cg_solve(fine_op64, dev(b), x, mfgpu.PRECOND_MG, mg, tol_rel=1e-8)
    assert info.final_residual <= 1e-8 * np.linalg.norm(b)
    xo, it_o, _ = O.mg_cg_solve(oms, oops, b, cheb_degree=4, tol_rel=1e-8, coarse_kind=1)
    print(f"mixed: iterations {info.iterations} (fp64 oracle {it_o}), err vs fp64 solve {rel(x, xo):.2e}")
    assert info.iterations <= it_o + 3
    assert rel(x, xo) < 1e-6


@pytest.mark.parametrize("coarse_kind,precisions", [(1, None), (0, None), (1, [mfgpu.FP64, mfgpu.FP32, mfgpu.FP32])])
def test_graphed_vcycle_equals_eager(cuda, coarse_kind, precisions):
    """The V-cycle's sweeps replayed as hipGraphs give bitwise the eager
    results: over several replays with eager work in between, after a new
    linearisation (re-setup and re-capture), and through MG-CG."""
    print(f"graphed V-cycle coarse_kind={coarse_kind} precisions={precisions}", flush=True)
    gms, oms, gops, oops, trs = hierarchy("solid", [0, 0, 0], [2.0, 1.0, 1.0], [2, 1, 1], 3, 3, precisions=precisions)
    eager = mfgpu.Multigrid(gops, trs, cheb_degree=5, coarse_kind=coarse_kind)
    graphed = mfgpu.Multigrid(gops, trs, cheb_degree=5, coarse_kind=coarse_kind)
    eager.use_graph(False)
    graphed.use_graph(True)
    n = gms[-1].n_dofs
    for relinearised in (False, True):
        if relinearised:
            for op, m in zip(gops, gms):  # another linearisation point on every level
                op.cache(dev(m.u0(amp=2e-3, noise=1e-5, seed=9), op.dtype))
        eager.setup()
        graphed.setup()
        for rep in range(4):
            src = mfgpu.fill_uniform(20 + rep, torch.empty(n, dtype=torch.float64, device="cuda"))
            want = torch.empty_like(src)
            got = torch.full_like(src, float("nan"))
            eager.vmult(want, src)
            graphed.vmult(got, src)
            assert torch.equal(got, want), (relinearised, rep)
            torch.cuda.synchronize()
            _ = float(src.sum())  # eager work between replays
        assert graphed.graph_nodes() > 0 and eager.graph_nodes() == 0
    fine = gops[-1] if precisions is None else mfgpu.NeoHookOperator(gms[-1], O.MU, O.NU, O.RHO, O.ALPHA1)
    if precisions is not None:
        fine.cache(dev(gms[-1].u0(amp=2e-3, noise=1e-5, seed=9)))
    b = mfgpu.fill_uniform(30, torch.empty(n, dtype=torch.float64, device="cuda"))
    x1, x2 = torch.zeros_like(b), torch.zeros_like(b)
    i1 = mfgpu.cg_solve(fine, b, x1, mfgpu.PRECOND_MG, eager, tol_rel=1e-8)
    i2 = mfgpu.cg_solve(fine, b, x2, mfgpu.PRECOND_MG, graphed, tol_rel=1e-8)
    assert i1.iterations == i2.iterations and torch.equal(x1, x2)


def test_coarse_cg_results_independent_of_history(cuda):
    """The V-cycle's coarse CG stops on the device and the host reads its
    state after runs of iterations sized by the previous coarse solve: the
    results must not depend on that history.  The same input gives bitwise the
    same output after another input (another coarse residual history) and on a
    fresh hierarchy whose first coarse solve reads back after one iteration."""
    print("coarse CG history independence", flush=True)
    gms, oms, gops, oops, trs = hierarchy("solid", [0, 0, 0], [2.0, 1.0, 1.0], [2, 1, 1], 3, 3)
    n = gms[-1].n_dofs
    x = mfgpu.fill_uniform(40, torch.empty(n, dtype=torch.float64, device="cuda"))
    y = torch.zeros_like(x)
    y[n // 3] = 1.0  # a point load
    mg = mfgpu.Multigrid(gops, trs, cheb_degree=5, coarse_kind=1, coarse_tolerance=1e-6)
    mg.setup()
    outs = []
    for src in (x, y, x, y, x):
        out = torch.full_like(src, float("nan"))
        mg.vmult(out, src)
        outs.append(out)
    fresh = mfgpu.Multigrid(gops, trs, cheb_degree=5, coarse_kind=1, coarse_tolerance=1e-6)
    fresh.setup()
    first = torch.full_like(x, float("nan"))
    fresh.vmult(first, x)
    assert torch.isfinite(outs[0]).all() and torch.isfinite(outs[1]).all()
    assert torch.equal(outs[0], outs[2]) and torch.equal(outs[0], outs[4]) and torch.equal(outs[0], first)
    assert torch.equal(outs[1], outs[3])
