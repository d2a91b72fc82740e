"""The benchmark configuration itself: the Q4 beam of 128 x 32 x 32 cells
(25,610,499 DoFs), cellwise node numbering, FP64 tensor2 tangent linearised at
1e-3 sin sin sin + 1e-5 noise.

- vmult against the oracle (lexicographic mesh, permuted), relative l2 <= 1e-12,
  for both tensor2 point storages (16 folded words and the 10-word gradient
  storage that the benchmark runs);
- size-independent properties: symmetry, linearity, identity rows of the
  Dirichlet dofs, diagonal entries equal to unit probes;
- mixed-precision MG-CG recovers a manufactured solution.
"""
import time

import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu

LO, HI, N, P = [0, 0, 0], [4.0, 1.0, 1.0], [128, 32, 32], 4


@pytest.fixture(scope="module")
def beam():
    t0 = time.perf_counter()
    lex = mfgpu.BoxMesh(LO, HI, N, P, 3, 1, 7)
    mesh = mfgpu.BoxMesh(LO, HI, N, P, 3, 1, 7, numbering="cellwise")
    perm = np.full(lex.n_nodes, -1, dtype=np.int64)
    perm[lex.cell_nodes()] = mesh.cell_nodes()
    dof = torch.from_numpy((perm[:, None] * 3 + np.arange(3)[None, :]).ravel()).cuda()  # cellwise dof of lex dof
    del lex
    op = mfgpu.NeoHookOperator(mesh, O.MU, O.NU, O.RHO, O.ALPHA1)
    op.cache(torch.from_numpy(mesh.u0(amp=1e-3, noise=1e-5, seed=0)).cuda())
    print(f"full-size beam set up in {time.perf_counter() - t0:.1f} s", flush=True)
    return mesh, op, dof


def test_fullsize_vmult_matches_oracle(cuda, beam):
    mesh, op, dof = beam
    t0 = time.perf_counter()
    om = O.Mesh(LO, HI, N, P, 3, 1, 7)
    oo = O.Solid(om, caching=1)
    oo.cache(om.u0(amp=1e-3, noise=1e-5, seed=0))
    x = O.fill_uniform(1, om.n_dofs)
    ref = oo.vmult(x)
    print(f"oracle set-up and vmult: {time.perf_counter() - t0:.1f} s", flush=True)
    xg = torch.empty(mesh.n_dofs, dtype=torch.float64, device="cuda")
    xg[dof] = torch.from_numpy(x).cuda()
    y = torch.empty_like(xg)
    op.vmult(y, xg)
    got = y[dof].cpu().numpy()
    r = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"full-size vmult relative difference to the oracle: {r:.2e}", flush=True)
    assert r <= 1e-12
    # the benchmark's operator: the same linearisation with the 10-word storage
    g = mfgpu.NeoHookOperator(mesh, O.MU, O.NU, O.RHO, O.ALPHA1)
    g.set_point_storage(mfgpu.T2_GRADIENT)
    g.cache(torch.from_numpy(mesh.u0(amp=1e-3, noise=1e-5, seed=0)).cuda())
    assert g.n_fields() == 10
    yg = torch.empty_like(xg)
    g.vmult(yg, xg)
    rg = np.linalg.norm(yg[dof].cpu().numpy() - ref) / np.linalg.norm(ref)
    rf = float((yg - y).norm() / y.norm())
    print(f"gradient storage: relative difference to the oracle {rg:.2e}, to the folded storage {rf:.2e}", flush=True)
    assert rg <= 1e-12 and rf <= 1e-14


def test_fullsize_symmetry_linearity_constraints(cuda, beam):
    mesh, op, _ = beam
    a = mfgpu.fill_uniform(2, torch.empty(mesh.n_dofs, dtype=torch.float64, device="cuda"))
    b = mfgpu.fill_uniform(3, torch.empty_like(a))
    con = torch.from_numpy(mesh.constrained().astype(np.int64)).cuda()
    a[con] = 0.0
    b[con] = 0.0
    Aa, Ab = torch.empty_like(a), torch.empty_like(a)
    op.vmult(Aa, a)
    op.vmult(Ab, b)
    ab, ba = float(a @ Ab), float(b @ Aa)
    assert abs(ab - ba) <= 1e-12 * abs(ab)
    c = 0.37 * a + b
    Ac = torch.empty_like(a)
    op.vmult(Ac, c)
    assert float((Ac - (0.37 * Aa + Ab)).norm()) <= 1e-13 * float(Ac.norm())
    # rows of Dirichlet dofs are the identity
    x = mfgpu.fill_uniform(4, torch.empty_like(a))
    y = torch.empty_like(a)
    op.vmult(y, x)
    assert torch.equal(y[con], x[con])


def test_fullsize_diagonal_equals_unit_probes(cuda, beam):
    mesh, op, _ = beam
    op.compute_diagonal()
    d = op.diagonal()
    e = torch.zeros(mesh.n_dofs, dtype=torch.float64, device="cuda")
    y = torch.empty_like(e)
    con = set(mesh.constrained().tolist())
    rng = np.random.default_rng(5)
    for i in rng.integers(0, mesh.n_dofs, 6).tolist() + [mesh.n_dofs // 2]:
        e.zero_()
        e[i] = 1.0
        op.vmult(y, e)
        want = 1.0 if i in con else float(y[i])
        assert abs(float(d[i]) - want) <= 1e-12 * abs(want), i


def _mg_solve(fine, mesh, level_precision, coarse_kind):
    """MG-CG on the 6-level beam hierarchy (levels >= 1 in level_precision, the
    coarse level FP64), rhs = A x_true (seed 2, Dirichlet entries zero)."""
    L = 6
    meshes = [mfgpu.BoxMesh(LO, HI, [c >> (L - 1 - l) for c in N], P, 3, 1, 7, numbering="cellwise")
              for l in range(L - 1)] + [mesh]
    levels = [mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1,
                                    precision=mfgpu.FP64 if l == 0 else level_precision)
              for l, m in enumerate(meshes[:-1])]
    levels.append(fine if level_precision == mfgpu.FP64 else
                  mfgpu.NeoHookOperator(mesh, O.MU, O.NU, O.RHO, O.ALPHA1, precision=mfgpu.FP32))
    trs = [mfgpu.Transfer(levels[l - 1], levels[l], mfgpu.box_children(meshes[l - 1], meshes[l])) for l in range(1, L)]
    u = [None] * L
    u[-1] = torch.from_numpy(meshes[-1].u0(amp=1e-3, noise=1e-5, seed=0)).cuda().to(levels[-1].dtype)
    for l in range(L - 1, 0, -1):
        u[l - 1] = torch.empty(meshes[l - 1].n_dofs, dtype=levels[l - 1].dtype, device="cuda")
        trs[l - 1].interpolate(u[l - 1], u[l])
    for op, uu in zip(levels, u):
        if op is not fine:
            op.cache(uu)
    mg = mfgpu.Multigrid(levels, trs, coarse_kind=coarse_kind)
    mg.setup()
    xt = mfgpu.fill_uniform(2, torch.empty(mesh.n_dofs, dtype=torch.float64, device="cuda"))
    xt[torch.from_numpy(mesh.constrained().astype(np.int64)).cuda()] = 0.0
    b = torch.empty_like(xt)
    fine.vmult(b, xt)
    x = torch.zeros_like(b)
    info = mfgpu.cg_solve(fine, b, x, mfgpu.PRECOND_MG, mg, tol_rel=1e-8)
    r = torch.empty_like(b)
    fine.vmult(r, x)
    rel = float((r - b).norm() / b.norm())
    err = float((x - xt).norm() / xt.norm())
    return x, info, rel, err


def test_fullsize_mixed_mgcg_recovers_solution(cuda, beam):
    """FP32 V-cycle (FP64 coarse solve and outer CG) on the 6-level hierarchy."""
    mesh, fine, _ = beam
    x, info, rel, err = _mg_solve(fine, mesh, mfgpu.FP32, 0)
    print(f"full-size mixed MG-CG: {info.iterations} iterations, error to x_true {err:.2e}", flush=True)
    assert info.iterations <= 10 and rel <= 1e-8
    assert err <= 1e-6


def test_fullsize_fp64_and_mixed_mgcg_with_coarse_cg(cuda, beam):
    """The MG-CG of the benchmark's two legs (coarse FP64 CG to 1e-3): plain
    FP64 levels and FP32 levels both reach the FP64 residual 1e-8 (checked by
    an independent operator application), within one iteration of each other,
    and their solutions agree far below the solver tolerance."""
    mesh, fine, _ = beam
    x64, i64, rel64, err64 = _mg_solve(fine, mesh, mfgpu.FP64, 1)
    x32, i32, rel32, err32 = _mg_solve(fine, mesh, mfgpu.FP32, 1)
    d = float((x32 - x64).norm() / x64.norm())
    print(f"full-size MG-CG, coarse CG: FP64 {i64.iterations} its (residual {rel64:.2e}, error {err64:.2e}); "
          f"mixed {i32.iterations} its (residual {rel32:.2e}); mixed vs FP64 {d:.2e}", flush=True)
    assert rel64 <= 1e-8 and rel32 <= 1e-8
    assert abs(i64.iterations - i32.iterations) <= 1
    assert err64 <= 1e-6 and d <= 1e-8


def test_fullsize_newton_step_mixed_equals_fp64(cuda, beam):
    """One Newmark step of the benchmark's Newton leg on the full beam (6
    levels, body force (0, 0, -9.81)), FP64 levels against FP32 levels: both
    converge with the same Newton count and the same displacement to well
    below the Newton tolerance."""
    mesh = beam[0]
    L = 6
    meshes = [mfgpu.BoxMesh(LO, HI, [c >> (L - 1 - l) for c in N], P, 3, 1, 7, numbering="cellwise")
              for l in range(L - 1)] + [mesh]
    u0 = torch.from_numpy(mesh.u0(amp=1e-3, noise=1e-5, seed=0)).cuda()
    out = {}
    for name, prec in (("fp64", mfgpu.FP64), ("mixed", mfgpu.FP32)):
        prob = mfgpu.SolidProblem(meshes, O.MU, O.NU, O.RHO, dt=1e-3, body=(0.0, 0.0, -9.81), level_precision=prec,
                                  tol_f=1e-8, tol_lin=1e-8, coarse_kind=1)
        prob.set_state(u=u0)
        info = prob.step()
        out[name] = (info, prob.get_state()[0])
        print(f"full-size Newton step {name}: {info.newton_iterations} Newton, {info.linear_iterations} CG, "
              f"{info.seconds:.3f} s, residual reduction {info.residual_final / info.residual_0:.2e}", flush=True)
        del prob
        torch.cuda.empty_cache()
    (i64, u64), (i32, u32) = out["fp64"], out["mixed"]
    assert i64.converged and i32.converged
    assert i64.newton_iterations == i32.newton_iterations
    assert i32.residual_final <= 1e-8 * i32.residual_0
    d = float((u32 - u64).norm() / u64.norm())
    print(f"mixed vs FP64 displacement: {d:.2e}", flush=True)
    assert d <= 1e-10
