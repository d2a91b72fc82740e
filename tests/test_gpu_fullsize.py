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


def test_fullsize_mixed_mgcg_recovers_solution(cuda, beam):
    """FP32 V-cycle (FP64 coarse solve and outer CG) on the 6-level hierarchy."""
    mesh, fine, _ = beam
    L = 6
    meshes = [mfgpu.BoxMesh(LO, HI, [c >> (L - 1 - l) for c in N], P, 3, 1, 7, numbering="cellwise")
              for l in range(L)]
    levels = [mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1,
                                    precision=mfgpu.FP64 if l == 0 else mfgpu.FP32) for l, m in enumerate(meshes)]
    trs = [mfgpu.Transfer(levels[l - 1], levels[l], mfgpu.box_children(meshes[l - 1], meshes[l])) for l in range(1, L)]
    u = [None] * L
    u[-1] = torch.from_numpy(meshes[-1].u0(amp=1e-3, noise=1e-5, seed=0)).cuda().float()
    for l in range(L - 1, 0, -1):
        u[l - 1] = torch.empty(meshes[l - 1].n_dofs, dtype=levels[l - 1].dtype, device="cuda")
        trs[l - 1].interpolate(u[l - 1], u[l])
    for op, uu in zip(levels, u):
        op.cache(uu)
    mg = mfgpu.Multigrid(levels, trs)
    mg.setup()
    xt = mfgpu.fill_uniform(2, torch.empty(mesh.n_dofs, dtype=torch.float64, device="cuda"))
    xt[torch.from_numpy(mesh.constrained().astype(np.int64)).cuda()] = 0.0
    b = torch.empty_like(xt)
    fine.vmult(b, xt)
    x = torch.zeros_like(b)
    info = mfgpu.cg_solve(fine, b, x, mfgpu.PRECOND_MG, mg, tol_rel=1e-8)
    err = float((x - xt).norm() / xt.norm())
    print(f"full-size mixed MG-CG: {info.iterations} iterations, error to x_true {err:.2e}", flush=True)
    assert info.iterations <= 10
    assert err <= 1e-6
