"""Heat time step on the device: quadrature points, right-hand side with
inhomogeneous Dirichlet data (LaplaceProblem::assemble_rhs) and the L2 error
(LaplaceProblem::compute_error).

- quadrature points: closed form on a box (lo + (cell + Gauss point) h);
- right-hand side: the oracle's int (u_old + dt f) phi - A u_D (mass operator,
  load vector and the operator without constraints), relative 1e-13;
- two backward-Euler steps reproduce u = 1 + x^2 + a y^2 + g z^2 + b t exactly
  on Q2 (partitioned_heat.h), nodal and L2 error below 1e-10;
- L2 error of a one-dof perturbation d e_i is |d| sqrt(M_ii).
"""
import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()


@pytest.mark.parametrize("p", [1, 2, 4])
def test_quadrature_points_closed_form(cuda, p):
    lo, hi, n = np.array([0.5, -1.0, 2.0]), np.array([1.5, 1.0, 2.6]), [3, 2, 4]
    m = mfgpu.BoxMesh(lo, hi, n, p, 1, 63, 1, numbering="cellwise")
    op = mfgpu.LaplaceOperator(m, 0.1)
    x = op.quadrature_points().cpu().numpy()
    gauss = (np.polynomial.legendre.leggauss(p + 1)[0] + 1) / 2
    h = (hi - lo) / n
    N = p + 1
    for c in (0, 5, m.n_cells - 1):
        cc = [c % n[0], (c // n[0]) % n[1], c // (n[0] * n[1])]
        for q in (0, 7 % N ** 3, N ** 3 - 1):
            ijk = [q % N, (q // N) % N, q // (N * N)]
            want = [lo[d] + (cc[d] + gauss[ijk[d]]) * h[d] for d in range(3)]
            assert np.abs(x[c * N ** 3 + q] - want).max() < 1e-14


@pytest.mark.parametrize("p", [1, 3])
def test_rhs_matches_oracle(cuda, p):
    lo, hi, n, dt, f = [0, 0, 0], [1.0, 0.8, 1.2], [3, 2, 3], 0.05, -1.7
    m = mfgpu.BoxMesh(lo, hi, n, p, 1, 63, 1)
    om = O.Mesh(lo, hi, n, p, 1, 63, 1)
    coef = np.random.default_rng(p).uniform(0.5, 2.0, om.n_cells * (p + 1) ** 3)
    op = mfgpu.LaplaceOperator(m, dt, coef)
    A_full = O.Heat(om, dt=dt, coef=coef)
    M = O.Heat(om, dt=0.0)
    con = om.constrained()
    u_old = O.fill_uniform(3, om.n_dofs)
    u_d = np.zeros(om.n_dofs)
    u_d[con] = O.fill_uniform(4, om.n_dofs)[con]
    ref = M.vmult_noconstraints(u_old) + dt * M.load(f) - A_full.vmult_noconstraints(u_d)
    ref[con] = 0.0
    rhs = torch.full((om.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
    fq = torch.full((op.n_qp,), f, dtype=torch.float64, device="cuda")
    op.assemble_rhs(rhs, dev(u_old), dev(u_d), fq)
    got = rhs.cpu().numpy()
    assert np.all(got[con] == 0.0)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def test_backward_euler_reproduces_quadratic_solution(cuda):
    a, b, g, dt, p = 3.0, 1.3, 0.7, 0.1, 2
    m = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [3, 3, 3], p, 1, 63, 1, numbering="cellwise")
    op = mfgpu.LaplaceOperator(m, dt)
    op.compute_diagonal()
    X = m.node_coords()
    xq = op.quadrature_points().cpu().numpy()
    con = torch.from_numpy(m.constrained().astype(np.int64)).cuda()

    def exact(x, t):
        return 1 + x[:, 0] ** 2 + a * x[:, 1] ** 2 + g * x[:, 2] ** 2 + b * t

    fq = torch.full((op.n_qp,), b - 2 - 2 * a - 2 * g, dtype=torch.float64, device="cuda")
    u = dev(exact(X, 0.0))
    rhs = torch.empty_like(u)
    for step in range(1, 3):
        t = step * dt
        u_d = torch.zeros_like(u)
        u_d[con] = dev(exact(X, t))[con]
        op.assemble_rhs(rhs, u, u_d, fq)
        w = torch.zeros_like(u)
        info = mfgpu.cg_solve(op, rhs, w, mfgpu.PRECOND_JACOBI, omega=1.0, tol_rel=1e-13, max_iter=2000)
        assert info.iterations > 0
        u = w + u_d
        nodal = float((u - dev(exact(X, t))).abs().max())
        l2 = op.l2_error(u, dev(exact(xq, t)))
        print(f"step {step}: {info.iterations} CG iterations, nodal error {nodal:.1e}, L2 error {l2:.1e}")
        assert nodal < 1e-10 and l2 < 1e-10


def test_l2_error_closed_forms(cuda):
    p, lo, hi, n = 3, [0, 0, 0], [1.0, 0.5, 0.7], [2, 3, 2]
    m = mfgpu.BoxMesh(lo, hi, n, p, 1, 0, 1)
    om = O.Mesh(lo, hi, n, p, 1, 0, 1)
    op = mfgpu.LaplaceOperator(m, 0.1)
    xq = op.quadrature_points().cpu().numpy()
    X = m.node_coords()
    lin = lambda x: 0.5 + 2 * x[:, 0] - x[:, 1] + 3 * x[:, 2]
    assert op.l2_error(dev(lin(X)), dev(lin(xq))) < 1e-13  # the element space holds it
    zero = torch.zeros(op.n_qp, dtype=torch.float64, device="cuda")
    mass_diag = O.Heat(om, dt=0.0).diagonal()
    for i, d in ((0, 0.3), (om.n_dofs // 2, -1.1), (om.n_dofs - 1, 2e-3)):
        e = np.zeros(om.n_dofs)
        e[i] = d
        want = abs(d) * np.sqrt(mass_diag[i])  # || d phi_i ||_L2
        assert abs(op.l2_error(dev(e), zero) - want) <= 1e-13 * want
