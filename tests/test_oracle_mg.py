"""Pins of the oracle's multigrid pieces: transfer by polynomial reproduction
and exact transposition, Chebyshev eigenvalue estimates against a dense
eigen-solve, and MG-CG by recovering a known solution."""
import numpy as np
import pytest

import oracle as O


def poly(X, p, comp=0):
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    return (1 + comp) * (x ** p - 0.5 * y ** p * x + 0.3 * z ** p + x * y * z)


@pytest.mark.parametrize("p,nc", [(1, 1), (2, 3), (3, 1), (4, 3)])
def test_prolongation_reproduces_degree_p_polynomials(p, nc):
    cm = O.Mesh([0, 0, 0], [2.0, 1.0, 1.0], [2, 1, 1], p, nc, face_mask=0)
    fm = O.Mesh([0, 0, 0], [2.0, 1.0, 1.0], [4, 2, 2], p, nc, face_mask=0)
    Xc, Xf = cm.node_coords(), fm.node_coords()
    c = np.zeros(cm.n_dofs)
    f_exact = np.zeros(fm.n_dofs)
    for k in range(nc):
        c[k::nc] = poly(Xc, p, k)
        f_exact[k::nc] = poly(Xf, p, k)
    f = np.zeros(fm.n_dofs)
    O.prolongate_add(cm, fm, c, f)
    assert np.abs(f - f_exact).max() < 1e-12
    # nodal interpolation back to the coarse level is exact as well
    assert np.abs(O.interpolate(cm, fm, f_exact) - c).max() < 1e-12


@pytest.mark.parametrize("fm_mask", [0, 1, 63])
def test_restriction_is_transpose_of_prolongation(fm_mask):
    p, nc = 3, 3
    cm = O.Mesh([0, 0, 0], [1, 1, 1], [2, 2, 1], p, nc, face_mask=fm_mask)
    fm = O.Mesh([0, 0, 0], [1, 1, 1], [4, 4, 2], p, nc, face_mask=fm_mask)
    rng = np.random.default_rng(1)
    c = rng.uniform(-1, 1, cm.n_dofs)
    f = rng.uniform(-1, 1, fm.n_dofs)
    Pc = np.zeros(fm.n_dofs)
    O.prolongate_add(cm, fm, c, Pc)
    Rf = np.zeros(cm.n_dofs)
    O.restrict_add(cm, fm, f, Rf)
    assert abs(f @ Pc - c @ Rf) < 1e-12 * abs(f @ Pc)
    # constrained entries: prolongation leaves fine ones untouched, restriction coarse ones
    assert np.all(Pc[fm.constrained()] == 0)
    assert np.all(Rf[cm.constrained()] == 0)


def test_lanczos_eigenvalue_estimate_matches_dense_eigensolve():
    m = O.Mesh([0, 0, 0], [1, 1, 1], [2, 1, 1], 2, 3, face_mask=1)
    s = O.Solid(m, caching=1)
    s.cache(m.u0(amp=1e-2))
    n = m.n_dofs
    A = np.stack([s.vmult(e) for e in np.eye(n)], axis=1)
    d = np.diag(A)
    ev = np.linalg.eigvals(A / d[:, None]).real
    lmax, lmin = O.cheb_eigenvalues(s, n)
    # Ritz values bound the spectrum from inside and converge to its ends
    assert lmax <= ev.max() * (1 + 1e-12) and abs(lmax - ev.max()) < 1e-5 * ev.max()
    assert lmin >= ev.min() * (1 - 1e-12) and abs(lmin - ev.min()) < 1e-5 * ev.max()
    # few iterations give a lower bound of the largest eigenvalue
    l10, _ = O.cheb_eigenvalues(s, 10)
    assert l10 <= ev.max() * (1 + 1e-12) and l10 > 0.5 * ev.max()


@pytest.mark.parametrize("kind", ["solid", "heat"])
def test_mg_cg_recovers_known_solution(kind):
    if kind == "solid":
        meshes = O.box_levels([0, 0, 0], [1, 1, 1], [1, 1, 1], 3, 2, 3)
        ops = [O.Solid(m, caching=1) for m in meshes]
        for m, op in zip(meshes, ops):
            op.cache(m.u0(amp=1e-3, noise=0.0))
        prm = dict(cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30)
    else:
        meshes = O.box_levels([0, 0, 0], [1, 1, 1], [1, 1, 1], 3, 2, 1, face_mask=63, comp_mask=1)
        ops = [O.Heat(m, dt=0.1) for m in meshes]
        prm = dict(cheb_degree=5, smoothing_range=15.0, eig_cg_iterations=10)
    fine, A = meshes[-1], ops[-1]
    xt = O.fill_uniform(2, fine.n_dofs)
    xt[fine.constrained()] = 0.0
    b = A.vmult(xt)
    x, it, rel = O.mg_cg_solve(meshes, ops, b, tol_rel=1e-10, **prm)
    assert it > 0 and it < 40 and rel <= 1e-10
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-6
    # far fewer iterations than Jacobi-CG
    _, it_j, _ = O.cg_solve(A, b, precond=1, tol_rel=1e-10)
    assert it_j > 2 * it
