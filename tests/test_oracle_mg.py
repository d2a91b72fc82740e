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


# ---------------------------------------------------------------------------
# Chebyshev smoother and V-cycle composition, pinned by closed forms.
# deal.II's PreconditionChebyshev (configured at mf_elasticity.h:1149-1180,
# 1209-1216) applies the Chebyshev polynomial of the first kind on the
# interval [theta - delta, theta + delta] to D^-1 A: after a degree-k
# application the error of every eigenvector of D^-1 A with eigenvalue lam is
# multiplied by T_k((theta - lam)/delta) / T_k(theta/delta).

def cheb_T(k, x):
    return np.polynomial.chebyshev.chebval(x, [0.0] * k + [1.0])


def cheb_residual_factor(k, theta, delta, lam):
    return cheb_T(k, (theta - lam) / delta) / cheb_T(k, theta / delta)


def spd_matrix(n, seed, cond=200.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.geomspace(1.0, cond, n)
    A = (Q * lam) @ Q.T
    s = rng.uniform(0.5, 3.0, n)  # uneven diagonal, so D^-1 A differs from A
    return s[:, None] * A * s[None, :]


def eig_DA(A):
    """Eigenvalues and eigenvectors of D^-1 A, D = diag(A), through the
    symmetric similar matrix D^-1/2 A D^-1/2."""
    d = np.diag(A)
    M = A / np.sqrt(d)[:, None] / np.sqrt(d)[None, :]
    lam, Q = np.linalg.eigh(M)
    return lam, Q / np.sqrt(d)[:, None]


@pytest.mark.parametrize("degree", [2, 4, 5])
def test_chebyshev_error_propagator_is_the_closed_form_on_every_eigenvector(degree):
    n, rng_range = 14, 20.0
    A = spd_matrix(n, seed=degree)
    lam, V = eig_DA(A)
    _, info = O.cheb_dense(A, np.zeros(n), degree, rng_range, n)
    # the interval: [1.2 lmax / range, 1.2 lmax], lmax the Lanczos estimate
    # (n steps: exact to rounding on this spectrum)
    assert abs(info["lmax"] - lam.max()) < 1e-9 * lam.max()
    max_ev = 1.2 * info["lmax"]
    assert abs(info["theta"] - 0.5 * (max_ev + max_ev / rng_range)) < 1e-12 * max_ev
    assert abs(info["delta"] - 0.5 * (max_ev - max_ev / rng_range)) < 1e-12 * max_ev
    assert info["degree"] == degree
    th, de = info["theta"], info["delta"]
    for i in range(n):
        v = V[:, i]
        x, _ = O.cheb_dense(A, A @ v, degree, rng_range, n)  # zero initial guess
        expect = cheb_residual_factor(degree, th, de, lam[i])
        assert np.linalg.norm((v - x) - expect * v) < 1e-11 * np.linalg.norm(v), (i, lam[i], expect)
    # step from a non-zero guess: error e0 = x_true - x0 maps to r(D^-1 A) e0
    rng = np.random.default_rng(7)
    xt, x0 = rng.standard_normal(n), rng.standard_normal(n)
    x1, _ = O.cheb_dense(A, A @ xt, degree, rng_range, n, x0=x0)
    c = np.linalg.solve(V, xt - x0)
    e1 = V @ (cheb_residual_factor(degree, th, de, lam) * c)
    assert np.linalg.norm((xt - x1) - e1) < 1e-11 * np.linalg.norm(xt - x0)


def test_chebyshev_solver_mode_degree_meets_the_reduction_bound():
    """Level 0 as a solver (mf_elasticity.h:1209-1216): range 1e-3 is the
    target reduction and the degree is chosen from the eigenvalue estimate.
    With kappa = 1.2 lmax / lmin the bound max |T_k((theta - lam)/delta)| /
    T_k(theta/delta) = 1/T_k(theta/delta) must fall below the target at the
    chosen k and not at k - 1 (the least such degree)."""
    eps = 1e-3
    for seed, cond in [(1, 20.0), (2, 50.0), (3, 120.0)]:
        n = 12
        A = spd_matrix(n, seed=seed, cond=cond)
        lam, V = eig_DA(A)
        _, info = O.cheb_dense(A, np.zeros(n), 0, eps, n)
        # Ritz values lie inside the spectrum
        assert lam.min() * (1 - 1e-12) <= info["lmin"] and info["lmax"] <= lam.max() * (1 + 1e-12)
        max_ev, alpha = 1.2 * info["lmax"], min(0.9 * info["lmax"], info["lmin"])
        th, de, k = info["theta"], info["delta"], info["degree"]
        assert abs(th - 0.5 * (max_ev + alpha)) < 1e-12 * max_ev and abs(de - 0.5 * (max_ev - alpha)) < 1e-12 * max_ev
        assert cheb_T(k, th / de) > 1.0 / eps >= cheb_T(k - 1, th / de), (k, cond)
        # inside the interval the error propagator is the closed form and
        # reduces every eigencomponent at least to the target (outside it the
        # polynomial grows and rounding of the degree-k recurrence dominates)
        inside = [i for i in range(n) if alpha <= lam[i] <= max_ev]
        assert len(inside) >= n - 1
        for i in inside:
            v = V[:, i]
            x, _ = O.cheb_dense(A, A @ v, 0, eps, n)
            r = cheb_residual_factor(k, th, de, lam[i])
            assert abs(r) <= eps * (1 + 1e-9)
            assert np.linalg.norm((v - x) - r * v) < 1e-10 * np.linalg.norm(v), (k, i)


def _dense_levels(meshes, ops):
    mats, P = [], []
    for m, op in zip(meshes, ops):
        mats.append(np.stack([op.vmult(e) for e in np.eye(m.n_dofs)], axis=1))
    for l in range(1, len(meshes)):
        cols = []
        for e in np.eye(meshes[l - 1].n_dofs):
            f = np.zeros(meshes[l].n_dofs)
            O.prolongate_add(meshes[l - 1], meshes[l], e, f)
            cols.append(f)
        P.append(np.stack(cols, axis=1))
    return mats, P


def _cheb_matrix(A, k, theta, delta):
    """S = p(D^-1 A) D^-1 with 1 - lam p(lam) = T_k((theta-lam)/delta)/T_k(theta/delta),
    built from the eigen-decomposition of D^-1/2 A D^-1/2 (symmetric)."""
    d = np.diag(A)
    M = A / np.sqrt(d)[:, None] / np.sqrt(d)[None, :]
    lam, Q = np.linalg.eigh(M)
    p = (1.0 - cheb_residual_factor(k, theta, delta, lam)) / lam
    return (Q * p) @ Q.T / np.sqrt(d)[:, None] / np.sqrt(d)[None, :]


def _vcycle_hierarchy():
    meshes = O.box_levels([0, 0, 0], [1, 1, 1], [1, 1, 1], 3, 2, 3)
    ops = [O.Solid(m, caching=1) for m in meshes]
    for m, op in zip(meshes, ops):
        op.cache(m.u0(amp=1e-2, noise=0.0))
    return meshes, ops


@pytest.mark.parametrize("degree", [4, 5])
def test_vcycle_equals_independent_dense_composition(degree):
    """The oracle's V-cycle (Chebyshev coarse solver) against one composed here
    from dense level matrices, dense prolongation P (restriction P^T) and the
    closed-form smoother polynomials: pre-smoothing from a zero guess, residual,
    restriction, coarse solve, prolongation, post-smoothing step."""
    meshes, ops = _vcycle_hierarchy()
    mats, P = _dense_levels(meshes, ops)
    b = O.fill_uniform(5, meshes[-1].n_dofs)
    x_orc, eig = O.mg_vmult(meshes, ops, b, cheb_degree=degree, smoothing_range=20.0, eig_cg_iterations=30,
                            coarse_kind=0)
    S = []
    for l, (lmax, lmin, k) in enumerate(eig):
        max_ev = 1.2 * lmax
        alpha = min(0.9 * lmax, lmin) if l == 0 else max_ev / 20.0
        assert int(k) == (degree if l > 0 else k)
        S.append(_cheb_matrix(mats[l], int(k), 0.5 * (max_ev + alpha), 0.5 * (max_ev - alpha)))

    def vcycle(l, rhs):
        if l == 0:
            return S[0] @ rhs
        x = S[l] @ rhs
        xc = vcycle(l - 1, P[l - 1].T @ (rhs - mats[l] @ x))
        x = x + P[l - 1] @ xc
        return x + S[l] @ (rhs - mats[l] @ x)

    x_np = vcycle(len(meshes) - 1, b)
    assert np.linalg.norm(x_orc - x_np) < 1e-11 * np.linalg.norm(x_np)


def test_vcycle_with_chebyshev_coarse_solver_is_symmetric():
    meshes, ops = _vcycle_hierarchy()
    n = meshes[-1].n_dofs
    x, y = O.fill_uniform(11, n), O.fill_uniform(12, n)
    Bx, _ = O.mg_vmult(meshes, ops, x, cheb_degree=5, coarse_kind=0)
    By, _ = O.mg_vmult(meshes, ops, y, cheb_degree=5, coarse_kind=0)
    assert abs(x @ By - y @ Bx) < 1e-13 * np.linalg.norm(x) * np.linalg.norm(By)
    # and positive definite on these vectors
    assert x @ Bx > 0 and y @ By > 0


PINS = ["test_chebyshev_error_propagator_is_the_closed_form_on_every_eigenvector",
        "test_chebyshev_solver_mode_degree_meets_the_reduction_bound",
        "test_vcycle_equals_independent_dense_composition",
        "test_vcycle_with_chebyshev_coarse_solver_is_symmetric"]


@pytest.mark.parametrize("plant,caught_by", [
    ("ORC_PLANT_CHEB_F2", ["test_chebyshev_error_propagator_is_the_closed_form_on_every_eigenvector",
                           "test_chebyshev_solver_mode_degree_meets_the_reduction_bound",
                           "test_vcycle_equals_independent_dense_composition"]),
    ("ORC_PLANT_POST_ZERO_GUESS", ["test_vcycle_equals_independent_dense_composition"]),
    ("ORC_PLANT_NO_POST_SMOOTH", ["test_vcycle_equals_independent_dense_composition",
                                  "test_vcycle_with_chebyshev_coarse_solver_is_symmetric"]),
])
def test_pins_fail_on_a_planted_mistake(plant, caught_by, tmp_path):
    """Builds a copy of the oracle with a planted mistake in the smoother
    recurrence (half the coefficient of the preconditioned residual) or in the
    V-cycle composition (post-smoothing from a zero guess, which drops the
    coarse correction; no post-smoothing, which breaks symmetry) and checks
    that the pins above fail on it."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    odir = os.path.join(here, "..", "oracle")
    lib = str(tmp_path / "liboracle_planted.so")
    r = subprocess.run(["make", "-s", "-C", odir, "-B", f"LIB={lib}", f"PLANT=-D{plant}", lib], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, MFG_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-rf",
                        os.path.join(here, "test_oracle_mg.py"), "-k", " or ".join(PINS)],
                       capture_output=True, text=True, env=env, cwd=here)
    assert r.returncode != 0, r.stdout
    failed = {line.split("::")[1].split("[")[0].split(" ")[0] for line in r.stdout.splitlines()
              if line.startswith("FAILED")}
    for name in caught_by:
        assert name in failed, (name, r.stdout[-3000:])
