"""Pins of the oracle's solid and heat operators.

Every check here is independent of the oracle's own sum-factorised code path:
dense quadrature assembly, closed forms, the finite-difference derivative of
the residual, an energy computed in numpy, and linear elasticity at u = 0.
"""
import numpy as np
import pytest

import oracle as O


def unit_mesh(n, p, nc, face_mask=1):
    return O.Mesh([0, 0, 0], [1, 1, 1], [n, n, n], p, nc, face_mask, 7)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_solid_sumfact_matches_dense_quadrature(p):
    m = O.Mesh([0, 0, 0], [1.0, 0.7, 0.5], [2, 2, 1], p, 3)
    u = m.u0(amp=5e-2, noise=1e-3)
    s = O.Solid(m, caching=1)
    s.cache(u)
    x = O.fill_uniform(1, m.n_dofs)
    y, yb = s.vmult(x), s.vmult_bruteforce(x)
    assert np.linalg.norm(y - yb) / np.linalg.norm(yb) < 1e-13


def test_cachings_define_the_same_operator():
    m = unit_mesh(2, 2, 3)
    u = m.u0(amp=1e-1, noise=1e-3)
    x = O.fill_uniform(7, m.n_dofs)
    ys = []
    for caching in (0, 1, 2, 3):
        s = O.Solid(m, caching=caching)
        s.cache(u)
        ys.append(s.vmult(x))
    for y in ys[1:]:
        assert np.linalg.norm(y - ys[0]) / np.linalg.norm(ys[0]) < 1e-13


def distorted(m, eps=0.08):
    """The mesh with its vertices moved by a smooth map: cells are no longer affine."""
    v = m.cell_vertices().reshape(-1, 3)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    w = v + eps * np.stack([np.sin(np.pi * y) * z, np.sin(np.pi * x) * z, x * y], axis=1)
    m.set_cell_vertices(w.ravel())
    return m


def test_distorted_mesh_nodes_follow_the_vertices():
    m = distorted(O.Mesh([0, 0, 0], [1, 1, 1], [2, 2, 2], 2, 3))
    cn = m.cell_nodes().reshape(-1, 27)
    x = m.node_coords()
    v = m.cell_vertices().reshape(-1, 8, 3)
    corner = [0, 2, 6, 8, 18, 20, 24, 26]  # local index of vertex i + 2j + 4k
    assert np.abs(x[cn[:, corner]] - v).max() == 0.0
    # an edge midpoint is the mean of its two vertices (trilinear map restricted to an edge)
    assert np.abs(x[cn[:, 1]] - 0.5 * (v[:, 0] + v[:, 1])).max() < 1e-15


@pytest.mark.parametrize("p", [1, 2, 3])
def test_distorted_mesh_sumfact_and_cachings(p):
    """Non-affine cells: sum factorisation equals the dense quadrature, and the
    four cachings (reference-frame geometry from the vertices, deformed-frame
    geometry from the displaced support points) define the same operator."""
    m = distorted(O.Mesh([0, 0, 0], [1.0, 0.7, 0.5], [2, 2, 1], p, 3))
    u = m.u0(amp=5e-2, noise=1e-3)
    x = O.fill_uniform(1, m.n_dofs)
    ys = []
    for caching in (0, 1, 2, 3):
        s = O.Solid(m, caching=caching)
        s.cache(u)
        ys.append(s.vmult(x))
        if caching == 1:
            yb = s.vmult_bruteforce(x)
            assert np.linalg.norm(ys[-1] - yb) / np.linalg.norm(yb) < 1e-13
    for y in ys[1:]:
        assert np.linalg.norm(y - ys[0]) / np.linalg.norm(ys[0]) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 3])
def test_distorted_mesh_heat_sumfact_matches_dense(p):
    m = distorted(O.Mesh([0, 0, 0], [1.0, 2.0, 0.5], [2, 1, 2], p, 1, face_mask=0))
    coef = np.linspace(0.5, 2.0, m.n_cells * (p + 1) ** 3)
    h = O.Heat(m, dt=0.1, coef=coef)
    x = O.fill_uniform(3, m.n_dofs)
    yb = h.vmult_bruteforce(x)
    assert np.linalg.norm(h.vmult(x) - yb) < 1e-13 * np.linalg.norm(yb)


@pytest.mark.parametrize("form", [0, 1])
def test_tangent_is_symmetric(form):
    m = unit_mesh(2, 2, 3)
    s = O.Solid(m, formulation=form, caching=1)
    s.cache(m.u0(amp=1e-1, noise=1e-3))
    a, b = O.fill_uniform(3, m.n_dofs), O.fill_uniform(4, m.n_dofs)
    ab, ba = a @ s.vmult(b), b @ s.vmult(a)
    assert abs(ab - ba) < 1e-13 * abs(ab)


@pytest.mark.parametrize("form", [0, 1])
def test_tangent_is_derivative_of_residual(form):
    m = unit_mesh(2, 2, 3)
    u = m.u0(amp=1e-1, noise=1e-3)
    s = O.Solid(m, formulation=form, caching=1)
    s.cache(u)
    v = O.fill_uniform(5, m.n_dofs)
    con = m.constrained()
    v[con] = 0.0
    eps = 1e-7

    def R(uu):  # Newmark acceleration a = alpha_1 (u - u_n) + const
        return s.residual(uu, O.ALPHA1 * uu)

    fd = -(R(u + eps * v) - R(u - eps * v)) / (2 * eps)
    Av = s.vmult(v)
    Av[con] = 0.0
    assert np.linalg.norm(fd - Av) / np.linalg.norm(Av) < 1e-8


def _numpy_energy(m, u, mu, lam, cell_mat=None):
    """int Psi dX for Psi = mu/2 (tr C - 3 - 2 lnJ) + lam (lnJ)^2, evaluated in
    numpy with its own tensor-product quadrature (one box cell layout);
    cell_mat: optional (mu, lam) per cell instead."""
    p, N = m.p, m.p + 1
    z = O.gll_points(p)
    xq, wq = O.gauss_rule(N)
    S, D = O.lagrange(z, xq)
    nodes = m.cell_nodes().reshape(m.n_cells, N, N, N)  # [cell][c][b][a]
    verts = m.cell_vertices().reshape(m.n_cells, 8, 3)
    E = 0.0
    for c in range(m.n_cells):
        h = verts[c, 7] - verts[c, 0]
        U = np.stack([u[3 * nodes[c] + d] for d in range(3)])  # [d][kc][jb][ia]
        # grad_xi u: [d][f] at (q2,q1,q0)
        g0 = np.einsum("ia,jb,kc,dcba->dkji", D, S, S, U)
        g1 = np.einsum("ia,jb,kc,dcba->dkji", S, D, S, U)
        g2 = np.einsum("ia,jb,kc,dcba->dkji", S, S, D, U)
        G = np.stack([g0 / h[0], g1 / h[1], g2 / h[2]], axis=1)  # [d][e]
        F = G + np.eye(3)[:, :, None, None, None]
        F = np.moveaxis(F, (0, 1), (-2, -1))
        J = np.linalg.det(F)
        trC = np.einsum("...ij,...ij->...", F, F)
        mc, lc = cell_mat[c] if cell_mat is not None else (mu, lam)
        psi = mc / 2 * (trC - 3 - 2 * np.log(J)) + lc * np.log(J) ** 2
        W = np.einsum("k,j,i->kji", wq, wq, wq) * np.prod(h)
        E += np.sum(psi * W)
    return E


def test_residual_is_minus_energy_gradient():
    m = O.Mesh([0, 0, 0], [1.0, 0.8, 0.6], [1, 1, 1], 2, 3, face_mask=0)
    u = m.u0(amp=1e-1, noise=1e-2)
    s = O.Solid(m, caching=1)
    lam = 2 * O.MU * O.NU / (1 - 2 * O.NU)
    R = s.residual(u, np.zeros_like(u))
    rng = np.random.default_rng(0)
    for _ in range(3):
        v = rng.uniform(-1, 1, m.n_dofs)
        eps = 1e-6
        dE = (_numpy_energy(m, u + eps * v, O.MU, lam) - _numpy_energy(m, u - eps * v, O.MU, lam)) / (2 * eps)
        assert abs(dE + R @ v) < 1e-6 * abs(dE)


# the inclusion material: the reference builds it with 100 mu (mf_elasticity.h:470-476); the density differs
# too here, so that the inertia terms are checked as well
INCL = (100 * O.MU, O.NU, 2 * O.RHO)


def test_residual_is_minus_energy_gradient_two_materials():
    """Cells of material id 1 carry the inclusion material: the residual is the
    gradient of the energy with the material chosen per cell."""
    m = O.Mesh([0, 0, 0], [1.0, 0.8, 0.6], [2, 1, 1], 2, 3, face_mask=0)
    u = m.u0(amp=1e-1, noise=1e-2)
    ids = np.array([0, 1], dtype=np.int32)
    s = O.Solid(m, caching=1, inclusion=INCL, material_ids=ids)
    lame = lambda mu, nu: 2 * mu * nu / (1 - 2 * nu)
    cell_mat = [(O.MU, lame(O.MU, O.NU)), (INCL[0], lame(INCL[0], INCL[1]))]
    R = s.residual(u, np.zeros_like(u))
    rng = np.random.default_rng(1)
    for _ in range(3):
        v = rng.uniform(-1, 1, m.n_dofs)
        eps = 1e-6
        dE = (_numpy_energy(m, u + eps * v, 0, 0, cell_mat) - _numpy_energy(m, u - eps * v, 0, 0, cell_mat)) / (2 * eps)
        assert abs(dE + R @ v) < 1e-6 * abs(dE)


@pytest.mark.parametrize("form,caching", [(1, 1), (1, 0), (1, 2), (1, 3), (0, 1)])
def test_inclusion_is_chosen_cell_by_cell(form, caching):
    """Two materials by cell (set_material(material, material_inclusion),
    mf_nh_operator.h:404, 658): swapping the materials cell by cell and adding
    gives the sum of the two one-material operators (vmult, diagonal and
    residual), and with every cell of id 1 the operator is the one-material
    operator of the inclusion."""
    m = O.Mesh([0, 0, 0], [1.0, 0.7, 0.5], [3, 2, 2], 2, 3)
    u = m.u0(amp=5e-2, noise=1e-3)
    ids = (np.random.default_rng(3).uniform(size=m.n_cells) < 0.4).astype(np.int32)
    assert 0 < ids.sum() < m.n_cells
    base = (O.MU, O.NU, O.RHO)

    def make(mat, inclusion=None, material_ids=None):
        s = O.Solid(m, mu=mat[0], nu=mat[1], rho=mat[2], formulation=form, caching=caching, inclusion=inclusion,
                    material_ids=material_ids)
        s.cache(u)
        return s

    A, B = make(base, INCL, ids), make(base, INCL, 1 - ids)
    H0, H1 = make(base), make(INCL)
    x = O.fill_uniform(11, m.n_dofs)
    acc = O.fill_uniform(12, m.n_dofs)
    body = (0.5, -9.81, 2.0)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(A.vmult(x) + B.vmult(x), H0.vmult(x) + H1.vmult(x)) < 1e-13
    assert rel(A.diagonal() + B.diagonal(), H0.diagonal() + H1.diagonal()) < 1e-13
    assert rel(A.residual(u, acc, body) + B.residual(u, acc, body),
               H0.residual(u, acc, body) + H1.residual(u, acc, body)) < 1e-13
    ones = make(base, INCL, np.ones(m.n_cells, dtype=np.int32))
    assert rel(ones.vmult(x), H1.vmult(x)) < 1e-15
    assert rel(ones.residual(u, acc, body), H1.residual(u, acc, body)) < 1e-15
    # the two materials really differ on this mesh
    assert rel(A.vmult(x), H0.vmult(x)) > 1e-2 and rel(A.vmult(x), H1.vmult(x)) > 1e-2


INCLUSION_PINS = ["test_residual_is_minus_energy_gradient_two_materials", "test_inclusion_is_chosen_cell_by_cell"]


@pytest.mark.parametrize("plant,caught_by", [
    ("ORC_PLANT_INCL_CACHE_MAT0", ["test_inclusion_is_chosen_cell_by_cell"]),
    ("ORC_PLANT_INCL_RESIDUAL_MAT0", INCLUSION_PINS),
])
def test_inclusion_pins_fail_on_a_planted_mistake(plant, caught_by, tmp_path):
    """A copy of the oracle whose cache or residual ignores the cell's material
    fails the inclusion pins above."""
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
                        os.path.join(here, "test_oracle_operators.py"), "-k", " or ".join(INCLUSION_PINS)],
                       capture_output=True, text=True, env=env, cwd=here)
    assert r.returncode != 0, r.stdout
    failed = {line.split("::")[1].split("[")[0].split(" ")[0] for line in r.stdout.splitlines()
              if line.startswith("FAILED")}
    for name in caught_by:
        assert name in failed, (name, r.stdout[-3000:])


def _linear_elasticity_dense(m, mu, lam):
    """Dense linear elasticity stiffness with Lame pair (mu, lam), numpy."""
    p, N = m.p, m.p + 1
    z = O.gll_points(p)
    xq, wq = O.gauss_rule(N)
    S, D = O.lagrange(z, xq)
    K = np.zeros((m.n_dofs, m.n_dofs))
    nodes = m.cell_nodes()
    verts = m.cell_vertices().reshape(m.n_cells, 8, 3)
    for c in range(m.n_cells):
        h = verts[c, 7] - verts[c, 0]
        cn = nodes[c * N ** 3:(c + 1) * N ** 3]
        for q2 in range(N):
            for q1 in range(N):
                for q0 in range(N):
                    w = wq[q0] * wq[q1] * wq[q2] * np.prod(h)
                    grads = np.zeros((N ** 3, 3))
                    for kc in range(N):
                        for jb in range(N):
                            for ia in range(N):
                                a = ia + N * (jb + N * kc)
                                grads[a] = [D[q0, ia] * S[q1, jb] * S[q2, kc] / h[0],
                                            S[q0, ia] * D[q1, jb] * S[q2, kc] / h[1],
                                            S[q0, ia] * S[q1, jb] * D[q2, kc] / h[2]]
                    # B matrix rows: strain components
                    for a in range(N ** 3):
                        for b in range(N ** 3):
                            ga, gb = grads[a], grads[b]
                            # k_ab[d][e] = mu (ga.gb delta_de + gb_d ga_e) + lam ga_d gb_e
                            kab = mu * (ga @ gb * np.eye(3) + np.outer(gb, ga)) + lam * np.outer(ga, gb)
                            for d in range(3):
                                for e in range(3):
                                    K[3 * cn[a] + d, 3 * cn[b] + e] += w * kab[d, e]
    return K


@pytest.mark.parametrize("form,lame_factor", [(1, 2.0), (0, 1.0)])
def test_u0_reduces_to_linear_elasticity(form, lame_factor):
    m = O.Mesh([0, 0, 0], [1.0, 0.5, 0.5], [2, 1, 1], 1, 3, face_mask=0)
    s = O.Solid(m, formulation=form, caching=1, alpha1=0.0)
    s.cache(np.zeros(m.n_dofs))
    lam = 2 * O.MU * O.NU / (1 - 2 * O.NU)
    K = _linear_elasticity_dense(m, O.MU, lame_factor * lam)
    A = np.stack([s.vmult(e) for e in np.eye(m.n_dofs)], axis=1)
    assert np.abs(A - K).max() < 1e-10 * np.abs(K).max()


def test_mass_closed_form_any_linearisation_point():
    # constant fields have zero strain: e_d^T A e_d = rho alpha_1 |Omega|
    m = O.Mesh([0, 0, 0], [4.0, 1.0, 1.0], [4, 1, 1], 2, 3, face_mask=0)
    s = O.Solid(m, caching=1)
    s.cache(m.u0(amp=1e-1, noise=1e-3))
    for d in range(3):
        e = np.zeros(m.n_dofs)
        e[d::3] = 1.0
        assert abs(e @ s.vmult(e) - O.RHO * O.ALPHA1 * 4.0) < 1e-11 * O.RHO * O.ALPHA1 * 4.0


def test_diagonal_matches_unit_probes():
    m = unit_mesh(2, 2, 3)
    s = O.Solid(m, caching=1)
    s.cache(m.u0(amp=1e-1, noise=1e-3))
    d = s.diagonal()
    con = set(m.constrained().tolist())
    for i in range(0, m.n_dofs, 7):
        e = np.zeros(m.n_dofs)
        e[i] = 1.0
        ref = 1.0 if i in con else s.vmult_bruteforce(e)[i]
        assert abs(d[i] - ref) < 1e-12 * abs(ref)


def test_constrained_rows_are_identity():
    m = unit_mesh(2, 3, 3)
    s = O.Solid(m, caching=1)
    s.cache(m.u0())
    x = O.fill_uniform(1, m.n_dofs)
    y = s.vmult(x)
    con = m.constrained()
    assert np.array_equal(y[con], x[con])


@pytest.mark.parametrize("p", [1, 2, 3])
def test_heat_sumfact_matches_dense_and_mass(p):
    m = O.Mesh([0, 0, 0], [1.0, 2.0, 0.5], [2, 1, 2], p, 1, face_mask=0)
    rng = np.random.default_rng(p)
    coef = rng.uniform(0.5, 2.0, m.n_cells * (p + 1) ** 3)
    h = O.Heat(m, dt=0.1, coef=coef)
    x = O.fill_uniform(1, m.n_dofs)
    assert np.linalg.norm(h.vmult(x) - h.vmult_bruteforce(x)) < 1e-13 * np.linalg.norm(h.vmult_bruteforce(x))
    one = np.ones(m.n_dofs)
    assert abs(one @ h.vmult(one) - 1.0) < 1e-13  # |Omega| = 1


def test_heat_manufactured_quadratic_solution_is_exact():
    """u = 1 + x^2 + a y^2 + g z^2 + b t solves u_t - lap u = b - 2 - 2a - 2g
    and is reproduced exactly by Q2 + backward Euler (partitioned_heat.h:22-81)."""
    a, b, g, dt = 3.0, 1.3, 0.7, 0.1
    m = O.Mesh([0, 0, 0], [1, 1, 1], [3, 3, 3], 2, 1, face_mask=63, comp_mask=1)
    mf = O.Mesh([0, 0, 0], [1, 1, 1], [3, 3, 3], 2, 1, face_mask=0, comp_mask=1)
    X = m.node_coords()

    def exact(t):
        return 1 + X[:, 0] ** 2 + a * X[:, 1] ** 2 + g * X[:, 2] ** 2 + b * t

    A = O.Heat(m, dt=dt)
    A_full = O.Heat(mf, dt=dt)
    M = O.Heat(mf, dt=0.0)
    con = m.constrained()
    u = exact(0.0)
    for step in range(1, 3):
        t = step * dt
        ud = np.zeros(m.n_dofs)
        ud[con] = exact(t)[con]
        rhs = M.vmult_noconstraints(u) + dt * A.load(b - 2 - 2 * a - 2 * g) - A_full.vmult_noconstraints(ud)
        rhs[con] = 0.0
        w, it, rel = O.cg_solve(A, rhs, precond=1, tol_rel=1e-14, max_iter=2000)
        assert it > 0
        u = w + ud
        assert np.abs(u - exact(t)).max() < 1e-11


def test_interface_traction_closed_forms():
    """The oracle's interface traction (Nanson's formula from the deformed
    tangents) against what the kinematics fix: with u = 0 the traction
    integrates to t times the face area; a uniform stretch lambda along x keeps
    the x faces' area and scales the others by lambda; a shear u = g y e_x
    scales the x faces by sqrt(1 + g^2) and keeps the y faces; the first moment
    of a constant traction is t times the face's first moment."""
    lo, hi, n, p = [0, 0, 0], [2.0, 0.6, 0.4], [3, 2, 2], 2
    m = O.Mesh(lo, hi, n, p, 3, 0, 7)
    X = m.node_coords()
    tv = np.array([1.5, -0.7, 0.3])
    area = {0: 0.6 * 0.4, 1: 0.6 * 0.4, 2: 2.0 * 0.4, 3: 2.0 * 0.4, 4: 2.0 * 0.6, 5: 2.0 * 0.6}

    def total(faces, u):
        t = np.tile(tv, len(faces) // 2 * (p + 1) ** 2)
        return m.interface_traction(faces, u, t).reshape(-1, 3)

    zero = np.zeros(m.n_dofs)
    for bid in range(6):
        faces = m.boundary_faces(bid)
        r = total(faces, zero)
        assert np.allclose(r.sum(axis=0), tv * area[bid], rtol=1e-13, atol=0), bid
        on_face = np.isclose(X[:, bid // 2], (lo if bid % 2 == 0 else hi)[bid // 2])
        assert np.all(r[~on_face] == 0.0)
        # first moment: the nodal basis reproduces linear functions
        tang = [d for d in range(3) if d != bid // 2][0]
        centroid = (lo[tang] + hi[tang]) / 2
        assert np.allclose((r * X[:, tang:tang + 1]).sum(axis=0), tv * area[bid] * centroid, rtol=1e-13, atol=0)
    lam, g = 1.3, 0.45
    stretch = np.zeros((m.n_nodes, 3))
    stretch[:, 0] = (lam - 1) * X[:, 0]
    shear = np.zeros((m.n_nodes, 3))
    shear[:, 0] = g * X[:, 1]
    for bid in range(6):
        faces = m.boundary_faces(bid)
        s = 1.0 if bid // 2 == 0 else lam
        assert np.allclose(total(faces, stretch.ravel()).sum(axis=0), s * tv * area[bid], rtol=1e-13, atol=0), bid
        s = np.sqrt(1 + g * g) if bid // 2 == 0 else 1.0
        assert np.allclose(total(faces, shear.ravel()).sum(axis=0), s * tv * area[bid], rtol=1e-13, atol=0), bid


def test_face_values_and_heat_flux_closed_forms():
    """Face values reproduce a field the element space holds (a trilinear-free
    linear function, exact at every face point); a constant heat flux q
    integrates to -dt q times the face area, with the face's first moment."""
    lo, hi, n, p = [0, 0, 0], [2.0, 0.6, 0.4], [3, 2, 2], 3
    N = p + 1
    gauss = (np.polynomial.legendre.leggauss(N)[0] + 1) / 2
    h = (np.array(hi) - np.array(lo)) / n
    for nc in (1, 3):
        m = O.Mesh(lo, hi, n, p, nc, 0, 7)
        X = m.node_coords()
        coef = np.array([[0.5, 2.0, -1.0, 3.0], [1.0, -0.5, 0.25, 0.0], [-2.0, 0.0, 1.5, 1.0]])[:nc]
        u = np.stack([c[0] + X @ c[1:] for c in coef], axis=1).ravel()
        for bid in range(6):
            faces = m.boundary_faces(bid)
            vals = m.face_values(faces, u)
            axis, side = bid // 2, bid % 2
            ta, tb = [d for d in range(3) if d != axis]
            pts = []
            for cell in faces[0::2]:
                cc = [cell % n[0], (cell // n[0]) % n[1], cell // (n[0] * n[1])]
                for q in range(N * N):
                    x = np.empty(3)
                    x[axis] = lo[axis] + (cc[axis] + side) * h[axis]
                    x[ta] = lo[ta] + (cc[ta] + gauss[q % N]) * h[ta]
                    x[tb] = lo[tb] + (cc[tb] + gauss[q // N]) * h[tb]
                    pts.append(x)
            pts = np.array(pts)
            want = np.stack([c[0] + pts @ c[1:] for c in coef], axis=1)
            assert np.abs(vals - want).max() <= 1e-13 * np.abs(want).max(), (nc, bid)
    m = O.Mesh(lo, hi, n, p, 1, 0, 1)
    X = m.node_coords()
    area = {0: 0.24, 1: 0.24, 2: 0.8, 3: 0.8, 4: 1.2, 5: 1.2}
    q, dt = 2.5, 0.1
    for bid in range(6):
        faces = m.boundary_faces(bid)
        r = m.interface_flux(faces, np.full(len(faces) // 2 * N * N, q), dt)
        assert abs(r.sum() + dt * q * area[bid]) <= 1e-13 * dt * q * area[bid]
        tang = [d for d in range(3) if d != bid // 2][0]
        cen = (lo[tang] + hi[tang]) / 2
        assert abs(r @ X[:, tang] + dt * q * area[bid] * cen) <= 1e-13 * dt * q * area[bid]


@pytest.mark.parametrize("p,n", [(2, [8, 8, 8]), (4, [5, 3, 4]), (4, [6, 6, 6])])
def test_simd_baseline_equals_oracle_vmult(p, n):
    """The benchmark's CPU baseline (cells in SIMD batches of 8, collocation
    sum factorisation, OpenMP over coloured batches) is the oracle's operator:
    the same vmult to rounding, including ragged last batches and Dirichlet
    rows."""
    m = O.Mesh([0, 0, 0], [1.0, 0.7, 0.9], n, p, 3, 1 | 16, 7)
    s = O.Solid(m, caching=1)
    s.cache(m.u0(amp=2e-2, noise=1e-4, seed=3))
    x = O.fill_uniform(7, m.n_dofs)
    ref = s.vmult(x)
    y = O.SimdSolid(s).vmult(x)
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    assert np.array_equal(y[m.constrained()], x[m.constrained()])
