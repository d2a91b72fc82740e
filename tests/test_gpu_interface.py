"""Coupling interface of the solid on the device (Solid::assemble_residual's
traction term, mf_elasticity.h:1849-1897, and the face quadrature points of
the coupling mesh, include/adapter/quad_interface.h:198-340).

- traction: the device's det F F^-T N pull-back against the oracle's Nanson
  form (deformed tangents, no inverse), every degree, all six box boundary
  ids, random displacement and traction, relative l2 <= 1e-12; Dirichlet
  dofs untouched; cellwise numbering equals lexicographic permuted;
- face points: closed form on a box (lo + (cell + Gauss point) h on the
  tangential axes, the face's coordinate on the normal one).
"""
import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu

LO, HI = [0, 0, 0], [1.0, 0.6, 0.4]


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()


@pytest.mark.parametrize("p,n", [(1, [3, 2, 2]), (2, [2, 3, 2]), (3, [2, 2, 2]), (4, [2, 2, 1]), (5, [1, 2, 1]),
                                 (6, [1, 1, 2])])
def test_traction_matches_oracle(cuda, p, n):
    print(f"interface traction p={p} n={n}", flush=True)
    om = O.Mesh(LO, HI, n, p, 3, 1, 7)
    gm = mfgpu.BoxMesh(LO, HI, n, p, 3, 1, 7)
    op = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1)
    u = om.u0(amp=5e-2, noise=1e-3)
    con = om.constrained()
    for bid in range(6):
        faces = gm.boundary_faces(bid)
        assert np.array_equal(faces.ravel(), om.boundary_faces(bid))
        t = O.fill_uniform(10 + bid, len(faces) * (p + 1) ** 2 * 3)
        ref = om.interface_traction(faces.ravel(), u, t)
        base = O.fill_uniform(20 + bid, om.n_dofs)
        rhs = dev(base)
        op.add_interface_traction(rhs, dev(u), faces, dev(t))
        got = rhs.cpu().numpy() - base
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref), bid
        assert np.array_equal(rhs.cpu().numpy()[con], base[con])  # Dirichlet dofs untouched


def test_traction_cellwise_numbering(cuda):
    p, n = 4, [3, 2, 2]
    lex = mfgpu.BoxMesh(LO, HI, n, p, 3, 1, 7)
    cw = mfgpu.BoxMesh(LO, HI, n, p, 3, 1, 7, numbering="cellwise")
    perm = np.full(lex.n_nodes, -1, dtype=np.int64)
    perm[lex.cell_nodes()] = cw.cell_nodes()
    dof = (perm[:, None] * 3 + np.arange(3)[None, :]).ravel()
    ol = mfgpu.NeoHookOperator(lex, O.MU, O.NU, O.RHO, O.ALPHA1)
    oc = mfgpu.NeoHookOperator(cw, O.MU, O.NU, O.RHO, O.ALPHA1)
    u = lex.u0(amp=2e-2, noise=1e-4)
    uc = np.empty_like(u)
    uc[dof] = u
    faces = np.concatenate([lex.boundary_faces(b) for b in (1, 3, 4)])
    t = dev(O.fill_uniform(3, len(faces) * (p + 1) ** 2 * 3))
    rl = torch.zeros(lex.n_dofs, dtype=torch.float64, device="cuda")
    rc = torch.zeros_like(rl)
    ol.add_interface_traction(rl, dev(u), faces, t)
    oc.add_interface_traction(rc, dev(uc), faces, t)
    a, b = rl.cpu().numpy(), rc.cpu().numpy()[dof]
    assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(a)


def test_face_points_closed_form(cuda):
    p, n = 3, [3, 2, 2]
    gm = mfgpu.BoxMesh(LO, HI, n, p, 3, 1, 7)
    op = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1)
    gauss = (np.polynomial.legendre.leggauss(p + 1)[0] + 1) / 2
    h = (np.array(HI) - np.array(LO)) / n
    N = p + 1
    for bid in range(6):
        faces = gm.boundary_faces(bid)
        x = op.face_quadrature_points(faces).cpu().numpy().reshape(len(faces), N * N, 3)
        axis, side = bid // 2, bid % 2
        ta, tb = [d for d in range(3) if d != axis]
        for f, (cell, face) in enumerate(faces):
            cc = [cell % n[0], (cell // n[0]) % n[1], cell // (n[0] * n[1])]
            for q in range(N * N):
                want = np.empty(3)
                want[axis] = LO[axis] + (cc[axis] + side) * h[axis]
                want[ta] = LO[ta] + (cc[ta] + gauss[q % N]) * h[ta]
                want[tb] = LO[tb] + (cc[tb] + gauss[q // N]) * h[tb]
                assert np.abs(x[f, q] - want).max() < 1e-14


def test_traction_errors(cuda):
    gm = mfgpu.BoxMesh(LO, HI, [2, 2, 2], 2, 3, 1, 7)
    u = torch.zeros(gm.n_dofs, dtype=torch.float64, device="cuda")
    t = torch.zeros(4 * 9 * 3, dtype=torch.float64, device="cuda")
    op32 = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, precision=mfgpu.FP32)
    with pytest.raises(mfgpu.MfgError):
        op32.add_interface_traction(u, u, gm.boundary_faces(1), t)
    op = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1)
    with pytest.raises(mfgpu.MfgError):
        op.add_interface_traction(u, u, np.array([[8, 1]]), t)  # no cell 8
    with pytest.raises(mfgpu.MfgError):
        op.add_interface_traction(u, u, np.array([[0, 6]]), t)  # no face 6
    with pytest.raises(mfgpu.MfgError):
        gm.boundary_faces(6)


@pytest.mark.parametrize("p,n", [(1, [3, 2, 2]), (2, [2, 3, 2]), (4, [2, 2, 1]), (6, [1, 1, 2])])
def test_face_values_and_heat_flux_match_oracle(cuda, p, n):
    """Face values of the solid displacement and of a temperature field, and
    the heat interface flux term, against the oracle (full tensor-product
    evaluation; cross-product area element) at relative 1e-13 / 1e-12."""
    print(f"face values / heat flux p={p} n={n}", flush=True)
    om = O.Mesh(LO, HI, n, p, 3, 1, 7)
    gm = mfgpu.BoxMesh(LO, HI, n, p, 3, 1, 7)
    sop = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1)
    u = om.u0(amp=5e-2, noise=1e-3)
    hm = O.Mesh(LO, HI, n, p, 1, 63 - 2, 1)  # every face but x = 1 Dirichlet
    hg = mfgpu.BoxMesh(LO, HI, n, p, 1, 63 - 2, 1)
    dt = 0.05
    hop = mfgpu.LaplaceOperator(hg, dt)
    T = O.fill_uniform(7, hm.n_dofs)
    con = hm.constrained()
    for bid in range(6):
        faces = gm.boundary_faces(bid)
        ref = om.face_values(faces.ravel(), u)
        got = sop.face_values(faces, dev(u)).cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), bid
        ref = hm.face_values(faces.ravel(), T)
        got = hop.face_values(faces, dev(T)).cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), bid
        q = O.fill_uniform(30 + bid, len(faces) * (p + 1) ** 2)
        ref = hm.interface_flux(faces.ravel(), q, dt)
        base = O.fill_uniform(40 + bid, hm.n_dofs)
        rhs = dev(base)
        hop.add_interface_flux(rhs, faces, dev(q))
        got = rhs.cpu().numpy() - base
        if np.linalg.norm(ref) > 0:
            assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref), bid
        else:
            assert np.abs(got).max() <= 1e-15
        assert np.array_equal(rhs.cpu().numpy()[con], base[con])
