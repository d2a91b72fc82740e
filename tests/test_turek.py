"""End-to-end pin of the oracle by the reference's own goldens: the Turek flap
(3D variant, see turek_case.py) run through the oracle's Newton-Newmark step
with the interface traction reproduces the Newton counts and the watchpoint
displacements of turek_gmg_tensor2_serial.output to 6 significant digits."""
import numpy as np

import mfgpu
import oracle as O
import turek_case as T


def test_turek_mesh_constraints_closed_form_and_builders_agree():
    for l in range(T.LEVELS):
        n = [c << l for c in T.COARSE]
        om = O.Mesh(T.LO, T.HI, n, 1, 3, face_comp=T.FACE_COMP)
        gm = mfgpu.BoxMesh(T.LO, T.HI, n, 1, 3, face_comp=T.FACE_COMP)
        NN = [c + 1 for c in n]
        # x = lo: 3 components of NN_y NN_z nodes; z faces: the z component of
        # the other nodes on them
        assert om.n_constrained == 3 * NN[1] * NN[2] + 2 * (NN[0] - 1) * NN[1]
        assert np.array_equal(om.constrained(), gm.constrained())
        assert np.array_equal(om.cell_nodes(), gm.cell_nodes())


def test_turek_goldens_oracle():
    meshes = O.box_levels(T.LO, T.HI, T.COARSE, T.LEVELS, 1, 3, face_comp=T.FACE_COMP)
    fm = meshes[-1]
    a1 = 1.0 / (T.BETA * T.DT * T.DT)
    ops = [O.Solid(m, mu=T.MU, nu=T.NU, rho=T.RHO, alpha1=a1, caching=1) for m in meshes]
    faces = np.concatenate([fm.boundary_faces(b) for b in T.INTERFACE_IDS])
    t = np.tile(T.TRACTION, len(faces) * 4)  # 2 x 2 Gauss points per face
    w = T.watch_nodes(fm.node_coords())
    assert len(w) == fm.n[2] + 1
    u, v, a = np.zeros(fm.n_dofs), np.zeros(fm.n_dofs), np.zeros(fm.n_dofs)
    cg = []
    for step in (1, 2):
        it, li = O.newmark_step(meshes, ops, u, v, a, beta=T.BETA, gamma=T.GAMMA, dt=T.DT, body=T.BODY,
                                faces=faces, traction=t, **T.newton_params(fm.n[2]), **T.mg_params())
        d = u.reshape(-1, 3)[w]
        print(f"step {step}: Newton {it}, CG {li}, watchpoint {d[0, :2]}")
        assert it == T.GOLDEN_NEWTON
        # plane strain: every z layer of the watch line moves alike, no z motion
        assert np.abs(d - d[0]).max() <= 1e-12 * np.abs(d).max()
        assert np.abs(u.reshape(-1, 3)[:, 2]).max() <= 1e-12 * np.abs(u).max()
        for k in range(2):
            assert T.six_digits(d[0, k], T.GOLDEN[step][k]), (step, k, d[0, k], T.GOLDEN[step][k])
        cg.append(li / it)
    # average CG iterations per linear solve: 29 in 2D; the 3D hierarchy (thin
    # cells, 7:1 aspect ratio in x-z) takes a few more -- a soft figure
    print(f"average CG per solve {np.mean(cg):.1f} (2D golden {T.GOLDEN_AVG_CG})")
    assert abs(np.mean(cg) - T.GOLDEN_AVG_CG) <= 5
