"""The reference's Turek flap goldens through the device Newton-Newmark step
(mfg_solid_problem_step with the interface traction), for the four
formulation-1 cachings and for FP64 and FP32 multigrid levels (the
reference's levels are float): 4 Newton iterations per step and the
watchpoint displacements of turek_gmg_*_serial.output to 6 significant digits
(the 3D variant and its tolerances: turek_case.py)."""
import numpy as np
import pytest
import torch

import mfgpu
import turek_case as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("caching", ["tensor2", "tensor4", "tensor4_ns", "scalar_referential"])
@pytest.mark.parametrize("level_precision", [mfgpu.FP64, mfgpu.FP32])
def test_turek_goldens_device(cuda, caching, level_precision):
    meshes = [mfgpu.BoxMesh(T.LO, T.HI, [c << l for c in T.COARSE], 1, 3, face_comp=T.FACE_COMP)
              for l in range(T.LEVELS)]
    fm = meshes[-1]
    prob = mfgpu.SolidProblem(meshes, T.MU, T.NU, T.RHO, beta=T.BETA, gamma=T.GAMMA, dt=T.DT, body=T.BODY,
                              level_precision=level_precision, caching=caching, **T.newton_params(fm.n_cells_dir[2]),
                              **T.mg_params())
    faces = np.concatenate([fm.boundary_faces(b) for b in T.INTERFACE_IDS])
    traction = torch.tensor(np.tile(T.TRACTION, len(faces) * 4), dtype=torch.float64, device=cuda)
    prob.set_interface(faces, traction)
    w = T.watch_nodes(fm.node_coords())
    cg = []
    for step in (1, 2):
        info = prob.step()
        u = prob.get_state()[0].cpu().numpy().reshape(-1, 3)
        d = u[w]
        print(f"{caching} levels {'fp64' if level_precision == mfgpu.FP64 else 'fp32'} step {step}: "
              f"Newton {info.newton_iterations}, CG {info.linear_iterations}, condition {info.condition_estimate:.3g}, "
              f"watchpoint {d[0, :2]}", flush=True)
        assert info.converged and info.newton_iterations == T.GOLDEN_NEWTON
        assert info.condition_estimate > 1.0
        assert np.abs(d - d[0]).max() <= 1e-10 * np.abs(d).max()
        for k in range(2):
            assert T.six_digits(d[0, k], T.GOLDEN[step][k]), (step, k, d[0, k], T.GOLDEN[step][k])
        cg.append(info.linear_iterations / info.newton_iterations)
    assert abs(np.mean(cg) - T.GOLDEN_AVG_CG) <= 5


@pytest.mark.parametrize("level_precision", [mfgpu.FP64, mfgpu.FP32])
def test_turek_jacobi_goldens_device(cuda, level_precision):
    """turek_jacobi_tensor2: the same flap solved with PreconditionJacobi
    (relaxation 0.65, at most 2 n_dofs CG iterations; turek_jacobi_tensor2.prm
    against turek_gmg_tensor2.prm).  Its golden output holds the same Newton
    counts and watchpoint displacements as the multigrid cases
    (turek_jacobi_tensor2_serial.output:10-17); the multigrid levels are not
    set up (update_matrix_free, mf_elasticity.h:1320-1324), so the level
    precision must not matter."""
    meshes = [mfgpu.BoxMesh(T.LO, T.HI, [c << l for c in T.COARSE], 1, 3, face_comp=T.FACE_COMP)
              for l in range(T.LEVELS)]
    fm = meshes[-1]
    kw = T.newton_params(fm.n_cells_dir[2])
    kw["max_lin"] = 2 * fm.n_dofs
    prob = mfgpu.SolidProblem(meshes, T.MU, T.NU, T.RHO, beta=T.BETA, gamma=T.GAMMA, dt=T.DT, body=T.BODY,
                              level_precision=level_precision, preconditioner="jacobi", relaxation=0.65, **kw,
                              **T.mg_params())
    faces = np.concatenate([fm.boundary_faces(b) for b in T.INTERFACE_IDS])
    traction = torch.tensor(np.tile(T.TRACTION, len(faces) * 4), dtype=torch.float64, device=cuda)
    prob.set_interface(faces, traction)
    w = T.watch_nodes(fm.node_coords())
    cg = []
    for step in (1, 2):
        info = prob.step()
        d = prob.get_state()[0].cpu().numpy().reshape(-1, 3)[w]
        print(f"jacobi levels {'fp64' if level_precision == mfgpu.FP64 else 'fp32'} step {step}: "
              f"Newton {info.newton_iterations}, CG {info.linear_iterations}, watchpoint {d[0, :2]}", flush=True)
        assert info.converged and info.newton_iterations == T.GOLDEN_NEWTON
        for k in range(2):
            assert T.six_digits(d[0, k], T.GOLDEN[step][k]), (step, k, d[0, k], T.GOLDEN[step][k])
        cg.append(info.linear_iterations / info.newton_iterations)
    # far more CG iterations than with multigrid (the 2D golden: 701 against 29)
    assert np.mean(cg) > 5 * T.GOLDEN_AVG_CG


def test_newton_preconditioner_arguments(cuda):
    meshes = [mfgpu.BoxMesh(T.LO, T.HI, [c << l for c in T.COARSE], 1, 3, face_comp=T.FACE_COMP) for l in range(2)]
    with pytest.raises(mfgpu.MfgError):
        mfgpu.SolidProblem(meshes, T.MU, T.NU, T.RHO, preconditioner="jacobi", relaxation=0.0)
