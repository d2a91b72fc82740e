"""Device Newton-Newmark driver and residual against the oracle."""
import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu

BODY = (0.0, 0.0, -9.81)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=torch.float64)


@pytest.mark.parametrize("form,p", [(1, 2), (0, 2), (1, 4)])
def test_residual_parity(cuda, form, p):
    lo, hi, n = [0, 0, 0], [1.0, 0.5, 0.5], [2, 2, 1]
    gm = mfgpu.BoxMesh(lo, hi, n, p, 3, 1, 7)
    om = O.Mesh(lo, hi, n, p, 3, 1, 7)
    u = om.u0(amp=5e-2, noise=1e-3)
    acc = O.fill_uniform(4, om.n_dofs)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, formulation=form)
    oo = O.Solid(om, formulation=form, caching=1)
    R, nrm = mfgpu.solid_residual(go, dev(u), dev(acc), BODY)
    ref = oo.residual(u, acc, BODY)
    assert np.linalg.norm(R.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)
    assert abs(nrm - np.linalg.norm(ref)) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("level_precision", [mfgpu.FP64, mfgpu.FP32])
def test_newmark_steps_match_oracle(cuda, level_precision):
    lo, hi = [0, 0, 0], [1, 1, 1]
    gms = [mfgpu.BoxMesh(lo, hi, [1 << l] * 3, 2, 3, 1, 7) for l in range(3)]
    oms = O.box_levels(lo, hi, [1, 1, 1], 3, 2, 3)
    oops = [O.Solid(m, caching=1) for m in oms]
    prob = mfgpu.SolidProblem(gms, O.MU, O.NU, O.RHO, body=BODY, level_precision=level_precision, tol_f=1e-8,
                              coarse_kind=1 if level_precision == mfgpu.FP32 else 0)
    u = oms[-1].u0(amp=1e-3, noise=1e-5)
    v = np.zeros_like(u)
    a = np.zeros_like(u)
    prob.set_state(dev(u), dev(v), dev(a))
    for step in range(2):
        info = prob.step()
        it, li = O.newmark_step(oms, oops, u, v, a, body=BODY, tol_f=1e-8,
                                coarse_kind=1 if level_precision == mfgpu.FP32 else 0)
        gu, gv, ga = (x.cpu().numpy() for x in prob.get_state())
        print(f"step {step}: newton gpu {info.newton_iterations} oracle {it}; linear gpu {info.linear_iterations} "
              f"oracle {li}; |du| rel {np.linalg.norm(gu - u) / np.linalg.norm(u):.2e}")
        assert info.converged and info.newton_iterations == it
        assert np.linalg.norm(gu - u) <= 1e-7 * np.linalg.norm(u)
        assert np.linalg.norm(gv - v) <= 1e-5 * np.linalg.norm(v)
        # the converged state satisfies the equations of motion
        R = oops[-1].residual(gu, ga, BODY)
        assert np.linalg.norm(R) <= 1e-8 * info.residual_0 * 10
    assert prob.memory() > 0


@pytest.mark.parametrize("level_precision", [mfgpu.FP64, mfgpu.FP32])
def test_coupled_steps_with_interface_traction(cuda, level_precision):
    """Newmark steps with a traction on the x = 1 face (the coupling interface
    of mf_elasticity.h:1849-1897): every converged state satisfies the
    oracle's equations of motion including its (Nanson-form) traction term;
    an interface that is set and then removed gives the uncoupled step bit for
    bit."""
    lo, hi = [0, 0, 0], [1, 1, 1]
    gms = [mfgpu.BoxMesh(lo, hi, [1 << l] * 3, 2, 3, 1, 7) for l in range(3)]
    oms = O.box_levels(lo, hi, [1, 1, 1], 3, 2, 3)
    oop = O.Solid(oms[-1], caching=1)
    kw = dict(body=BODY, level_precision=level_precision, tol_f=1e-8,
              coarse_kind=1 if level_precision == mfgpu.FP32 else 0)
    faces = gms[-1].boundary_faces(1)
    t = np.tile([0.0, -5e3, 2e3], len(faces) * 9)
    tdev = dev(t)
    u = oms[-1].u0(amp=1e-3, noise=1e-5)
    zero = np.zeros_like(u)
    prob = mfgpu.SolidProblem(gms, O.MU, O.NU, O.RHO, **kw)
    prob.set_interface(faces, tdev)
    prob.set_state(dev(u), dev(zero), dev(zero))
    plain = mfgpu.SolidProblem(gms, O.MU, O.NU, O.RHO, **kw)
    plain.set_state(dev(u), dev(zero), dev(zero))
    cleared = mfgpu.SolidProblem(gms, O.MU, O.NU, O.RHO, **kw)
    cleared.set_interface(faces, tdev)
    cleared.set_interface(np.zeros((0, 2), dtype=np.int32), None)
    cleared.set_state(dev(u), dev(zero), dev(zero))
    for step in range(2):
        info = prob.step()
        assert info.converged
        gu, gv, ga = (x.cpu().numpy() for x in prob.get_state())
        R = oop.residual(gu, ga, BODY) + oms[-1].interface_traction(faces.ravel(), gu, t)
        print(f"step {step}: newton {info.newton_iterations}, |R| {np.linalg.norm(R):.2e} vs R0 {info.residual_0:.2e}",
              flush=True)
        assert np.linalg.norm(R) <= 1e-8 * info.residual_0 * 10
        pi, ci = plain.step(), cleared.step()
        pu, cu = plain.get_state()[0], cleared.get_state()[0]
        assert torch.equal(pu, cu) and pi.newton_iterations == ci.newton_iterations
        assert float((pu - dev(gu)).norm()) > 1e-3 * float(pu.norm())  # the traction moves the solid
