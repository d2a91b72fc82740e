"""Committed fixtures of the small FP64 configuration (tests/golden/small_cube.npz,
written by tests/golden/make_small_cube.py from the oracle).  The CPU tests pin
the oracle against the fixture (a change in the oracle's layout, operator or
solver shows here); the GPU tests hold the device path to the same numbers:
vmult to 1e-12, MG-CG iterations +-1 and solution to 1e-8."""
import hashlib
import os

import numpy as np
import pytest

import oracle as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_cube.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(FIX)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def levels():
    return O.box_levels([0, 0, 0], [1, 1, 1], [2, 2, 2], 3, 2, 3, 1, 7)


def test_fixture_layout_and_inputs(gold):
    fine = levels()[-1]
    assert int(gold["n_dofs"]) == fine.n_dofs == 14739
    assert int(gold["n_constrained"]) == len(fine.constrained()) == 867
    assert str(gold["cell_nodes_sha"]) == sha(fine.cell_nodes().astype(np.int32))
    assert str(gold["constrained_sha"]) == sha(fine.constrained().astype(np.int32))
    assert np.array_equal(fine.u0(amp=1e-3, noise=1e-5, seed=0), gold["u0"])


def test_oracle_vmult_matches_fixture(gold):
    fine = levels()[-1]
    op = O.Solid(fine, caching=1)
    op.cache(gold["u0"])
    dst = op.vmult(O.fill_uniform(1, fine.n_dofs))
    assert np.linalg.norm(dst - gold["dst"]) <= 1e-14 * np.linalg.norm(gold["dst"])


def test_oracle_mgcg_matches_fixture(gold):
    meshes = levels()
    us = [gold["u0"]]
    for l in range(len(meshes) - 1, 0, -1):
        us.insert(0, O.interpolate(meshes[l - 1], meshes[l], us[0]))
    ops = [O.Solid(m, caching=1) for m in meshes]
    for op, u in zip(ops, us):
        op.cache(u)
    xt = O.fill_uniform(2, meshes[-1].n_dofs)
    xt[meshes[-1].constrained()] = 0.0
    x, its, rel = O.mg_cg_solve(meshes, ops, ops[-1].vmult(xt), cheb_degree=4, smoothing_range=20.0,
                                eig_cg_iterations=30, coarse_kind=0, tol_rel=1e-8)
    assert its == int(gold["mg_iterations"])
    assert np.linalg.norm(x - gold["mg_x"]) <= 1e-12 * np.linalg.norm(gold["mg_x"])


@pytest.mark.gpu
def test_gpu_matches_fixture(cuda, gold):
    import torch
    import mfgpu
    meshes = [mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [2 << l] * 3, 2, 3, 1, 7) for l in range(3)]
    ops = [mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1) for m in meshes]
    trs = [mfgpu.Transfer(ops[l - 1], ops[l], mfgpu.box_children(meshes[l - 1], meshes[l])) for l in (1, 2)]
    u = torch.from_numpy(gold["u0"]).cuda()
    ug = [None, None, u]
    for l in (2, 1):
        ug[l - 1] = torch.empty(meshes[l - 1].n_dofs, dtype=torch.float64, device="cuda")
        trs[l - 1].interpolate(ug[l - 1], ug[l])
    for op, uu in zip(ops, ug):
        op.cache(uu)
    n = meshes[-1].n_dofs
    src = mfgpu.fill_uniform(1, torch.empty(n, dtype=torch.float64, device="cuda"))
    dst = torch.empty_like(src)
    ops[-1].vmult(dst, src)
    ref = gold["dst"]
    assert np.linalg.norm(dst.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)
    xt = mfgpu.fill_uniform(2, torch.empty(n, dtype=torch.float64, device="cuda"))
    xt[torch.from_numpy(meshes[-1].constrained()).long().cuda()] = 0.0
    b = torch.empty_like(xt)
    ops[-1].vmult(b, xt)
    mg = mfgpu.Multigrid(ops, trs, cheb_degree=4, smoothing_range=20.0, eig_cg_iterations=30, coarse_kind=0)
    mg.setup()
    x = torch.zeros_like(xt)
    info = mfgpu.cg_solve(ops[-1], b, x, mfgpu.PRECOND_MG, mg, tol_rel=1e-8)
    assert abs(info.iterations - int(gold["mg_iterations"])) <= 1
    gx = gold["mg_x"]
    assert np.linalg.norm(x.cpu().numpy() - gx) <= 1e-8 * np.linalg.norm(gx)
