"""mfg_cg_solve_noalloc: the allocation-free CG (scratch in a file-scope
device array) returns what mfg_cg_solve returns on the same operator and
preconditioner -- iterations, residuals and solution bitwise -- on two calls
in a row, the second with a smaller operator, and allocates no device memory
after its set-up."""
import ctypes

import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu


def _noalloc():
    f = mfgpu.lib().mfg_cg_solve_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_void_p]
    return f


def _problem(n, p, dev):
    lo, hi = [0, 0, 0], [1, 1, 1]
    meshes = [mfgpu.BoxMesh(lo, hi, [c >> (1 - l) for c in n], p, 3, 1, 7) for l in range(2)]
    ops = [mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1) for m in meshes]
    u = torch.from_numpy(meshes[-1].u0(amp=1e-3, noise=1e-5, seed=0)).to(dev)
    ops[-1].cache(u)
    tr = mfgpu.Transfer(ops[0], ops[1], mfgpu.box_children(meshes[0], meshes[1]))
    uc = torch.empty(meshes[0].n_dofs, dtype=torch.float64, device=dev)
    tr.interpolate(uc, u)
    ops[0].cache(uc)
    for op in ops:
        op.compute_diagonal()
    mg = mfgpu.Multigrid(ops, [tr], cheb_degree=4)
    mg.setup()
    xt = mfgpu.fill_uniform(2, torch.empty(meshes[-1].n_dofs, dtype=torch.float64, device=dev))
    xt[torch.from_numpy(meshes[-1].constrained()).long().to(dev)] = 0.0
    b = torch.empty_like(xt)
    ops[-1].vmult(b, xt)
    return ops[-1], mg, b, (meshes, ops, tr)


def _solve(fn, A, pre, mg, b, x):
    x.zero_()
    info = mfgpu.SolveInfo()
    st = fn(A.h, pre, None if mg is None else mg.h, 1.0, ctypes.c_void_p(b.data_ptr()),
            ctypes.c_void_p(x.data_ptr()), 1e-8, 1000, ctypes.byref(info))
    torch.cuda.synchronize()
    assert st == 0, mfgpu.lib().mfg_last_error()
    return x, info


def test_noalloc_cg_equals_cg_solve(cuda):
    """The first call on the device is its set-up: it looks up the scratch
    array's device address, which loads the module that holds the array
    (the array's memory then stays with the process).  No call after it
    changes the device's free memory."""
    f = _noalloc()
    ref_fn = mfgpu.lib().mfg_cg_solve
    first = True
    for n, p in (([8, 8, 8], 2), ([4, 4, 4], 2)):  # the second call is smaller
        A, mg, b, keep = _problem(n, p, cuda)
        for pre, m in ((mfgpu.PRECOND_JACOBI, None), (mfgpu.PRECOND_MG, mg)):
            x_ref, i_ref = _solve(ref_fn, A, pre, m, b, torch.empty_like(b))
            x = torch.empty_like(b)
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            x, info = _solve(f, A, pre, m, b, x)
            free1 = torch.cuda.mem_get_info()[0]
            print(f"n={n} p={p} precond={pre}: {info.iterations} iterations, "
                  f"residual {info.final_residual:.3e}, free bytes {free0} -> {free1}", flush=True)
            assert first or free1 == free0, "the allocation-free CG allocated device memory"
            first = False
            assert info.iterations == i_ref.iterations > 0
            assert info.final_residual == i_ref.final_residual
            assert info.initial_residual == i_ref.initial_residual
            assert info.condition_estimate == i_ref.condition_estimate
            assert torch.equal(x, x_ref)
            r = torch.empty_like(b)
            A.vmult(r, x)
            assert float(torch.linalg.norm(r - b) / torch.linalg.norm(b)) <= 1e-8 * 1.0001


def test_noalloc_cg_rejects_operators_beyond_the_scratch(cuda):
    f = _noalloc()
    m = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [24, 24, 24], 3, 3, 1, 7)  # 73^3 * 3 > 2^20 rows
    assert m.n_dofs > (1 << 20)
    A = mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1)
    A.cache(torch.zeros(m.n_dofs, dtype=torch.float64, device=cuda))
    A.compute_diagonal()
    b = torch.ones(m.n_dofs, dtype=torch.float64, device=cuda)
    x = torch.zeros_like(b)
    st = f(A.h, mfgpu.PRECOND_JACOBI, None, 1.0, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
           1e-8, 10, None)
    assert st == 1 and b"2^20" in mfgpu.lib().mfg_last_error()
