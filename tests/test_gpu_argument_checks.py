"""Vectors the kernels would index out of bounds or misread are refused
before any launch: wrong dtype, wrong length, host memory, non-contiguous,
and dst aliasing src (the colour scatter reads dst while other workgroups
still gather src)."""
import ctypes

import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def op(cuda):
    m = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [2, 2, 2], 2, 3, 1, 7)
    A = mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1)
    A.cache(torch.zeros(m.n_dofs, dtype=torch.float64, device=cuda))
    A.compute_diagonal()
    return A


def test_rejections(op, cuda):
    n = op.m()
    good = torch.ones(n, dtype=torch.float64, device=cuda)
    dst = torch.empty_like(good)
    bad = {
        "float32": torch.ones(n, dtype=torch.float32, device=cuda),
        "short": torch.ones(n - 3, dtype=torch.float64, device=cuda),
        "host": torch.ones(n, dtype=torch.float64),
        "strided": torch.ones(2 * n, dtype=torch.float64, device=cuda)[::2],
    }
    for name, t in bad.items():
        for fn in (op.vmult, op.vmult_add, op.Tvmult, op.Tvmult_add):
            with pytest.raises(ValueError):
                fn(dst, t)
            with pytest.raises(ValueError):
                fn(t, good)
        with pytest.raises(ValueError):
            op.precondition_Jacobi(dst, t, 1.0)
        with pytest.raises(ValueError):
            mfgpu.cg_solve(op, t, dst)
    with pytest.raises(ValueError):
        op.vmult(good, good)
    with pytest.raises(ValueError):
        mfgpu.cg_solve(op, good, good)
    # the library refuses aliasing on its own (callers that bypass the wrapper)
    st = mfgpu.lib().mfg_operator_vmult(op.h, ctypes.c_void_p(good.data_ptr()), ctypes.c_void_p(good.data_ptr()))
    assert st == 1 and b"different" in mfgpu.lib().mfg_last_error()
    # and the checked calls still work
    op.vmult(dst, good)
    torch.cuda.synchronize()
    assert torch.isfinite(dst).all()
