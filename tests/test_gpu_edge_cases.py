"""Edge cases of the device operators against the oracle: a single cell
(every degree), thin and odd-sized boxes, no Dirichlet boundary, every dof
constrained (the operator is the identity), a zero source."""
import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else a
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=torch.float64)


def solid(lo, hi, n, p, face_mask, amp=5e-2):
    gm = mfgpu.BoxMesh(lo, hi, n, p, 3, face_mask, 7)
    om = O.Mesh(lo, hi, n, p, 3, face_mask, 7)
    u = om.u0(amp=amp, noise=amp / 50)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1)
    go.cache(dev(u))
    oo = O.Solid(om, caching=1)
    oo.cache(u)
    return om, go, oo


def heat(lo, hi, n, p, face_mask):
    gm = mfgpu.BoxMesh(lo, hi, n, p, 1, face_mask, 1)
    om = O.Mesh(lo, hi, n, p, 1, face_mask, 1)
    coef = np.random.default_rng(3).uniform(0.5, 2.0, om.n_cells * (p + 1) ** 3)
    return om, mfgpu.LaplaceOperator(gm, 0.1, coef), O.Heat(om, dt=0.1, coef=coef)


def check(om, go, oo, seed=1):
    x = O.fill_uniform(seed, om.n_dofs)
    y = torch.full((om.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
    go.vmult(y, dev(x))
    ref = oo.vmult(x)
    assert rel(y, ref) < TOL64
    go.compute_diagonal()
    assert rel(go.diagonal(), oo.diagonal()) < TOL64
    return y


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_single_cell(cuda, p):
    print(f"single cell p={p}", flush=True)
    check(*solid([0, 0, 0], [0.7, 0.5, 0.3], [1, 1, 1], p, 1))
    check(*heat([0, 0, 0], [0.7, 0.5, 0.3], [1, 1, 1], p, 1))


@pytest.mark.parametrize("n,p", [([5, 1, 3], 3), ([1, 7, 1], 2), ([9, 2, 1], 4)])
def test_thin_and_odd_boxes(cuda, n, p):
    print(f"box {n} p={p}", flush=True)
    # amplitude well below the smallest cell size, so that det F stays positive
    check(*solid([0, 0, 0], [1.0, 0.4, 0.6], n, p, 1 | 8, amp=2e-3))
    check(*heat([0, 0, 0], [1.0, 0.4, 0.6], n, p, 1 | 32))


def test_no_dirichlet_boundary(cuda):
    om, go, oo = solid([0, 0, 0], [1.0, 0.6, 0.4], [3, 2, 2], 2, 0)
    assert len(om.constrained()) == 0
    check(om, go, oo)
    om, go, oo = heat([0, 0, 0], [1.0, 0.6, 0.4], [3, 2, 2], 3, 0)
    check(om, go, oo)


def test_every_dof_constrained_is_the_identity(cuda):
    """A Q1 heat cell whose eight nodes all lie on Dirichlet faces: the
    operator is the identity and its diagonal is one."""
    gm = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [1, 1, 1], 1, 1, 63, 1)
    assert gm.n_constrained == gm.n_dofs == 8
    op = mfgpu.LaplaceOperator(gm, 0.1)
    x = mfgpu.fill_uniform(5, torch.empty(gm.n_dofs, dtype=torch.float64, device="cuda"))
    y = torch.empty_like(x)
    op.vmult(y, x)
    assert torch.equal(y, x)
    op.compute_diagonal()
    assert torch.equal(op.diagonal().cpu(), torch.ones(8, dtype=torch.float64))


def test_zero_source(cuda):
    om, go, oo = solid([0, 0, 0], [1.0, 0.6, 0.4], [2, 2, 2], 2, 1)
    y = torch.full((om.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
    go.vmult(y, torch.zeros(om.n_dofs, dtype=torch.float64, device="cuda"))
    assert torch.count_nonzero(y) == 0
