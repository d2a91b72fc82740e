"""Two materials by cell: the inclusion material of set_material(material,
material_inclusion) on the cells whose material id is not 0
(mf_nh_operator.h:121-128, 404, 656-658; residual mf_elasticity.h:1781-1784).
Device operators, residual and Newmark steps against the oracle with the
same material ids; bars as for one material (FP64 1e-12, FP32 1e-5)."""
import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5
# the reference builds the inclusion with 100 mu (mf_elasticity.h:470-476); the density differs too here
INCL = (100 * O.MU, O.NU, 2 * O.RHO)
BODY = (0.0, 0.0, -9.81)
LO, HI = [0, 0, 0], [1.0, 0.6, 0.4]


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else a
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def dev(x, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype)


def cell_ids(n, seed):
    ids = (np.random.default_rng(seed).uniform(size=int(np.prod(n))) < 0.4).astype(np.int32)
    ids[0], ids[-1] = 1, 0  # both materials present
    return ids


def pair(n, p, ids, form=1, caching="tensor2", precision=mfgpu.FP64, storage=None):
    gm = mfgpu.BoxMesh(LO, HI, n, p, 3, 1, 7)
    gm.set_material_ids(ids)
    om = O.Mesh(LO, HI, n, p, 3, 1, 7)
    u = om.u0(amp=5e-2, noise=1e-3)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, formulation=form, caching=caching,
                               precision=precision, inclusion=INCL)
    if storage is not None:
        go.set_point_storage(storage)
    go.cache(dev(u, go.dtype))
    oo = O.Solid(om, formulation=form, caching=mfgpu.CACHING[caching], inclusion=INCL, material_ids=ids)
    oo.cache(u)
    return gm, om, go, oo, u


@pytest.mark.parametrize("p", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("caching,form", [("tensor2", 1), ("tensor2", 0), ("scalar_referential", 1),
                                          ("tensor4", 1), ("tensor4_ns", 1)])
def test_vmult_and_diagonal(cuda, p, caching, form):
    print(f"two materials p={p} caching={caching} formulation={form}", flush=True)
    n = [3, 2, 2]
    ids = cell_ids(n, 5 + p)
    gm, om, go, oo, u = pair(n, p, ids, form, caching)
    x = O.fill_uniform(1, om.n_dofs)
    ref = oo.vmult(x)
    for sched in (mfgpu.CELLS_GATHER, mfgpu.CELLS_COLOURS):
        go.set_cell_schedule(sched)
        y = torch.full((om.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
        go.vmult(y, dev(x))
        assert rel(y, ref) < TOL64, sched
        z = dev(ref).clone()
        go.vmult_add(z, dev(x))
        assert rel(z, 2 * ref) < TOL64, sched
    go.compute_diagonal()
    assert rel(go.diagonal(), oo.diagonal()) < TOL64
    # the ids matter: the one-material operator is far from this one
    one = O.Solid(om, formulation=form, caching=mfgpu.CACHING[caching])
    one.cache(u)
    assert rel(ref, one.vmult(x)) > 1e-2


@pytest.mark.parametrize("p", [2, 4])
def test_gradient_storage_and_schedules_bitwise(cuda, p):
    """The 10-word point storage, both cell schedules: bitwise equal to each
    other (the launch order is the stored cell order) and to the oracle."""
    n = [4, 2, 2]
    ids = cell_ids(n, 21)
    gm, om, go, oo, u = pair(n, p, ids, storage=mfgpu.T2_GRADIENT)
    x = O.fill_uniform(2, om.n_dofs)
    ref = oo.vmult(x)
    out = {}
    for sched in (mfgpu.CELLS_COLOURS, mfgpu.CELLS_GATHER):
        go.set_cell_schedule(sched)
        y = torch.empty(om.n_dofs, dtype=torch.float64, device="cuda")
        go.vmult(y, dev(x))
        assert rel(y, ref) < TOL64
        out[sched] = y
    assert torch.equal(out[mfgpu.CELLS_COLOURS], out[mfgpu.CELLS_GATHER])


@pytest.mark.parametrize("p", [2, 3])
def test_fp32(cuda, p):
    n = [3, 2, 2]
    ids = cell_ids(n, 31)
    gm, om, go, oo, u = pair(n, p, ids, precision=mfgpu.FP32)
    x = O.fill_uniform(3, om.n_dofs)
    y = torch.empty(om.n_dofs, dtype=torch.float32, device="cuda")
    go.vmult(y, dev(x, torch.float32))
    assert rel(y.double(), oo.vmult(x)) < TOL32


def test_ids_without_inclusion_material_are_one_material(cuda):
    """inclusion_mu = 0: cells of any id take the first material, bit for bit
    the operator of a mesh without ids."""
    n, p = [3, 2, 2], 2
    ids = cell_ids(n, 41)
    a = mfgpu.BoxMesh(LO, HI, n, p, 3, 1, 7)
    a.set_material_ids(ids)
    b = mfgpu.BoxMesh(LO, HI, n, p, 3, 1, 7)
    u = dev(O.Mesh(LO, HI, n, p, 3, 1, 7).u0(amp=5e-2, noise=1e-3))
    x = mfgpu.fill_uniform(4, torch.empty(a.n_dofs, dtype=torch.float64, device="cuda"))
    ys = []
    for m in (a, b):
        op = mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1)
        op.cache(u)
        y = torch.empty_like(x)
        op.vmult(y, x)
        ys.append(y)
    assert torch.equal(ys[0], ys[1])


@pytest.mark.parametrize("form,p", [(1, 2), (0, 2), (1, 4)])
def test_residual(cuda, form, p):
    n = [3, 2, 2]
    ids = cell_ids(n, 51)
    gm, om, go, oo, u = pair(n, p, ids, form)
    acc = O.fill_uniform(4, om.n_dofs)
    R, nrm = mfgpu.solid_residual(go, dev(u), dev(acc), BODY)
    ref = oo.residual(u, acc, BODY)
    assert np.linalg.norm(R.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)
    assert abs(nrm - np.linalg.norm(ref)) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("level_precision", [mfgpu.FP64, mfgpu.FP32])
def test_newmark_steps(cuda, level_precision):
    """An inclusion set on the coarse mesh and inherited by the refined
    levels (deal.II's refinement keeps a cell's material id): two Newmark
    steps, Newton counts equal to the oracle's and the same state."""
    coarse_n, L, p = [2, 1, 1], 3, 2
    lo, hi = [0, 0, 0], [2.0, 1.0, 1.0]
    coarse_ids = np.array([0, 1], dtype=np.int32)
    gms, oops = [], []
    oms = O.box_levels(lo, hi, coarse_n, L, p, 3)
    for l in range(L):
        n = [c << l for c in coarse_n]
        ids = mfgpu.refined_material_ids(coarse_ids, coarse_n, l)
        g = mfgpu.BoxMesh(lo, hi, n, p, 3, 1, 7)
        g.set_material_ids(ids)
        gms.append(g)
        oops.append(O.Solid(oms[l], caching=1, inclusion=INCL, material_ids=ids))
    kw = dict(coarse_kind=1 if level_precision == mfgpu.FP32 else 0)
    prob = mfgpu.SolidProblem(gms, O.MU, O.NU, O.RHO, body=BODY, level_precision=level_precision, tol_f=1e-8,
                              inclusion=INCL, **kw)
    u = oms[-1].u0(amp=1e-3, noise=1e-5)
    v = np.zeros_like(u)
    a = np.zeros_like(u)
    prob.set_state(dev(u), dev(v), dev(a))
    for step in range(2):
        info = prob.step()
        it, li = O.newmark_step(oms, oops, u, v, a, body=BODY, tol_f=1e-8, **kw)
        gu, gv, ga = (x.cpu().numpy() for x in prob.get_state())
        print(f"step {step}: newton gpu {info.newton_iterations} oracle {it}; linear gpu {info.linear_iterations} "
              f"oracle {li}", flush=True)
        assert info.converged and info.newton_iterations == it
        assert np.linalg.norm(gu - u) <= 1e-7 * np.linalg.norm(u)
        R = oops[-1].residual(gu, ga, BODY)
        assert np.linalg.norm(R) <= 1e-8 * info.residual_0 * 10


def test_batches_refuse_two_material_operators(cuda):
    n, p = [2, 2, 2], 2
    gm, om, go, oo, u = pair(n, p, cell_ids(n, 61))
    with pytest.raises(mfgpu.MfgError):
        mfgpu.Batch([go])


def test_bad_inclusion_material_is_refused(cuda):
    m = mfgpu.BoxMesh(LO, HI, [2, 2, 2], 2, 3, 1, 7)
    with pytest.raises(mfgpu.MfgError):
        mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1, inclusion=(-1.0, 0.3, 1.0))
    with pytest.raises(ValueError):
        m.set_material_ids(np.zeros(3, dtype=np.int32))
