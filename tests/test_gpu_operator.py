"""Device operator parity: libmfgpu (through the C ABI) against the oracle on
identical meshes and seeded inputs.  Bars: FP64 vmult relative l2 <= 1e-12,
FP32 operators relative l2 <= 1e-5 (single-precision arithmetic, ~1e-7 per
operation accumulated over the sum-factorisation sweeps)."""
import numpy as np
import pytest
import torch

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else a
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def dev(x, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype)


def solid_pair(lo, hi, n, p, form=1, face_mask=1, amp=5e-2, precision=mfgpu.FP64):
    gm = mfgpu.BoxMesh(lo, hi, n, p, 3, face_mask, 7)
    om = O.Mesh(lo, hi, n, p, 3, face_mask, 7)
    u = om.u0(amp=amp, noise=1e-3)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, formulation=form, precision=precision)
    go.cache(dev(u, go.dtype))
    oo = O.Solid(om, formulation=form, caching=1)
    oo.cache(u)
    return gm, om, go, oo, u


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_solid_vmult_fp64_all_degrees(cuda, p):
    print(f"solid vmult p={p}", flush=True)
    gm, om, go, oo, u = solid_pair([0, 0, 0], [1.0, 0.6, 0.4], [3, 2, 2], p)
    x = O.fill_uniform(1, om.n_dofs)
    y = torch.empty(om.n_dofs, dtype=torch.float64, device="cuda")
    go.vmult(y, dev(x))
    ref = oo.vmult(x)
    assert rel(y, ref) < TOL64
    # constrained rows are exactly identity
    con = om.constrained()
    assert np.array_equal(y.cpu().numpy()[con], x[con])


def folded_tensor2(ref, with_jxw_ref):
    """The device's tensor2 (formulation 1) fields from the reference's 18:
    Jinv_cur, JxW c2 / 2, JxW T (and JxW c1 = JxW_ref on non-affine meshes)."""
    jxw, c2, c1, T = ref[9], ref[10], ref[11], ref[12:18]
    out = list(ref[:9]) + [jxw * c2 / 2] + [jxw * t for t in T]
    return out + [jxw * c1] if with_jxw_ref else out


@pytest.mark.parametrize("p", [2, 4])
def test_solid_cache_matches_oracle(cuda, p):
    gm, om, go, oo, u = solid_pair([0, 0, 0], [1.0, 0.6, 0.4], [2, 2, 1], p, amp=1e-1)
    # the gradient storage: Grad_xi u and s = mu - 2 lambda ln J per point.
    # (J_ref + Grad_xi u)^-1 is the oracle's current-frame Jinv, and s is its
    # c2 / (2 c1) with c2 = 2 s / J, c1 = 1 / J.
    assert go.point_storage() == mfgpu.T2_GRADIENT  # the default on all-affine meshes
    go.set_point_storage(mfgpu.T2_GRADIENT)  # no change: the cache stays
    go.cache(dev(u))
    assert go.n_fields() == 10
    g = go.qp_data().cpu().numpy()
    ref18 = oo.tensor2_data()
    Jref = np.diag([0.5, 0.3, 0.4])
    Jc = Jref[None, :, :] + g[:9].T.reshape(-1, 3, 3)
    Jinv = np.linalg.inv(Jc).reshape(-1, 9).T
    for f in range(9):
        assert np.abs(Jinv[f] - ref18[f]).max() <= 1e-12 * np.abs(ref18[f]).max(), f
    s_ref = ref18[10] / (2 * ref18[11])
    assert np.abs(g[9] - s_ref).max() <= 1e-12 * np.abs(s_ref).max()
    # the folded storage: 16 words, JxW_ref per cell
    go.set_point_storage(mfgpu.T2_FOLDED)
    go.cache(dev(u))
    assert go.n_fields() == 16
    qd = go.qp_data().cpu().numpy()
    ref = folded_tensor2(oo.tensor2_data(), False)
    for f in range(16):
        scale = np.abs(ref[f]).max()
        assert np.abs(qd[f] - ref[f]).max() <= 1e-12 * scale, f


class DistortedMesh:
    """A box mesh's level description with other cell vertices."""

    def __init__(self, base, vertices):
        self.base, self.vert = base, np.ascontiguousarray(vertices, dtype=np.float64)
        self.cn, self.con = base.cell_nodes(), base.constrained()
        d = mfgpu.LevelDesc()
        d.degree, d.n_q_1d, d.n_components = base.degree, base.degree + 1, base.n_components
        d.n_cells, d.n_nodes = base.n_cells, base.n_nodes
        d.cell_nodes = self.cn.ctypes.data_as(mfgpu._c_i32p)
        d.cell_vertices = self.vert.ctypes.data_as(mfgpu._c_dp)
        d.constrained_dofs = self.con.ctypes.data_as(mfgpu._c_i32p)
        d.n_constrained = len(self.con)
        self.desc = d
        self.n_cells, self.n_nodes, self.n_dofs, self.degree = base.n_cells, base.n_nodes, base.n_dofs, base.degree


def distort(v, eps=0.08):
    v = v.reshape(-1, 3)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    return (v + eps * np.stack([np.sin(np.pi * y) * z, np.sin(np.pi * x) * z, x * y], axis=1)).ravel()


@pytest.mark.parametrize("p,precision", [(2, mfgpu.FP64), (4, mfgpu.FP64), (3, mfgpu.FP32)])
def test_solid_tensor2_non_affine_cells(cuda, p, precision):
    """Vertices moved by a smooth map: no cell is affine, the tensor2 cache
    keeps JxW_ref per point (17 fields); vmult, diagonal and cache agree with
    the oracle on the same distorted geometry."""
    print(f"non-affine tensor2 p={p} precision={precision}", flush=True)
    lo, hi, n = [0, 0, 0], [1.0, 0.6, 0.4], [3, 2, 2]
    om = O.Mesh(lo, hi, n, p, 3, 1, 7)
    v = distort(om.cell_vertices())
    om.set_cell_vertices(v)
    gm = DistortedMesh(mfgpu.BoxMesh(lo, hi, n, p, 3, 1, 7), v)
    u = om.u0(amp=5e-2, noise=1e-3)
    oo = O.Solid(om, caching=1)
    oo.cache(u)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, precision=precision)
    go.cache(dev(u, go.dtype))
    assert go.n_fields() == 17
    x = O.fill_uniform(5, om.n_dofs)
    y = torch.empty(om.n_dofs, dtype=go.dtype, device="cuda")
    go.vmult(y, dev(x, go.dtype))
    tol = TOL64 if precision == mfgpu.FP64 else TOL32
    assert rel(y.double(), oo.vmult(x)) < tol
    go.compute_diagonal()
    assert rel(go.diagonal().double(), oo.diagonal()) < tol
    if precision == mfgpu.FP64:
        qd = go.qp_data().cpu().numpy()
        ref = folded_tensor2(oo.tensor2_data(), True)
        for f in range(17):
            assert np.abs(qd[f] - ref[f]).max() <= 1e-12 * np.abs(ref[f]).max(), f


@pytest.mark.parametrize("p", [1, 2, 4])
def test_solid_formulation0(cuda, p):
    gm, om, go, oo, u = solid_pair([0, 0, 0], [1, 1, 1], [2, 2, 2], p, form=0)
    x = O.fill_uniform(2, om.n_dofs)
    y = torch.empty_like(dev(x))
    go.vmult(y, dev(x))
    assert rel(y, oo.vmult(x)) < TOL64


def test_small_config_vmult_and_diagonal(cuda):
    """8^3 Q2 unit cube, 14,739 dofs, u0 = 1e-3 sin sin sin + 1e-5 noise."""
    gm = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [8, 8, 8], 2, 3, 1, 7)
    om = O.Mesh([0, 0, 0], [1, 1, 1], [8, 8, 8], 2, 3, 1, 7)
    u = om.u0(amp=1e-3, noise=1e-5, seed=0)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1)
    go.cache(dev(u))
    oo = O.Solid(om, caching=1)
    oo.cache(u)
    x = O.fill_uniform(1, om.n_dofs)
    y = torch.empty(om.n_dofs, dtype=torch.float64, device="cuda")
    go.vmult(y, dev(x))
    assert rel(y, oo.vmult(x)) < TOL64
    go.compute_diagonal()
    d = go.diagonal()
    dref = oo.diagonal()
    assert rel(d, dref) < TOL64
    inv = go.get_matrix_diagonal_inverse().cpu().numpy()
    assert np.abs(inv * dref - 1).max() < 1e-13
    assert go.el(17, 17) == pytest.approx(dref[17], rel=1e-13)
    with pytest.raises(mfgpu.MfgError):
        go.el(1, 2)


def test_vmult_add_tvmult_jacobi(cuda):
    gm, om, go, oo, u = solid_pair([0, 0, 0], [1, 1, 1], [3, 3, 2], 2)
    x = O.fill_uniform(3, om.n_dofs)
    z = O.fill_uniform(4, om.n_dofs)
    ref = oo.vmult(x)
    y = dev(z)
    go.vmult_add(y, dev(x))
    assert rel(y, z + ref) < TOL64
    y2 = torch.empty_like(y)
    go.Tvmult(y2, dev(x))
    assert rel(y2, ref) < TOL64
    y3 = dev(z)
    go.Tvmult_add(y3, dev(x))
    assert rel(y3, z + ref) < TOL64
    go.compute_diagonal()
    j = torch.empty_like(y)
    go.precondition_Jacobi(j, dev(x), 0.65)
    assert rel(j, 0.65 * x / oo.diagonal()) < 1e-14
    assert go.m() == go.n() == om.n_dofs
    assert go.memory_consumption() >= go.n_fields() * 8 * om.n_cells * 27


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_heat_vmult_and_diagonal(cuda, p):
    print(f"heat p={p}", flush=True)
    lo, hi, n = [0, 0, 0], [1.0, 2.0, 0.5], [3, 2, 2]
    gm = mfgpu.BoxMesh(lo, hi, n, p, 1, 63, 1)
    om = O.Mesh(lo, hi, n, p, 1, 63, 1)
    rng = np.random.default_rng(p)
    coef = rng.uniform(0.5, 2.0, om.n_cells * (p + 1) ** 3)
    go = mfgpu.LaplaceOperator(gm, 0.1, coef)
    oo = O.Heat(om, dt=0.1, coef=coef)
    x = O.fill_uniform(5, om.n_dofs)
    y = torch.empty(om.n_dofs, dtype=torch.float64, device="cuda")
    go.vmult(y, dev(x))
    assert rel(y, oo.vmult(x)) < TOL64
    go.compute_diagonal()
    assert rel(go.diagonal(), oo.diagonal()) < TOL64
    assert go.n_fields() == 1  # affine cells: geometry per cell, the coefficient per point


@pytest.mark.parametrize("p,precision", [(2, mfgpu.FP64), (4, mfgpu.FP64), (3, mfgpu.FP32)])
def test_heat_non_affine_cells(cuda, p, precision):
    """Distorted vertices: the heat op keeps Jinv, JxW and c per point (11 fields)."""
    print(f"non-affine heat p={p} precision={precision}", flush=True)
    lo, hi, n = [0, 0, 0], [1.0, 0.8, 0.6], [3, 2, 2]
    om = O.Mesh(lo, hi, n, p, 1, 63, 1)
    v = distort(om.cell_vertices())
    om.set_cell_vertices(v)
    gm = DistortedMesh(mfgpu.BoxMesh(lo, hi, n, p, 1, 63, 1), v)
    coef = np.random.default_rng(p).uniform(0.5, 2.0, om.n_cells * (p + 1) ** 3)
    go = mfgpu.LaplaceOperator(gm, 0.1, coef, precision=precision)
    oo = O.Heat(om, dt=0.1, coef=coef)
    assert go.n_fields() == 11
    x = O.fill_uniform(6, om.n_dofs)
    y = torch.empty(om.n_dofs, dtype=go.dtype, device="cuda")
    go.vmult(y, dev(x, go.dtype))
    tol = TOL64 if precision == mfgpu.FP64 else TOL32
    assert rel(y.double(), oo.vmult(x)) < tol
    go.compute_diagonal()
    assert rel(go.diagonal().double(), oo.diagonal()) < tol


@pytest.mark.parametrize("p", [2, 4])
def test_fp32_operators(cuda, p):
    gm, om, go, oo, u = solid_pair([0, 0, 0], [1, 1, 1], [2, 2, 2], p, precision=mfgpu.FP32)
    x = O.fill_uniform(1, om.n_dofs)
    y = torch.empty(om.n_dofs, dtype=torch.float32, device="cuda")
    go.vmult(y, dev(x, torch.float32))
    assert rel(y.double(), oo.vmult(x)) < TOL32
    go.compute_diagonal()
    assert rel(go.diagonal().double(), oo.diagonal()) < TOL32
    hm = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [3, 3, 3], p, 1, 1, 1)
    hg = mfgpu.LaplaceOperator(hm, 0.1, None, precision=mfgpu.FP32)
    ho = O.Heat(O.Mesh([0, 0, 0], [1, 1, 1], [3, 3, 3], p, 1, 1, 1), dt=0.1)
    xh = O.fill_uniform(2, hm.n_dofs)
    yh = torch.empty(hm.n_dofs, dtype=torch.float32, device="cuda")
    hg.vmult(yh, dev(xh, torch.float32))
    assert rel(yh.double(), ho.vmult(xh)) < TOL32


def test_ragged_and_single_cell_meshes(cuda):
    for lo, hi, n, p, fm in [([0, 0, 0], [1, 1, 1], [1, 1, 1], 3, 0),
                             ([0, 0, 0], [2, 1, 3], [5, 1, 3], 2, 63),
                             ([0, 0, 0], [1, 1, 1], [7, 1, 1], 1, 1)]:
        gm, om, go, oo, u = solid_pair(lo, hi, n, p, face_mask=fm)
        x = O.fill_uniform(9, om.n_dofs)
        y = torch.empty(om.n_dofs, dtype=torch.float64, device="cuda")
        go.vmult(y, dev(x))
        assert rel(y, oo.vmult(x)) < TOL64


def test_colouring_and_bit_reproducibility(cuda):
    gm, om, go, oo, u = solid_pair([0, 0, 0], [1, 1, 1], [5, 4, 3], 2)
    assert go.n_colours() == 8
    x = dev(O.fill_uniform(3, om.n_dofs))
    y1, y2 = torch.empty_like(x), torch.empty_like(x)
    go.vmult(y1, x)
    go.vmult(y2, x)
    assert torch.equal(y1, y2)
    one = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [1, 1, 1], 2, 3, 1, 7)
    assert mfgpu.NeoHookOperator(one, O.MU, O.NU, O.RHO, O.ALPHA1).n_colours() == 1


ORACLE_CACHING = {"scalar_referential": 0, "tensor2": 1, "tensor4": 2, "tensor4_ns": 3}


@pytest.mark.parametrize("caching", ["scalar_referential", "tensor4", "tensor4_ns"])
@pytest.mark.parametrize("p", [1, 2, 4])
def test_solid_other_cachings(cuda, caching, p):
    """The other MF cachings against the oracle's implementation of the same
    caching, against the device tensor2 operator (all four define one
    operator), diagonal included; FP32 to the FP32 bar."""
    print(f"caching {caching} p={p}", flush=True)
    lo, hi, n = [0, 0, 0], [1.0, 0.6, 0.4], [3, 2, 2]
    gm = mfgpu.BoxMesh(lo, hi, n, p, 3, 1, 7)
    om = O.Mesh(lo, hi, n, p, 3, 1, 7)
    u = om.u0(amp=5e-2, noise=1e-3)
    oo = O.Solid(om, caching=ORACLE_CACHING[caching])
    oo.cache(u)
    x = O.fill_uniform(3, om.n_dofs)
    ref = oo.vmult(x)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, caching=caching)
    go.cache(dev(u))
    y = torch.empty(om.n_dofs, dtype=torch.float64, device="cuda")
    go.vmult(y, dev(x))
    assert rel(y, ref) < TOL64
    t2 = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, caching="tensor2")
    t2.cache(dev(u))
    y2 = torch.empty_like(y)
    t2.vmult(y2, dev(x))
    assert rel(y, y2.cpu().numpy()) < TOL64
    go.compute_diagonal()
    assert rel(go.diagonal(), oo.diagonal()) < TOL64
    g32 = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, caching=caching, precision=mfgpu.FP32)
    g32.cache(dev(u, torch.float32))
    y32 = torch.empty(om.n_dofs, dtype=torch.float32, device="cuda")
    g32.vmult(y32, dev(x, torch.float32))
    assert rel(y32.double(), ref) < TOL32


@pytest.mark.parametrize("p,n", [(1, [4, 3, 2]), (2, [5, 2, 3]), (3, [3, 3, 3]), (4, [4, 2, 2])])
def test_cell_schedules(cuda, p, n):
    """The GATHER schedule against the colour schedule (bitwise) and the
    oracle; vmult_add too; repeated applications are deterministic."""
    gm, om, go, oo, u = solid_pair([0, 0, 0], [1.0, 0.6, 0.4], n, p)
    x = O.fill_uniform(4, om.n_dofs)
    ref = oo.vmult(x)
    out = {}
    for sched in (mfgpu.CELLS_COLOURS, mfgpu.CELLS_GATHER):
        go.set_cell_schedule(sched)
        assert go.cell_schedule() == sched
        y = torch.full((om.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
        go.vmult(y, dev(x))
        assert rel(y, ref) < TOL64, sched
        z = dev(ref).clone()
        go.vmult_add(z, dev(x))
        assert rel(z, 2 * ref) < TOL64, sched
        out[sched] = y
        y2 = torch.empty_like(y)
        go.vmult(y2, dev(x))
        assert torch.equal(y, y2)  # deterministic
    assert torch.equal(out[mfgpu.CELLS_GATHER], out[mfgpu.CELLS_COLOURS])
    with pytest.raises(mfgpu.MfgError):  # the removed segment schedules are refused
        go.set_cell_schedule(1)


@pytest.mark.parametrize("kind,p,precision,caching,form", [
    ("solid", 2, mfgpu.FP64, "tensor2", 1), ("solid", 4, mfgpu.FP64, "tensor2", 1), ("solid", 3, mfgpu.FP32, "tensor2", 1),
    ("solid", 2, mfgpu.FP64, "tensor2", 0), ("solid", 3, mfgpu.FP64, "scalar_referential", 1),
    ("solid", 2, mfgpu.FP64, "tensor4", 1), ("heat", 3, mfgpu.FP64, None, 1), ("heat", 5, mfgpu.FP32, None, 1)])
def test_gather_schedule_equals_colours(cuda, kind, p, precision, caching, form):
    """GATHER sums each node's cell contributions in colour order: vmult and
    vmult_add are bitwise the colour schedule's, Dirichlet rows included."""
    print(f"gather schedule {kind} p={p} precision={precision} caching={caching} formulation={form}", flush=True)
    n = [3, 2, 3]
    if kind == "solid":
        m = mfgpu.BoxMesh([0, 0, 0], [1.0, 0.6, 0.4], n, p, 3, 1 | 8, 7, numbering="cellwise")
        op = mfgpu.NeoHookOperator(m, O.MU, O.NU, O.RHO, O.ALPHA1, formulation=form, caching=caching,
                                   precision=precision)
        op.cache(torch.from_numpy(m.u0(amp=1e-2, noise=1e-5)).to("cuda", op.dtype))
    else:
        m = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], n, p, 1, 63, 1)
        coef = np.random.default_rng(p).uniform(0.5, 2.0, m.n_cells * (p + 1) ** 3)
        op = mfgpu.LaplaceOperator(m, 0.1, coef, precision=precision)
    x = mfgpu.fill_uniform(7, torch.empty(m.n_dofs, dtype=op.dtype, device="cuda"))
    z0 = mfgpu.fill_uniform(8, torch.empty_like(x))
    res = {}
    for sched in (mfgpu.CELLS_COLOURS, mfgpu.CELLS_GATHER, mfgpu.CELLS_COLOURS):
        op.set_cell_schedule(sched)
        assert op.cell_schedule() == sched
        y = torch.full_like(x, float("nan"))
        op.vmult(y, x)
        z = z0.clone()
        op.vmult_add(z, x)
        if sched in res:
            assert torch.equal(y, res[sched][0]) and torch.equal(z, res[sched][1])
        res[sched] = (y, z)
    assert torch.equal(res[mfgpu.CELLS_GATHER][0], res[mfgpu.CELLS_COLOURS][0])
    assert torch.equal(res[mfgpu.CELLS_GATHER][1], res[mfgpu.CELLS_COLOURS][1])
    con = torch.from_numpy(m.constrained().astype(np.int64)).cuda()
    assert torch.equal(res[mfgpu.CELLS_GATHER][0][con], x[con])


def test_error_paths(cuda):
    gm = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], [2, 2, 2], 2, 3, 1, 7)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1)
    y = torch.zeros(gm.n_dofs, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(y)
    with pytest.raises(mfgpu.MfgError) as e:
        go.vmult(y, x)  # before cache()
    assert e.value.code == 5
    # det F <= 0: a displacement that inverts cells
    om = O.Mesh([0, 0, 0], [1, 1, 1], [2, 2, 2], 2, 3, 1, 7)
    X = om.node_coords()
    u = np.zeros(om.n_dofs)
    u[0::3] = -3.0 * X[:, 0]
    with pytest.raises(mfgpu.MfgError) as e:
        go.cache(dev(u))
    assert e.value.code == 3 and "det_F" in str(e.value)
    with pytest.raises(mfgpu.MfgError):  # formulation 0 exists for tensor2 only (mf_nh_operator.h:706)
        mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, formulation=0, caching="tensor4")


@pytest.mark.parametrize("p,precision,n", [(1, mfgpu.FP64, [4, 3, 2]), (2, mfgpu.FP64, [3, 2, 3]),
                                           (4, mfgpu.FP64, [4, 2, 2]), (5, mfgpu.FP64, [2, 2, 2]),
                                           (3, mfgpu.FP32, [3, 3, 3]), (4, mfgpu.FP32, [4, 2, 2])])
def test_gradient_point_storage(cuda, p, precision, n):
    """The 10-word tensor2 storage (Grad_xi u and s per point, the rest rebuilt
    from per-cell J_ref) gives the folded layout's operator: vmult and diagonal
    against the oracle at the usual bars and against the folded storage to
    1e-14 (FP64), on every cell schedule; switching back reproduces the folded
    results bit for bit."""
    print(f"gradient point storage p={p} precision={precision} n={n}", flush=True)
    lo, hi = [0, 0, 0], [1.0, 0.6, 0.4]
    gm = mfgpu.BoxMesh(lo, hi, n, p, 3, 1 | 8, 7, numbering="cellwise")
    lex = mfgpu.BoxMesh(lo, hi, n, p, 3, 1 | 8, 7)
    om = O.Mesh(lo, hi, n, p, 3, 1 | 8, 7)
    perm = np.full(lex.n_nodes, -1, dtype=np.int64)
    perm[lex.cell_nodes()] = gm.cell_nodes()
    dof = (perm[:, None] * 3 + np.arange(3)[None, :]).ravel()  # cellwise dof of each lexicographic dof
    u = om.u0(amp=2e-2, noise=1e-4)
    oo = O.Solid(om, caching=1)
    oo.cache(u)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, precision=precision)
    ug = np.empty(om.n_dofs)
    ug[dof] = u
    x = O.fill_uniform(5, om.n_dofs)
    xg = np.empty(om.n_dofs)
    xg[dof] = x
    tol = TOL64 if precision == mfgpu.FP64 else TOL32
    ref, dref = oo.vmult(x), oo.diagonal()
    res = {}
    for storage in (mfgpu.T2_FOLDED, mfgpu.T2_GRADIENT, mfgpu.T2_FOLDED):
        go.set_point_storage(storage)
        assert go.point_storage() == storage
        go.cache(dev(ug, go.dtype))
        assert go.n_fields() == (10 if storage == mfgpu.T2_GRADIENT else 16)
        go.compute_diagonal()
        d = go.diagonal()
        assert rel(d.double().cpu().numpy()[dof], dref) < tol
        ys = {}
        for sched in (mfgpu.CELLS_COLOURS, mfgpu.CELLS_GATHER):
            go.set_cell_schedule(sched)
            y = torch.full((om.n_dofs,), float("nan"), dtype=go.dtype, device="cuda")
            go.vmult(y, dev(xg, go.dtype))
            assert rel(y.double().cpu().numpy()[dof], ref) < tol, sched
            ys[sched] = y
        if storage in res:
            assert torch.equal(d, res[storage][0])
            assert all(torch.equal(ys[k], res[storage][1][k]) for k in ys)
        res[storage] = (d, ys)
    if precision == mfgpu.FP64:
        fd, fy = res[mfgpu.T2_FOLDED]
        gd, gy = res[mfgpu.T2_GRADIENT]
        assert rel(gd, fd.cpu().numpy()) < 1e-14
        for k in fy:
            assert rel(gy[k], fy[k].cpu().numpy()) < 1e-14, k


@pytest.mark.parametrize("p,precision", [(3, mfgpu.FP64), (4, mfgpu.FP64), (4, mfgpu.FP32), (2, mfgpu.FP64)])
def test_gradient_point_storage_sheared_affine_cells(cuda, p, precision):
    """An affine map with off-diagonal terms (shear and tilt) of the box: every
    cell affine, no reference Jacobian diagonal, so the 10-word storage takes
    its general per-cell tables (boxes take the diagonal-only form).  vmult
    and diagonal against the oracle on the same geometry; colours and GATHER
    bitwise equal."""
    print(f"sheared affine 10-word p={p} precision={precision}", flush=True)
    lo, hi, n = [0, 0, 0], [1.0, 0.6, 0.4], [3, 2, 2]
    A = np.array([[1.0, 0.3, 0.1], [0.05, 1.0, 0.2], [0.1, -0.15, 1.0]])
    om = O.Mesh(lo, hi, n, p, 3, 1, 7)
    v = (om.cell_vertices().reshape(-1, 3) @ A.T).ravel()
    om.set_cell_vertices(v)
    gm = DistortedMesh(mfgpu.BoxMesh(lo, hi, n, p, 3, 1, 7), v)
    u = om.u0(amp=5e-2, noise=1e-3)
    oo = O.Solid(om, caching=1)
    oo.cache(u)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, precision=precision)
    go.set_point_storage(mfgpu.T2_GRADIENT)
    go.cache(dev(u, go.dtype))
    assert go.n_fields() == 10
    tol = TOL64 if precision == mfgpu.FP64 else TOL32
    x = O.fill_uniform(5, om.n_dofs)
    ref = oo.vmult(x)
    ys = []
    for sched in (mfgpu.CELLS_COLOURS, mfgpu.CELLS_GATHER):
        go.set_cell_schedule(sched)
        y = torch.full((om.n_dofs,), float("nan"), dtype=go.dtype, device="cuda")
        go.vmult(y, dev(x, go.dtype))
        assert rel(y.double(), ref) < tol, sched
        ys.append(y)
    assert torch.equal(ys[0], ys[1])
    go.compute_diagonal()
    assert rel(go.diagonal().double(), oo.diagonal()) < tol


def test_gradient_point_storage_errors(cuda):
    lo, hi, n, p = [0, 0, 0], [1.0, 0.6, 0.4], [3, 2, 2], 2
    gm = mfgpu.BoxMesh(lo, hi, n, p, 3, 1, 7)
    go = mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1)
    assert go.point_storage() == mfgpu.T2_GRADIENT  # the default on all-affine meshes
    go.set_point_storage(mfgpu.T2_FOLDED)
    u = torch.from_numpy(gm.u0(amp=1e-2, noise=1e-5)).cuda()
    go.cache(u)
    mem16 = go.memory_consumption()
    go.set_point_storage(mfgpu.T2_GRADIENT)
    y = torch.zeros(gm.n_dofs, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(y)
    with pytest.raises(mfgpu.MfgError) as e:  # the change discards the cache
        go.vmult(y, x)
    assert e.value.code == 5
    go.cache(u)
    assert go.memory_consumption() < mem16
    with pytest.raises(mfgpu.MfgError):
        go.set_point_storage(7)
    # non-affine cells keep the folded storage
    om = O.Mesh(lo, hi, n, p, 3, 1, 7)
    dm = DistortedMesh(gm, distort(om.cell_vertices()))
    dop = mfgpu.NeoHookOperator(dm, O.MU, O.NU, O.RHO, O.ALPHA1)
    assert dop.point_storage() == mfgpu.T2_FOLDED
    with pytest.raises(mfgpu.MfgError):
        dop.set_point_storage(mfgpu.T2_GRADIENT)
    # other cachings, formulation 0 and heat have no such storage
    for kw in ({"caching": "tensor4"}, {"formulation": 0}):
        with pytest.raises(mfgpu.MfgError):
            mfgpu.NeoHookOperator(gm, O.MU, O.NU, O.RHO, O.ALPHA1, **kw).set_point_storage(mfgpu.T2_GRADIENT)
    with pytest.raises(mfgpu.MfgError):
        mfgpu.LaplaceOperator(mfgpu.BoxMesh(lo, hi, n, p, 1, 63, 1), 0.1).set_point_storage(mfgpu.T2_GRADIENT)
