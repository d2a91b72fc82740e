"""Batched instances through the C ABI: one launch applies many uneven heat
and solid operators (degrees 1..6, FP64 and FP32), the batched V-cycle and the
lockstep CG.  Every instance performs the operation sequence of its own
per-instance call, so the bar is equality with the per-instance results
(bitwise, with the per-instance operators on the colour cell schedule that
the batched kernels use; the per-instance paths are themselves pinned to the
oracle by the operator and multigrid tests)."""
import numpy as np
import pytest
import torch

import mfgpu

pytestmark = pytest.mark.gpu


def make_ops(specs, precision, storage=mfgpu.T2_FOLDED):
    """specs: (kind, n_cells, degree, formulation); solids cached at u0, the
    formulation-1 ones with the given tensor2 point storage."""
    ops, meshes = [], []
    for k, (kind, n, p, form) in enumerate(specs):
        if kind == "solid":
            m = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], n, p, 3, 1, 7)
            op = mfgpu.NeoHookOperator(m, 384615.38461538462, 0.3, 1000.0, 4e6, formulation=form,
                                       precision=precision)
            if form == 1:
                op.set_point_storage(storage)
            u = m.u0(amp=1e-3, noise=1e-5, seed=k)
            op.cache(torch.from_numpy(u).to("cuda", op.dtype))
        else:
            m = mfgpu.BoxMesh([0, 0, 0], [1, 1, 1], n, p, 1, 63, 1)
            coef = np.full(m.n_cells * (p + 1) ** 3, 0.5 + 0.1 * k)
            op = mfgpu.LaplaceOperator(m, 0.1, coef, precision=precision)
        op.set_cell_schedule(mfgpu.CELLS_COLOURS)  # the batched kernels' cell order: bitwise comparable
        ops.append(op)
        meshes.append(m)
    return ops, meshes


SPECS = [("heat", [3, 4, 2], 1, 1), ("solid", [2, 3, 2], 2, 1), ("heat", [4, 2, 3], 3, 1),
         ("solid", [2, 2, 3], 4, 0), ("solid", [3, 2, 2], 1, 1), ("heat", [2, 2, 2], 6, 1),
         ("solid", [2, 1, 2], 5, 1), ("heat", [2, 3, 2], 5, 1), ("solid", [1, 2, 1], 6, 0),
         ("heat", [2, 2, 1], 4, 1), ("solid", [2, 2, 2], 3, 0), ("heat", [5, 3, 2], 2, 1)]


@pytest.mark.parametrize("precision,storage", [(mfgpu.FP64, mfgpu.T2_FOLDED), (mfgpu.FP32, mfgpu.T2_FOLDED),
                                               (mfgpu.FP64, mfgpu.T2_GRADIENT)])
@pytest.mark.parametrize("schedule", [mfgpu.SCHEDULE_COLOUR, mfgpu.SCHEDULE_QUEUE, mfgpu.SCHEDULE_STREAMS,
                                      mfgpu.SCHEDULE_GATHER])
def test_batched_vmult_equals_per_instance(cuda, precision, storage, schedule):
    print(f"batched vmult precision={precision} storage={storage} schedule={schedule}", flush=True)
    ops, meshes = make_ops(SPECS, precision, storage)
    dt = ops[0].dtype
    srcs = [mfgpu.fill_uniform(100 + i, torch.empty(m.n_dofs, dtype=dt, device="cuda")) for i, m in enumerate(meshes)]
    ref = [torch.empty_like(s) for s in srcs]
    for op, s, r in zip(ops, srcs, ref):
        op.vmult(r, s)
    b = mfgpu.Batch(ops)
    b.set_schedule(schedule)
    for rep in range(3):  # repeated launches: the queue counters must reset
        out = [torch.full_like(s, float("nan")) for s in srcs]
        b.vmult(out, srcs)
        b.synchronize()
        for i, (o, r) in enumerate(zip(out, ref)):
            assert torch.equal(o, r), (rep, i, SPECS[i])
    # vmult_add adds A src to what dst holds
    acc = [r.clone() for r in ref]
    b.vmult_add(acc, srcs)
    b.synchronize()
    for i, (a, r) in enumerate(zip(acc, ref)):
        # in the norm: entries of A src that are sums of cancelling cell
        # contributions round differently from r + r
        err = float((a.double() - 2 * r.double()).norm() / (2 * r.double()).norm())
        assert err < (1e-6 if precision == mfgpu.FP32 else 1e-14), (i, err)


def test_batched_cells_per_task_and_tuner(cuda):
    ops, meshes = make_ops(SPECS, mfgpu.FP32)
    srcs = [mfgpu.fill_uniform(200 + i, torch.empty(m.n_dofs, dtype=torch.float32, device="cuda"))
            for i, m in enumerate(meshes)]
    ref = [torch.empty_like(s) for s in srcs]
    for op, s, r in zip(ops, srcs, ref):
        op.vmult(r, s)
    b = mfgpu.Batch(ops)
    out = [torch.empty_like(s) for s in srcs]
    for cells in (1, 2):
        for p in range(1, 7):
            for kind in (mfgpu.BATCH_SOLID_F1, mfgpu.BATCH_SOLID_F0, mfgpu.BATCH_HEAT):
                b.set_cells_per_task(kind, p, min(cells, 512 // (p + 1) ** 3))
        b.vmult(out, srcs)
        b.synchronize()
        assert all(torch.equal(o, r) for o, r in zip(out, ref)), cells
    with pytest.raises(mfgpu.MfgError):
        b.set_cells_per_task(mfgpu.BATCH_HEAT, 4, 5)  # 5 * 125 > 512 threads
    t = b.tune(srcs, out, repetitions=2)
    d = t.as_dict()
    print("tuning", d)
    assert d["schedule"] in ("queue", "colour", "streams", "gather") and t.waves_per_simd in (2, 4)
    assert t.ms_streams > 0 and t.ms_gather > 0
    for p in range(1, 7):
        assert 1 <= d["cells_per_task"]["heat"][p - 1] <= 512 // (p + 1) ** 3
    out = [torch.empty_like(s) for s in srcs]
    b.vmult(out, srcs)
    b.synchronize()
    assert all(torch.equal(o, r) for o, r in zip(out, ref))


def instances(specs, level_precision=mfgpu.FP32):
    out = []
    for k, (kind, n, p) in enumerate(specs):
        out.append(mfgpu.Instance(kind, n, p, level_precision=level_precision, conductivity=0.5 + 0.25 * k,
                                  coarse_eig_iterations=40, seed=k, cell_schedule=mfgpu.CELLS_COLOURS))
    return out


ISPECS = [("solid", [4, 2, 2], 2), ("heat", [4, 4, 4], 2), ("solid", [3, 2, 2], 1), ("heat", [2, 4, 2], 3),
          ("solid", [2, 2, 4], 3), ("heat", [3, 3, 3], 1)]


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("schedule", [mfgpu.SCHEDULE_COLOUR, mfgpu.SCHEDULE_QUEUE, mfgpu.SCHEDULE_STREAMS])
def test_batched_vcycle_equals_per_instance(cuda, graph, schedule):
    print(f"batched V-cycle graph={graph} schedule={schedule}", flush=True)
    inst = instances(ISPECS)
    assert sorted({len(i.meshes) for i in inst}) == [1, 2, 3]  # uneven hierarchies
    srcs = [mfgpu.fill_uniform(300 + k, torch.empty(i.n_dofs, dtype=torch.float64, device="cuda"))
            for k, i in enumerate(inst)]
    ref = [torch.empty_like(s) for s in srcs]
    for i, s, r in zip(inst, srcs, ref):
        i.mg.vmult(r, s)
    bs = mfgpu.BatchSolver([i.fine for i in inst], [i.mg for i in inst])
    bs.set_schedule(schedule)
    bs.use_graph(graph)
    for rep in range(3):  # graph replays with eager work in between
        out = [torch.full_like(s, float("nan")) for s in srcs]
        bs.precondition(out, srcs)
        for k, (o, r) in enumerate(zip(out, ref)):
            assert torch.equal(o, r), (rep, k, ISPECS[k])
        torch.cuda.synchronize()
        _ = sum(float(s.sum()) for s in srcs)
    if schedule == mfgpu.SCHEDULE_STREAMS:  # every instance's own V-cycle: no batched program
        return
    launches, nodes = bs.info()
    print(f"V-cycle program: {launches} launches, graph nodes {nodes}")
    assert launches > 0 and (nodes > 0 if graph else True)


@pytest.mark.parametrize("schedule", [None, mfgpu.SCHEDULE_COLOUR, mfgpu.SCHEDULE_STREAMS])
def test_batched_cg_equals_per_instance(cuda, schedule):
    """None: the tuner's choice."""
    inst = instances(ISPECS)
    bs_ = []
    for k, i in enumerate(inst):
        xt = mfgpu.fill_uniform(400 + k, torch.empty(i.n_dofs, dtype=torch.float64, device="cuda"))
        xt[torch.from_numpy(i.constrained).long().cuda()] = 0.0
        b = torch.empty_like(xt)
        i.fine.vmult(b, xt)
        bs_.append(b)
    ref, ref_info = [], []
    for i, b in zip(inst, bs_):
        x = torch.zeros_like(b)
        ref_info.append(mfgpu.cg_solve(i.fine, b, x, mfgpu.PRECOND_MG, i.mg, tol_rel=1e-8))
        ref.append(x)
    solver = mfgpu.BatchSolver([i.fine for i in inst], [i.mg for i in inst])
    fine_t, level_t = solver.tune(repetitions=2)
    print("tuned fine", fine_t.as_dict(), "levels", level_t.as_dict())
    if schedule is not None:
        solver.set_schedule(schedule)
    xs = [torch.zeros_like(b) for b in bs_]
    infos, status = solver.solve(bs_, xs, tol_rel=1e-8)
    for k in range(len(inst)):
        assert status[k] == 0
        assert infos[k].iterations == ref_info[k].iterations, (k, infos[k].iterations, ref_info[k].iterations)
        assert infos[k].final_residual == ref_info[k].final_residual
        assert torch.equal(xs[k], ref[k]), k
    # a second solve from the solution converges at once
    infos2, status2 = solver.solve(bs_, xs, tol_rel=1e-8)
    assert all(s == 0 for s in status2) and all(i.iterations == 0 for i in infos2)


def test_batched_errors(cuda):
    inst = instances(ISPECS[:2])
    with pytest.raises(mfgpu.MfgError):  # fine operators of different precisions
        f32 = mfgpu.NeoHookOperator(inst[0].meshes[-1], 384615.38461538462, 0.3, 1000.0, 4e6, precision=mfgpu.FP32)
        mfgpu.BatchSolver([f32, inst[1].fine], [inst[0].mg, inst[1].mg])
    m = inst[1]
    mg1 = mfgpu.Multigrid(m.levels, m.transfers, coarse_kind=1)
    mg1.setup()
    with pytest.raises(mfgpu.MfgError):  # the batched V-cycle needs the Chebyshev coarse solver
        mfgpu.BatchSolver([m.fine], [mg1])
    ops, meshes = make_ops(SPECS[:2], mfgpu.FP64)
    b = mfgpu.Batch(ops)
    s = torch.zeros(meshes[0].n_dofs, dtype=torch.float64, device="cuda")
    t = torch.zeros(meshes[1].n_dofs, dtype=torch.float64, device="cuda")
    with pytest.raises(mfgpu.MfgError):  # aliasing source and destination
        b.vmult([s, t], [s, t])


@pytest.mark.parametrize("schedule", [mfgpu.SCHEDULE_COLOUR, mfgpu.SCHEDULE_GATHER, mfgpu.SCHEDULE_STREAMS])
def test_batched_vmult_follows_operator_changes(cuda, schedule):
    """An operator changed after the batch was built (a new point storage
    reallocates its cache): the batch refuses it until it is cached again,
    then rebuilds its tables and matches the per-instance results again."""
    print(f"batch after an operator change, schedule={schedule}", flush=True)
    ops, meshes = make_ops(SPECS, mfgpu.FP64)
    srcs = [mfgpu.fill_uniform(700 + i, torch.empty(m.n_dofs, dtype=torch.float64, device="cuda"))
            for i, m in enumerate(meshes)]
    b = mfgpu.Batch(ops)
    b.set_schedule(schedule)
    out = [torch.empty_like(s) for s in srcs]
    b.vmult(out, srcs)
    b.synchronize()
    k = next(i for i, sp in enumerate(SPECS) if sp[0] == "solid" and sp[3] == 1 and sp[2] >= 3)
    ops[k].set_point_storage(mfgpu.T2_GRADIENT)  # discards the cache
    with pytest.raises(mfgpu.MfgError) as e:
        b.vmult(out, srcs)
    assert e.value.code == 5
    m = meshes[k]
    ops[k].cache(torch.from_numpy(m.u0(amp=1e-3, noise=1e-5, seed=k)).to("cuda", torch.float64))
    ref = [torch.empty_like(s) for s in srcs]
    for op, s, r in zip(ops, srcs, ref):
        op.vmult(r, s)
    for rep in range(2):
        out = [torch.full_like(s, float("nan")) for s in srcs]
        b.vmult(out, srcs)
        b.synchronize()
        for i, (o, r) in enumerate(zip(out, ref)):
            assert torch.equal(o, r), (rep, i, SPECS[i])


def test_batched_solver_refuses_changed_operators(cuda):
    inst = instances(ISPECS[:3])
    solver = mfgpu.BatchSolver([i.fine for i in inst], [i.mg for i in inst])
    bs_ = [mfgpu.fill_uniform(900 + k, torch.empty(i.n_dofs, dtype=torch.float64, device="cuda"))
           for k, i in enumerate(inst)]
    for k, i in enumerate(inst):
        bs_[k][torch.from_numpy(i.constrained).long().cuda()] = 0.0
    xs = [torch.zeros_like(b) for b in bs_]
    infos, status = solver.solve(bs_, xs, tol_rel=1e-6)
    assert all(st == 0 for st in status)
    k = next(j for j, sp in enumerate(ISPECS[:3]) if sp[0] == "solid")
    fine = inst[k].fine
    fine.set_point_storage(mfgpu.T2_GRADIENT if fine.point_storage() == mfgpu.T2_FOLDED else mfgpu.T2_FOLDED)
    fine.cache(torch.from_numpy(inst[k].meshes[-1].u0(amp=1e-3, noise=1e-5, seed=k)).to("cuda", torch.float64))
    with pytest.raises(mfgpu.MfgError) as e:
        solver.solve(bs_, [torch.zeros_like(b) for b in bs_], tol_rel=1e-6)
    assert e.value.code == 5
