"""Serving loop: continuous batching of Newmark steps over device-resident
sessions.  Each session's results must be identical alone (sequential) and in
the concurrent mix, also when HBM-budget admission forces evictions."""
import numpy as np
import pytest

import mfgpu
import oracle as O

pytestmark = pytest.mark.gpu


def specs():
    out = []
    for k, (nx, p) in enumerate([(8, 2), (12, 2), (8, 3), (16, 2), (4, 3), (8, 4)]):
        out.append(mfgpu.session_spec([0, 0, 0], [4.0, 1.0, 1.0], [nx, max(nx // 4, 1), max(nx // 4, 1)], p, O.MU,
                                      O.NU, O.RHO, body=(0.0, 0.0, -9.81), seed=3 + k, level_precision=mfgpu.FP32))
    return out


def trace(n_sessions, n_requests, rate, seed=100):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    return [(int(rng.integers(n_sessions)), float(x)) for x in t]


def test_mix_equals_alone(cuda):
    sp = specs()
    reqs = trace(len(sp), 18, rate=200.0)
    alone, st1, fin1 = mfgpu.serve(sp, reqs, max_concurrent=1, real_time=False, final_states=True)
    mix, st4, fin4 = mfgpu.serve(sp, reqs, max_concurrent=4, real_time=True, final_states=True)
    print(f"alone: {st1.steps_per_s:.1f} steps/s p50 {st1.latency_p50_s*1e3:.1f} ms; "
          f"mix: {st4.steps_per_s:.1f} steps/s p50 {st4.latency_p50_s*1e3:.1f} ms p99 {st4.latency_p99_s*1e3:.1f} ms")
    assert st1.completed == st4.completed == len(reqs) and st4.failed == 0
    for a, b in zip(alone, mix):
        assert a.status == b.status == 0
        assert (a.newton_iterations, a.linear_iterations) == (b.newton_iterations, b.linear_iterations)
        assert a.residual_final == b.residual_final
    for x, y in zip(fin1, fin4):
        assert np.array_equal(x, y)
    # sessions that received requests moved away from their initial state
    assert any(np.abs(x).max() > 0 for x in fin4)


def test_hbm_budget_forces_evictions(cuda):
    sp = specs()[:4]
    reqs = trace(len(sp), 12, rate=500.0, seed=7)
    ref, _, fin_ref = mfgpu.serve(sp, reqs, max_concurrent=1, real_time=False, final_states=True)
    # budget large enough for any one session but not for all of them
    sizes = []
    for s in sp:
        r, st, _ = mfgpu.serve([s], [(0, 0.0)], max_concurrent=1, real_time=False)
        sizes.append(st.peak_bytes)
    budget = int(max(sizes) * 1.5)
    res, st, fin = mfgpu.serve(sp, reqs, max_concurrent=2, hbm_budget_bytes=budget, real_time=True,
                               final_states=True)
    print(f"budget {budget} B: admissions {st.admissions} evictions {st.evictions} peak {st.peak_bytes}")
    assert st.evictions > 0 and st.peak_bytes <= budget and st.failed == 0
    for a, b in zip(ref, res):
        assert (a.newton_iterations, a.linear_iterations, a.residual_final) == \
               (b.newton_iterations, b.linear_iterations, b.residual_final)
    for x, y in zip(fin_ref, fin):
        assert np.array_equal(x, y)


def test_session_larger_than_budget_is_rejected(cuda):
    sp = specs()[:1]
    with pytest.raises(mfgpu.MfgError):
        mfgpu.serve(sp, [(0, 0.0)], hbm_budget_bytes=1024)


@pytest.mark.timeout(120)
def test_failed_admission_releases_its_budget(cuda):
    """A session whose problem cannot be built (bad material, refused when it
    is made resident) fails its requests; the bytes reserved for it are
    released, so the other session, too large to share the budget with it,
    is still admitted and served, and the call returns an error instead of
    waiting for budget that never frees."""
    good = specs()[:1]
    bad = mfgpu.session_spec([0, 0, 0], [4.0, 1.0, 1.0], [8, 2, 2], 2, -1.0, O.NU, O.RHO, seed=9,
                             level_precision=mfgpu.FP32)
    _, st, _ = mfgpu.serve(good, [(0, 0.0)], max_concurrent=1, real_time=False)
    budget = int(st.peak_bytes * 1.5)  # room for one of the two sessions at a time
    reqs = [(1, 0.0), (1, 0.0), (0, 0.0), (0, 0.0)]
    with pytest.raises(mfgpu.MfgError) as e:
        mfgpu.serve(good + [bad], reqs, max_concurrent=2, hbm_budget_bytes=budget, real_time=False)
    assert "bad material" in str(e.value)


def test_bad_newmark_parameters_rejected_before_serving(cuda):
    bad = mfgpu.session_spec([0, 0, 0], [4.0, 1.0, 1.0], [8, 2, 2], 2, O.MU, O.NU, O.RHO, dt=0.0)
    with pytest.raises(mfgpu.MfgError) as e:
        mfgpu.serve([bad], [(0, 0.0), (0, 0.0)])
    assert "Newmark" in str(e.value)
