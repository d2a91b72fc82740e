// Serving loop: continuous batching of Newmark time steps over many
// device-resident solid sessions (see mfg_serve in include/mfgpu.h).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>

#include "common.hpp"
#include "problem.hpp"

using namespace mfg;

namespace {

using Clock = std::chrono::steady_clock;

struct Session {
  mfg_session_spec spec{};
  std::vector<mfg_mesh> meshes;
  std::vector<mfg_level_desc> descs;
  std::vector<std::vector<int32_t>> children;
  std::vector<const int32_t *> child_ptrs;
  std::unique_ptr<SolidProblem> prob;  // non-null while resident
  std::vector<double> hu, hv, ha;      // host copy of the state while evicted
  size_t bytes = 0;
  bool busy = false, built = false;
  double last_used = -1;
  int64_t n_dofs = 0;
  ~Session() {
    prob.reset();
    for (auto m : meshes) mfg_box_mesh_destroy(m);
  }

  void build_hierarchy() {
    if (built) return;
    int n[3] = {spec.n_cells[0], spec.n_cells[1], spec.n_cells[2]};
    int levels = 1;
    while (n[0] % 2 == 0 && n[1] % 2 == 0 && n[2] % 2 == 0) {
      n[0] /= 2;
      n[1] /= 2;
      n[2] /= 2;
      ++levels;
    }
    for (int l = 0; l < levels; ++l) {
      const int32_t nl[3] = {n[0] << l, n[1] << l, n[2] << l};
      mfg_mesh m = nullptr;
      if (mfg_box_mesh_create(spec.lo, spec.hi, nl, spec.degree, 3, 1, 7, &m)) throw Error(MFG_ERR_ARGUMENT, mfg_last_error());
      meshes.push_back(m);
      if (mfg_box_mesh_renumber(m, MFG_NUMBERING_CELLWISE)) throw Error(MFG_ERR_ARGUMENT, mfg_last_error());
      mfg_level_desc d;
      mfg_box_mesh_describe(m, &d);
      descs.push_back(d);
    }
    children.resize(levels > 1 ? levels - 1 : 0);
    for (int l = 1; l < levels; ++l) {
      children[l - 1].resize((size_t)descs[l - 1].n_cells * 8);
      if (mfg_box_mesh_children(meshes[l - 1], meshes[l], children[l - 1].data())) throw Error(MFG_ERR_ARGUMENT, mfg_last_error());
      child_ptrs.push_back(children[l - 1].data());
    }
    n_dofs = descs.back().n_nodes * 3;
    bytes = estimate_solid_problem_bytes(levels, descs.data(), spec.level_precision);
    hu.assign(n_dofs, 0.0);
    hv.assign(n_dofs, 0.0);
    ha.assign(n_dofs, 0.0);
    mfg_box_mesh_fill_u0(meshes.back(), spec.u0_amplitude, spec.u0_noise, spec.seed, hu.data());
    built = true;
  }

  // prob becomes non-null only once the problem and its state are on the
  // device, so a failure here leaves the session non-resident
  void make_resident(hipStream_t s) {
    auto P = make_solid_problem((int)descs.size(), descs.data(), child_ptrs.empty() ? nullptr : child_ptrs.data(),
                                spec.material, spec.newmark, spec.mg, spec.level_precision, s);
    MFG_HIP(hipMemcpyAsync(P->u.ptr, hu.data(), n_dofs * 8, hipMemcpyHostToDevice, s));
    MFG_HIP(hipMemcpyAsync(P->v_old.ptr, hv.data(), n_dofs * 8, hipMemcpyHostToDevice, s));
    MFG_HIP(hipMemcpyAsync(P->a_old.ptr, ha.data(), n_dofs * 8, hipMemcpyHostToDevice, s));
    MFG_HIP(hipStreamSynchronize(s));
    prob = std::move(P);
  }

  void save_state(hipStream_t s) {
    MFG_HIP(hipMemcpyAsync(hu.data(), prob->u.ptr, n_dofs * 8, hipMemcpyDeviceToHost, s));
    MFG_HIP(hipMemcpyAsync(hv.data(), prob->v_old.ptr, n_dofs * 8, hipMemcpyDeviceToHost, s));
    MFG_HIP(hipMemcpyAsync(ha.data(), prob->a_old.ptr, n_dofs * 8, hipMemcpyDeviceToHost, s));
    MFG_HIP(hipStreamSynchronize(s));
  }
};

struct Server {
  std::vector<std::unique_ptr<Session>> sessions;
  std::vector<mfg_request> requests;
  std::vector<mfg_request_result> results;
  mfg_server_params prm{};
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int> ready;         // released request indices, arrival order
  std::vector<char> released;
  size_t resident_bytes = 0, peak_bytes = 0;
  int admissions = 0, evictions = 0, max_resident = 0, resident = 0, done = 0;
  Clock::time_point t0;
  std::string first_error;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  // Pick a runnable request (lock held).  Returns -1 if none can run now.
  // On success the session is marked busy and its bytes are reserved;
  // `evict` lists sessions to evict (made idle and non-resident here).
  int pick(std::vector<Session *> &evict) {
    for (auto it = ready.begin(); it != ready.end(); ++it) {
      const int r = *it;
      Session &S = *sessions[requests[r].session];
      if (S.busy) continue;
      // earlier released request of the same session still waiting? keep order
      bool earlier = false;
      for (auto jt = ready.begin(); jt != it; ++jt)
        if (requests[*jt].session == requests[r].session) earlier = true;
      if (earlier) continue;
      if (!S.prob) {
        size_t need = S.bytes;
        std::vector<Session *> victims;
        size_t freed = 0;
        if (resident_bytes + need > prm.hbm_budget_bytes) {
          std::vector<Session *> idle;
          for (auto &o : sessions)
            if (o->prob && !o->busy) idle.push_back(o.get());
          std::sort(idle.begin(), idle.end(), [](Session *a, Session *b) { return a->last_used < b->last_used; });
          for (Session *v : idle) {
            if (resident_bytes - freed + need <= prm.hbm_budget_bytes) break;
            victims.push_back(v);
            freed += v->bytes;
          }
          if (resident_bytes - freed + need > prm.hbm_budget_bytes && resident > (int)victims.size()) continue;
        }
        for (Session *v : victims) {
          v->busy = true;  // being evicted by the picking worker
          evict.push_back(v);
        }
        resident_bytes = resident_bytes - freed + need;
        peak_bytes = std::max(peak_bytes, resident_bytes);
        resident += 1 - (int)victims.size();
        max_resident = std::max(max_resident, resident);
        evictions += (int)victims.size();
        ++admissions;
      }
      S.busy = true;
      ready.erase(it);
      return r;
    }
    return -1;
  }

  void worker() {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
      std::lock_guard<std::mutex> g(mu);
      first_error = "stream creation failed";
      done = (int)requests.size();
      cv.notify_all();
      return;
    }
    for (;;) {
      std::vector<Session *> evict;
      int r = -1;
      bool admitted = false;  // pick() reserved the session's bytes
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (done >= (int)requests.size()) break;
          r = pick(evict);
          if (r >= 0) break;
          cv.wait_for(lk, std::chrono::milliseconds(1));
        }
        if (r < 0) break;
        admitted = !sessions[requests[r].session]->prob;
      }
      Session &S = *sessions[requests[r].session];
      mfg_request_result &res = results[r];
      res.start_s = now();
      try {
        for (Session *v : evict) {
          v->save_state(s);
          v->prob.reset();
        }
        if (!S.prob) S.make_resident(s);
        S.prob->set_stream(s);
        mfg_step_info info{};
        S.prob->step(info);
        res.newton_iterations = info.newton_iterations;
        res.linear_iterations = info.linear_iterations;
        res.residual_final = info.residual_final;
        res.status = MFG_OK;
      } catch (const Error &e) {
        res.status = e.code;
        std::lock_guard<std::mutex> g(mu);
        if (first_error.empty()) first_error = e.what();
      } catch (const std::exception &e) {
        res.status = MFG_ERR_ARGUMENT;
        std::lock_guard<std::mutex> g(mu);
        if (first_error.empty()) first_error = e.what();
      }
      res.end_s = now();
      {
        std::lock_guard<std::mutex> g(mu);
        // pick() counted the evictions and the admission as done; undo what a
        // failure above left undone, so the budget keeps matching the
        // sessions that are actually resident
        for (Session *v : evict)
          if (v->prob) {
            resident_bytes += v->bytes;
            ++resident;
            --evictions;
          }
        if (admitted && !S.prob) {
          resident_bytes -= S.bytes;
          --resident;
          --admissions;
        }
        for (Session *v : evict) v->busy = false;
        S.busy = false;
        S.last_used = res.end_s;
        ++done;
      }
      cv.notify_all();
    }
    (void)hipStreamDestroy(s);
  }

  void run() {
    const int n = (int)requests.size();
    results.assign(n, mfg_request_result{});
    released.assign(n, 0);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return requests[a].arrival_s < requests[b].arrival_s; });
    // latency runs from the release: the arrival time in real time, t = 0
    // when every request is released at once
    for (int i = 0; i < n; ++i) results[i].arrival_s = prm.real_time ? requests[i].arrival_s : 0.0;
    t0 = Clock::now();
    std::vector<std::thread> workers;
    const int W = std::max(1, prm.max_concurrent);
    for (int w = 0; w < W; ++w) workers.emplace_back([this] { worker(); });
    // release requests at their arrival times
    size_t next = 0;
    while (next < order.size()) {
      const double due = prm.real_time ? requests[order[next]].arrival_s : 0.0;
      const double t = now();
      if (t < due) {
        std::this_thread::sleep_for(std::chrono::duration<double>(std::min(due - t, 0.002)));
        continue;
      }
      {
        std::lock_guard<std::mutex> g(mu);
        while (next < order.size() && (!prm.real_time || requests[order[next]].arrival_s <= now()))
          ready.push_back(order[next++]);
      }
      cv.notify_all();
    }
    for (auto &t : workers) t.join();
  }
};

}  // namespace

extern "C" {

int64_t mfg_session_n_dofs(const mfg_session_spec *spec) {
  if (!spec) return -1;
  const int p = spec->degree;
  return (int64_t)(spec->n_cells[0] * p + 1) * (spec->n_cells[1] * p + 1) * (spec->n_cells[2] * p + 1) * 3;
}

int mfg_serve(const mfg_session_spec *sessions, int32_t n_sessions, const mfg_request *requests, int32_t n_requests,
              const mfg_server_params *params, mfg_request_result *results, double *const *final_u,
              mfg_serve_stats *stats) {
  return guarded([&] {
    MFG_REQUIRE(sessions && n_sessions > 0 && requests && n_requests >= 0 && params && results, MFG_ERR_ARGUMENT,
                "null argument");
    MFG_REQUIRE(params->max_concurrent >= 1 && params->max_concurrent <= 64, MFG_ERR_ARGUMENT,
                "max_concurrent must be in 1..64");
    Server srv;
    srv.prm = *params;
    for (int i = 0; i < n_sessions; ++i) {
      const mfg_newmark_params &nm = sessions[i].newmark;
      MFG_REQUIRE(nm.dt > 0 && nm.beta > 0 && nm.gamma >= 0 && nm.max_newton >= 1 && nm.max_lin >= 1 &&
                      nm.linear_preconditioner >= MFG_NEWTON_GMG && nm.linear_preconditioner <= MFG_NEWTON_NONE &&
                      (nm.linear_preconditioner != MFG_NEWTON_JACOBI || nm.jacobi_relaxation > 0),
                  MFG_ERR_ARGUMENT, "bad Newmark parameters");
      auto s = std::make_unique<Session>();
      s->spec = sessions[i];
      s->build_hierarchy();
      MFG_REQUIRE(s->bytes <= params->hbm_budget_bytes, MFG_ERR_ARGUMENT,
                  "a session does not fit the HBM budget on its own");
      srv.sessions.push_back(std::move(s));
    }
    for (int r = 0; r < n_requests; ++r)
      MFG_REQUIRE(requests[r].session >= 0 && requests[r].session < n_sessions, MFG_ERR_ARGUMENT,
                  "request for an unknown session");
    srv.requests.assign(requests, requests + n_requests);
    srv.run();
    std::copy(srv.results.begin(), srv.results.end(), results);
    if (final_u)
      for (int i = 0; i < n_sessions; ++i) {
        Session &S = *srv.sessions[i];
        if (!final_u[i]) continue;
        if (S.prob) {
          MFG_HIP(hipMemcpy(final_u[i], S.prob->u.ptr, S.n_dofs * 8, hipMemcpyDeviceToHost));
        } else {
          std::memcpy(final_u[i], S.hu.data(), S.n_dofs * 8);
        }
      }
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      std::vector<double> lat;
      double end = 0, first = 1e300;
      for (auto &r : srv.results) {
        if (r.status == MFG_OK) {
          lat.push_back(r.end_s - r.arrival_s);
          ++stats->completed;
        } else {
          ++stats->failed;
        }
        end = std::max(end, r.end_s);
        first = std::min(first, r.arrival_s);
      }
      std::sort(lat.begin(), lat.end());
      auto pct = [&](double q) {
        if (lat.empty()) return 0.0;
        const double x = q * (lat.size() - 1);
        const size_t i = (size_t)x;
        const double f = x - i;
        return i + 1 < lat.size() ? lat[i] * (1 - f) + lat[i + 1] * f : lat[i];
      };
      stats->makespan_s = end - (first < 1e300 ? std::min(first, 0.0) : 0.0);
      stats->steps_per_s = stats->makespan_s > 0 ? stats->completed / stats->makespan_s : 0.0;
      stats->latency_p50_s = pct(0.5);
      stats->latency_p99_s = pct(0.99);
      double m = 0;
      for (double x : lat) m += x;
      stats->latency_mean_s = lat.empty() ? 0.0 : m / lat.size();
      stats->admissions = srv.admissions;
      stats->evictions = srv.evictions;
      stats->max_resident = srv.max_resident;
      stats->peak_bytes = srv.peak_bytes;
    }
    if (!srv.first_error.empty()) throw Error(MFG_ERR_NUMERIC, "a request failed: " + srv.first_error);
  });
}

}  // extern "C"
