// Batched operator application and MG-preconditioned CG over many
// independent instances (heat and solid, degrees 1..6, uneven meshes).
//
// Host side of the batched kernels: task tables per set of instances, the
// per-degree cells-per-task autotuner, the batched V-cycle as a fixed
// program of launches (optionally captured once into a hipGraph and
// replayed), and CG in lockstep over all instances with one host
// synchronisation per reduction.  Every instance runs exactly the operation
// sequence of its own MGData::vmult / cg_solve, so batched results equal the
// per-instance results.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batched_kernels.hpp"
#include "common.hpp"
#include "operator.hpp"
#include "solver.hpp"

namespace mfg {

static_assert(kSolidF1 == MFG_BATCH_SOLID_F1 && kSolidF0 == MFG_BATCH_SOLID_F0 && kHeat == MFG_BATCH_HEAT &&
                  kSolidF1G == MFG_BATCH_SOLID_F1_GRADIENT && kNumKinds == 4,
              "instance kinds are the ABI's MFG_BATCH_* numbers");
static int code_of(const OperatorData &op) {
  const int kind = op.kind == 1 ? kHeat : op.t2_gradient() ? kSolidF1G : (op.sp.formulation == 1 ? kSolidF1 : kSolidF0);
  return kind * 8 + op.p;
}
static int max_cpb(int p) { return kBatchThreads / ((p + 1) * (p + 1) * (p + 1)); }

struct Tuning {
  int cpb[8 * kNumKinds] = {};  // per code; 0 = the largest that fits kBatchThreads
  int queue = 1;     // schedule: MFG_SCHEDULE_COLOUR, _QUEUE or _STREAMS
  int waves = 2;     // register budget of the kernel: 2 or 4 waves per SIMD (FP32 only)
  double ms_colour = 0, ms_queue = 0, ms_streams = 0, ms_gather = 0;
  int cpb_for(int code) const {
    const int m = max_cpb(code & 7), c = cpb[code];
    return c > 0 && c <= m ? c : m;
  }
};

template <class T>
static void upload_vec(DeviceBuffer &buf, const std::vector<T> &v, hipStream_t s) {
  upload(buf, v.data(), v.size(), s);
}

static int device_cus() {
  static const int cus = [] {
    int dev = 0, n = 0;
    MFG_HIP(hipGetDevice(&dev));
    MFG_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    return n;
  }();
  return cus;
}

template <typename Number, int WAVES, int KINDS>
static int queue_capacity() {
  static const int cap = [] {
    int nb = 0;
    MFG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batched_apply_kernel<Number, WAVES, KINDS>,
                                                         kBatchThreads, 0));
    return std::max(1, nb) * device_cus();
  }();
  return cap;
}

// Streams of the STREAMS schedule: instances are spread over them by work
// (longest first onto the least-loaded stream), each running its own
// operator's launches, so small instances overlap instead of queueing behind
// each other.  Forked from and joined back to the caller's stream with
// events (which also works under graph capture).  Default priority.
struct StreamPool {
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> done;
  hipEvent_t fork_ev = nullptr;
  std::vector<std::vector<int>> assign;  // instances per stream
  StreamPool() = default;
  StreamPool(const StreamPool &) = delete;
  StreamPool &operator=(const StreamPool &) = delete;
  ~StreamPool() {
    for (hipStream_t x : streams) (void)hipStreamDestroy(x);
    for (hipEvent_t e : done) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
  }
  void build(const std::vector<double> &work, int max_streams) {
    const int n = (int)work.size();
    const int S = std::max(1, std::min(max_streams, n));
    while ((int)streams.size() < S) {
      hipStream_t x;
      hipEvent_t e;
      MFG_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
      streams.push_back(x);
      MFG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      done.push_back(e);
    }
    if (!fork_ev) MFG_HIP(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    std::vector<double> load(S, 0.0);
    assign.assign(S, {});
    for (int i : order) {
      const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[j] += work[i];
      assign[j].push_back(i);
    }
  }
  // run(i, stream) for every instance on its stream, ordered after the
  // caller's work and before the caller's next
  template <class F>
  void run(hipStream_t caller, F &&f) {
    MFG_HIP(hipEventRecord(fork_ev, caller));
    for (size_t j = 0; j < assign.size(); ++j) {
      if (assign[j].empty()) continue;
      MFG_HIP(hipStreamWaitEvent(streams[j], fork_ev, 0));
      for (int i : assign[j]) f(i, streams[j]);
      MFG_HIP(hipEventRecord(done[j], streams[j]));
      MFG_HIP(hipStreamWaitEvent(caller, done[j], 0));
    }
  }
};

// work estimate of an operator application: quadrature points times the
// words read per point
static double apply_work(const OperatorData &op) { return (double)op.nqp * (op.n_fields + 2.0 * op.nc + 1.0); }

// ------------------------------------------------------- apply groups ---
// Task tables over a set of instances of one precision: one instance table,
// and the tasks of each instance kind present (each kind is its own launch,
// so that a launch carries the registers of its kind's code paths only).
struct ApplyGroup {
  struct KindTasks {
    int kind = 0;
    DeviceBuffer tasks, counters;
    std::vector<int> colour_tasks;  // tasks of colour k: [colour_tasks[k], colour_tasks[k+1])
    int n_tasks = 0, n_slots = 0;
  };
  int prec = MFG_FP64;
  std::vector<const OperatorData *> ops;
  std::vector<int> ids;
  DeviceBuffer inst;
  std::vector<KindTasks> kinds;

  ApplyGroup(std::vector<const OperatorData *> o, std::vector<int> id) : ops(std::move(o)), ids(std::move(id)) {
    MFG_REQUIRE(!ops.empty(), MFG_ERR_ARGUMENT, "empty batch");
    prec = ops[0]->precision;
    for (const OperatorData *op : ops) {
      MFG_REQUIRE(op && op->ready, MFG_ERR_STATE, "batched operator used before cache()");
      MFG_REQUIRE(op->precision == prec, MFG_ERR_ARGUMENT, "batched operators must share one precision");
      MFG_REQUIRE(op->p >= 1 && op->p <= kMaxDegree, MFG_ERR_ARGUMENT, "unsupported degree");
      MFG_REQUIRE(op->kind == 1 || op->sp.caching == MFG_CACHING_TENSOR2, MFG_ERR_ARGUMENT,
                  "batched solid operators use the tensor2 caching");
      MFG_REQUIRE(!op->has_inclusion(), MFG_ERR_ARGUMENT, "batched solid operators have one material");
    }
  }

  template <typename Number>
  void build_t(const Tuning &tu, hipStream_t s) {
    const int n = (int)ops.size();
    std::vector<BatchInst<Number>> hi(n);
    int n_col = 0;
    for (int i = 0; i < n; ++i) {
      const OperatorData &op = *ops[i];
      BatchInst<Number> &e = hi[i];
      e.cell_nodes = op.cell_nodes.as<uint32_t>();
      e.qd = op.qd.as<Number>();
      e.constrained = op.constrained.as<int32_t>();
      e.shape = op.tab_shape.as<Number>();
      e.colloc = op.tab_colloc.as<Number>();
      e.nqp = op.nqp;
      e.n_constrained = op.n_constrained;
      e.code = code_of(op);
      e.cpb = tu.cpb_for(e.code);
      e.id = ids[i];
      e.prm[0] = e.prm[1] = e.prm[2] = Number(0);
      e.cell_jxw = op.affine_ref && op.kind == 0 ? op.cell_jxw.as<Number>() : nullptr;
      e.cell_geo = op.affine_ref && op.kind == 1 ? op.cell_geo.as<Number>() : nullptr;
      e.cell_jac = op.t2_gradient() ? op.cell_jac.as<Number>() : nullptr;
      e.wq = op.wq3.as<Number>();
      if (op.kind == 1) {
        e.prm[0] = Number(op.dt);
      } else if (op.sp.formulation == 1) {
        e.prm[0] = Number(2.0 * op.lambda);
        e.prm[1] = Number(op.rho_alpha);
        e.prm[2] = Number(op.sp.mu);
      } else {
        e.prm[0] = Number(op.kappa);
        e.prm[1] = Number(op.c_1);
        e.prm[2] = Number(op.rho_alpha);
      }
      n_col = std::max(n_col, (int)op.colour_start.size() - 1);
    }
    upload_vec(inst, hi, s);
    kinds.clear();
    for (int kind : {kSolidF1, kSolidF1G, kSolidF0, kHeat}) {
      bool present = false;
      for (int i = 0; i < n; ++i) present |= (hi[i].code >> 3) == kind;
      if (!present) continue;
      kinds.emplace_back();
      build_kind(kinds.back(), kind, hi, n_col, s);
    }
  }

  template <typename Number>
  void build_kind(KindTasks &kt, int kind, const std::vector<BatchInst<Number>> &hi, int n_col, hipStream_t s) {
    const int n = (int)ops.size();
    kt.kind = kind;
    std::vector<BatchTask> tv;
    std::vector<int> last_slot(n, -1), last_count(n, 0);
    kt.colour_tasks.assign(1, 0);
    int n_slots = 0;
    for (int col = 0; col < n_col; ++col) {
      for (int i = 0; i < n; ++i) {
        if ((hi[i].code >> 3) != kind) continue;
        const OperatorData &op = *ops[i];
        const int ncol = (int)op.colour_start.size() - 1;
        const int cb = col < ncol ? op.colour_start[col] : 0, ce = col < ncol ? op.colour_start[col + 1] : 0;
        const int cpb = hi[i].cpb;
        int groups = (ce - cb + cpb - 1) / cpb;
        const bool fix_only = col == 0 && groups == 0 && op.n_constrained > 0;
        if (fix_only) groups = 1;
        if (groups == 0) continue;
        const int slot = n_slots++;
        for (int g = 0; g < groups; ++g) {
          BatchTask t;
          t.inst = i;
          t.cell_begin = std::min(ce, cb + g * cpb);
          t.cell_end = std::min(ce, cb + (g + 1) * cpb);
          t.fix_begin = t.fix_end = 0;
          if (col == 0) {
            t.fix_begin = (int32_t)(op.n_constrained * g / groups);
            t.fix_end = (int32_t)(op.n_constrained * (g + 1) / groups);
          }
          t.wait_slot = last_slot[i];
          t.wait_count = last_count[i];
          t.done_slot = slot;
          tv.push_back(t);
        }
        last_slot[i] = slot;
        last_count[i] = groups;
      }
      kt.colour_tasks.push_back((int)tv.size());
    }
    kt.n_tasks = (int)tv.size();
    kt.n_slots = n_slots;
    upload_vec(kt.tasks, tv, s);
    kt.counters.allocate((size_t)(2 + n_slots) * sizeof(int));
    MFG_HIP(hipMemsetAsync(kt.counters.ptr, 0, kt.counters.bytes, s));
    MFG_HIP(hipStreamSynchronize(s));
  }

  void build(const Tuning &tu, hipStream_t s) {
    if (prec == MFG_FP64) build_t<double>(tu, s);
    else build_t<float>(tu, s);
  }

  template <typename Number, int WAVES, int KINDS>
  void launch_kind(const KindTasks &kt, hipStream_t s, const void *vecs, const int *active, int fix_mode, int assign,
                   int queue, int *error) const {
    BatchApplyArgs<Number> a;
    a.inst = inst.as<BatchInst<Number>>();
    a.tasks = kt.tasks.as<BatchTask>();
    a.vecs = (const BatchVecs<Number> *)vecs;
    a.active = active;
    a.queue = queue;
    a.fix_mode = fix_mode;
    a.assign = assign;
    a.n_slots = kt.n_slots;
    a.counters = kt.counters.as<int>();
    a.error = error;
    if (kt.n_tasks == 0) return;
    if (queue) {
      a.task_begin = 0;
      a.task_end = kt.n_tasks;
      const int grid = std::min(kt.n_tasks, queue_capacity<Number, WAVES, KINDS>());
      batched_apply_kernel<Number, WAVES, KINDS><<<grid, kBatchThreads, 0, s>>>(a);
      MFG_HIP(hipGetLastError());
      return;
    }
    for (size_t col = 0; col + 1 < kt.colour_tasks.size(); ++col) {
      a.task_begin = kt.colour_tasks[col];
      a.task_end = kt.colour_tasks[col + 1];
      if (a.task_end == a.task_begin) continue;
      batched_apply_kernel<Number, WAVES, KINDS><<<a.task_end - a.task_begin, kBatchThreads, 0, s>>>(a);
      MFG_HIP(hipGetLastError());
    }
  }

  template <typename Number, int WAVES>
  void launch_t(hipStream_t s, const void *vecs, const int *active, int fix_mode, int assign, int queue,
                int *error) const {
    for (const KindTasks &kt : kinds) {
      if (kt.kind == kSolidF1)
        launch_kind<Number, WAVES, kind_bit(kSolidF1)>(kt, s, vecs, active, fix_mode, assign, queue, error);
      else if (kt.kind == kSolidF1G)
        launch_kind<Number, WAVES, kind_bit(kSolidF1G)>(kt, s, vecs, active, fix_mode, assign, queue, error);
      else if (kt.kind == kSolidF0)
        launch_kind<Number, WAVES, kind_bit(kSolidF0)>(kt, s, vecs, active, fix_mode, assign, queue, error);
      else
        launch_kind<Number, WAVES, kind_bit(kHeat)>(kt, s, vecs, active, fix_mode, assign, queue, error);
    }
  }

  // kernel launches of one application
  int launches(const Tuning &tu) const {
    int l = 0;
    for (const KindTasks &kt : kinds) l += tu.queue == MFG_SCHEDULE_QUEUE ? 1 : (int)kt.colour_tasks.size() - 1;
    return l;
  }

  // vmult (fix_mode 1, first-touch stores) or vmult_add (fix_mode 2, adds)
  void launch(hipStream_t s, const void *vecs, const int *active, bool add, const Tuning &tu, int *error) const {
    const int fix = add ? 2 : 1, assign = add ? 0 : 1;
    const int queue = tu.queue == MFG_SCHEDULE_QUEUE ? 1 : 0;
    if (prec == MFG_FP64) launch_t<double, 2>(s, vecs, active, fix, assign, queue, error);
    else if (tu.waves == 4) launch_t<float, 4>(s, vecs, active, fix, assign, queue, error);
    else launch_t<float, 2>(s, vecs, active, fix, assign, queue, error);
  }

  // Table of (src, dst) per entry, in entry order.
  void make_vecs(DeviceBuffer &buf, const std::vector<const void *> &src, const std::vector<void *> &dst,
                 hipStream_t s) const {
    std::vector<const void *> v(2 * ops.size());
    for (size_t i = 0; i < ops.size(); ++i) {
      MFG_REQUIRE(src[i] && dst[i], MFG_ERR_ARGUMENT, "null vector");
      MFG_REQUIRE(src[i] != dst[i], MFG_ERR_ARGUMENT, "vmult needs distinct source and destination");
      v[2 * i] = src[i];
      v[2 * i + 1] = dst[i];
    }
    upload_vec(buf, v, s);
  }

  // Picks cells per task for every code of the group, then the schedule, by
  // timing applications restricted (through the active mask) to the
  // instances of one code.
  void tune(Tuning &tu, const void *vecs, int reps, hipStream_t s, int *error) {
    const int n = (int)ops.size();
    int max_id = 0;
    for (int id : ids) max_id = std::max(max_id, id);
    DeviceBuffer mask((size_t)(max_id + 1) * sizeof(int));
    std::vector<int> hm(max_id + 1, 0);
    hipEvent_t e0, e1;
    MFG_HIP(hipEventCreate(&e0));
    MFG_HIP(hipEventCreate(&e1));
    auto time_it = [&](const int *active, const Tuning &t) {
      for (int w = 0; w < 2; ++w) launch(s, vecs, active, false, t, error);
      MFG_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r) launch(s, vecs, active, false, t, error);
      MFG_HIP(hipEventRecord(e1, s));
      MFG_HIP(hipEventSynchronize(e1));
      float ms = 0;
      MFG_HIP(hipEventElapsedTime(&ms, e0, e1));
      return (double)ms / reps;
    };
    try {
      std::vector<int> codes;
      for (int i = 0; i < n; ++i) {
        const int c = code_of(*ops[i]);
        if (std::find(codes.begin(), codes.end(), c) == codes.end()) codes.push_back(c);
      }
      for (int code : codes) {
        std::fill(hm.begin(), hm.end(), 0);
        for (int i = 0; i < n; ++i)
          if (code_of(*ops[i]) == code) hm[ids[i]] = 1;
        MFG_HIP(hipMemcpy(mask.ptr, hm.data(), hm.size() * sizeof(int), hipMemcpyHostToDevice));
        std::vector<int> cand;
        for (int c = max_cpb(code & 7); c >= 1; c /= 2) cand.push_back(c);
        if (cand.size() > 4) cand.resize(4);
        double best = 1e300;
        int best_c = cand[0];
        for (int c : cand) {
          tu.cpb[code] = c;
          build(tu, s);
          const double ms = time_it(mask.as<int>(), tu);
          if (ms < best) {
            best = ms;
            best_c = c;
          }
        }
        tu.cpb[code] = best_c;
      }
      build(tu, s);
      // register budget (FP32), then the schedule, on the whole batch
      if (prec == MFG_FP32) {
        Tuning t2 = tu, t4 = tu;
        t2.waves = 2;
        t4.waves = 4;
        tu.waves = time_it(nullptr, t4) < time_it(nullptr, t2) ? 4 : 2;
      }
      Tuning tc = tu, tq = tu;
      tc.queue = 0;
      tq.queue = 1;
      tu.ms_colour = time_it(nullptr, tc);
      tu.ms_queue = time_it(nullptr, tq);
      tu.queue = tu.ms_queue <= tu.ms_colour ? 1 : 0;
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    MFG_HIP(hipEventDestroy(e0));
    MFG_HIP(hipEventDestroy(e1));
  }
};

// -------------------------------------------------- segmented vectors ---
struct SegTable {
  DeviceBuffer segs;
  int n = 0, prec = MFG_FP64, prec_a = MFG_FP64;
  int64_t chunks = 0;
  void build(std::vector<VecSeg> v, hipStream_t s) {
    int64_t c = 0;
    for (VecSeg &g : v) {
      g.chunk_begin = c;
      c += (g.n + kSegChunk - 1) / kSegChunk;
    }
    chunks = c;
    n = (int)v.size();
    upload_vec(segs, v, s);
  }
};

template <typename T, typename TA, int OP>
static void seg_launch_t(const SegTable &t, const int *active, const double *dyn, hipStream_t s) {
  if (t.chunks == 0) return;
  seg_kernel<T, TA, OP><<<(unsigned)t.chunks, kSegThreads, 0, s>>>(t.segs.as<VecSeg>(), t.n, active, dyn);
  MFG_HIP(hipGetLastError());
}

template <int OP>
static void seg_launch(const SegTable &t, const int *active, const double *dyn, hipStream_t s) {
  if (OP == kSegConvert) {
    if (t.prec == MFG_FP64 && t.prec_a == MFG_FP64) seg_launch_t<double, double, OP>(t, active, dyn, s);
    else if (t.prec == MFG_FP64) seg_launch_t<double, float, OP>(t, active, dyn, s);
    else if (t.prec_a == MFG_FP64) seg_launch_t<float, double, OP>(t, active, dyn, s);
    else seg_launch_t<float, float, OP>(t, active, dyn, s);
  } else if (t.prec == MFG_FP64) {
    seg_launch_t<double, double, OP>(t, active, dyn, s);
  } else {
    seg_launch_t<float, float, OP>(t, active, dyn, s);
  }
}

static VecSeg make_seg(void *y, const void *a, const void *b, const void *c, const void *d, int64_t n, int inst,
                       double s0 = 0, double s1 = 0) {
  VecSeg g;
  std::memset(&g, 0, sizeof(g));
  g.y = y;
  g.a = a;
  g.b = b;
  g.c = c;
  g.d = d;
  g.n = n;
  g.inst = inst;
  g.s0 = s0;
  g.s1 = s1;
  return g;
}

// Per-instance sums with the single-vector reduction order; one table entry
// per instance, in instance order.
struct SegReducer {
  DeviceBuffer partials, sums;
  double *host = nullptr;
  int n = 0;
  void ensure(int ni) {
    if (ni <= n) return;
    partials.allocate((size_t)ni * kRedBlocks * sizeof(double));
    sums.allocate((size_t)ni * sizeof(double));
    if (host) (void)hipHostFree(host);
    host = nullptr;
    MFG_HIP(hipHostMalloc((void **)&host, (size_t)ni * sizeof(double), hipHostMallocDefault));
    n = ni;
  }
  ~SegReducer() {
    if (host) (void)hipHostFree(host);
  }
  // launches only; fetch() reads the sums back
  void dot(const SegTable &t, const int *active, hipStream_t s) {
    ensure(t.n);
    const dim3 grid(kRedBlocks, t.n);
    if (t.prec == MFG_FP64)
      seg_dot_partial_kernel<double><<<grid, kRedThreads, 0, s>>>(t.segs.as<VecSeg>(), active, partials.as<double>());
    else
      seg_dot_partial_kernel<float><<<grid, kRedThreads, 0, s>>>(t.segs.as<VecSeg>(), active, partials.as<double>());
    seg_sum_partials_kernel<<<t.n, kRedThreads, 0, s>>>(partials.as<double>(), sums.as<double>());
    MFG_HIP(hipGetLastError());
  }
  void cg_update(const SegTable &t, const int *active, const double *alpha, hipStream_t s) {
    ensure(t.n);
    const dim3 grid(kRedBlocks, t.n);
    if (t.prec == MFG_FP64)
      seg_cg_update_kernel<double><<<grid, kRedThreads, 0, s>>>(t.segs.as<VecSeg>(), active, alpha,
                                                                partials.as<double>());
    else
      seg_cg_update_kernel<float><<<grid, kRedThreads, 0, s>>>(t.segs.as<VecSeg>(), active, alpha,
                                                               partials.as<double>());
    seg_sum_partials_kernel<<<t.n, kRedThreads, 0, s>>>(partials.as<double>(), sums.as<double>());
    MFG_HIP(hipGetLastError());
  }
  std::vector<double> fetch(int ni, hipStream_t s) {
    MFG_HIP(hipMemcpyAsync(host, sums.ptr, (size_t)ni * sizeof(double), hipMemcpyDeviceToHost, s));
    MFG_HIP(hipStreamSynchronize(s));
    return std::vector<double>(host, host + ni);
  }
};

// --------------------------------------------------- batched transfers ---
// Restriction (coarse colours in turn) or prolongation between one pair of
// levels of many instances.
struct TransferGroup {
  int kind = 0;  // 0 prolongate_add, 1 restrict_add
  int cprec = MFG_FP32, fprec = MFG_FP32;
  DeviceBuffer tab, tasks;
  std::vector<int> colour_tasks;

  template <typename CN, typename FN>
  void build_t(const std::vector<TransferData *> &tr, const std::vector<MGLevel *> &coarse,
               const std::vector<MGLevel *> &fine, const std::vector<int> &ids, hipStream_t s) {
    const int n = (int)tr.size();
    std::vector<BatchTransfer<CN, FN>> h(n);
    int n_col = 0;
    for (int i = 0; i < n; ++i) {
      const TransferData &t = *tr[i];
      BatchTransfer<CN, FN> &e = h[i];
      std::memset(&e, 0, sizeof(e));
      e.args.coarse_nodes = t.coarse->cell_nodes.as<uint32_t>();
      e.args.lattice = t.lattice.as<uint32_t>();
      e.args.E = t.E.as<double>();
      if (kind == 0) {
        e.args.coarse_in = coarse[i]->sol.as<CN>();
        e.args.fine_out = fine[i]->sol.as<FN>();
      } else {
        e.args.coarse_out = coarse[i]->def.as<CN>();
        e.args.fine_in = fine[i]->tmp.as<FN>();
      }
      e.code = t.p + (t.nc == 3 ? 8 : 0);
      e.inst = ids[i];
      n_col = std::max(n_col, (int)t.coarse->colour_start.size() - 1);
    }
    std::vector<BatchCellTask> tv;
    colour_tasks.assign(1, 0);
    if (kind == 0) {  // every fine node has one owner: no colouring needed
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < tr[i]->coarse->n_cells; ++c) tv.push_back({i, c});
      colour_tasks.push_back((int)tv.size());
    } else {
      for (int col = 0; col < n_col; ++col) {
        for (int i = 0; i < n; ++i) {
          const std::vector<int> &cs = tr[i]->coarse->colour_start;
          if (col + 1 >= (int)cs.size()) continue;
          for (int c = cs[col]; c < cs[col + 1]; ++c) tv.push_back({i, c});
        }
        colour_tasks.push_back((int)tv.size());
      }
    }
    upload_vec(tab, h, s);
    upload_vec(tasks, tv, s);
  }

  void build(int k, const std::vector<TransferData *> &tr, const std::vector<MGLevel *> &coarse,
             const std::vector<MGLevel *> &fine, const std::vector<int> &ids, hipStream_t s) {
    kind = k;
    cprec = coarse[0]->op->precision;
    fprec = fine[0]->op->precision;
    for (size_t i = 0; i < tr.size(); ++i)
      MFG_REQUIRE(coarse[i]->op->precision == cprec && fine[i]->op->precision == fprec, MFG_ERR_ARGUMENT,
                  "batched levels must share one precision per level");
    if (cprec == MFG_FP64 && fprec == MFG_FP64) build_t<double, double>(tr, coarse, fine, ids, s);
    else if (cprec == MFG_FP64) build_t<double, float>(tr, coarse, fine, ids, s);
    else if (fprec == MFG_FP64) build_t<float, double>(tr, coarse, fine, ids, s);
    else build_t<float, float>(tr, coarse, fine, ids, s);
  }

  template <typename CN, typename FN>
  void launch_t(hipStream_t s, const int *active) const {
    for (size_t col = 0; col + 1 < colour_tasks.size(); ++col) {
      const int b = colour_tasks[col], e = colour_tasks[col + 1];
      if (e == b) continue;
      if (kind == 0)
        batched_transfer_kernel<0, CN, FN><<<e - b, 256, 0, s>>>(tab.as<BatchTransfer<CN, FN>>(),
                                                                  tasks.as<BatchCellTask>(), b, e, active);
      else
        batched_transfer_kernel<1, CN, FN><<<e - b, 256, 0, s>>>(tab.as<BatchTransfer<CN, FN>>(),
                                                                  tasks.as<BatchCellTask>(), b, e, active);
      MFG_HIP(hipGetLastError());
    }
  }
  void launch(hipStream_t s, const int *active) const {
    if (cprec == MFG_FP64 && fprec == MFG_FP64) launch_t<double, double>(s, active);
    else if (cprec == MFG_FP64) launch_t<double, float>(s, active);
    else if (fprec == MFG_FP64) launch_t<float, double>(s, active);
    else launch_t<float, float>(s, active);
  }
};

// ------------------------------------------------ batch GATHER schedule ---
// The instances of one (kind, degree) are packed into units of about equal
// work (a large instance alone, small ones together); a unit is one cells
// launch over all its instances' cells (batched_cells_kernel) followed, on the
// same stream, by one segmented per-node sum over its instances
// (batched_gather_kernel).  Units are spread over streams by work, so the sums
// of one unit overlap the cells of others.  Each instance's GATHER tables and
// contribution buffer are its operator's own (built on first use, as
// mfg_operator_set_cell_schedule builds them): a batch and a GATHER
// application of one of its operators on another stream must not overlap.
struct BatchGather {
  struct Unit {
    int code = 0;  // kind * 8 + degree
    std::vector<int> inst;
    DeviceBuffer tasks, segs, block_begin;
    int n_tasks = 0;
    int64_t blocks = 0;
    double work = 0;
  };
  std::vector<OperatorData *> ops;
  int prec = MFG_FP64;
  std::vector<Unit> units;
  StreamPool pool;
  DeviceBuffer vecs;
  std::vector<const void *> last_src;
  std::vector<void *> last_dst;

  template <typename Number>
  void build_t(hipStream_t s) {
    const int n = (int)ops.size();
    double total = 0;
    std::map<int, std::vector<int>> by_code;
    for (int i = 0; i < n; ++i) {
      OperatorData &op = *ops[i];
      if (!op.contrib.ptr) op.build_gather();
      by_code[code_of(op)].push_back(i);
      total += apply_work(op);
    }
    // units of at most cap work (one instance at least), largest first; the
    // floor (about 2^18 solid quadrature points) keeps small batches at one
    // unit per (kind, degree), where launches rather than work set the time
    const double cap = std::max(total / (2.0 * kBatchStreams), 5.0e6);
    units.clear();
    for (auto &kv : by_code) {
      std::vector<int> v = kv.second;
      std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return apply_work(*ops[a]) > apply_work(*ops[b]); });
      Unit *u = nullptr;
      for (int i : v) {
        const double w = apply_work(*ops[i]);
        if (!u || u->work + w > cap) {
          units.emplace_back();
          u = &units.back();
          u->code = kv.first;
        }
        u->inst.push_back(i);
        u->work += w;
      }
    }
    std::vector<double> work;
    for (Unit &u : units) {
      std::vector<BatchTask> tv;
      for (int i : u.inst) {
        const OperatorData &op = *ops[i];
        const int p = op.p;
        const int cpb = p == 1 ? DefaultCPB<1, Number>::value : p == 2 ? DefaultCPB<2, Number>::value : 1;
        for (int c = 0; c < op.n_cells; c += cpb) {
          BatchTask t{};
          t.inst = i;
          t.cell_begin = c;
          t.cell_end = std::min(op.n_cells, c + cpb);
          tv.push_back(t);
        }
      }
      u.n_tasks = (int)tv.size();
      upload_vec(u.tasks, tv, s);
      work.push_back(u.work);
    }
    pool.build(work, kBatchStreams);
  }

  BatchGather(const std::vector<OperatorData *> &o, hipStream_t s) : ops(o) {
    prec = ops[0]->precision;
    if (prec == MFG_FP64) build_t<double>(s);
    else build_t<float>(s);
  }

  template <typename Number>
  void set_vectors(const void *const *src, void *const *dst, hipStream_t s) {
    const size_t n = ops.size();
    std::vector<const void *> sv(src, src + n);
    std::vector<void *> dv(dst, dst + n);
    if (sv == last_src && dv == last_dst && vecs.ptr) return;
    std::vector<BatchVecs<Number>> v(n);
    for (size_t i = 0; i < n; ++i) {
      MFG_REQUIRE(src[i] && dst[i], MFG_ERR_ARGUMENT, "null vector");
      MFG_REQUIRE(src[i] != dst[i], MFG_ERR_ARGUMENT, "vmult needs distinct source and destination");
      v[i].src = (const Number *)src[i];
      v[i].dst = ops[i]->contrib.as<Number>();
    }
    upload_vec(vecs, v, s);
    for (Unit &u : units) {
      std::vector<GatherSeg<Number>> g;
      std::vector<int64_t> bb(1, 0);  // first workgroup of every segment, and the grid size
      for (int i : u.inst) {
        const OperatorData &op = *ops[i];
        GatherSeg<Number> e;
        e.contrib = op.contrib.as<Number>();
        e.offsets = op.gather_offsets.as<int32_t>();
        e.points = op.gather_points.as<int32_t>();
        e.cmask = op.gather_cmask.as<uint8_t>();
        e.src = (const Number *)src[i];
        e.dst = (Number *)dst[i];
        e.n_nodes = op.n_nodes;
        g.push_back(e);
        bb.push_back(bb.back() + (op.n_nodes + kGatherThreads - 1) / kGatherThreads);
      }
      MFG_REQUIRE(bb.back() < INT32_MAX, MFG_ERR_ARGUMENT, "batch unit too large for one gather launch");
      upload_vec(u.segs, g, s);
      upload_vec(u.block_begin, bb, s);
      u.blocks = bb.back();
    }
    last_src = sv;
    last_dst = dv;
  }

  template <int P, int KIND, typename Number>
  void launch_cells(const Unit &u, const void *inst_table, hipStream_t s) const {
    constexpr int CPB = DefaultCPB<P, Number>::value;
    BatchCellsArgs<Number> a;
    a.inst = (const BatchInst<Number> *)inst_table;
    a.tasks = u.tasks.as<BatchTask>();
    a.vecs = vecs.as<BatchVecs<Number>>();
    batched_cells_kernel<P, CPB, KIND, Number><<<u.n_tasks, block_threads<P, CPB>(), 0, s>>>(a);
    MFG_HIP(hipGetLastError());
  }

  template <int KIND, typename Number>
  void launch_kind(const Unit &u, const void *inst_table, hipStream_t s) const {
    switch (u.code & 7) {
      case 1: launch_cells<1, KIND, Number>(u, inst_table, s); break;
      case 2: launch_cells<2, KIND, Number>(u, inst_table, s); break;
      case 3: launch_cells<3, KIND, Number>(u, inst_table, s); break;
      case 4: launch_cells<4, KIND, Number>(u, inst_table, s); break;
      case 5: launch_cells<5, KIND, Number>(u, inst_table, s); break;
      case 6: launch_cells<6, KIND, Number>(u, inst_table, s); break;
      default: MFG_REQUIRE(false, MFG_ERR_ARGUMENT, "unsupported degree");
    }
  }

  template <typename Number>
  void apply_t(const void *inst_table, const void *const *src, void *const *dst, bool add, hipStream_t s) {
    set_vectors<Number>(src, dst, s);
    const int fix = add ? 2 : 1;
    pool.run(s, [&](int ui, hipStream_t us) {
      const Unit &u = units[ui];
      const int kind = u.code >> 3;
      switch (kind) {
        case kSolidF1: launch_kind<kSolidF1, Number>(u, inst_table, us); break;
        case kSolidF1G: launch_kind<kSolidF1G, Number>(u, inst_table, us); break;
        case kSolidF0: launch_kind<kSolidF0, Number>(u, inst_table, us); break;
        default: launch_kind<kHeat, Number>(u, inst_table, us); break;
      }
      const int n_segs = (int)u.inst.size();
      if (kind == kHeat)
        batched_gather_kernel<Number, 1><<<(unsigned)u.blocks, kGatherThreads, 0, us>>>(
            u.segs.as<GatherSeg<Number>>(), u.block_begin.as<int64_t>(), n_segs, fix);
      else
        batched_gather_kernel<Number, 3><<<(unsigned)u.blocks, kGatherThreads, 0, us>>>(
            u.segs.as<GatherSeg<Number>>(), u.block_begin.as<int64_t>(), n_segs, fix);
      MFG_HIP(hipGetLastError());
    });
  }

  void apply(const void *inst_table, const void *const *src, void *const *dst, bool add, hipStream_t s) {
    if (prec == MFG_FP64) apply_t<double>(inst_table, src, dst, add, s);
    else apply_t<float>(inst_table, src, dst, add, s);
  }
};

// ---------------------------------------------------- batched vmult API ---
struct BatchVmult {
  std::unique_ptr<ApplyGroup> group;
  std::vector<OperatorData *> mops;  // the same operators (the GATHER schedule builds their tables)
  std::unique_ptr<BatchGather> gather;
  Tuning tu;
  hipStream_t stream = nullptr;
  DeviceBuffer vecs, error;
  std::vector<const void *> last_src;
  std::vector<void *> last_dst;

  // each operator's buffers_version and instance code when the tables were built
  std::vector<uint64_t> versions;
  std::vector<int> codes;

  BatchVmult(const std::vector<OperatorData *> &mo) : mops(mo) {
    const std::vector<const OperatorData *> ops(mo.begin(), mo.end());
    std::vector<int> ids(ops.size());
    for (size_t i = 0; i < ids.size(); ++i) ids[i] = (int)i;
    group.reset(new ApplyGroup(ops, ids));
    stream = ops[0]->stream;
    error.allocate(sizeof(int));
    MFG_HIP(hipMemset(error.ptr, 0, sizeof(int)));
    MFG_HIP(hipDeviceSynchronize());  // operator set-up on other streams is complete
    group->build(tu, stream);
    record_versions();
  }
  void record_versions() {
    versions.clear();
    codes.clear();
    for (const OperatorData *op : mops) {
      versions.push_back(op->buffers_version);
      codes.push_back(code_of(*op));
    }
  }
  // The tables hold the operators' device pointers: an operator whose buffers
  // were reallocated since they were built (a new point storage, say) makes
  // the batch rebuild them; an operator without a cache is refused.
  void validate() {
    bool changed = false;
    for (size_t i = 0; i < mops.size(); ++i) {
      MFG_REQUIRE(mops[i]->ready, MFG_ERR_STATE, "batched operator used before cache()");
      changed |= mops[i]->buffers_version != versions[i] || code_of(*mops[i]) != codes[i];
    }
    if (!changed) return;
    MFG_HIP(hipStreamSynchronize(stream));
    const std::vector<const OperatorData *> ops(mops.begin(), mops.end());
    std::unique_ptr<ApplyGroup> g(new ApplyGroup(ops, group->ids));
    g->build(tu, stream);
    group = std::move(g);
    gather.reset();
    last_src.clear();
    last_dst.clear();
    record_versions();
  }
  void set_vectors(const void *const *src, void *const *dst) {
    const size_t n = group->ops.size();
    std::vector<const void *> s(src, src + n);
    std::vector<void *> d(dst, dst + n);
    if (s == last_src && d == last_dst && vecs.ptr) return;
    group->make_vecs(vecs, s, d, stream);
    last_src = s;
    last_dst = d;
  }
  StreamPool pool;
  // STREAMS: every instance's own apply (its default cell schedule) on the
  // pool's streams; bitwise the per-instance result
  void apply_streams(const void *const *src, void *const *dst, bool add) {
    const size_t n = group->ops.size();
    for (size_t i = 0; i < n; ++i) {
      MFG_REQUIRE(src[i] && dst[i], MFG_ERR_ARGUMENT, "null vector");
      MFG_REQUIRE(src[i] != dst[i], MFG_ERR_ARGUMENT, "vmult needs distinct source and destination");
    }
    if (pool.assign.empty()) {
      std::vector<double> w;
      for (const OperatorData *op : group->ops) w.push_back(apply_work(*op));
      pool.build(w, kBatchStreams);
    }
    pool.run(stream, [&](int i, hipStream_t s) { group->ops[i]->apply(src[i], dst[i], add, s); });
  }
  void apply_gather(const void *const *src, void *const *dst, bool add) {
    if (!gather) {
      gather.reset(new BatchGather(mops, stream));
      record_versions();  // building the operators' GATHER tables counts as a change of their buffers
    }
    gather->apply(group->inst.ptr, src, dst, add, stream);
  }
  void apply(const void *const *src, void *const *dst, bool add) {
    validate();
    if (tu.queue == MFG_SCHEDULE_STREAMS) {
      apply_streams(src, dst, add);
      return;
    }
    if (tu.queue == MFG_SCHEDULE_GATHER) {
      apply_gather(src, dst, add);
      return;
    }
    set_vectors(src, dst);
    group->launch(stream, vecs.ptr, nullptr, add, tu, error.as<int>());
  }
  // the fused schedules' tuning, then the STREAMS and GATHER schedules timed
  // against the best of them
  void tune(const void *const *src, void *const *dst, int reps) {
    validate();
    set_vectors(src, dst);
    group->tune(tu, vecs.ptr, reps, stream, error.as<int>());
    hipEvent_t e0, e1;
    MFG_HIP(hipEventCreate(&e0));
    MFG_HIP(hipEventCreate(&e1));
    auto time_it = [&](auto &&fn) {
      for (int w = 0; w < 2; ++w) fn();
      MFG_HIP(hipEventRecord(e0, stream));
      for (int r = 0; r < reps; ++r) fn();
      MFG_HIP(hipEventRecord(e1, stream));
      MFG_HIP(hipEventSynchronize(e1));
      float ms = 0;
      MFG_HIP(hipEventElapsedTime(&ms, e0, e1));
      return (double)ms / reps;
    };
    try {
      tu.ms_streams = time_it([&] { apply_streams(src, dst, false); });
      tu.ms_gather = time_it([&] { apply_gather(src, dst, false); });
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double fused = std::min(tu.ms_colour, tu.ms_queue);
    if (tu.ms_streams < fused && tu.ms_streams <= tu.ms_gather) tu.queue = MFG_SCHEDULE_STREAMS;
    else if (tu.ms_gather < fused) tu.queue = MFG_SCHEDULE_GATHER;
  }
  void synchronize() {
    MFG_HIP(hipStreamSynchronize(stream));
    int e = 0;
    MFG_HIP(hipMemcpy(&e, error.ptr, sizeof(int), hipMemcpyDeviceToHost));
    MFG_REQUIRE(e == 0, MFG_ERR_HIP, "batched kernel reported a wait timeout");
  }
};

// --------------------------------------------------- batched MG and CG ---
struct BatchSolver {
  int n = 0;
  std::vector<OperatorData *> fine;
  std::vector<MGData *> mg;
  hipStream_t stream = nullptr;
  Tuning tu_fine, tu_mg;
  bool use_graph = true;
  // STREAMS schedule: every instance's own MGData::vmult / cg_solve, the
  // instances spread over worker threads with a stream each
  bool concurrent = false;
  double ms_lockstep = 0, ms_concurrent = 0;  // one preconditioner application of all instances
  StreamPool pool;

  std::unique_ptr<ApplyGroup> fine_group;
  std::map<std::string, std::unique_ptr<ApplyGroup>> level_groups;
  // The V-cycle of a set of instances as a fixed list of launches, with the
  // tables it reads; one program per set of instances still iterating, so
  // converged instances launch nothing.  Captured into a hipGraph on its
  // second use and replayed from then on.
  struct Program {
    std::vector<std::function<void(hipStream_t)>> steps;
    std::vector<std::unique_ptr<DeviceBuffer>> tables;
    std::vector<std::unique_ptr<SegTable>> segs;
    std::vector<std::unique_ptr<TransferGroup>> transfers;
    int launches = 0, uses = 0;
    size_t nodes = 0;
    hipGraphExec_t graph = nullptr;
    ~Program() {
      if (graph) (void)hipGraphExecDestroy(graph);
    }
  };
  std::map<std::vector<int>, std::unique_ptr<Program>> programs;
  Program *cur = nullptr;  // the program being built

  // outer-CG workspace (fine precision), one slice per instance
  DeviceBuffer ws_r, ws_z, ws_p, ws_h;
  std::vector<int64_t> off;
  DeviceBuffer active, dyn, error;
  int *host_active = nullptr;
  double *host_dyn = nullptr;
  SegReducer red;
  // tables over the workspace built once
  DeviceBuffer vec_hp;
  SegTable seg_pz, seg_xpby, dot_rr, dot_rz, dot_ph;

  int fprec() const { return fine[0]->precision; }
  size_t fsize() const { return fine[0]->elem_size(); }
  void *slice(DeviceBuffer &b, int i) const { return (char *)b.ptr + off[i] * fsize(); }

  BatchSolver(const std::vector<OperatorData *> &f, const std::vector<MGData *> &m) : fine(f), mg(m) {
    n = (int)fine.size();
    MFG_REQUIRE(n > 0 && mg.size() == fine.size(), MFG_ERR_ARGUMENT, "empty batch");
    for (int i = 0; i < n; ++i) {
      MFG_REQUIRE(fine[i] && fine[i]->ready, MFG_ERR_STATE, "fine operator used before cache()");
      MFG_REQUIRE(fine[i]->precision == fine[0]->precision, MFG_ERR_ARGUMENT,
                  "fine operators must share one precision");
      MFG_REQUIRE(mg[i] && mg[i]->ready, MFG_ERR_STATE, "multigrid not set up");
      MFG_REQUIRE(mg[i]->coarse_kind() == 0, MFG_ERR_ARGUMENT,
                  "the batched V-cycle needs the Chebyshev coarse solver (coarse_kind 0)");
      MFG_REQUIRE(mg[i]->levels.back().op->n_dofs == fine[i]->n_dofs, MFG_ERR_ARGUMENT,
                  "finest multigrid level and fine operator differ in size");

    }
    MFG_HIP(hipDeviceSynchronize());
    MFG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    off.resize(n + 1, 0);
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + ((fine[i]->n_dofs + 63) / 64) * 64;
    for (DeviceBuffer *b : {&ws_r, &ws_z, &ws_p, &ws_h}) b->allocate((size_t)off[n] * fsize());
    active.allocate((size_t)n * sizeof(int));
    dyn.allocate((size_t)n * sizeof(double));
    error.allocate(sizeof(int));
    MFG_HIP(hipMemset(error.ptr, 0, sizeof(int)));
    MFG_HIP(hipHostMalloc((void **)&host_active, (size_t)n * sizeof(int), hipHostMallocDefault));
    MFG_HIP(hipHostMalloc((void **)&host_dyn, (size_t)n * sizeof(double), hipHostMallocDefault));
    for (int i = 0; i < n; ++i) host_active[i] = 1;
    upload_active();
    red.ensure(n);
    build();
    versions = current_versions();
  }
  // buffers_version of every fine and level operator: the programs and tables
  // hold their device pointers
  std::vector<uint64_t> versions;
  std::vector<uint64_t> current_versions() const {
    std::vector<uint64_t> v;
    for (int i = 0; i < n; ++i) {
      v.push_back(fine[i]->buffers_version);
      for (const MGLevel &L : mg[i]->levels) v.push_back(L.op->buffers_version);
    }
    return v;
  }
  void check_unchanged() const {
    MFG_REQUIRE(current_versions() == versions, MFG_ERR_STATE,
                "an operator's device buffers changed after the batched solver was created: create it again");
    for (int i = 0; i < n; ++i)
      MFG_REQUIRE(fine[i]->ready && mg[i]->ready, MFG_ERR_STATE, "batched solver operator or multigrid not set up");
  }
  ~BatchSolver() {
    if (stream) (void)hipStreamSynchronize(stream);
    programs.clear();
    if (stream) (void)hipStreamDestroy(stream);
    if (host_active) (void)hipHostFree(host_active);
    if (host_dyn) (void)hipHostFree(host_dyn);
  }

  void upload_active() {
    MFG_HIP(hipMemcpyAsync(active.ptr, host_active, (size_t)n * sizeof(int), hipMemcpyHostToDevice, stream));
  }
  void upload_dyn() {
    MFG_HIP(hipMemcpyAsync(dyn.ptr, host_dyn, (size_t)n * sizeof(double), hipMemcpyHostToDevice, stream));
  }
  void check_error() {
    int e = 0;
    MFG_HIP(hipMemcpy(&e, error.ptr, sizeof(int), hipMemcpyDeviceToHost));
    MFG_REQUIRE(e == 0, MFG_ERR_HIP, "batched kernel reported a wait timeout");
  }

  SegTable *new_seg(std::vector<VecSeg> v, int prec, int prec_a = -1) {
    cur->segs.emplace_back(new SegTable());
    SegTable *t = cur->segs.back().get();
    t->prec = prec;
    t->prec_a = prec_a < 0 ? prec : prec_a;
    t->build(std::move(v), stream);
    return t;
  }
  const void *new_table(const ApplyGroup &g, const std::vector<const void *> &src, const std::vector<void *> &dst) {
    cur->tables.emplace_back(new DeviceBuffer());
    g.make_vecs(*cur->tables.back(), src, dst, stream);
    return cur->tables.back()->ptr;
  }

  // apply group for the instances `ins` at global level g
  ApplyGroup &level_group(int g, const std::vector<int> &ins) {
    std::string key = std::to_string(g) + ":";
    for (int i : ins) key += std::to_string(i) + ",";
    auto it = level_groups.find(key);
    if (it != level_groups.end()) return *it->second;
    std::vector<const OperatorData *> ops;
    for (int i : ins) ops.push_back(mg[i]->levels[g].op);
    std::unique_ptr<ApplyGroup> grp(new ApplyGroup(ops, ins));
    grp->build(tu_mg, stream);
    ApplyGroup &ref = *grp;
    level_groups[key] = std::move(grp);
    return ref;
  }

  // t = A x at level g (x = sol, or given by `src_tmp`: tmp = A sol)
  void add_apply(int g, const std::vector<int> &ins, bool to_tmp) {
    if (ins.empty()) return;
    // split by precision of the level operators
    std::map<int, std::vector<int>> by_prec;
    for (int i : ins) by_prec[mg[i]->levels[g].op->precision].push_back(i);
    for (auto &kv : by_prec) {
      ApplyGroup &grp = level_group(g, kv.second);
      std::vector<const void *> src;
      std::vector<void *> dst;
      for (int i : kv.second) {
        MGLevel &L = mg[i]->levels[g];
        src.push_back(L.sol.ptr);
        dst.push_back(to_tmp ? L.tmp.ptr : L.t.ptr);
      }
      const void *vt = new_table(grp, src, dst);
      const ApplyGroup *gp = &grp;
      cur->steps.push_back([this, gp, vt](hipStream_t s) {
        gp->launch(s, vt, active.as<int>(), false, tu_mg, error.as<int>());
      });
      cur->launches += gp->launches(tu_mg);
    }
  }

  template <int OP>
  void add_seg_op(std::vector<std::pair<VecSeg, int>> items) {  // (segment, precision)
    std::map<int, std::vector<VecSeg>> by_prec;
    for (auto &it : items) by_prec[it.second].push_back(it.first);
    for (auto &kv : by_prec) {
      SegTable *t = new_seg(kv.second, kv.first);
      cur->steps.push_back([this, t](hipStream_t s) { seg_launch<OP>(*t, active.as<int>(), dyn.as<double>(), s); });
      ++cur->launches;
    }
  }

  // MGLevel::chebyshev for the instances `ins` at global level g, x = sol, b = def
  void add_chebyshev(int g, const std::vector<int> &ins, bool zero_guess) {
    std::vector<std::pair<VecSeg, int>> first;
    for (int i : ins) {
      MGLevel &L = mg[i]->levels[g];
      const int pr = L.op->precision;
      const double inv_theta = 1.0 / L.theta;
      if (zero_guess)
        first.push_back({make_seg(L.sol.ptr, L.d.ptr, L.def.ptr, L.op->inv_diag.ptr, nullptr, L.op->n_dofs, i,
                                  inv_theta),
                         pr});
      else
        first.push_back({make_seg(L.sol.ptr, L.d.ptr, L.def.ptr, L.t.ptr, L.op->inv_diag.ptr, L.op->n_dofs, i,
                                  inv_theta),
                         pr});
    }
    if (zero_guess) {
      add_seg_op<kSegChebFirst>(first);
    } else {
      add_apply(g, ins, false);
      add_seg_op<kSegChebFirstStep>(first);
    }
    // recurrence coefficients per instance, as MGLevel::chebyshev computes them
    std::vector<int> q;
    std::vector<std::vector<std::pair<double, double>>> coef;
    int kmax = 0;
    for (int i : ins) {
      MGLevel &L = mg[i]->levels[g];
      if (L.degree < 2 || std::abs(L.delta) < 1e-40) continue;
      std::vector<std::pair<double, double>> c;
      double rhok = L.delta / L.theta;
      const double sigma = L.theta / L.delta;
      for (int k = 0; k < L.degree - 1; ++k) {
        const double rhokp = 1.0 / (2.0 * sigma - rhok);
        c.push_back({rhokp * rhok, 2.0 * rhokp / L.delta});
        rhok = rhokp;
      }
      kmax = std::max(kmax, (int)c.size());
      q.push_back(i);
      coef.push_back(c);
    }
    for (int k = 0; k < kmax; ++k) {
      std::vector<int> qk;
      std::vector<std::pair<VecSeg, int>> upd;
      for (size_t j = 0; j < q.size(); ++j) {
        if (k >= (int)coef[j].size()) continue;
        const int i = q[j];
        MGLevel &L = mg[i]->levels[g];
        qk.push_back(i);
        upd.push_back({make_seg(L.sol.ptr, L.d.ptr, L.def.ptr, L.t.ptr, L.op->inv_diag.ptr, L.op->n_dofs, i,
                                coef[j][k].first, coef[j][k].second),
                       L.op->precision});
      }
      add_apply(g, qk, false);
      add_seg_op<kSegChebUpdate>(upd);
    }
  }

  void add_transfer(int g, const std::vector<int> &ins, int kind) {
    std::vector<TransferData *> tr;
    std::vector<MGLevel *> c, f;
    for (int i : ins) {
      tr.push_back(mg[i]->transfers[g - 1]);
      c.push_back(&mg[i]->levels[g - 1]);
      f.push_back(&mg[i]->levels[g]);
    }
    cur->transfers.emplace_back(new TransferGroup());
    TransferGroup *t = cur->transfers.back().get();
    t->build(kind, tr, c, f, ins, stream);
    cur->steps.push_back([this, t](hipStream_t s) { t->launch(s, active.as<int>()); });
    cur->launches += (int)t->colour_tasks.size() - 1;
  }

  // The V-cycle (MGData::vmult from r to z) of the instances `ins`,
  // bottom-aligned: global level g is level g of every instance that has it.
  Program &program_for(const std::vector<int> &ins) {
    auto it = programs.find(ins);
    if (it != programs.end()) return *it->second;
    std::unique_ptr<Program> pr(new Program());
    cur = pr.get();
    try {
      build_vcycle(ins);
    } catch (...) {
      cur = nullptr;
      throw;
    }
    cur = nullptr;
    Program &ref = *pr;
    programs[ins] = std::move(pr);
    return ref;
  }

  void build_vcycle(const std::vector<int> &ins) {
    int depth = 0;
    for (int i : ins) depth = std::max(depth, (int)mg[i]->levels.size());
    auto has = [&](int g) {
      std::vector<int> v;
      for (int i : ins)
        if ((int)mg[i]->levels.size() > g) v.push_back(i);
      return v;
    };
    {  // copy_to_mg: def of the finest level = r
      std::map<int, std::vector<VecSeg>> by;
      for (int i : ins) {
        MGLevel &T = mg[i]->levels.back();
        by[T.op->precision].push_back(make_seg(T.def.ptr, slice(ws_r, i), nullptr, nullptr, nullptr, T.op->n_dofs, i));
      }
      for (auto &kv : by) {
        SegTable *t = new_seg(kv.second, kv.first, fprec());
        cur->steps.push_back([this, t](hipStream_t s) { seg_launch<kSegConvert>(*t, active.as<int>(), nullptr, s); });
        ++cur->launches;
      }
    }
    for (int g = depth - 1; g >= 1; --g) {  // down: pre-smooth, residual, restrict
      const std::vector<int> P = has(g);
      add_chebyshev(g, P, true);
      add_apply(g, P, true);
      std::vector<std::pair<VecSeg, int>> res, fill;
      for (int i : P) {
        MGLevel &L = mg[i]->levels[g], &C = mg[i]->levels[g - 1];
        res.push_back({make_seg(L.tmp.ptr, L.def.ptr, nullptr, nullptr, nullptr, L.op->n_dofs, i), L.op->precision});
        fill.push_back({make_seg(C.def.ptr, nullptr, nullptr, nullptr, nullptr, C.op->n_dofs, i, 0.0),
                        C.op->precision});
      }
      add_seg_op<kSegResidual>(res);
      add_seg_op<kSegFill>(fill);
      add_transfer(g, P, 1);
    }
    add_chebyshev(0, has(0), true);  // coarse: Chebyshev solver mode
    for (int g = 1; g < depth; ++g) {  // up: prolongate, post-smooth
      const std::vector<int> P = has(g);
      add_transfer(g, P, 0);
      add_chebyshev(g, P, false);
    }
    {  // copy_from_mg: z = sol of the finest level
      std::map<int, std::vector<VecSeg>> by;
      for (int i : ins) {
        MGLevel &T = mg[i]->levels.back();
        by[T.op->precision].push_back(make_seg(slice(ws_z, i), T.sol.ptr, nullptr, nullptr, nullptr, T.op->n_dofs, i));
      }
      for (auto &kv : by) {
        SegTable *t = new_seg(kv.second, fprec(), kv.first);
        cur->steps.push_back([this, t](hipStream_t s) { seg_launch<kSegConvert>(*t, active.as<int>(), nullptr, s); });
        ++cur->launches;
      }
    }
  }

  // Programs and level groups depend on the tuning: drop them after a change.
  void reset_programs() {
    MFG_HIP(hipStreamSynchronize(stream));
    programs.clear();
    level_groups.clear();
  }

  std::vector<int> all() const {
    std::vector<int> v(n);
    for (int i = 0; i < n; ++i) v[i] = i;
    return v;
  }
  std::vector<int> active_list() const {
    std::vector<int> v;
    for (int i = 0; i < n; ++i)
      if (host_active[i]) v.push_back(i);
    return v;
  }

  void build() {
    std::vector<const OperatorData *> f(fine.begin(), fine.end());
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i) ids[i] = i;
    fine_group.reset(new ApplyGroup(f, ids));
    fine_group->build(tu_fine, stream);
    std::vector<const void *> src;
    std::vector<void *> dst;
    std::vector<VecSeg> pz, xp, rr, rz, ph;
    for (int i = 0; i < n; ++i) {
      const int64_t m = fine[i]->n_dofs;
      src.push_back(slice(ws_p, i));
      dst.push_back(slice(ws_h, i));
      pz.push_back(make_seg(slice(ws_p, i), slice(ws_z, i), nullptr, nullptr, nullptr, m, i));
      xp.push_back(make_seg(slice(ws_p, i), slice(ws_z, i), nullptr, nullptr, nullptr, m, i));
      rr.push_back(make_seg(nullptr, slice(ws_r, i), slice(ws_r, i), nullptr, nullptr, m, i));
      rz.push_back(make_seg(nullptr, slice(ws_r, i), slice(ws_z, i), nullptr, nullptr, m, i));
      ph.push_back(make_seg(nullptr, slice(ws_p, i), slice(ws_h, i), nullptr, nullptr, m, i));
    }
    fine_group->make_vecs(vec_hp, src, dst, stream);
    for (auto &pr : std::vector<std::pair<SegTable *, std::vector<VecSeg> *>>{
             {&seg_pz, &pz}, {&seg_xpby, &xp}, {&dot_rr, &rr}, {&dot_rz, &rz}, {&dot_ph, &ph}}) {
      pr.first->prec = pr.first->prec_a = fprec();
      pr.first->build(*pr.second, stream);
    }
    reset_programs();
  }

  void run_vcycle(const std::vector<int> &ins) {
    if (ins.empty()) return;
    Program &p = program_for(ins);
    ++p.uses;
    if (!use_graph || (p.uses < 2 && !p.graph)) {
      for (auto &f : p.steps) f(stream);
      return;
    }
    if (!p.graph) {
      hipGraph_t g = nullptr;
      MFG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        for (auto &f : p.steps) f(stream);
      } catch (...) {
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      MFG_HIP(hipStreamEndCapture(stream, &g));
      size_t nn = 0;
      (void)hipGraphGetNodes(g, nullptr, &nn);
      p.nodes = nn;
      const hipError_t e = hipGraphInstantiate(&p.graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      MFG_HIP(e);
    }
    MFG_HIP(hipGraphLaunch(p.graph, stream));
  }

  void tune(int reps) {
    check_unchanged();
    MFG_HIP(hipDeviceSynchronize());
    // fine operators on the workspace vectors
    fine_group->tune(tu_fine, vec_hp.ptr, reps, stream, error.as<int>());
    // level operators: the finest level of every instance (the bulk of the V-cycle)
    std::vector<const OperatorData *> ops;
    std::vector<int> ids;
    std::vector<const void *> src;
    std::vector<void *> dst;
    for (int i = 0; i < n; ++i) {
      MGLevel &T = mg[i]->levels.back();
      if (!ops.empty() && T.op->precision != ops[0]->precision) continue;
      ops.push_back(T.op);
      ids.push_back(i);
      src.push_back(T.sol.ptr);
      dst.push_back(T.t.ptr);
    }
    ApplyGroup top(ops, ids);
    top.build(tu_mg, stream);
    DeviceBuffer vt;
    top.make_vecs(vt, src, dst, stream);
    top.tune(tu_mg, vt.ptr, reps, stream, error.as<int>());
    fine_group->build(tu_fine, stream);
    reset_programs();
    MFG_HIP(hipStreamSynchronize(stream));
    check_error();
    // then the schedule of whole preconditioner applications: the lockstep
    // program (graph replays) against every instance's own V-cycle on the
    // worker threads, on the workspace vectors r -> z
    std::vector<const void *> r(n);
    std::vector<void *> z(n);
    for (int i = 0; i < n; ++i) {
      r[i] = slice(ws_r, i);
      z[i] = slice(ws_z, i);
    }
    auto time_pc = [&](bool conc) {
      concurrent = conc;
      for (int w = 0; w < 3; ++w) precondition(r.data(), z.data());
      const auto t0 = std::chrono::steady_clock::now();
      for (int k = 0; k < reps; ++k) precondition(r.data(), z.data());
      return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
    };
    ms_lockstep = time_pc(false);
    ms_concurrent = time_pc(true);
    concurrent = ms_concurrent < ms_lockstep;
    const int sched = concurrent ? MFG_SCHEDULE_STREAMS : (tu_mg.queue == MFG_SCHEDULE_QUEUE ? 1 : 0);
    tu_mg.ms_streams = tu_fine.ms_streams = ms_concurrent;
    if (concurrent) tu_fine.queue = tu_mg.queue = sched;
  }

  // z = MG(r) for every instance
  void precondition(const void *const *src, void *const *dst) {
    check_unchanged();
    if (concurrent) {
      std::string err;
      std::vector<int32_t> st(n, MFG_OK);
      run_concurrent([&](int i) { mg[i]->vmult(dst[i], src[i], fprec()); }, st.data(), err);
      for (int i = 0; i < n; ++i)
        if (st[i] != MFG_OK) throw Error(st[i], "batched preconditioner: " + err);
      return;
    }
    for (int i = 0; i < n; ++i) host_active[i] = 1;
    upload_active();
    std::vector<VecSeg> in, out;
    for (int i = 0; i < n; ++i) {
      in.push_back(make_seg(slice(ws_r, i), src[i], nullptr, nullptr, nullptr, fine[i]->n_dofs, i));
      out.push_back(make_seg(dst[i], slice(ws_z, i), nullptr, nullptr, nullptr, fine[i]->n_dofs, i));
    }
    SegTable ti, to;
    ti.prec = ti.prec_a = to.prec = to.prec_a = fprec();
    ti.build(in, stream);
    to.build(out, stream);
    seg_launch<kSegConvert>(ti, nullptr, nullptr, stream);
    run_vcycle(all());
    seg_launch<kSegConvert>(to, nullptr, nullptr, stream);
    MFG_HIP(hipStreamSynchronize(stream));
    check_error();
  }

  // Runs f(i) for every instance on the pool's worker threads, instance i's
  // operators (fine and levels) switched to its worker's stream meanwhile.
  // Errors of one instance go to status[i] (when given) and stop nothing else.
  void ensure_pool() {
    if (!pool.assign.empty()) return;
    std::vector<double> w;
    for (int i = 0; i < n; ++i) {
      double x = apply_work(*fine[i]);
      for (MGLevel &L : mg[i]->levels) x += 11.0 * apply_work(*L.op);  // operator applications per cycle
      w.push_back(x);
    }
    pool.build(w, kBatchStreams);
  }
  template <class F>
  void run_concurrent(F &&f, int32_t *status, std::string &first_error) {
    ensure_pool();
    MFG_HIP(hipDeviceSynchronize());  // the callers' vectors are complete on every stream
    std::mutex mu;
    std::vector<std::thread> workers;
    for (size_t j = 0; j < pool.assign.size(); ++j) {
      workers.emplace_back([&, j] {
        hipStream_t st = pool.streams[j];
        for (int i : pool.assign[j]) {
          std::vector<OperatorData *> ops{fine[i]};
          for (MGLevel &L : mg[i]->levels)
            if (L.op != fine[i]) ops.push_back(L.op);
          std::vector<hipStream_t> saved;
          for (OperatorData *op : ops) {
            saved.push_back(op->stream);
            op->stream = st;
          }
          int code = MFG_OK;
          std::string msg;
          try {
            f(i);
            MFG_HIP(hipStreamSynchronize(st));
          } catch (const Error &e) {
            code = e.code;
            msg = e.what();
            (void)hipStreamSynchronize(st);
          } catch (const std::exception &e) {
            code = MFG_ERR_ARGUMENT;
            msg = e.what();
            (void)hipStreamSynchronize(st);
          }
          for (size_t k = 0; k < ops.size(); ++k) ops[k]->stream = saved[k];
          if (status) status[i] = code;
          if (code != MFG_OK) {
            std::lock_guard<std::mutex> g(mu);
            if (first_error.empty()) first_error = msg;
          }
        }
      });
    }
    for (auto &t : workers) t.join();
  }

  // every instance's own cg_solve (MFG_PRECOND_MG), concurrently
  void solve_concurrent(const void *const *b, void *const *x, double tol_rel, int max_iter, mfg_solve_info *info,
                        int32_t *status) {
    ensure_pool();
    // one CG workspace and reducer per worker
    std::vector<std::unique_ptr<CGWorkspace>> ws(pool.assign.size());
    std::vector<std::unique_ptr<Reducer>> red(pool.assign.size());
    std::vector<int> worker_of(n, 0);
    for (size_t j = 0; j < pool.assign.size(); ++j) {
      ws[j].reset(new CGWorkspace());
      red[j].reset(new Reducer());
      for (int i : pool.assign[j]) worker_of[i] = (int)j;
    }
    std::string err;
    run_concurrent(
        [&](int i) {
          const int j = worker_of[i];
          std::memset(&info[i], 0, sizeof(info[i]));
          cg_solve(*fine[i], MFG_PRECOND_MG, mg[i], 1.0, b[i], x[i], tol_rel, max_iter, &info[i], *ws[j], *red[j]);
        },
        status, err);
    for (int i = 0; i < n; ++i)
      if (status[i] != MFG_OK && status[i] != MFG_ERR_NOCONV && status[i] != MFG_ERR_NUMERIC)
        throw Error(status[i], "batched solve: " + err);
  }

  // cg_solve with the multigrid preconditioner, all instances in lockstep
  void solve(const void *const *b, void *const *x, double tol_rel, int max_iter, mfg_solve_info *info,
             int32_t *status) {
    check_unchanged();
    if (concurrent) {
      solve_concurrent(b, x, tol_rel, max_iter, info, status);
      return;
    }
    for (int i = 0; i < n; ++i) {
      MFG_REQUIRE(b[i] && x[i], MFG_ERR_ARGUMENT, "null vector");
      host_active[i] = 1;
    }
    upload_active();
    // per-solve tables over the caller's vectors
    DeviceBuffer vec_hx;
    std::vector<const void *> src(x, x + n);
    std::vector<void *> dst;
    std::vector<VecSeg> rb, bb, upd;
    for (int i = 0; i < n; ++i) {
      const int64_t m = fine[i]->n_dofs;
      dst.push_back(slice(ws_h, i));
      rb.push_back(make_seg(slice(ws_r, i), b[i], slice(ws_h, i), nullptr, nullptr, m, i, -1.0));
      bb.push_back(make_seg(nullptr, b[i], b[i], nullptr, nullptr, m, i));
      upd.push_back(make_seg(x[i], slice(ws_r, i), slice(ws_p, i), slice(ws_h, i), nullptr, m, i));
    }
    fine_group->make_vecs(vec_hx, src, dst, stream);
    SegTable t_rb, t_bb, t_upd;
    for (SegTable *t : {&t_rb, &t_bb, &t_upd}) t->prec = t->prec_a = fprec();
    t_rb.build(rb, stream);
    t_bb.build(bb, stream);
    t_upd.build(upd, stream);
    const int *act = active.as<int>();
    // r = b - A x
    fine_group->launch(stream, vec_hx.ptr, act, false, tu_fine, error.as<int>());
    seg_launch<kSegAddScaled>(t_rb, act, nullptr, stream);
    red.dot(t_bb, act, stream);
    const std::vector<double> bn = red.fetch(n, stream);
    red.dot(dot_rr, act, stream);
    const std::vector<double> r0 = red.fetch(n, stream);
    std::vector<double> tol(n), res(n), gh(n);
    std::vector<std::vector<double>> alphas(n), betas(n);
    int n_active = 0;
    for (int i = 0; i < n; ++i) {
      tol[i] = tol_rel * std::sqrt(bn[i]);
      res[i] = std::sqrt(r0[i]);
      info[i].initial_residual = res[i];
      info[i].iterations = 0;
      info[i].condition_estimate = 1.0;
      status[i] = MFG_OK;
      host_active[i] = res[i] > tol[i] ? 1 : 0;
      n_active += host_active[i];
    }
    if (n_active) {
      upload_active();
      run_vcycle(active_list());
      red.dot(dot_rz, act, stream);
      const std::vector<double> g0 = red.fetch(n, stream);
      for (int i = 0; i < n; ++i) gh[i] = g0[i];
      seg_launch<kSegConvert>(seg_pz, act, nullptr, stream);
      for (int it = 1; it <= max_iter && n_active; ++it) {
        fine_group->launch(stream, vec_hp.ptr, act, false, tu_fine, error.as<int>());
        red.dot(dot_ph, act, stream);
        const std::vector<double> ph = red.fetch(n, stream);
        bool changed = false;
        for (int i = 0; i < n; ++i) {
          if (!host_active[i]) continue;
          if (!(ph[i] > 0)) {  // breakdown of this instance
            status[i] = MFG_ERR_NUMERIC;
            host_active[i] = 0;
            --n_active;
            changed = true;
            continue;
          }
          host_dyn[i] = gh[i] / ph[i];
          alphas[i].push_back(host_dyn[i]);
        }
        if (changed) upload_active();
        if (!n_active) break;
        upload_dyn();
        red.cg_update(t_upd, act, dyn.as<double>(), stream);
        const std::vector<double> rr = red.fetch(n, stream);
        changed = false;
        for (int i = 0; i < n; ++i) {
          if (!host_active[i]) continue;
          res[i] = std::sqrt(rr[i]);
          info[i].iterations = it;
          if (res[i] <= tol[i]) {
            host_active[i] = 0;
            --n_active;
            changed = true;
          }
        }
        if (!n_active) break;
        if (changed) upload_active();
        run_vcycle(active_list());
        red.dot(dot_rz, act, stream);
        const std::vector<double> ghn = red.fetch(n, stream);
        for (int i = 0; i < n; ++i) {
          if (!host_active[i]) continue;
          host_dyn[i] = ghn[i] / gh[i];
          gh[i] = ghn[i];
          betas[i].push_back(host_dyn[i]);
        }
        upload_dyn();
        seg_launch<kSegXpby>(seg_xpby, act, dyn.as<double>(), stream);
      }
    }
    MFG_HIP(hipStreamSynchronize(stream));
    check_error();
    for (int i = 0; i < n; ++i) {
      info[i].final_residual = res[i];
      if (!alphas[i].empty()) {
        std::vector<double> &al = alphas[i];
        if (al.size() > betas[i].size() + 1) al.resize(betas[i].size() + 1);
        double lmin = 0, lmax = 0;
        lanczos_estimate(al, betas[i], lmin, lmax);
        info[i].condition_estimate = lmin > 0 ? lmax / lmin : 0.0;
      }
      if (status[i] == MFG_OK && res[i] > tol[i]) status[i] = MFG_ERR_NOCONV;
    }
  }
};

}  // namespace mfg

using namespace mfg;

struct mfg_batch_t {
  std::unique_ptr<BatchVmult> b;
};
struct mfg_batch_solver_t {
  std::unique_ptr<BatchSolver> s;
};

static void export_tuning(const Tuning &tu, mfg_batch_tuning *out) {
  if (!out) return;
  std::memset(out, 0, sizeof(*out));
  out->schedule = tu.queue;
  for (int kind = 0; kind < kNumKinds; ++kind)
    for (int p = 1; p <= kMaxDegree; ++p) out->cells_per_task[kind][p] = tu.cpb_for(kind * 8 + p);
  out->waves_per_simd = tu.waves;
  out->ms_colour = tu.ms_colour;
  out->ms_queue = tu.ms_queue;
  out->ms_streams = tu.ms_streams;
  out->ms_gather = tu.ms_gather;
}

extern "C" {

int mfg_batch_create(int32_t n, const mfg_operator *ops, mfg_batch *out) {
  return guarded([&] {
    MFG_REQUIRE(out && ops && n > 0, MFG_ERR_ARGUMENT, "null argument");
    std::vector<OperatorData *> v;
    for (int i = 0; i < n; ++i) {
      MFG_REQUIRE(ops[i], MFG_ERR_ARGUMENT, "null operator");
      v.push_back(ops[i]);
    }
    std::unique_ptr<mfg_batch_t> b(new mfg_batch_t());
    b->b.reset(new BatchVmult(v));
    *out = b.release();
  });
}

int mfg_batch_destroy(mfg_batch b) {
  return guarded([&] { delete b; });
}

int mfg_batch_set_stream(mfg_batch b, void *stream) {
  return guarded([&] {
    MFG_REQUIRE(b, MFG_ERR_ARGUMENT, "null batch");
    b->b->stream = (hipStream_t)stream;
  });
}

int mfg_batch_set_schedule(mfg_batch b, int32_t schedule) {
  return guarded([&] {
    MFG_REQUIRE(b, MFG_ERR_ARGUMENT, "null batch");
    MFG_REQUIRE(schedule == MFG_SCHEDULE_COLOUR || schedule == MFG_SCHEDULE_QUEUE ||
                    schedule == MFG_SCHEDULE_STREAMS || schedule == MFG_SCHEDULE_GATHER,
                MFG_ERR_ARGUMENT, "unknown schedule");
    b->b->tu.queue = schedule;
  });
}

int mfg_batch_set_cells_per_task(mfg_batch b, int32_t kind, int32_t degree, int32_t cells) {
  return guarded([&] {
    MFG_REQUIRE(b, MFG_ERR_ARGUMENT, "null batch");
    MFG_REQUIRE(kind >= 0 && kind < kNumKinds && degree >= 1 && degree <= kMaxDegree, MFG_ERR_ARGUMENT,
                "unknown kind or degree");
    MFG_REQUIRE(cells >= 0 && cells <= max_cpb(degree), MFG_ERR_ARGUMENT, "cells per task out of range");
    b->b->tu.cpb[kind * 8 + degree] = cells;
    b->b->group->build(b->b->tu, b->b->stream);
  });
}

int mfg_batch_vmult(mfg_batch b, void *const *dst, const void *const *src) {
  return guarded([&] {
    MFG_REQUIRE(b && dst && src, MFG_ERR_ARGUMENT, "null argument");
    b->b->apply(src, dst, false);
  });
}

int mfg_batch_vmult_add(mfg_batch b, void *const *dst, const void *const *src) {
  return guarded([&] {
    MFG_REQUIRE(b && dst && src, MFG_ERR_ARGUMENT, "null argument");
    b->b->apply(src, dst, true);
  });
}

int mfg_batch_synchronize(mfg_batch b) {
  return guarded([&] {
    MFG_REQUIRE(b, MFG_ERR_ARGUMENT, "null batch");
    b->b->synchronize();
  });
}

int mfg_batch_tune(mfg_batch b, const void *const *src, void *const *dst, int32_t repetitions,
                   mfg_batch_tuning *result) {
  return guarded([&] {
    MFG_REQUIRE(b && src && dst && repetitions > 0, MFG_ERR_ARGUMENT, "null argument");
    BatchVmult &v = *b->b;
    v.tune(src, dst, repetitions);
    v.synchronize();
    export_tuning(v.tu, result);
  });
}

int mfg_batch_get_tuning(mfg_batch b, mfg_batch_tuning *result) {
  return guarded([&] {
    MFG_REQUIRE(b && result, MFG_ERR_ARGUMENT, "null argument");
    export_tuning(b->b->tu, result);
  });
}

int mfg_batch_solver_create(int32_t n, const mfg_operator *fine, const mfg_mg *mgs, mfg_batch_solver *out) {
  return guarded([&] {
    MFG_REQUIRE(out && fine && mgs && n > 0, MFG_ERR_ARGUMENT, "null argument");
    std::vector<OperatorData *> f;
    std::vector<MGData *> m;
    for (int i = 0; i < n; ++i) {
      MFG_REQUIRE(fine[i] && mgs[i], MFG_ERR_ARGUMENT, "null operator or multigrid");
      f.push_back(fine[i]);
      m.push_back(mgs[i]);
    }
    std::unique_ptr<mfg_batch_solver_t> s(new mfg_batch_solver_t());
    s->s.reset(new BatchSolver(f, m));
    *out = s.release();
  });
}

int mfg_batch_solver_destroy(mfg_batch_solver s) {
  return guarded([&] { delete s; });
}

int mfg_batch_solver_use_graph(mfg_batch_solver s, int32_t on) {
  return guarded([&] {
    MFG_REQUIRE(s, MFG_ERR_ARGUMENT, "null solver");
    s->s->use_graph = on != 0;
  });
}

int mfg_batch_solver_tune(mfg_batch_solver s, int32_t repetitions, mfg_batch_tuning *fine,
                          mfg_batch_tuning *levels) {
  return guarded([&] {
    MFG_REQUIRE(s && repetitions > 0, MFG_ERR_ARGUMENT, "null argument");
    s->s->tune(repetitions);
    export_tuning(s->s->tu_fine, fine);
    export_tuning(s->s->tu_mg, levels);
  });
}

int mfg_batch_solver_set_schedule(mfg_batch_solver s, int32_t schedule) {
  return guarded([&] {
    MFG_REQUIRE(s, MFG_ERR_ARGUMENT, "null solver");
    MFG_REQUIRE(schedule == MFG_SCHEDULE_COLOUR || schedule == MFG_SCHEDULE_QUEUE ||
                    schedule == MFG_SCHEDULE_STREAMS,
                MFG_ERR_ARGUMENT, "unknown schedule");
    s->s->concurrent = schedule == MFG_SCHEDULE_STREAMS;
    if (!s->s->concurrent) {
      s->s->tu_fine.queue = s->s->tu_mg.queue = schedule;
      s->s->build();
    } else {
      s->s->tu_fine.queue = s->s->tu_mg.queue = schedule;
    }
  });
}

int mfg_batch_solver_precondition(mfg_batch_solver s, void *const *dst, const void *const *src) {
  return guarded([&] {
    MFG_REQUIRE(s && dst && src, MFG_ERR_ARGUMENT, "null argument");
    s->s->precondition(src, dst);
  });
}

int mfg_batch_solver_solve(mfg_batch_solver s, const void *const *b, void *const *x, double tol_rel,
                           int32_t max_iter, mfg_solve_info *info, int32_t *status) {
  return guarded([&] {
    MFG_REQUIRE(s && b && x && info && status, MFG_ERR_ARGUMENT, "null argument");
    MFG_REQUIRE(tol_rel > 0 && max_iter > 0, MFG_ERR_ARGUMENT, "invalid tolerance or iteration limit");
    s->s->solve(b, x, tol_rel, max_iter, info, status);
  });
}

int mfg_batch_solver_info(mfg_batch_solver s, int32_t *vcycle_launches, int32_t *graph_nodes) {
  return guarded([&] {
    MFG_REQUIRE(s, MFG_ERR_ARGUMENT, "null solver");
    auto &p = s->s->program_for(s->s->all());
    if (vcycle_launches) *vcycle_launches = p.launches;
    if (graph_nodes) *graph_nodes = (int32_t)p.nodes;
  });
}

}  // extern "C"
