// Kernels that serve many independent operators (instances) in one launch:
// heat and solid, degrees 1..6, uneven meshes.
//
// Cell work is cut into tasks (one instance, one colour, up to cpb cells of
// that colour).  Tasks are stored colour-major across instances, so one
// workgroup size (kBatchThreads) fits every degree: a task of degree p uses
// cpb * (p+1)^3 <= kBatchThreads threads and the per-degree cpb is a tuning
// parameter.  Two schedules share the kernel:
//  - per colour: one launch per colour, a workgroup per task, no waiting;
//  - queue: one persistent launch per operator application; workgroups take
//    tasks in order from an atomic ticket, and a task of colour c waits until
//    all tasks of the instance's previous non-empty colour have finished.
//    Every wait is on tasks with smaller tickets, which were taken by running
//    workgroups, so the launch cannot deadlock whatever its grid size; the
//    wait is also bounded (kSpinLimit) and reports an error instead of
//    hanging.  The last workgroup to leave clears the counters, so the next
//    launch (or graph replay) starts from zero without a fill.
// The arithmetic of a cell is the one of the single-operator kernels
// (cell_group), so batched results equal per-instance results bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cell_kernels.hpp"
#include "transfer_kernels.hpp"
#include "vector_kernels.hpp"

namespace mfg {

constexpr int kBatchThreads = 512;
constexpr int kBatchStreams = 8;  // streams of the STREAMS schedule (and worker threads of its solves)
constexpr int kSpinLimit = 1 << 22;

// instance kinds in the dispatch code (code = kind * 8 + p); kSolidF1G is
// formulation 1 from the 10-word point storage
enum : int { kSolidF1 = 0, kSolidF0 = 1, kHeat = 2, kSolidF1G = 3 };
constexpr int kNumKinds = 4;
constexpr int kind_bit(int kind) { return 1 << kind; }

template <typename Number>
struct BatchInst {
  const uint32_t *cell_nodes;
  const Number *qd;
  const int32_t *constrained;
  const Number *shape, *colloc;
  int64_t nqp, n_constrained;
  int code, cpb;
  int id;  // batch-wide instance number (active mask)
  Number prm[3];  // F1: 2 lambda, rho alpha_1 (and mu for F1G); F0: kappa, c_1, rho alpha_1; heat: dt
  const Number *cell_jxw, *wq;  // F1: per-cell reference JxW (or null) and tensor weights
  const Number *cell_geo;       // heat: per-cell Jinv and det J (or null)
  const Number *cell_jac;       // F1G: per-cell J_ref and J_ref^{-1}
};

struct BatchTask {
  int32_t inst;        // index into the instance table
  int32_t cell_begin;  // stored (colour-major) cell range of the task
  int32_t cell_end;
  int32_t fix_begin;   // share of the constrained list handled by this task
  int32_t fix_end;
  int32_t wait_slot;   // counter to wait on (queue schedule), -1 = none
  int32_t wait_count;
  int32_t done_slot;   // counter this task increments when done
};

template <typename Number>
struct BatchVecs {
  const Number *src;
  Number *dst;
};

template <typename Number>
struct BatchApplyArgs {
  const BatchInst<Number> *inst;
  const BatchTask *tasks;
  const BatchVecs<Number> *vecs;  // per entry of the instance table
  const int *active;              // per batch-wide instance number (nullable): 0 = skip
  int task_begin, task_end;
  int queue;
  int fix_mode;  // 1: dst[c] = src[c], 2: dst[c] += src[c] on constrained rows
  int assign;    // first touch stores (vmult) or always adds (vmult_add)
  int n_slots;
  int *counters;  // [0] ticket, [1] exited workgroups, [2 + slot] done counts
  int *error;
};

template <int P, typename Number, class QOp>
__device__ inline void batch_task_cells(const QOp &qop, const BatchInst<Number> &in, const BatchVecs<Number> &v,
                                        const BatchTask &tk, int assign, Number *s_tab, Number *s_tile) {
  constexpr int N = P + 1;
  cell_group<P, Number, QOp>(qop, in.cell_nodes, v.src, v.dst, tk.cell_begin, tk.cell_end, in.cpb, assign, in.shape,
                             in.colloc, s_tab, s_tab + N * N, s_tile);
}

// KINDS: the instance kinds this kernel is compiled for (kind_bit mask).  A
// launch serves the tasks of one kind, so its register allocation is that of
// the kind's paths only, not of the union.
template <int P, int KINDS, typename Number>
__device__ inline void batch_task_degree(int kind, const BatchInst<Number> &in, const BatchVecs<Number> &v,
                                         const BatchTask &tk, int assign, Number *s_tab, Number *s_tile) {
  if constexpr ((KINDS & kind_bit(kSolidF1)) != 0) {
    if (kind == kSolidF1) {
      QSolidT2F1<Number> q{in.qd, in.nqp, in.prm[0], in.prm[1], in.cell_jxw, in.wq};
      batch_task_cells<P>(q, in, v, tk, assign, s_tab, s_tile);
    }
  }
  if constexpr ((KINDS & kind_bit(kSolidF1G)) != 0) {
    if (kind == kSolidF1G) {
      QSolidT2F1G<Number> q{in.qd, in.nqp, in.prm[2], in.prm[0], in.prm[1], in.cell_jac, in.cell_jxw, in.wq};
      batch_task_cells<P>(q, in, v, tk, assign, s_tab, s_tile);
    }
  }
  if constexpr ((KINDS & kind_bit(kSolidF0)) != 0) {
    if (kind == kSolidF0) {
      QSolidT2F0<Number> q{in.qd, in.nqp, in.prm[0], in.prm[1], in.prm[2]};
      batch_task_cells<P>(q, in, v, tk, assign, s_tab, s_tile);
    }
  }
  if constexpr ((KINDS & kind_bit(kHeat)) != 0) {
    if (kind == kHeat) {
      QHeat<Number> q{in.qd, in.nqp, in.prm[0], in.cell_geo, in.wq};
      batch_task_cells<P>(q, in, v, tk, assign, s_tab, s_tile);
    }
  }
}

// WAVES: minimum waves per SIMD the register allocation must allow (2: up to
// 256 VGPRs; 4: up to 128, which makes the fused FP32 kernel spill a little
// but doubles its occupancy).  The tuner picks.
template <typename Number, int WAVES, int KINDS>
__global__ void __launch_bounds__(kBatchThreads, WAVES) batched_apply_kernel(BatchApplyArgs<Number> a) {
  __shared__ Number s_tab[2 * 7 * 7];
  // 3 NC slots of cpb*N3 <= kBatchThreads points
  __shared__ Number s_tile[((KINDS & ~kind_bit(kHeat)) != 0 ? 9 : 3) * kBatchThreads];
  __shared__ int s_task, s_last;
  const int t = threadIdx.x;
  for (int it = 0;; ++it) {
    if (t == 0) {
      if (a.queue) s_task = a.task_begin + atomicAdd(&a.counters[0], 1);
      else s_task = a.task_begin + xcd_remap(blockIdx.x, gridDim.x) + it * gridDim.x;
    }
    __syncthreads();
    // uniform across the wave: the task and instance records go to scalar registers
    const int task = __builtin_amdgcn_readfirstlane(s_task);
    if (task >= a.task_end) break;
    const BatchTask tk = a.tasks[task];
    if (a.queue && tk.wait_slot >= 0) {
      if (t == 0) {
        int spins = 0;
        while (__hip_atomic_load(&a.counters[2 + tk.wait_slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
               tk.wait_count) {
          __builtin_amdgcn_s_sleep(2);
          if (++spins > kSpinLimit) {
            atomicOr(a.error, 2);
            break;
          }
        }
        __threadfence();  // acquire: the previous colour's stores are visible
      }
      __syncthreads();
    }
    const BatchInst<Number> in = a.inst[tk.inst];
    if (!a.active || a.active[in.id]) {
      const BatchVecs<Number> v = a.vecs[tk.inst];
      if (a.fix_mode) {
        for (int idx = tk.fix_begin + t; idx < tk.fix_end; idx += kBatchThreads) {
          const int64_t dof = in.constrained[idx];
          if (a.fix_mode == 1) v.dst[dof] = v.src[dof];
          else v.dst[dof] += v.src[dof];
        }
      }
      const int kind = in.code >> 3;
      switch (in.code & 7) {
        case 1: batch_task_degree<1, KINDS, Number>(kind, in, v, tk, a.assign, s_tab, s_tile); break;
        case 2: batch_task_degree<2, KINDS, Number>(kind, in, v, tk, a.assign, s_tab, s_tile); break;
        case 3: batch_task_degree<3, KINDS, Number>(kind, in, v, tk, a.assign, s_tab, s_tile); break;
        case 4: batch_task_degree<4, KINDS, Number>(kind, in, v, tk, a.assign, s_tab, s_tile); break;
        case 5: batch_task_degree<5, KINDS, Number>(kind, in, v, tk, a.assign, s_tab, s_tile); break;
        case 6: batch_task_degree<6, KINDS, Number>(kind, in, v, tk, a.assign, s_tab, s_tile); break;
        default: break;
      }
    }
    if (a.queue) {
      __threadfence();  // release: this wave's stores reach device scope
      __syncthreads();
      if (t == 0) atomicAdd(&a.counters[2 + tk.done_slot], 1);
    }
    __syncthreads();  // s_task and the LDS tile are reused by the next task
  }
  if (a.queue) {
    if (t == 0) {
      __threadfence();
      s_last = atomicAdd(&a.counters[1], 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (s_last) {  // every other workgroup has left: clear for the next launch
      __threadfence();
      for (int i = t; i < 2 + a.n_slots; i += kBatchThreads) a.counters[i] = 0;
    }
  }
}

// ------------------------------------------------------ GATHER schedule ---
// One launch per (kind, degree) present: every cell of every instance of that
// kind and degree, a workgroup per task (DefaultCPB cells of one instance),
// with the register allocation and the cell arithmetic of the single-operator
// kernels.  Each point stores its contribution into its instance's
// contribution buffer: no colours and no waits between workgroups.  Then
// batched_gather_kernel sums every node's contributions in colour order, the
// order in which the colour launches add them, so the result is the
// per-instance result bit for bit.
template <int KIND, bool kEarlyCell, typename Number>
__device__ inline auto batch_qop(const BatchInst<Number> &in) {
  if constexpr (KIND == kSolidF1) {
    return QSolidT2F1<Number>{in.qd, in.nqp, in.prm[0], in.prm[1], in.cell_jxw, in.wq};
  } else if constexpr (KIND == kSolidF1G) {
    return QSolidT2F1G<Number, kEarlyCell>{in.qd,       in.nqp,      in.prm[2], in.prm[0],
                                           in.prm[1],   in.cell_jac, in.cell_jxw, in.wq};
  } else if constexpr (KIND == kSolidF0) {
    return QSolidT2F0<Number>{in.qd, in.nqp, in.prm[0], in.prm[1], in.prm[2]};
  } else {
    return QHeat<Number>{in.qd, in.nqp, in.prm[0], in.cell_geo, in.wq};
  }
}

template <typename Number>
struct BatchCellsArgs {
  const BatchInst<Number> *inst;
  const BatchTask *tasks;         // this launch's tasks: inst, cell_begin, cell_end
  const BatchVecs<Number> *vecs;  // per instance: src and the instance's contribution buffer
};

template <int P, int CPB, int KIND, typename Number>
__global__ void __launch_bounds__((block_threads<P, CPB>())) batched_cells_kernel(BatchCellsArgs<Number> a) {
  constexpr int N = P + 1, N3 = N * N * N, NC = KIND == kHeat ? 1 : 3;
  constexpr bool kSpare = spare_slots<P, CPB, Number, NC>();  // as apply_cells: the same arithmetic
  __shared__ Number s_shape[N * N], s_colloc[N * N];
  __shared__ Number s_tile[CPB * (kSpare ? 5 : 3) * NC * N3];
  const BatchTask tk = a.tasks[xcd_remap(blockIdx.x, gridDim.x)];
  const BatchInst<Number> in = a.inst[tk.inst];
  const BatchVecs<Number> v = a.vecs[tk.inst];
  const auto q = batch_qop<KIND, CPB == 1>(in);
  cell_group<P, Number, decltype(q), true, kSpare>(q, in.cell_nodes, v.src, v.dst, tk.cell_begin, tk.cell_end, CPB,
                                                   1, in.shape, in.colloc, s_shape, s_colloc, s_tile);
}

template <typename Number>
struct GatherSeg {
  const Number *contrib;
  const int32_t *offsets, *points;
  const uint8_t *cmask;
  const Number *src;
  Number *dst;
  int64_t n_nodes;
};

constexpr int kGatherThreads = 256;

// One workgroup per kGatherThreads nodes of one instance (segment);
// block_begin[s] is segment s's first workgroup (n_segs + 1 entries, the last
// one the grid size).
template <typename Number, int NC>
__global__ void __launch_bounds__(kGatherThreads)
    batched_gather_kernel(const GatherSeg<Number> *__restrict__ segs, const int64_t *__restrict__ block_begin,
                          int n_segs, int fix_mode) {
  const int64_t blk = blockIdx.x;
  int lo = 0, hi = n_segs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (block_begin[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  const GatherSeg<Number> g = segs[lo];
  const int64_t node = (blk - block_begin[lo]) * kGatherThreads + threadIdx.x;
  if (node >= g.n_nodes) return;
  gather_node<Number, NC>(g.contrib, g.offsets, g.points, g.cmask, g.src, g.dst, node, fix_mode);
}

// --------------------------------------------------- segmented vectors ---
// One segment per instance vector; a workgroup handles one chunk of
// kSegChunk entries of one segment (found by binary search on chunk_begin).
constexpr int kSegThreads = 256;
constexpr int kSegChunk = 4 * kSegThreads;

struct VecSeg {
  void *y;
  const void *a, *b, *c, *d;
  int64_t n;
  int64_t chunk_begin;
  double s0, s1;
  int inst;  // for the active mask and per-instance dynamic scalars
  int pad_;
};

__device__ inline int find_segment(const VecSeg *segs, int n_segs, int64_t chunk) {
  int lo = 0, hi = n_segs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].chunk_begin <= chunk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Elementwise operations (the same expressions as the single-vector kernels).
enum : int {
  kSegFill = 0,        // y = s0
  kSegConvert,         // y = (T) a           (TA = source type)
  kSegResidual,        // y = a - y
  kSegChebFirst,       // y(x) = d(a) = invD(c) * b * s0 ; stored to both
  kSegChebFirstStep,   // d(a) = invD(d) * (b - c(t)) * s0 ; x(y) += d
  kSegChebUpdate,      // d(a) = s0 d + s1 invD(d) (b - c(t)) ; x(y) += d
  kSegAddScaled,       // y = a + s0 b
  kSegXpby,            // y = a + dyn[inst] y
};

template <typename T, typename TA, int OP>
__global__ void __launch_bounds__(kSegThreads)
    seg_kernel(const VecSeg *segs, int n_segs, const int *active, const double *dyn) {
  const int64_t chunk = blockIdx.x;
  const VecSeg sg = segs[find_segment(segs, n_segs, chunk)];
  if (active && !active[sg.inst]) return;
  const int64_t base = (chunk - sg.chunk_begin) * kSegChunk;
  T *y = (T *)sg.y;
  for (int64_t i = base + threadIdx.x; i < sg.n && i < base + kSegChunk; i += kSegThreads) {
    if (OP == kSegFill) {
      y[i] = (T)sg.s0;
    } else if (OP == kSegConvert) {
      y[i] = (T)((const TA *)sg.a)[i];
    } else if (OP == kSegResidual) {
      y[i] = ((const T *)sg.a)[i] - y[i];
    } else if (OP == kSegChebFirst) {
      const T *b = (const T *)sg.b, *invD = (const T *)sg.c;
      const T v = (T)((double)invD[i] * (double)b[i] * sg.s0);
      ((T *)sg.a)[i] = v;
      y[i] = v;
    } else if (OP == kSegChebFirstStep) {
      const T *b = (const T *)sg.b, *tt = (const T *)sg.c, *invD = (const T *)sg.d;
      const T v = cheb_first_step_value<T>(invD[i], b[i], tt[i], sg.s0);
      ((T *)sg.a)[i] = v;
      y[i] += v;
    } else if (OP == kSegChebUpdate) {
      T *d = (T *)sg.a;
      const T *b = (const T *)sg.b, *tt = (const T *)sg.c, *invD = (const T *)sg.d;
      const T v = cheb_update_value<T>(d[i], invD[i], b[i], tt[i], sg.s0, sg.s1);
      d[i] = v;
      y[i] += v;
    } else if (OP == kSegAddScaled) {
      y[i] = (T)((double)((const T *)sg.a)[i] + sg.s0 * (double)((const T *)sg.b)[i]);
    } else if (OP == kSegXpby) {
      y[i] = (T)((double)((const T *)sg.a)[i] + dyn[sg.inst] * (double)y[i]);
    }
  }
}

// Reductions with the partition of the single-vector kernels (kRedBlocks
// workgroups striding over the vector), one row of workgroups per segment:
// each instance's sum is bit-identical to dot_partial + sum_partials.
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
    seg_dot_partial_kernel(const VecSeg *segs, const int *active, double *partials) {
  const VecSeg sg = segs[blockIdx.y];
  double *out = partials + (int64_t)blockIdx.y * kRedBlocks;
  double v = 0;
  if (!active || active[sg.inst]) {
    const T *x = (const T *)sg.a, *z = (const T *)sg.b;
    T xv[kRedBatch], zv[kRedBatch];
    red_batches(
        sg.n,
        [&](int64_t i, int u) {
          xv[u] = x[i];
          zv[u] = z[i];
        },
        [&](int64_t, int u) { v += (double)xv[u] * (double)zv[u]; });
  }
  block_reduce_store<T>(v, out);
}

// x(y) += alpha p(b) ; r(a) -= alpha h(c) ; partial r.r  with alpha = dyn[inst]
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
    seg_cg_update_kernel(const VecSeg *segs, const int *active, const double *dyn, double *partials) {
  const VecSeg sg = segs[blockIdx.y];
  double *out = partials + (int64_t)blockIdx.y * kRedBlocks;
  double v = 0;
  if (!active || active[sg.inst]) {
    T *x = (T *)sg.y, *r = (T *)sg.a;
    const T *p = (const T *)sg.b, *h = (const T *)sg.c;
    const double alpha = dyn[sg.inst];
    T rv[kRedBatch], xv[kRedBatch], pv[kRedBatch], hv[kRedBatch];
    red_batches(
        sg.n,
        [&](int64_t i, int u) {
          rv[u] = r[i];
          xv[u] = x[i];
          pv[u] = p[i];
          hv[u] = h[i];
        },
        [&](int64_t i, int u) {
          x[i] = (T)((double)xv[u] + alpha * (double)pv[u]);
          const T ri = (T)((double)rv[u] - alpha * (double)hv[u]);
          r[i] = ri;
          v += (double)ri * (double)ri;
        });
  }
  block_reduce_store<T>(v, out);
}

__global__ void __launch_bounds__(kRedThreads) seg_sum_partials_kernel(const double *partials, double *out) {
  __shared__ double s[kRedThreads];
  const double *p = partials + (int64_t)blockIdx.x * kRedBlocks;
  double v = 0;
  for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) v += p[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// ----------------------------------------------------- batched transfer ---
// One workgroup per (instance, coarse cell) task; code = p + 8 * (nc == 3).
template <typename CN, typename FN>
struct BatchTransfer {
  TransferArgs<CN, FN> args;  // cell_begin/cell_end unused
  int code, inst;
};

struct BatchCellTask {
  int32_t entry;  // index into the transfer table
  int32_t cell;
};

template <int KIND, int P, int NC, typename CN, typename FN>
__device__ inline void transfer_cell(const TransferArgs<CN, FN> &a, int cell, double *lds) {
  if (KIND == 0) prolongate_cell<P, NC, CN, FN>(a, cell, lds);
  else restrict_cell<P, NC, CN, FN>(a, cell, lds);
}

// KIND 0: fine += P coarse ; KIND 1: coarse += P^T fine (one colour per launch)
template <int KIND, typename CN, typename FN>
__global__ void __launch_bounds__(256)
    batched_transfer_kernel(const BatchTransfer<CN, FN> *tab, const BatchCellTask *tasks, int task_begin,
                            int task_end, const int *active) {
  __shared__ double lds[transfer_lds<6>()];
  const int task = task_begin + blockIdx.x;
  if (task >= task_end) return;
  const BatchCellTask tk = tasks[task];
  const BatchTransfer<CN, FN> &e = tab[tk.entry];
  if (active && !active[e.inst]) return;
  const TransferArgs<CN, FN> a = e.args;
  switch (e.code) {
    case 1: transfer_cell<KIND, 1, 1, CN, FN>(a, tk.cell, lds); break;
    case 2: transfer_cell<KIND, 2, 1, CN, FN>(a, tk.cell, lds); break;
    case 3: transfer_cell<KIND, 3, 1, CN, FN>(a, tk.cell, lds); break;
    case 4: transfer_cell<KIND, 4, 1, CN, FN>(a, tk.cell, lds); break;
    case 5: transfer_cell<KIND, 5, 1, CN, FN>(a, tk.cell, lds); break;
    case 6: transfer_cell<KIND, 6, 1, CN, FN>(a, tk.cell, lds); break;
    case 9: transfer_cell<KIND, 1, 3, CN, FN>(a, tk.cell, lds); break;
    case 10: transfer_cell<KIND, 2, 3, CN, FN>(a, tk.cell, lds); break;
    case 11: transfer_cell<KIND, 3, 3, CN, FN>(a, tk.cell, lds); break;
    case 12: transfer_cell<KIND, 4, 3, CN, FN>(a, tk.cell, lds); break;
    case 13: transfer_cell<KIND, 5, 3, CN, FN>(a, tk.cell, lds); break;
    case 14: transfer_cell<KIND, 6, 3, CN, FN>(a, tk.cell, lds); break;
    default: break;
  }
}

}  // namespace mfg
