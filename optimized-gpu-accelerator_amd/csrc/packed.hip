// Packed uneven batches: many unrelated LPs (different A, sizes, scalings)
// advance together, one restarted-PDHG iteration of all of them per pair of
// launches.  Their scaled matrices are concatenated (block-diagonal CSR with
// global column indices), every instance keeps its own restart / primal
// weight / step state, and the rows are cut into tasks: one workgroup's worth
// of rows of one instance, grouped by row length.  The kernels are persistent
// over eight device work queues (one per XCD): each workgroup draws task
// tickets with an atomic until its queue is empty, skips instances that have
// converged, and exits; the workgroup that draws a queue's last ticket resets
// it for the next launch.  No workgroup ever waits for another, so every wave
// reaches its exit.  When few live tasks remain they run as direct launches
// (one workgroup per task), and the last few live LPs finish in single-LP
// runs of the batched engine.
//
// The reference solves such LPs one at a time (solvePdhg, src/pdhg/pdhg.h:73);
// the per-instance algorithm is the same (src/pdhg/pdhg.cpp:483-687).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>

#include "../../include/pdhg_hip.h"
#include "device.h"
#include "engine.h"
#include "kcommon.h"
#include "prepare.h"

namespace pdhgamd {

namespace {

struct PTask {
  int inst, start, count, logS;  // rows perm[start .. start+count), 1 << logS lanes per row
};

struct PScal {  // per-instance scalars
  const double *eta, *beta, *gamma, *offset;
};

// Rows of one task: up to kTaskGroups workgroup-sized groups of one bucket of
// one instance, processed one group after the other.
constexpr int kTaskGroups = 4;

// Eight work queues, one per XCD: workgroups are dispatched round-robin over
// the XCDs, so queue q (blockIdx % 8 == q) serves one XCD and covers the
// contiguous task range beg[q] .. beg[q+1] (neighbouring instances share that
// XCD's L2).  Each queue's counter has its own 128-byte line.
struct Queues {
  int beg[9];
  int direct;  // 1: one workgroup per task (task = blockIdx), no tickets
};
constexpr int kQueueStride = 32;  // ints between counters
constexpr int kQueueSet = 8 * kQueueStride;

// Draws the next task for the whole workgroup from its queue (-1: empty).
// Every workgroup of the grid draws exactly one exhausted ticket per launch,
// so the exhausted tickets of queue q run from len to len + nwg - 1 and the
// holder of the last one resets the counter for the next launch.
__device__ inline int next_ticket(int* counters, const Queues& qs) {
  __shared__ int ticket;
  const int q = blockIdx.x & 7;
  const int len = qs.beg[q + 1] - qs.beg[q];
  const int nwg = (static_cast<int>(gridDim.x) - q + 7) >> 3;
  __syncthreads();  // previous iteration's readers are done
  if (threadIdx.x == 0) {
    int* ctr = counters + q * kQueueStride;
    const int t = atomicAdd(ctr, 1);
    if (t == len + nwg - 1) atomicExch(ctr, 0);
    ticket = t < len ? qs.beg[q] + t : -1;
  }
  __syncthreads();
  return ticket;
}

// The first and the following tasks of a workgroup: tickets from its queue,
// or in a direct launch (a grid of exactly beg[8] workgroups) its own task.
__device__ inline int first_task(int* counters, const Queues& qs) {
  if (qs.direct) return static_cast<int>(blockIdx.x) < qs.beg[8] ? static_cast<int>(blockIdx.x) : -1;
  return next_ticket(counters, qs);
}
__device__ inline int next_task(int* counters, const Queues& qs) { return qs.direct ? -1 : next_ticket(counters, qs); }

struct PSlot {
  int row, sub, logS;
};
__device__ inline int task_groups(const PTask& t) {
  const int per = kWG >> t.logS;
  return (t.count + per - 1) / per;
}
__device__ inline PSlot task_slot(const PTask& t, const int* perm, int g) {
  PSlot s;
  s.logS = t.logS;
  const int r = g * (kWG >> t.logS) + (threadIdx.x >> t.logS);
  s.sub = threadIdx.x & ((1 << t.logS) - 1);
  s.row = r < t.count ? perm[t.start + r] : -1;
  return s;
}

template <typename T, bool CHECK>
__global__ __launch_bounds__(kWG) void k_pcol(const PTask* __restrict__ tasks, int ntasks, const int* __restrict__ perm,
                                              int* queue, Queues qs, const int* __restrict__ ptr,
                                              const int* __restrict__ idx, const T* __restrict__ val,
                                              const T* __restrict__ y, const T* __restrict__ yp, T* __restrict__ x,
                                              T* __restrict__ xp, T* __restrict__ x0, const T* __restrict__ c,
                                              const T* __restrict__ lo, const T* __restrict__ up, Inst st, int tloc,
                                              double rho, double* __restrict__ part) {
  for (int tk = first_task(queue, qs); tk >= 0; tk = next_task(queue, qs)) {
    const PTask task = tasks[tk];
    const int b = task.inst;
    if (st.done[b]) continue;
    const long long it = st.iter[b] + tloc;
    const bool rst = st.applyAt[b] == it;
    const T* __restrict__ ysrc = rst ? yp : y;
    const T tau = static_cast<T>(st.tau[b]);
    const double kk = static_cast<double>(it - st.restartAt[b]);
    const double lam = (kk + 1.0) / (kk + 2.0);
    double q[2] = {0.0, 0.0};
    const int ng = task_groups(task);
    for (int g = 0; g < ng; ++g) {
      const PSlot s = task_slot(task, perm, g);
      T acc = 0;
      if (s.row >= 0) acc = row_dot<0, T>(idx, val, ysrc, ptr[s.row] + s.sub, ptr[s.row + 1], 1 << s.logS, 0);
      acc = reduce_sub<0>(acc, s.logS);
      if (s.row >= 0 && s.sub == 0) {
        const int j = s.row;
        const T xin = rst ? xp[j] : x[j];
        const T xo = rst ? xin : x0[j];
        if (rst) x0[j] = xin;
        const T v = clamp_to<T>(xin - tau * (c[j] - acc), lo[j], up[j]);
        x[j] = static_cast<T>(lam) * (static_cast<T>(1.0 + rho) * v - static_cast<T>(rho) * xin) +
               static_cast<T>(1.0 - lam) * xo;
        xp[j] = v;
        if (CHECK) {
          const double d = static_cast<double>(xin) - static_cast<double>(v);
          const double e = static_cast<double>(v) - static_cast<double>(xo);
          q[0] += d * d;
          q[1] += e * e;
        }
      }
    }
    if (CHECK) block_partials<0, 2>(q, part, 0, ntasks, tk);
  }
}

template <typename T, bool CHECK>
__global__ __launch_bounds__(kWG) void k_prow(const PTask* __restrict__ tasks, int ntasks, const int* __restrict__ perm,
                                              int* queue, Queues qs, const int* __restrict__ ptr,
                                              const int* __restrict__ idx, const T* __restrict__ val,
                                              const double* __restrict__ val64, const T* __restrict__ xp,
                                              T* __restrict__ y, T* __restrict__ yp, T* __restrict__ y0,
                                              T* __restrict__ ax, T* __restrict__ axp, T* __restrict__ ax0,
                                              const T* __restrict__ rl, const T* __restrict__ ru,
                                              const double* __restrict__ rl64, const double* __restrict__ ru64,
                                              const double* __restrict__ rfac, Inst st, int tloc, double rho,
                                              double* __restrict__ part) {
  for (int tk = first_task(queue, qs); tk >= 0; tk = next_task(queue, qs)) {
    const PTask task = tasks[tk];
    const int b = task.inst;
    if (st.done[b]) continue;
    const long long it = st.iter[b] + tloc;
    const bool rst = st.applyAt[b] == it;
    const T sg = static_cast<T>(st.sigma[b]);
    const double kk = static_cast<double>(it - st.restartAt[b]);
    const double lam = (kk + 1.0) / (kk + 2.0);
    const T lamT = static_cast<T>(lam), omT = static_cast<T>(1.0 - lam);
    const T r1 = static_cast<T>(1.0 + rho), r0 = static_cast<T>(rho);
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    const int ng = task_groups(task);
    for (int g = 0; g < ng; ++g) {
      const PSlot s = task_slot(task, perm, g);
      T acc = 0;
      double acc64 = 0;
      if (s.row >= 0) {
        const int e0 = ptr[s.row] + s.sub, e1 = ptr[s.row + 1], step = 1 << s.logS;
        acc = row_dot<0, T>(idx, val, xp, e0, e1, step, 0);
        if (CHECK) acc64 = row_dot<0, double>(idx, val64, xp, e0, e1, step, 0);
      }
      acc = reduce_sub<0>(acc, s.logS);
      if (CHECK) acc64 = reduce_sub<0>(acc64, s.logS);
      if (s.row >= 0 && s.sub == 0) {
        const int i = s.row;
        const T yin = rst ? yp[i] : y[i];
        const T axin = rst ? axp[i] : ax[i];
        const T yo = rst ? yin : y0[i];
        const T axo = rst ? axin : ax0[i];
        if (rst) {
          y0[i] = yin;
          ax0[i] = axin;
        }
        // Reflected dual step in the cancellation-free form y+ = y - sigma (g - w).
        const T gr = static_cast<T>(2) * acc - axin;
        const T tq = yin / sg;
        const T dlo = gr - rl[i], dhi = gr - ru[i];
        const T yplus = dlo < tq ? yin - sg * dlo : (dhi > tq ? yin - sg * dhi : static_cast<T>(0));
        y[i] = lamT * (r1 * yplus - r0 * yin) + omT * yo;
        ax[i] = lamT * (r1 * acc - r0 * axin) + omT * axo;
        if (CHECK) {
          yp[i] = yplus;
          axp[i] = acc;
          const double L = rl64[i], U = ru64[i];
          const double res = (fmax(L - acc64, 0.0) + fmax(acc64 - U, 0.0)) * rfac[i];
          const double yy = static_cast<double>(yplus);
          q[0] += res * res;
          if (yy > 0 && L > -1e300) q[1] += yy * L;
          else if (yy < 0 && U < 1e300) q[1] += yy * U;
          const double d = static_cast<double>(yin) - yy, e = yy - static_cast<double>(yo);
          q[2] += d * d;
          q[3] += e * e;
        }
      }
    }
    if (CHECK) block_partials<0, 4>(q, part, 0, ntasks, tk);
  }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_pcolcheck(const PTask* __restrict__ tasks, int ntasks,
                                                   const int* __restrict__ perm, int* queue, Queues qs,
                                                   const int* __restrict__ ptr, const int* __restrict__ idx,
                                                   const double* __restrict__ val64, const T* __restrict__ yp,
                                                   const T* __restrict__ xp, const double* __restrict__ c64,
                                                   const double* __restrict__ lo64, const double* __restrict__ up64,
                                                   const double* __restrict__ cfac, Inst st,
                                                   double* __restrict__ part) {
  for (int tk = first_task(queue, qs); tk >= 0; tk = next_task(queue, qs)) {
    const PTask task = tasks[tk];
    if (st.done[task.inst]) continue;
    double q[3] = {0.0, 0.0, 0.0};
    const int ng = task_groups(task);
    for (int g = 0; g < ng; ++g) {
      const PSlot s = task_slot(task, perm, g);
      double acc = 0;
      if (s.row >= 0) acc = row_dot<0, double>(idx, val64, yp, ptr[s.row] + s.sub, ptr[s.row + 1], 1 << s.logS, 0);
      acc = reduce_sub<0>(acc, s.logS);
      if (s.row >= 0 && s.sub == 0) {
        const int j = s.row;
        const double cj = c64[j], l = lo64[j], u = up64[j];
        const double d = cj - acc;
        const bool hl = l > -1e300, hu = u < 1e300;
        double res = 0;
        if (!hl && !hu) res = fabs(d);
        else if (!hl) res = fmax(d, 0.0);
        else if (!hu) res = fmax(-d, 0.0);
        res *= cfac[j];
        q[0] += res * res;
        q[1] += cj * static_cast<double>(xp[j]);
        if (d > 0 && hl) q[2] += d * l;
        else if (d < 0 && hu) q[2] += d * u;
      }
    }
    block_partials<0, 3>(q, part, 2, ntasks, tk);
  }
}

// One workgroup per instance: fixed-order sums over the instance's own task
// ranges, then the shared FP64 check logic with that instance's scalars.
__global__ __launch_bounds__(kWG) void k_pfinalize(Inst st, const double* __restrict__ colPart, int nColTasks,
                                                   const int* __restrict__ colBeg, const double* __restrict__ rowPart,
                                                   int nRowTasks, const int* __restrict__ rowBeg, PScal ps,
                                                   double tol, long long maxIter, int blockIters, int checkEvery) {
  const int b = blockIdx.x;
  if (st.done[b]) return;
  double v[9];
  for (int q = 0; q < 5; ++q) {
    const double* p = colPart + static_cast<size_t>(q) * nColTasks;
    double s = 0;
    for (int k = colBeg[b] + threadIdx.x; k < colBeg[b + 1]; k += kWG) s += p[k];
    v[q] = s;
  }
  for (int q = 0; q < 4; ++q) {
    const double* p = rowPart + static_cast<size_t>(q) * nRowTasks;
    double s = 0;
    for (int k = rowBeg[b] + threadIdx.x; k < rowBeg[b + 1]; k += kWG) s += p[k];
    v[5 + q] = s;
  }
  __shared__ double sh[9][kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = 0; q < 9; ++q) {
    double s = v[q];
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) sh[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double S[9];
  for (int q = 0; q < 9; ++q) {
    S[q] = sh[q][0];
    for (int w = 1; w < kWaves; ++w) S[q] += sh[q][w];
  }
  RunScalars sc;
  sc.eta = ps.eta[b];
  sc.beta = ps.beta[b];
  sc.gamma = ps.gamma[b];
  sc.offset = ps.offset[b];
  sc.tol = tol;
  sc.maxIter = maxIter;
  finalize_instance(st, b, S, sc, blockIters, checkEvery);
}

// out = A in over all tasks (initial activities).
template <typename T>
__global__ __launch_bounds__(kWG) void k_pspmv(const PTask* __restrict__ tasks, int ntasks, const int* __restrict__ perm,
                                               int* queue, Queues qs, const int* __restrict__ ptr,
                                               const int* __restrict__ idx, const T* __restrict__ val,
                                               const T* __restrict__ in, T* __restrict__ out) {
  for (int tk = first_task(queue, qs); tk >= 0; tk = next_task(queue, qs)) {
    const PTask task = tasks[tk];
    const int ng = task_groups(task);
    for (int g = 0; g < ng; ++g) {
      const PSlot s = task_slot(task, perm, g);
      T acc = 0;
      if (s.row >= 0) acc = row_dot<0, T>(idx, val, in, ptr[s.row] + s.sub, ptr[s.row + 1], 1 << s.logS, 0);
      acc = reduce_sub<0>(acc, s.logS);
      if (s.row >= 0 && s.sub == 0) out[s.row] = acc;
    }
  }
}

// Tasks of a block-diagonal CSR: per instance, rows bucketed by the lane count
// nextpow2(ceil(len / nnzPerLane)) (capped at 64), each task up to
// kTaskGroups workgroups' worth of rows of one bucket of one instance.
// beg[b] .. beg[b+1] are the tasks of instance b.
void build_tasks(const std::vector<int>& ptr, const std::vector<int>& rowOff, int nnzPerLane, std::vector<PTask>& tasks,
                 std::vector<int>& perm, std::vector<int>& beg) {
  const int K = static_cast<int>(rowOff.size()) - 1;
  tasks.clear();
  perm.clear();
  beg.assign(K + 1, 0);
  for (int b = 0; b < K; ++b) {
    beg[b] = static_cast<int>(tasks.size());
    std::vector<std::vector<int>> bucket(7);
    for (int r = rowOff[b]; r < rowOff[b + 1]; ++r) {
      const int len = ptr[r + 1] - ptr[r];
      const int want = (len + nnzPerLane - 1) / std::max(1, nnzPerLane);
      bucket[std::min(6, ilog2(std::max(1, want)))].push_back(r);
    }
    for (int l = 0; l <= 6; ++l) {
      const int per = (kWG >> l) * kTaskGroups;
      for (size_t a = 0; a < bucket[l].size(); a += per) {
        PTask t;
        t.inst = b;
        t.start = static_cast<int>(perm.size());
        t.count = static_cast<int>(std::min<size_t>(per, bucket[l].size() - a));
        t.logS = l;
        perm.insert(perm.end(), bucket[l].begin() + a, bucket[l].begin() + a + t.count);
        tasks.push_back(t);
      }
    }
  }
  beg[K] = static_cast<int>(tasks.size());
}

// The live part of a task list: tasks of instances not done, in order, with
// their per-instance ranges (empty for finished instances) and the split of
// the list over the eight queues.
// When the live tasks cut into single workgroup groups number at most
// maxSplit, each is cut so: a few live instances then spread over many
// workgroups instead of a few workgroups walking their groups in turn.
void live_tasks(const std::vector<PTask>& all, const std::vector<int>& allBeg, const std::vector<int>& done,
                std::vector<PTask>& out, std::vector<int>& beg, Queues& qs, int maxSplit) {
  const int K = static_cast<int>(allBeg.size()) - 1;
  auto groups = [](const PTask& t) { return (t.count + (kWG >> t.logS) - 1) / (kWG >> t.logS); };
  long long pieces = 0;
  for (int b = 0; b < K; ++b)
    if (done.empty() || !done[b])
      for (int i = allBeg[b]; i < allBeg[b + 1]; ++i) pieces += groups(all[i]);
  const bool split = pieces <= maxSplit;
  out.clear();
  beg.assign(K + 1, 0);
  for (int b = 0; b < K; ++b) {
    beg[b] = static_cast<int>(out.size());
    if (!done.empty() && done[b]) continue;
    for (int i = allBeg[b]; i < allBeg[b + 1]; ++i) {
      if (!split) {
        out.push_back(all[i]);
        continue;
      }
      const int per = kWG >> all[i].logS;
      for (int g = 0; g * per < all[i].count; ++g) {
        PTask t = all[i];
        t.start += g * per;
        t.count = std::min(per, all[i].count - g * per);
        out.push_back(t);
      }
    }
  }
  beg[K] = static_cast<int>(out.size());
  const int n = beg[K];
  for (int q = 0; q <= 8; ++q) qs.beg[q] = static_cast<int>((static_cast<long long>(n) * q) / 8);
  qs.direct = split ? 1 : 0;
}

template <typename T>
class PackedRun {
 public:
  // A, AT: concatenated CSRs (global indices); rowOff / colOff: instance offsets.
  PackedRun(int device, const Csr& A, const Csr& AT, const std::vector<int>& rowOff, const std::vector<int>& colOff,
            const std::vector<double>& rfac, const std::vector<double>& cfac, int nnzPerLane, int checkEvery,
            double rho, bool useGraph)
      : M_(A.rows), N_(AT.rows), K_(static_cast<int>(rowOff.size()) - 1), checkEvery_(checkEvery), rho_(rho),
        useGraph_(useGraph) {
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&e0_));
    HIP_OK(hipEventCreate(&e1_));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, device));
    // At most four workgroups per CU, a multiple of 8 so every queue has
    // workgroups; set_tasks() shrinks the grid to the live task count.
    maxGrid_ = std::max(8, (std::max(1, prop.multiProcessorCount) * 4 + 7) / 8 * 8);
    grid_ = maxGrid_;
    std::vector<int> cperm, rperm;
    build_tasks(AT.ptr, colOff, nnzPerLane, colAll_, cperm, colBegAll_);
    build_tasks(A.ptr, rowOff, nnzPerLane, rowAll_, rperm, rowBegAll_);
    colTasks_ = up(colAll_);
    rowTasks_ = up(rowAll_);
    colPerm_ = up(cperm);
    rowPerm_ = up(rperm);
    dColBeg_ = up(colBegAll_);
    dRowBeg_ = up(rowBegAll_);
    set_tasks({});
    aPtr_ = up(A.ptr);
    aIdx_ = up(A.idx);
    tPtr_ = up(AT.ptr);
    tIdx_ = up(AT.idx);
    aVal64_ = up(A.val);
    tVal64_ = up(AT.val);
    if (sizeof(T) == 8) {
      aVal_ = reinterpret_cast<T*>(aVal64_);
      tVal_ = reinterpret_cast<T*>(tVal64_);
    } else {
      aVal_ = upT(A.val);
      tVal_ = upT(AT.val);
    }
    rfac_ = up(rfac);
    cfac_ = up(cfac);
    for (T** v : {&x_, &xp_, &x0_, &c_, &lo_, &up_}) *v = buf_.alloc<T>(N_);
    for (T** v : {&y_, &yp_, &y0_, &ax_, &axp_, &ax0_, &rl_, &ru_}) *v = buf_.alloc<T>(M_);
    for (double** v : {&c64_, &lo64_, &up64_}) *v = buf_.alloc<double>(N_);
    for (double** v : {&rl64_, &ru64_}) *v = buf_.alloc<double>(M_);
    colPart_ = buf_.alloc<double>(static_cast<size_t>(5) * colAll_.size());
    rowPart_ = buf_.alloc<double>(static_cast<size_t>(4) * rowAll_.size());
    queue_ = buf_.alloc<int>(4 * kQueueSet);
    std::vector<int> zq(4 * kQueueSet, 0);
    putv(queue_, zq);
    st_.tau = buf_.alloc<double>(K_);
    st_.sigma = buf_.alloc<double>(K_);
    st_.omega = buf_.alloc<double>(K_);
    st_.fpR = buf_.alloc<double>(K_);
    st_.fpPrev = buf_.alloc<double>(K_);
    st_.iter = buf_.alloc<long long>(K_);
    st_.since = buf_.alloc<long long>(K_);
    st_.restartAt = buf_.alloc<long long>(K_);
    st_.applyAt = buf_.alloc<long long>(K_);
    st_.doneIter = buf_.alloc<long long>(K_);
    st_.done = buf_.alloc<int>(K_);
    st_.restarts = buf_.alloc<int>(K_);
    cNorm_ = buf_.alloc<double>(K_);
    rhsNorm_ = buf_.alloc<double>(K_);
    st_.cNorm = cNorm_;
    st_.rhsNorm = rhsNorm_;
    st_.rec = buf_.alloc<CheckRecord>(K_);
    for (double** v : {&eta_, &beta_, &gamma_, &offset_}) *v = buf_.alloc<double>(K_);
    ps_ = PScal{eta_, beta_, gamma_, offset_};
  }
  ~PackedRun() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    (void)hipEventDestroy(e0_);
    (void)hipEventDestroy(e1_);
    (void)hipStreamDestroy(s_);
  }

  void upload(const std::vector<double>& c, const std::vector<double>& lo, const std::vector<double>& upv,
              const std::vector<double>& rl, const std::vector<double>& ru, const std::vector<double>& x,
              const std::vector<double>& y) {
    putT(c_, c);
    putT(lo_, lo);
    putT(up_, upv);
    putT(rl_, rl);
    putT(ru_, ru);
    put64(c64_, c);
    put64(lo64_, lo);
    put64(up64_, upv);
    put64(rl64_, rl);
    put64(ru64_, ru);
    for (T* v : {x_, x0_, xp_}) putT(v, x);
    for (T* v : {y_, y0_, yp_}) putT(v, y);
  }

  void init_state(const std::vector<double>& omega, const std::vector<double>& cNorm,
                  const std::vector<double>& rhsNorm, const std::vector<double>& eta, const std::vector<double>& beta,
                  const std::vector<double>& gamma, const std::vector<double>& offset, double tol, long long maxIter) {
    tol_ = tol;
    maxIter_ = maxIter;
    std::vector<double> tau(K_), sigma(K_), inf(K_, INFINITY);
    for (int b = 0; b < K_; ++b) {
      tau[b] = eta[b] / omega[b];
      sigma[b] = eta[b] * omega[b];
    }
    std::vector<long long> zero(K_, 0), neg(K_, -1);
    std::vector<int> zi(K_, 0);
    put64(st_.tau, tau);
    put64(st_.sigma, sigma);
    put64(st_.omega, omega);
    put64(st_.fpR, inf);
    put64(st_.fpPrev, inf);
    putv(st_.iter, zero);
    putv(st_.since, zero);
    putv(st_.restartAt, zero);
    putv(st_.applyAt, neg);
    putv(st_.doneIter, neg);
    putv(st_.done, zi);
    putv(st_.restarts, zi);
    put64(cNorm_, cNorm);
    put64(rhsNorm_, rhsNorm);
    put64(eta_, eta);
    put64(beta_, beta);
    put64(gamma_, gamma);
    put64(offset_, offset);
  }

  void start() {
    hipLaunchKernelGGL(k_pspmv<T>, dim3(rowGrid_), dim3(kWG), 0, s_, rowTasks_, nRow_, rowPerm_, queue_ + 3 * kQueueSet,
                       rowQ_, aPtr_, aIdx_, aVal_, x_, ax_);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(ax0_, ax_, sizeof(T) * M_, hipMemcpyDeviceToDevice, s_));
    HIP_OK(hipMemcpyAsync(axp_, ax_, sizeof(T) * M_, hipMemcpyDeviceToDevice, s_));
    HIP_OK(hipEventRecord(e0_, s_));
  }

  void run_block(int iters) {
    if (useGraph_ && iters == checkEvery_) {
      if (!exec_) {
        HIP_OK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        enqueue(iters);
        HIP_OK(hipStreamEndCapture(s_, &graph_));
        HIP_OK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
      }
      HIP_OK(hipGraphLaunch(exec_, s_));
      graphLaunches++;
    } else {
      enqueue(iters);
    }
    launches += 2 * iters + 2;
  }

  void fetch(std::vector<CheckRecord>& rec, std::vector<int>& done, std::vector<long long>& doneIter,
             std::vector<int>& restarts) {
    rec.resize(K_);
    done.resize(K_);
    doneIter.resize(K_);
    restarts.resize(K_);
    HIP_OK(hipMemcpyAsync(rec.data(), st_.rec, K_ * sizeof(CheckRecord), hipMemcpyDeviceToHost, s_));
    HIP_OK(hipMemcpyAsync(done.data(), st_.done, K_ * sizeof(int), hipMemcpyDeviceToHost, s_));
    HIP_OK(hipMemcpyAsync(doneIter.data(), st_.doneIter, K_ * sizeof(long long), hipMemcpyDeviceToHost, s_));
    HIP_OK(hipMemcpyAsync(restarts.data(), st_.restarts, K_ * sizeof(int), hipMemcpyDeviceToHost, s_));
    HIP_OK(hipStreamSynchronize(s_));
  }

  // Iterates, anchors, activities and restart state of slot q (columns
  // col0 .. col0+n, rows row0 .. row0+m) as a one-lane LaneState.
  void export_slot(int q, int col0, int n, int row0, int m, LaneState& ls) {
    PDHG_REQUIRE(q >= 0 && q < K_ && col0 >= 0 && n >= 0 && col0 + n <= N_ && row0 >= 0 && m >= 0 && row0 + m <= M_,
                 "packed: exported slot out of range");
    auto get = [&](const T* d, int off, int len, std::vector<double>& out) {
      std::vector<T> t(len);
      HIP_OK(hipMemcpyAsync(t.data(), d + off, len * sizeof(T), hipMemcpyDeviceToHost, s_));
      HIP_OK(hipStreamSynchronize(s_));
      out.assign(t.begin(), t.end());
    };
    get(x_, col0, n, ls.x);
    get(x0_, col0, n, ls.x0);
    get(xp_, col0, n, ls.xp);
    get(y_, row0, m, ls.y);
    get(y0_, row0, m, ls.y0);
    get(yp_, row0, m, ls.yp);
    get(ax_, row0, m, ls.ax);
    get(ax0_, row0, m, ls.ax0);
    get(axp_, row0, m, ls.axp);
    auto one = [&](const auto* d, auto& out) {
      using U = std::remove_cv_t<std::remove_pointer_t<decltype(d)>>;
      U v{};
      HIP_OK(hipMemcpyAsync(&v, d + q, sizeof(U), hipMemcpyDeviceToHost, s_));
      HIP_OK(hipStreamSynchronize(s_));
      out.assign(1, v);
    };
    one(st_.tau, ls.tau);
    one(st_.sigma, ls.sigma);
    one(st_.omega, ls.omega);
    one(st_.fpR, ls.fpR);
    one(st_.fpPrev, ls.fpPrev);
    one(st_.iter, ls.iter);
    one(st_.since, ls.since);
    one(st_.restartAt, ls.restartAt);
    one(st_.applyAt, ls.applyAt);
    one(st_.restarts, ls.restarts);
    ls.count = 1;
  }

  void download(bool primal, std::vector<double>& out) {
    const int len = primal ? N_ : M_;
    std::vector<T> t(len);
    HIP_OK(hipMemcpyAsync(t.data(), primal ? xp_ : yp_, len * sizeof(T), hipMemcpyDeviceToHost, s_));
    HIP_OK(hipStreamSynchronize(s_));
    out.assign(t.begin(), t.end());
  }

  double stop() {
    HIP_OK(hipEventRecord(e1_, s_));
    HIP_OK(hipEventSynchronize(e1_));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0_, e1_));
    return ms / 1000.0;
  }

  long long launches = 0, graphLaunches = 0, taskRebuilds = 0;
  int tasks() const { return nCol_ + nRow_; }

  // Drops the tasks of finished instances once they are at least half of the
  // current lists (the kernels skip them anyway, but every skipped task still
  // costs a ticket); the graph is captured again for the new lists.
  void drop_finished(const std::vector<int>& done) {
    int live = 0;  // tasks of the current lists that belong to live instances
    for (int b = 0; b < K_; ++b)
      if (!done[b]) live += (colBeg_[b + 1] - colBeg_[b]) + (rowBeg_[b + 1] - rowBeg_[b]);
    if (live == 0 || 2 * live > nCol_ + nRow_) return;
    set_tasks(done);
    if (exec_) {
      HIP_OK(hipGraphExecDestroy(exec_));
      exec_ = nullptr;
    }
    if (graph_) {
      HIP_OK(hipGraphDestroy(graph_));
      graph_ = nullptr;
    }
    ++taskRebuilds;
  }

 private:
  // Uploads the live task lists (all tasks when done is empty) into the
  // task buffers, which were sized for the full lists.
  void set_tasks(const std::vector<int>& done) {
    std::vector<PTask> ct, rt;
    std::vector<int> cb, rb;
    // The task buffers and partials hold as many tasks as the full lists.
    live_tasks(colAll_, colBegAll_, done, ct, cb, colQ_, std::min<int>(maxGrid_, colAll_.size()));
    live_tasks(rowAll_, rowBegAll_, done, rt, rb, rowQ_, std::min<int>(maxGrid_, rowAll_.size()));
    nCol_ = static_cast<int>(ct.size());
    nRow_ = static_cast<int>(rt.size());
    // No more workgroups than tasks: a launch over a few live instances
    // then starts a few workgroups instead of the full persistent grid.
    grid_ = std::min(maxGrid_, std::max(8, (std::max(nCol_, nRow_) + 7) / 8 * 8));
    // Direct launches (few live tasks): exactly one workgroup per task.
    colGrid_ = colQ_.direct ? std::max(1, nCol_) : grid_;
    rowGrid_ = rowQ_.direct ? std::max(1, nRow_) : grid_;
    if (!ct.empty()) putv(colTasks_, ct);
    if (!rt.empty()) putv(rowTasks_, rt);
    putv(dColBeg_, cb);
    putv(dRowBeg_, rb);
    colBeg_ = cb;
    rowBeg_ = rb;
  }

  void enqueue(int iters) {
    for (int t = 0; t < iters; ++t) {
      const bool chk = t == iters - 1;
      auto colK = chk ? k_pcol<T, true> : k_pcol<T, false>;
      auto rowK = chk ? k_prow<T, true> : k_prow<T, false>;
      hipLaunchKernelGGL(colK, dim3(colGrid_), dim3(kWG), 0, s_, colTasks_, nCol_, colPerm_, queue_, colQ_, tPtr_,
                         tIdx_, tVal_, y_, yp_, x_, xp_, x0_, c_, lo_, up_, st_, t, rho_, colPart_);
      hipLaunchKernelGGL(rowK, dim3(rowGrid_), dim3(kWG), 0, s_, rowTasks_, nRow_, rowPerm_, queue_ + kQueueSet, rowQ_,
                         aPtr_, aIdx_, aVal_, aVal64_, xp_, y_, yp_, y0_, ax_, axp_, ax0_, rl_, ru_, rl64_, ru64_,
                         rfac_, st_, t, rho_, rowPart_);
    }
    hipLaunchKernelGGL(k_pcolcheck<T>, dim3(colGrid_), dim3(kWG), 0, s_, colTasks_, nCol_, colPerm_,
                       queue_ + 2 * kQueueSet, colQ_, tPtr_,
                       tIdx_, tVal64_, yp_, xp_, c64_, lo64_, up64_, cfac_, st_, colPart_);
    hipLaunchKernelGGL(k_pfinalize, dim3(K_), dim3(kWG), 0, s_, st_, colPart_, nCol_, dColBeg_, rowPart_, nRow_,
                       dRowBeg_, ps_, tol_, maxIter_, iters, checkEvery_);
    HIP_OK(hipGetLastError());
  }
  // Host-to-device copies are ordered on the run's own (non-blocking)
  // stream, behind any launch still queued there, and waited for before the
  // host vector is released.
  template <typename U>
  U* up(const std::vector<U>& v) {
    U* d = buf_.alloc<U>(v.size());
    putv(d, v);
    return d;
  }
  T* upT(const std::vector<double>& v) {
    std::vector<T> t(v.begin(), v.end());
    return up(t);
  }
  void putT(T* d, const std::vector<double>& v) {
    std::vector<T> t(v.begin(), v.end());
    putv(d, t);
  }
  void put64(double* d, const std::vector<double>& v) { putv(d, v); }
  template <typename U>
  void putv(U* d, const std::vector<U>& v) {
    if (v.empty()) return;
    HIP_OK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice, s_));
    HIP_OK(hipStreamSynchronize(s_));
  }

  int M_, N_, K_, checkEvery_;
  double rho_;
  bool useGraph_;
  int grid_ = 8, maxGrid_ = 8, nCol_ = 0, nRow_ = 0, colGrid_ = 8, rowGrid_ = 8;
  std::vector<int> colBeg_, rowBeg_;  // per-instance ranges of the current task lists
  double tol_ = 1e-6;
  long long maxIter_ = 1LL << 62;
  hipStream_t s_ = nullptr;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  Buffers buf_;
  std::vector<PTask> colAll_, rowAll_;
  std::vector<int> colBegAll_, rowBegAll_;
  Queues colQ_{}, rowQ_{};
  PTask *colTasks_ = nullptr, *rowTasks_ = nullptr;
  int *colPerm_ = nullptr, *rowPerm_ = nullptr, *dColBeg_ = nullptr, *dRowBeg_ = nullptr, *queue_ = nullptr;
  int *aPtr_ = nullptr, *aIdx_ = nullptr, *tPtr_ = nullptr, *tIdx_ = nullptr;
  T *aVal_ = nullptr, *tVal_ = nullptr;
  double *aVal64_ = nullptr, *tVal64_ = nullptr, *rfac_ = nullptr, *cfac_ = nullptr;
  T *x_, *xp_, *x0_, *c_, *lo_, *up_, *y_, *yp_, *y0_, *ax_, *axp_, *ax0_, *rl_, *ru_;
  double *c64_, *lo64_, *up64_, *rl64_, *ru64_, *colPart_, *rowPart_, *cNorm_, *rhsNorm_;
  double *eta_, *beta_, *gamma_, *offset_;
  Inst st_{};
  PScal ps_{};
};

using Clock = std::chrono::steady_clock;

// Scaled data of a subset of the instances, concatenated block-diagonally
// (global row / column indices); slot q of the run is instance ids[q].
struct PackedData {
  std::vector<int> ids, rowOff, colOff;
  Csr A, AT;
  std::vector<double> rfac, cfac, c, lo, upv, rl, ru, x, y;
  std::vector<double> omega, cNorm, rhsNorm, eta, beta, gamma, offset;
};

// What one packed run leaves behind, per slot (iterates in the scaled space).
struct PackedOutcome {
  std::vector<int> done, restarts;
  std::vector<long long> doneIter;
  std::vector<CheckRecord> rec;
  std::vector<double> xs, ys;
  long long it = 0, launches = 0, graphLaunches = 0;
  bool expired = false, stalled = false;
  double kernel = 0;
  // Live slots handed to single-LP runs, with their state.
  std::vector<int> handed;
  std::vector<LaneState> handedState;
};

// xw / yw / ww (indexed by instance): a starting point in the scaled space and
// the primal weight to keep; null gives the reference's start, x0 = clamp(0),
// y0 = 0 and the weight ||c|| / ||b|| (src/pdhg/pdhg.cpp:509-545).
PackedData pack(const std::vector<Model>& lps, const std::vector<Prepared>& P, const std::vector<int>& ids,
                const std::vector<std::vector<double>>* xw, const std::vector<std::vector<double>>* yw,
                const std::vector<double>* ww) {
  PackedData d;
  const int K = static_cast<int>(ids.size());
  d.ids = ids;
  d.rowOff.assign(K + 1, 0);
  d.colOff.assign(K + 1, 0);
  for (int q = 0; q < K; ++q) {
    d.rowOff[q + 1] = d.rowOff[q] + P[ids[q]].m;
    d.colOff[q + 1] = d.colOff[q] + P[ids[q]].n;
  }
  const int M = d.rowOff[K], N = d.colOff[K];
  Csr& A = d.A;
  Csr& AT = d.AT;
  A.rows = M;
  A.cols = N;
  AT.rows = N;
  AT.cols = M;
  A.ptr.assign(1, 0);
  AT.ptr.assign(1, 0);
  std::vector<double>&rfac = d.rfac, &cfac = d.cfac, &c = d.c, &lo = d.lo, &upv = d.upv, &rl = d.rl, &ru = d.ru,
  &x = d.x, &y = d.y;
  y.reserve(M);
  std::vector<double>&omega = d.omega, &cNorm = d.cNorm, &rhsNorm = d.rhsNorm, &eta = d.eta, &beta = d.beta,
  &gamma = d.gamma, &offset = d.offset;
  for (std::vector<double>* v : {&cNorm, &rhsNorm, &eta, &beta, &gamma, &offset}) v->assign(K, 0.0);
  omega.assign(K, 1.0);
  const std::vector<int>& rowOff = d.rowOff;
  const std::vector<int>& colOff = d.colOff;
  for (int q = 0; q < K; ++q) {
    const int b = ids[q];
    const Prepared& p = P[b];
    const Model& lp = lps[b];
    for (int i = 0; i < p.m; ++i) {
      for (int e = p.A.ptr[i]; e < p.A.ptr[i + 1]; ++e) {
        A.idx.push_back(colOff[q] + p.A.idx[e]);
        A.val.push_back(p.A.val[e]);
      }
      A.ptr.push_back(static_cast<int>(A.idx.size()));
    }
    for (int j = 0; j < p.n; ++j) {
      for (int e = p.AT.ptr[j]; e < p.AT.ptr[j + 1]; ++e) {
        AT.idx.push_back(rowOff[q] + p.AT.idx[e]);
        AT.val.push_back(p.AT.val[e]);
      }
      AT.ptr.push_back(static_cast<int>(AT.idx.size()));
    }
    rfac.insert(rfac.end(), p.rfac.begin(), p.rfac.end());
    cfac.insert(cfac.end(), p.cfac.begin(), p.cfac.end());
    // Scaled case data and initial point (src/pdhg/pdhg.cpp:511-545).
    double cn0 = 0, bn0 = 0, cn = 0, bn = 0;
    for (int j = 0; j < p.n; ++j) {
      const double cj = lp.c[j] * p.C[j] / p.gamma;
      const double lj = finite_bound(lp.l[j]) ? lp.l[j] / p.C[j] / p.beta : -HUGE_VAL;
      const double uj = finite_bound(lp.u[j]) ? lp.u[j] / p.C[j] / p.beta : HUGE_VAL;
      c.push_back(cj);
      lo.push_back(lj);
      upv.push_back(uj);
      x.push_back(std::min(std::max(xw ? (*xw)[b][j] : 0.0, lj), uj));
      cn0 += lp.c[j] * lp.c[j];
      cn += cj * cj;
    }
    for (int i = 0; i < p.m; ++i) {
      const double li = finite_bound(lp.L[i]) ? lp.L[i] * p.R[i] / p.beta : -HUGE_VAL;
      const double ui = finite_bound(lp.U[i]) ? lp.U[i] * p.R[i] / p.beta : HUGE_VAL;
      rl.push_back(li);
      ru.push_back(ui);
      y.push_back(yw ? (*yw)[b][i] : 0.0);
      if (finite_bound(lp.L[i])) bn0 += lp.L[i] * lp.L[i];
      if (finite_bound(lp.U[i]) && lp.U[i] != lp.L[i]) bn0 += lp.U[i] * lp.U[i];
      if (finite_bound(li)) bn += li * li;
      if (finite_bound(ui) && ui != li) bn += ui * ui;
    }
    cNorm[q] = std::sqrt(cn0);
    rhsNorm[q] = std::sqrt(bn0);
    if (ww) omega[q] = (*ww)[b];
    else if (cn > 1e-20 && bn > 1e-20) omega[q] = std::sqrt(cn) / std::sqrt(bn);
    eta[q] = 0.998 / p.normA;
    beta[q] = p.beta;
    gamma[q] = p.gamma;
    offset[q] = p.offset;
  }
  PDHG_REQUIRE(A.idx.size() < static_cast<size_t>(INT32_MAX), "packed batch exceeds 2^31 nonzeros");
  return d;
}

// One packed run of at most maxIter iterations.  With stallRule (FP32
// iterates), an instance whose best relative KKT has not halved over the last
// 256 checks is stalled at the FP32 resolution floor; once half of the live
// instances are, the run ends with stalled = true.
template <typename T>
PackedOutcome run_packed(const PackedData& d, const pdhg_options& o, long long maxIter, Clock::time_point t0,
                         bool stallRule, int handoff) {
  const int K = static_cast<int>(d.ids.size());
  PackedRun<T> run(o.device, d.A, d.AT, d.rowOff, d.colOff, d.rfac, d.cfac, o.lanes_per_row > 0 ? o.lanes_per_row : 4,
                   o.check_every, o.reflection, o.use_graph != 0);
  run.upload(d.c, d.lo, d.upv, d.rl, d.ru, d.x, d.y);
  run.init_state(d.omega, d.cNorm, d.rhsNorm, d.eta, d.beta, d.gamma, d.offset, o.tolerance, maxIter);
  run.start();
  PackedOutcome r;
  std::vector<double> refKkt(K, INFINITY);
  std::vector<long long> lastGain(K, 0);
  long long checks = 0;
  auto lastProgress = Clock::now();
  while (r.it < maxIter) {
    const int nIt = static_cast<int>(std::min<long long>(o.check_every, maxIter - r.it));
    run.run_block(nIt);
    run.fetch(r.rec, r.done, r.doneIter, r.restarts);
    r.it += nIt;
    ++checks;
    int remaining = 0, stalled = 0;
    for (int q = 0; q < K; ++q) {
      if (r.done[q]) continue;
      ++remaining;
      const CheckRecord& c = r.rec[q];
      const double kkt = std::max(c.relP, std::max(c.relD, c.gap));
      if (kkt < 0.5 * refKkt[q]) {
        refKkt[q] = kkt;
        lastGain[q] = checks;
      }
      if (checks - lastGain[q] >= 256) ++stalled;
    }
    if (o.progress_seconds > 0 &&
        std::chrono::duration<double>(Clock::now() - lastProgress).count() >= o.progress_seconds) {
      lastProgress = Clock::now();
      std::fprintf(stderr, "  pdhg packed: iteration %lld, %d of %d live\n", r.it, remaining, K);
    }
    if (remaining == 0) break;
    if ((o.time_limit > 0 && std::chrono::duration<double>(Clock::now() - t0).count() >= o.time_limit) ||
        (o.should_stop && o.should_stop(o.stop_user))) {
      r.expired = true;
      break;
    }
    if (stallRule && stalled > 0 && 2 * stalled >= remaining) {
      r.stalled = true;
      break;
    }
    if (remaining <= handoff && remaining < K && r.it < maxIter) {
      // A few LPs left: each continues in a single-LP run, which iterates
      // faster than the packed kernels over a handful of tasks.
      for (int q = 0; q < K; ++q) {
        if (r.done[q]) continue;
        LaneState ls;
        run.export_slot(q, d.colOff[q], d.colOff[q + 1] - d.colOff[q], d.rowOff[q], d.rowOff[q + 1] - d.rowOff[q],
                        ls);
        r.handed.push_back(q);
        r.handedState.push_back(std::move(ls));
      }
      break;
    }
    run.drop_finished(r.done);
  }
  r.kernel = run.stop();
  run.download(true, r.xs);
  run.download(false, r.ys);
  r.launches = run.launches;
  r.graphLaunches = run.graphLaunches;
  return r;
}

// Results of the slots of one run, in the original space; iterations and
// restarts continue from itBase / restartBase (indexed by instance).
void write_results(const PackedData& d, const PackedOutcome& r, const std::vector<Prepared>& P, long long itBase,
                   const std::vector<int>& restartBase, pdhg_result* results) {
  for (size_t q = 0; q < d.ids.size(); ++q) {
    const int b = d.ids[q];
    pdhg_result& out = results[b];
    const CheckRecord& c = r.rec[q];
    out.status = r.done[q] ? PDHG_OPTIMAL : (r.expired ? PDHG_TIME_LIMIT : PDHG_ITERATION_LIMIT);
    out.iterations = itBase + (r.done[q] ? r.doneIter[q] : (r.expired ? r.it - 1 : r.it));
    out.restarts = restartBase[b] + r.restarts[q];
    out.objective = c.pobj;
    out.dual_objective = c.dobj;
    out.rel_primal = c.relP;
    out.rel_dual = c.relD;
    out.rel_gap = c.gap;
    out.relative_kkt = std::max(c.relP, std::max(c.relD, c.gap));
    out.primal_weight = c.omega;
    const Prepared& p = P[b];
    if (out.x)
      for (int j = 0; j < p.n; ++j) out.x[j] = r.xs[d.colOff[q] + j] * p.C[j] * p.beta;
    if (out.row_dual)
      for (int i = 0; i < p.m; ++i) out.row_dual[i] = r.ys[d.rowOff[q] + i] * p.R[i] * p.gamma;
  }
}

// The handed-off slots of a packed run, each continued in its own single-LP
// run (FP64 iterates, one lane) from the exported state, with the packed
// run's iteration limit maxIter on its own count; the runs' blocks are
// interleaved on their own streams.  Overwrites those slots' results
// (iterations and restarts offset by itBase / restartBase, as the packed
// run's); returns kernel seconds and adds the iterations the runs took beyond
// the packed run's.
double finish_handed(const PackedData& d, const PackedOutcome& r, const std::vector<Prepared>& P,
                     const pdhg_options& o, long long maxIter, long long itBase, const std::vector<int>& restartBase,
                     Clock::time_point t0, pdhg_result* results, long long* extraIters, int* extraRestarts) {
  struct Slot {
    int q, b;
    std::unique_ptr<DeviceRun> run;
    long long own = 0;
    bool finished = false, limit = false;
    CheckRecord rec{};
    int done = 0, restarts = 0;
    long long doneIter = -1;
  };
  std::vector<Slot> slots;
  for (size_t h = 0; h < r.handed.size(); ++h) {
    const int q = r.handed[h], b = d.ids[q];
    const Prepared& p = P[b];
    RunSetup rs;
    rs.m = p.m;
    rs.n = p.n;
    rs.B = 1;
    rs.precision = PDHG_FP64;
    rs.colMap.nnzPerLane = rs.rowMap.nnzPerLane = o.lanes_per_row > 0 ? o.lanes_per_row : 4;
    rs.rho = o.reflection;
    rs.checkEvery = o.check_every;
    rs.useGraph = o.use_graph != 0;
    Slot sl;
    sl.q = q;
    sl.b = b;
    sl.run = DeviceRun::create(o.device, p, rs);
    const int c0 = d.colOff[q], r0 = d.rowOff[q];
    auto cut = [](const std::vector<double>& v, int off, int len) {
      return std::vector<double>(v.begin() + off, v.begin() + off + len);
    };
    const auto c = cut(d.c, c0, p.n), lo = cut(d.lo, c0, p.n), up = cut(d.upv, c0, p.n);
    const auto rl = cut(d.rl, r0, p.m), ru = cut(d.ru, r0, p.m);
    sl.run->upload(c.data(), lo.data(), up.data(), rl.data(), ru.data(), r.handedState[h].x.data(),
                   r.handedState[h].y.data());
    RunScalars sc;
    sc.eta = d.eta[q];
    sc.beta = d.beta[q];
    sc.gamma = d.gamma[q];
    sc.offset = d.offset[q];
    sc.tol = o.tolerance;
    sc.maxIter = maxIter;
    const int live = 0;
    sl.run->init_state(sc, &d.omega[q], &d.cNorm[q], &d.rhsNorm[q], &live);
    sl.run->import_lanes(r.handedState[h]);
    sl.run->start(false);
    sl.own = r.handedState[h].iter[0];
    sl.restarts = r.handedState[h].restarts[0];
    slots.push_back(std::move(sl));
  }
  double kernel = 0;
  long long extra = 0;
  int restartsBefore = 0;
  for (const Slot& sl : slots) restartsBefore += sl.restarts;
  bool expired = false;
  for (;;) {
    bool any = false;
    for (Slot& sl : slots) {
      if (sl.finished) continue;
      const long long left = maxIter - sl.own;
      const int nIt = static_cast<int>(std::min<long long>(o.check_every - sl.own % o.check_every, left));
      sl.run->run_block(nIt);
      sl.own += nIt;
      any = true;
    }
    if (!any) break;
    for (Slot& sl : slots) {
      if (sl.finished) continue;
      sl.run->fetch(&sl.rec, &sl.done, &sl.doneIter, &sl.restarts);
      if (sl.done || sl.own >= maxIter) {
        sl.finished = true;
        sl.limit = !sl.done;
      }
    }
    if ((o.time_limit > 0 && std::chrono::duration<double>(Clock::now() - t0).count() >= o.time_limit) ||
        (o.should_stop && o.should_stop(o.stop_user))) {
      expired = true;
      break;
    }
  }
  for (Slot& sl : slots) {
    kernel += sl.run->stop();
    extra = std::max(extra, sl.own - r.it);
    pdhg_result& out = results[sl.b];
    const Prepared& p = P[sl.b];
    out.status = sl.done ? PDHG_OPTIMAL : (expired && !sl.limit ? PDHG_TIME_LIMIT : PDHG_ITERATION_LIMIT);
    out.iterations = itBase + (sl.done ? sl.doneIter : sl.own);
    out.restarts = restartBase[sl.b] + sl.restarts;
    const CheckRecord& c = sl.rec;
    out.objective = c.pobj;
    out.dual_objective = c.dobj;
    out.rel_primal = c.relP;
    out.rel_dual = c.relD;
    out.rel_gap = c.gap;
    out.relative_kkt = std::max(c.relP, std::max(c.relD, c.gap));
    out.primal_weight = c.omega;
    std::vector<double> xs(p.n), ys(p.m);
    sl.run->download(0, xs.data());
    sl.run->download(1, ys.data());
    if (out.x)
      for (int j = 0; j < p.n; ++j) out.x[j] = xs[j] * p.C[j] * p.beta;
    if (out.row_dual)
      for (int i = 0; i < p.m; ++i) out.row_dual[i] = ys[i] * p.R[i] * p.gamma;
  }
  int restartsAfter = 0;
  for (const Slot& sl : slots) restartsAfter += sl.restarts;
  *extraIters = extra;
  *extraRestarts = restartsAfter - restartsBefore;
  return kernel;
}

// LPs still live when at most this many remain move to single-LP runs.
constexpr int kHandoff = 4;

void solve_packed_all(const std::vector<Model>& lps, const pdhg_options& o, pdhg_result* results,
                      pdhg_stats* stats) {
  const auto t0 = Clock::now();
  const int K = static_cast<int>(lps.size());
  // Set-up of every instance (src/pdhg/pdhg.cpp:424-459), on host threads.
  std::vector<Prepared> P(K);
  {
    const int nt = std::max(1, std::min(16, static_cast<int>(std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    std::vector<std::string> err(nt);
    for (int t = 0; t < nt; ++t)
      pool.emplace_back([&, t] {
        try {
          for (int b = t; b < K; b += nt) P[b] = prepare(lps[b], o.ruiz_iterations, o.pock_chambolle != 0, o.rescale != 0);
        } catch (const std::exception& e) {
          err[t] = e.what();
        }
      });
    for (auto& th : pool) th.join();
    for (auto& e : err) PDHG_REQUIRE(e.empty(), e);
  }
  std::vector<int> all(K);
  for (int b = 0; b < K; ++b) all[b] = b;
  const PackedData d = pack(lps, P, all, nullptr, nullptr, nullptr);
  const double setup = std::chrono::duration<double>(Clock::now() - t0).count();
  const auto tSolve = Clock::now();
  std::vector<int> restartBase(K, 0);
  PackedOutcome r;
  if (o.precision == PDHG_FP64) r = run_packed<double>(d, o, o.max_iterations, t0, false, kHandoff);
  else r = run_packed<float>(d, o, o.max_iterations, t0, o.precision == PDHG_FP32_AUTO, 0);
  write_results(d, r, P, 0, restartBase, results);
  long long iterations = r.it, launches = r.launches, graphLaunches = r.graphLaunches;
  double kernel = r.kernel;
  int handedRestarts = 0;
  if (!r.handed.empty()) {
    long long extra = 0;
    kernel += finish_handed(d, r, P, o, o.max_iterations, 0, restartBase, t0, results, &extra, &handedRestarts);
    iterations += extra;
  }
  long long switchAt = -1;
  int restartsTotal = handedRestarts;
  for (int q = 0; q < K; ++q) restartsTotal += r.restarts[q];
  if (r.stalled) {
    // The live instances continue with FP64 iterates from their FP32 point and
    // primal weight; the move is a restart to that point (the anchor of the
    // new run is where FP32 stopped).
    switchAt = r.it;
    if (o.progress_seconds > 0)
      std::fprintf(stderr, "  pdhg packed: iteration %lld, FP32 progress stalled, continuing in FP64\n", r.it);
    std::vector<int> live;
    std::vector<std::vector<double>> xw(K), yw(K);
    std::vector<double> ww(K, 1.0);
    for (int q = 0; q < K; ++q) {
      if (r.done[q]) continue;
      const int b = d.ids[q];
      live.push_back(b);
      xw[b].assign(r.xs.begin() + d.colOff[q], r.xs.begin() + d.colOff[q + 1]);
      yw[b].assign(r.ys.begin() + d.rowOff[q], r.ys.begin() + d.rowOff[q + 1]);
      ww[b] = r.rec[q].omega;
      restartBase[b] = r.restarts[q];
    }
    const PackedData d2 = pack(lps, P, live, &xw, &yw, &ww);
    const PackedOutcome r2 = run_packed<double>(d2, o, o.max_iterations - r.it, t0, false, kHandoff);
    write_results(d2, r2, P, r.it, restartBase, results);
    iterations += r2.it;
    if (!r2.handed.empty()) {
      long long extra = 0;
      int more = 0;
      kernel += finish_handed(d2, r2, P, o, o.max_iterations - r.it, r.it, restartBase, t0, results, &extra, &more);
      iterations += extra;
      restartsTotal += more;
    }
    launches += r2.launches;
    graphLaunches += r2.graphLaunches;
    kernel += r2.kernel;
    for (size_t q = 0; q < live.size(); ++q) restartsTotal += r2.restarts[q];
  }
  if (stats) {
    pdhg_stats st{};
    st.iterations = iterations;
    st.restarts = restartsTotal;
    st.setup_seconds = setup;
    st.solve_seconds = std::chrono::duration<double>(Clock::now() - tSolve).count();
    st.kernel_seconds = kernel;
    st.kernel_launches = launches;
    st.graph_launches = graphLaunches;
    st.batch = K;
    st.precision = o.precision;
    st.fp64_switch_iteration = switchAt;
    std::snprintf(st.backend, sizeof(st.backend), "hip-gfx950-packed");
    std::snprintf(st.device_name, sizeof(st.device_name), "%s", device_info(o.device).name.c_str());
    *stats = st;
  }
}

}  // namespace

void solve_packed(const std::vector<Model>& lps, const pdhg_options& o, pdhg_result* results, pdhg_stats* stats) {
  PDHG_REQUIRE(!lps.empty(), "no LPs");
  check_options(o);
  const DeviceInfo& di = device_info(o.device);
  if (!di.available) throw Error("GPU unavailable: " + di.reason);
  solve_packed_all(lps, o, results, stats);
}

}  // namespace pdhgamd
