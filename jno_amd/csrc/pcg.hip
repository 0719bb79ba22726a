// Jacobi-preconditioned CG: plain FP64 (jax.scipy.sparse.linalg.cg semantics as called by
// jno.solve.cg, jno/solve.py:191-206) and FP32 inner iterations with FP64 residual replacement.
//
// One iteration is two or three kernels (the iteration mode is a launch-tuning choice):
//   K1  p_k = z_k + beta p_{k-1}  (z = dinv r, recomputed for every gathered neighbour, so p is
//       never a separate pass),  x += alpha_{k-1} p_{k-1} (the x update is deferred one iteration
//       and fused here),  q = A p_k,  partial p.q
//   K2  alpha_k = gamma_k / p.q,  r -= alpha_k q,  partials r.z and r.r
// or K0 (p_k and the deferred x update, own rows only), KS (q = A p_k, partial p.q), K2: more
// launches but a plain SpMV that gathers one vector instead of three, the faster choice on the
// 128^3 Kuhn lattice; or the single-reduction iteration KA KB further below.
// No value crosses workgroups inside a kernel: each workgroup writes its partial sums, and every
// workgroup of the next kernel reduces all of them itself, in one fixed order, before it starts.
// So the scalars are bitwise identical across workgroups and runs, and there are no atomics or
// cache-flushing fences on the iteration path.  Per-iteration scalars that must persist (k,
// gamma_{k-1}, alpha_{k-1}, stop flags) are written by workgroup 0 into fields the same kernel
// never reads.
// Convergence, the iteration cap and non-finite scalars are decided on the device; launches after
// the decision exit at once.  The host enqueues fixed chunks of iterations and polls a pinned copy
// of the state one chunk behind, so the queue never drains while it checks.
// Workgroups own contiguous row ranges, numbered so that each XCD (workgroups are dealt to the 8
// XCDs round-robin) gets one contiguous slab: gathered neighbour rows then mostly hit its own L2.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.h"
#include "pcg_common.h"
#include "sparse.h"

namespace jno {

constexpr int PCG_NT = 256;

constexpr int PCG_MAX_BLOCKS = 4096;  // workspace capacity for per-workgroup partials

// Launch shape of the iteration kernels, one per iteration precision (set by the autotuner through
// jno_pcg_set_tuning).  The defaults are among the fastest shapes of sweeps over workgroup size x
// grid cap x SpMV tile variant x iteration mode on a 2.1M-row 4-diagonal operator (128^3 Kuhn P1):
// FP64 iterations with p materialized (K0 KS K2), 1024 threads x 512 workgroups, two rows per
// thread in the SpMV kernel; FP32 inner iterations as one kernel per iteration on the Jacobi-scaled
// operator (mode 5), 512 threads x 512 workgroups (two per CU).
struct Tuning {
  int grid_cap;  // workgroups of the iteration kernels (a multiple of 8)
  int vec;       // DIA SpMV kernel: 0 two rows per thread, 1 VN-row 16-byte tiles, 2 four rows,
                 // 3 p staged in LDS (mode 4 K1; the others take 0).  Mode 5: 0 pcg_one (LDS-staged),
                 // 1 and 2 pcg_one_q (register-resident quads, far loads before / after the head)
  int block;     // threads per workgroup (256, 512, 1024)
  int split;     // iteration mode: 0 K1 K2, 1 K0 KS K2 (p materialized), 2 KA KB (single reduction),
                 // 3 K0 KS K2, 4 K1 K2 and 5 one kernel on the Jacobi-scaled operator (FP32 inner
                 // iterations on DIA)
};
// Iteration modes (Tuning::split): two-reduction CG as K1 K2 (p recomputed inside the SpMV
// kernel) or K0 KS K2 (p materialized), or single-reduction CG as KA KB.
// MODE_SCALED / MODE_SCALED_FUSED / MODE_ONE: the split / fused / one-kernel iteration on the
// Jacobi-scaled FP32 operator (FP32 inner iterations on a DIA operator; the unscaled split / fused /
// fused iteration elsewhere)
enum { MODE_FUSED = 0, MODE_SPLIT = 1, MODE_SINGLE = 2, MODE_SCALED = 3, MODE_SCALED_FUSED = 4, MODE_ONE = 5 };
// FP32: the register-resident one-kernel iteration with pipelined tiles (k1_vector 3), 256 threads x
// 768 workgroups (three tiles of ~930 rows each at 128^3); operators it does not take run the
// LDS-staged one (ONE_FALLBACK below)
constexpr Tuning ONE_FALLBACK = {512, 0, 512, 5};

// The tuning one solve runs: the launch shapes per iteration precision ([0] FP64 iterations, [1]
// FP32 inner iterations), the iterations per host poll and pacing.  Taken from the operator's own
// jno_tuning when it has one, else copied from the library defaults under g_def_mu at the start of
// the solve, so concurrent solves never share mutable launch state.
struct CallTuning {
  Tuning it[2];
  int chunk;
  int pace;
};
static std::mutex g_def_mu;
static CallTuning g_def = {{{512, 0, 1024, 1}, {768, 3, 256, 5}}, 16, 1};

// partial-sum areas (per workgroup): A = K1's p.q; B = K2's (r.z, r.r); P = prep's (rr_true, b.b)
struct Partials {
  double* A;   // [G]
  double* B;   // [2][G]
  double* P;   // [2][G]
  double* B2;  // [2][G]: single-reduction iterations read B and write B2 on even k, and swap on odd k
  double* C;   // [2][4][G]: one-kernel iterations, (r.r, p.q, q.r, q.q) by launch parity
};

template <typename T, typename V>
struct PcgArgs {
  int64_t n;
  int64_t rows_per_block;
  const int32_t* rowptr;
  const int32_t* col;
  const V* vals;
  int ndiag;
  int vec;    // DIA K1 tile variant
  int block;  // threads per workgroup of the launch (host side only)
  int split;  // three kernels per iteration (host side only)
  int unit;   // Jacobi-scaled DIA operator: unit main diagonal, z = r (host side only)
  DiaOffsets off;
  const T* dinv;
  T* r;
  const T* p_old;
  T* p_new;
  T* q;
  T* x;
  T* s;  // single-reduction iterations: s = A p
  T* u;  // single-reduction iterations: u = dinv r
  void* ext;  // format-specific scratch of the SpMV (the EBE product buffer)
  // one-kernel iterations: launch index and the other halves of the r and q pairs (r and q are read,
  // r_out and q_out written; the halves swap every launch)
  int64_t kk;
  T* r_out;
  T* q_out;
  PcgState* st;
  Partials part;
  // one-kernel iterations: host-pinned slot that this launch writes its final state into (the
  // last launch of each chunk of the pinned poller), or nullptr
  PcgState* post;
};

// The solve state for the host's poll, written at the end of the last launch of a chunk instead of
// a device-to-host copy command after it (the copies and the queue drain at each stop cost 0.35 ms
// of a 128^3 solve).  Only thread 0 of workgroup 0 writes state fields during a launch; the first
// lanes of its wave then copy the state into the slot, one 8-byte word each: the fence waits for
// the wave's earlier stores (thread 0's), the loads bypass the L1 (agent scope), and the stores go
// to coherent pinned host memory at system scope.  The host reads the slot after an event recorded
// behind the launch has completed.
__device__ __forceinline__ void post_state(const PcgState* st, PcgState* post) {
  static_assert(sizeof(PcgState) % 8 == 0 && sizeof(PcgState) / 8 <= 64, "one 8-byte word per lane of a wave");
  constexpr int W = (int)(sizeof(PcgState) / 8);
  if (post == nullptr || blockIdx.x != 0 || threadIdx.x >= W) return;
  __threadfence_block();
  const uint64_t v = __hip_atomic_load(reinterpret_cast<const uint64_t*>(st) + threadIdx.x, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(reinterpret_cast<uint64_t*>(post) + threadIdx.x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// sum of part[0..G) in a fixed order; every thread gets the total
template <int NT = PCG_NT>
__device__ __forceinline__ double reduce_partials(const double* __restrict__ part, int G, double* sm) {
  double s = 0.0;
  for (int i = threadIdx.x; i < G; i += NT) s += part[i];
  return block_sum<NT>(s, sm);
}

template <int NT = PCG_NT>
__device__ __forceinline__ void store_partial(double v, double* __restrict__ dst, double* sm) {
  v = block_sum<NT>(v, sm);
  if (threadIdx.x == 0) dst[blockIdx.x] = v;
}

// The iteration kernels' reductions, one barrier each (a kernel's fixed cost grows with the
// barriers its waves pass).
// Totals of the M consecutive partial arrays part[m*G .. m*G+G), identical in every thread: wave 0
// sums them (lane l takes entries l, l+64, ... in order, then a fixed butterfly) and broadcasts them
// through sm_in[M].
template <int M>
__device__ __forceinline__ void reduce_in(const double* __restrict__ part, int G, double* sm_in, double* out) {
  if (threadIdx.x < 64) {
    double s[M];
#pragma unroll
    for (int m = 0; m < M; ++m) s[m] = 0.0;
    // 8 entries per lane and array in flight at once (one round trip for G <= 512)
    for (int base = 0; base < G; base += 64 * 8) {
      double v[M][8];
#pragma unroll
      for (int m = 0; m < M; ++m)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int i = base + u * 64 + (int)threadIdx.x;
          v[m][u] = i < G ? part[(int64_t)m * G + i] : 0.0;
        }
#pragma unroll
      for (int m = 0; m < M; ++m)
#pragma unroll
        for (int u = 0; u < 8; ++u) s[m] += v[m][u];
    }
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const double t = wave_sum(s[m]);
      if (threadIdx.x == 0) sm_in[m] = t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) out[m] = sm_in[m];
}

// ---------------------------------------------------------------------------------------------
// K1: p update (+ deferred x update), q = A p, partial p.q
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T pnew_at(const T* __restrict__ dinv, const T* __restrict__ r, const T* __restrict__ p_old,
                                     T beta, int64_t j) {
  return dinv[j] * r[j] + beta * p_old[j];
}

// VN consecutive values as one 16-byte load (global_load_dwordx4; the address only needs the
// element alignment: gfx950 serves unaligned vector accesses)
template <typename T>
struct V16 {
  static constexpr int N = 16 / sizeof(T);
  T v[N];
};

template <typename T>
__device__ __forceinline__ V16<T> ld16(const T* p) {
  V16<T> r;
  __builtin_memcpy(&r, p, 16);
  return r;
}

// Plain (write-back) stores of the iteration vectors: non-temporal stores measured no faster here.
template <typename T>
__device__ __forceinline__ void sto(T* p, T v) {
  *p = v;
}

template <typename T>
__device__ __forceinline__ void st16(T* p, const V16<T>& r) {
  __builtin_memcpy(p, &r, 16);
}

// DIA row tile [i0, i0+VN): p_new = z + beta p_old on the tile, q = A p_new, deferred x update;
// returns the tile's p.q.  `interior` tiles need no bounds checks on neighbour rows.
// JAC = false: the Jacobi-scaled operator (z = r, unit main diagonal not read).
template <int ND, typename T, typename V, bool JAC = true>
__device__ __forceinline__ double dia_tile(const PcgArgs<T, V>& a, int64_t i0, T beta, T ap, bool upd_x) {
  constexpr int VN = V16<T>::N;
  const int64_t ld = dia_ld(a.n);
  const int nd = ND > 0 ? ND : a.ndiag;
  const V16<T> rv = ld16(a.r + i0), po = ld16(a.p_old + i0);
  V16<T> dv, pi, s;
  if (JAC) dv = ld16(a.dinv + i0);
#pragma unroll
  for (int j = 0; j < VN; ++j) pi.v[j] = (JAC ? dv.v[j] * rv.v[j] : rv.v[j]) + beta * po.v[j];
  if (JAC) {
    const V16<V> d0 = ld16(a.vals + i0);
#pragma unroll
    for (int j = 0; j < VN; ++j) s.v[j] = (T)d0.v[j] * pi.v[j];
  } else {
    s = pi;
  }
#pragma unroll
  for (int d = 1; d < nd; ++d) {
    const int64_t o = a.off.o[d];
    const V* vd = a.vals + (int64_t)d * ld;
    const V16<V> vu = ld16(vd + i0), vl = ld16(vd + i0 - o);
    const V16<T> ru = ld16(a.r + i0 + o), pu = ld16(a.p_old + i0 + o);
    const V16<T> rl = ld16(a.r + i0 - o), pl = ld16(a.p_old + i0 - o);
    V16<T> du, dl;
    if (JAC) {
      du = ld16(a.dinv + i0 + o);
      dl = ld16(a.dinv + i0 - o);
    }
#pragma unroll
    for (int j = 0; j < VN; ++j) {
      s.v[j] += (T)vu.v[j] * ((JAC ? du.v[j] * ru.v[j] : ru.v[j]) + beta * pu.v[j]);
      s.v[j] += (T)vl.v[j] * ((JAC ? dl.v[j] * rl.v[j] : rl.v[j]) + beta * pl.v[j]);
    }
  }
  st16(a.p_new + i0, pi);
  st16(a.q + i0, s);
  if (upd_x) {
    V16<T> xv = ld16(a.x + i0);
#pragma unroll
    for (int j = 0; j < VN; ++j) xv.v[j] += ap * po.v[j];
    st16(a.x + i0, xv);
  }
  double part = 0.0;
#pragma unroll
  for (int j = 0; j < VN; ++j) part += (double)pi.v[j] * (double)s.v[j];
  return part;
}

// one DIA row with bounds checks (tiles at the ends of the matrix)
template <typename T, typename V, bool JAC = true>
__device__ __forceinline__ double dia_row(const PcgArgs<T, V>& a, int64_t i, T beta, T ap, bool upd_x) {
  const int64_t ld = dia_ld(a.n);
  const T po = a.p_old[i];
  const T pi = (JAC ? a.dinv[i] * a.r[i] : a.r[i]) + beta * po;
  T s = JAC ? (T)a.vals[i] * pi : pi;
  for (int d = 1; d < a.ndiag; ++d) {
    const int64_t o = a.off.o[d];
    const V* v = a.vals + (int64_t)d * ld;
    if (i + o < a.n) s += (T)v[i] * ((JAC ? a.dinv[i + o] * a.r[i + o] : a.r[i + o]) + beta * a.p_old[i + o]);
    if (i - o >= 0) s += (T)v[i - o] * ((JAC ? a.dinv[i - o] * a.r[i - o] : a.r[i - o]) + beta * a.p_old[i - o]);
  }
  sto(&a.p_new[i], pi);
  sto(&a.q[i], s);
  if (upd_x) a.x[i] += ap * po;
  return (double)pi * (double)s;
}

// R interior DIA rows i, i + NT, ..., i + (R-1) NT of one thread (all neighbours in range) with a
// compile-time diagonal count: every load of the R rows is independent of every other, so they are
// all in flight before the first use (a runtime diagonal loop waits on each diagonal's loads).
template <int ND, int R, int NT, typename T, typename V, bool JAC = true>
__device__ __forceinline__ double dia_rows(const PcgArgs<T, V>& a, int64_t i, T beta, T ap, bool upd_x) {
  const int64_t ld = dia_ld(a.n);
  T dv[R], rv[R], po[R], xv[R];
  V d0[R], vu[R][ND - 1], vl[R][ND - 1];
  T du[R][ND - 1], ru[R][ND - 1], pu[R][ND - 1], dl[R][ND - 1], rl[R][ND - 1], pl[R][ND - 1];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t row = i + (int64_t)r * NT;
    if (JAC) dv[r] = a.dinv[row];
    rv[r] = a.r[row];
    po[r] = a.p_old[row];
    if (JAC) d0[r] = a.vals[row];
    xv[r] = upd_x ? a.x[row] : T(0);
#pragma unroll
    for (int d = 1; d < ND; ++d) {
      const int64_t o = a.off.o[d];
      vu[r][d - 1] = a.vals[(int64_t)d * ld + row];
      vl[r][d - 1] = a.vals[(int64_t)d * ld + row - o];
      if (JAC) du[r][d - 1] = a.dinv[row + o];
      ru[r][d - 1] = a.r[row + o];
      pu[r][d - 1] = a.p_old[row + o];
      if (JAC) dl[r][d - 1] = a.dinv[row - o];
      rl[r][d - 1] = a.r[row - o];
      pl[r][d - 1] = a.p_old[row - o];
    }
  }
  double part = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t row = i + (int64_t)r * NT;
    const T pi = (JAC ? dv[r] * rv[r] : rv[r]) + beta * po[r];
    T s = JAC ? (T)d0[r] * pi : pi;
#pragma unroll
    for (int d = 0; d < ND - 1; ++d) {
      s += (T)vu[r][d] * ((JAC ? du[r][d] * ru[r][d] : ru[r][d]) + beta * pu[r][d]);
      s += (T)vl[r][d] * ((JAC ? dl[r][d] * rl[r][d] : rl[r][d]) + beta * pl[r][d]);
    }
    sto(&a.p_new[row], pi);
    sto(&a.q[row], s);
    if (upd_x) sto(&a.x[row], xv[r] + ap * po[r]);
    part += (double)pi * (double)s;
  }
  return part;
}

// ND: compile-time DIA diagonal count (0 = runtime a.ndiag); L: lanes per CSR row
// Start of an iteration (K1, or K0 of the split iteration): reduce the r.z and r.r partials of the
// previous K2, take the stop decision of jax's cg loop (r.r <= atol2, k >= maxiter) plus the
// residual-replacement and non-finite stops, and form beta.  On a stop the deferred x update is
// finished, workgroup 0 records the stop in the state and false is returned.
template <int NT, typename T, typename V>
__device__ __forceinline__ bool iter_head(const PcgArgs<T, V>& a, double* sm, double& gamma, T& beta, T& ap,
                                          int64_t& k, int64_t& r0, int64_t& r1) {
  PcgState* st = a.st;
  const int G = gridDim.x;
  double in[2];
  reduce_in<2>(a.part.B, G, sm, in);
  gamma = in[0];
  const double rr = in[1];
  k = st->k;
  const double alpha_prev = st->alpha;
  const bool stop_conv = !(rr > st->atol2);
  const bool stop_cap = k >= st->maxiter;
  const bool stop_repl = rr <= st->rr_replace;
  const bool bad = !isfinite(gamma) || !isfinite(rr);
  ap = (T)alpha_prev;
  r0 = block_first_row(a.rows_per_block);
  r1 = min(a.n, r0 + a.rows_per_block);
  if (stop_conv || stop_cap || stop_repl || bad) {
    if (k > 0 && alpha_prev != 0.0)
      for (int64_t i = r0 + threadIdx.x; i < r1; i += NT) a.x[i] += ap * a.p_old[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done1 = 1;
      st->k_stop = k;
      st->rr_stop = rr;
      st->status = bad ? ST_NONFINITE : stop_conv ? ST_CONVERGED : stop_cap ? ST_MAXITER : ST_REPLACE;
    }
    return false;
  }
  beta = (T)((k == 0) ? 0.0 : gamma / st->gamma_prev);
  return true;
}

// JAC = false (DIA only): the Jacobi-scaled operator, z = r and a unit main diagonal
template <int NT, int L, bool DIA, int ND, typename T, typename V, bool JAC = true>
__global__ void __launch_bounds__(NT) pcg_k1(PcgArgs<T, V> a) {
  __shared__ double sm[4];                  // broadcast of reduced partial sums
  __shared__ double sm_out[2 * (NT / 64)];  // wave totals of this kernel's partial sums
  PcgState* st = a.st;
  if (st->done) return;
  double gamma;
  T beta, ap;
  int64_t k, r0, r1;
  if (!iter_head<NT, T, V>(a, sm, gamma, beta, ap, k, r0, r1)) return;
  double part = 0.0;
  if (DIA && a.vec == 1) {
    // each thread owns VN consecutive rows; the block's first row is VN-aligned
    constexpr int VN = V16<T>::N;
    int64_t omax = 0;
    for (int d = 1; d < a.ndiag; ++d) omax = max(omax, (int64_t)a.off.o[d]);
    for (int64_t i0 = r0 + (int64_t)threadIdx.x * VN; i0 < r1; i0 += (int64_t)VN * NT) {
      if (i0 + VN <= r1 && i0 - omax >= 0 && i0 + VN - 1 + omax < a.n) {
        part += dia_tile<ND, T, V, JAC>(a, i0, beta, ap, k > 0);
      } else {
        for (int64_t i = i0; i < min(r1, i0 + VN); ++i) part += dia_row<T, V, JAC>(a, i, beta, ap, k > 0);
      }
    }
  } else if (DIA && ND > 0) {
    // R rows per thread and step, NT apart (coalesced), all loads issued before the first use
    constexpr int R = 2;
    int64_t omax = 0;
#pragma unroll
    for (int d = 1; d < (ND > 0 ? ND : 1); ++d) omax = max(omax, (int64_t)a.off.o[d]);
    int64_t i = r0 + threadIdx.x;
    for (; i + (int64_t)(R - 1) * NT < r1; i += (int64_t)R * NT) {
      if (i - omax >= 0 && i + (int64_t)(R - 1) * NT + omax < a.n) {
        part += dia_rows<(ND > 0 ? ND : 1), R, NT, T, V, JAC>(a, i, beta, ap, k > 0);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) part += dia_row<T, V, JAC>(a, i + (int64_t)r * NT, beta, ap, k > 0);
      }
    }
    for (; i < r1; i += NT) part += dia_row<T, V, JAC>(a, i, beta, ap, k > 0);
  } else if (DIA) {
    for (int64_t i = r0 + threadIdx.x; i < r1; i += NT) part += dia_row<T, V, JAC>(a, i, beta, ap, k > 0);
  } else {
    const int sub = threadIdx.x % L;
    for (int64_t row = r0 + threadIdx.x / L; row < r1; row += NT / L) {
      const int b = a.rowptr[row], e = a.rowptr[row + 1];
      T s = 0;
      for (int j = b + sub; j < e; j += L) s += (T)a.vals[j] * pnew_at(a.dinv, a.r, a.p_old, beta, a.col[j]);
#pragma unroll
      for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
      if (sub == 0) {
        const T po = a.p_old[row];
        const T pi = a.dinv[row] * a.r[row] + beta * po;
        sto(&a.p_new[row], pi);
        sto(&a.q[row], s);
        if (k > 0) sto(&a.x[row], a.x[row] + ap * po);
        part += (double)pi * (double)s;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->gamma_cur = gamma;
  store_out<NT, 1>(&part, a.part.A, gridDim.x, sm_out);
}

// ---------------------------------------------------------------------------------------------
// LDS-staged kernels on the Jacobi-scaled DIA operator (unit main diagonal, z = r).  A workgroup's
// rows go in tiles [t0, t1): p_new is formed once per row for the tile and a halo of h rows on
// either side into LDS (h = the largest "near" offset), the own rows also to memory; after one
// barrier q = A p_new takes the near neighbours from LDS and recomputes only the far ones (offsets
// > LDS_HALO) from the iteration vectors.  NN = the number of near off-diagonals: the stored offsets
// ascend, so diagonals 1..NN are near and the rest far, fixed at compile time.
// Every global load is unconditional, from an index clamped to row 0 where it does not apply
// (ldc): a batch's loads then issue back to back and one wait covers them all, where guarded loads
// become branches and the compiler waits after each.  A value loaded that way is finite and is only
// ever multiplied by a zero coefficient: the stored upper diagonal is 0 past the last row, and the
// lower coefficient before the first row is selected to 0 (ldz), the one select that remains.
// ---------------------------------------------------------------------------------------------
constexpr int LDS_HALO = 256;
constexpr int LDS_STAGE_BYTES = 32768;

template <typename T>
__device__ __forceinline__ T ldc(const T* __restrict__ p, int i, bool ok) {
  return p[ok ? i : 0];
}

template <typename T>
__device__ __forceinline__ T ldz(const T* __restrict__ p, int i, bool ok) {
  const T v = p[ok ? i : 0];
  return ok ? v : T(0);
}

// K1 with p staged in LDS (mode 4, k1_vector = 3): p_new = r + beta p_old, the deferred x update
// and q = A p_new, partial p.q.  Per row about 13 global loads instead of 21, for the same bytes.
template <int NT, int ND, int NN, typename T, typename V>
__global__ void __launch_bounds__(NT) pcg_k1_lds(PcgArgs<T, V> a) {
  constexpr int TT = LDS_STAGE_BYTES / (int)sizeof(T) - 2 * LDS_HALO;  // rows per tile
  constexpr int U1 = 4;                                                 // staged rows per thread in flight
  constexpr int U2 = 2;                                                 // SpMV rows per thread in flight
  constexpr int NDA = ND > 1 ? ND : 2;
  __shared__ double sm[4];
  __shared__ double sm_out[2 * (NT / 64)];
  __shared__ T pt[TT + 2 * LDS_HALO];
  PcgState* st = a.st;
  if (st->done) return;
  double gamma;
  T beta, ap;
  int64_t k, r0l, r1l;
  if (!iter_head<NT, T, V>(a, sm, gamma, beta, ap, k, r0l, r1l)) return;
  // row indices fit 32 bits (CSR columns are int32): 32-bit offsets keep each address in one
  // register
  const bool upd_x = k > 0;
  const int n = (int)a.n, r0 = (int)min(r0l, a.n), r1 = (int)r1l;
  const V* vals = a.vals;
  const int ld = (int)dia_ld(a.n);
  int o[NDA];
#pragma unroll
  for (int d = 1; d < ND; ++d) o[d] = a.off.o[d];
  const int h = NN > 0 ? o[NN] : 0;
  double part = 0.0;
  for (int t0 = r0; t0 < r1; t0 += TT) {
    const int t1 = min(r1, t0 + TT), s0 = t0 - h, s1 = t1 + h;
    for (int j0 = s0 + (int)threadIdx.x; j0 < s1; j0 += U1 * NT) {
      T rv[U1], pv[U1], xv[U1];
#pragma unroll
      for (int u = 0; u < U1; ++u) {
        const int j = j0 + u * NT;
        const bool in = j < s1 && j >= 0 && j < n;
        rv[u] = ldc(a.r, j, in);
        pv[u] = ldc(a.p_old, j, in);
        xv[u] = ldc(a.x, j, in);
      }
#pragma unroll
      for (int u = 0; u < U1; ++u) {
        const int j = j0 + u * NT;
        if (j < s1) {
          const T pn = rv[u] + beta * pv[u];
          pt[j - s0] = pn;
          if (j >= t0 && j < t1) {
            sto(&a.p_new[j], pn);
            if (upd_x) sto(&a.x[j], xv[u] + ap * pv[u]);
          }
        }
      }
    }
    __syncthreads();
    for (int i0 = t0 + (int)threadIdx.x; i0 < t1; i0 += U2 * NT) {
      V vu[U2][NDA], vl[U2][NDA];
      T xu[U2][NDA], xl[U2][NDA];
#pragma unroll
      for (int u = 0; u < U2; ++u) {
        const int i = i0 + u * NT;
        const bool ok = i < t1;
#pragma unroll
        for (int d = 1; d < ND; ++d) {
          const V* vd = vals + d * ld;
          vu[u][d] = ldc(vd, i, ok);
          vl[u][d] = ldz(vd, i - o[d], ok && i - o[d] >= 0);
          if (d > NN) {
            const bool up = ok && i + o[d] < n, lo = ok && i - o[d] >= 0;
            xu[u][d] = ldc(a.r, i + o[d], up) + beta * ldc(a.p_old, i + o[d], up);
            xl[u][d] = ldc(a.r, i - o[d], lo) + beta * ldc(a.p_old, i - o[d], lo);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U2; ++u) {
        const int i = i0 + u * NT;
        if (i < t1) {
          const T pi = pt[i - s0];
          T s = pi;
#pragma unroll
          for (int d = 1; d < ND; ++d) {
            if (d <= NN) {
              xu[u][d] = pt[i + o[d] - s0];
              xl[u][d] = pt[i - o[d] - s0];
            }
            s += (T)vu[u][d] * xu[u][d];
            s += (T)vl[u][d] * xl[u][d];
          }
          sto(&a.q[i], s);
          part += (double)pi * (double)s;
        }
      }
    }
    if (t1 < r1) __syncthreads();  // the next tile rewrites pt
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->gamma_cur = gamma;
  store_out<NT, 1>(&part, a.part.A, gridDim.x, sm_out);
}

// The scalars of a one-kernel launch k from the reduced partials in = (r.r, p.q, q.r, q.q) of launch
// k-1 (see pcg_one below), and its stop decision, identical in every workgroup.  Returns false when
// the launch stops (block 0 records why and raises the stop flags).  special: 0 a plain step, 1 a
// first launch (alpha = beta = 0), 2 a rerun after a residual replacement (alpha = 0, beta from the
// recorded r.r two launches back).
__device__ __forceinline__ bool one_decide(PcgState* st, int64_t k, const double* in, double& alpha, double& beta,
                                           int& special) {
  const double rr = in[0];
  special = k == st->k_special ? st->special : 0;
  alpha = 0.0;
  beta = 0.0;
  if (special == 2) {
    beta = rr / st->g_hist[k & 1];
  } else if (special == 0) {
    alpha = rr / in[1];
    beta = (rr - 2.0 * alpha * in[2] + alpha * alpha * in[3]) / rr;
  }
  const bool stop_conv = special != 2 && !(rr > st->atol2);
  const bool stop_cap = k >= st->maxiter;
  const bool stop_repl = special == 0 && rr <= st->rr_replace;
  const bool bad = !isfinite(rr) || (!stop_conv && (!isfinite(alpha) || !isfinite(beta)));
  if (stop_conv || stop_cap || stop_repl || bad) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->k_stop = k;
      st->rr_stop = rr;
      st->status = bad ? ST_NONFINITE : stop_conv ? ST_CONVERGED : stop_cap ? ST_MAXITER : ST_REPLACE;
      st->done1 = 1;
      st->done = 1;  // every workgroup of this launch takes the same decision
    }
    return false;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && special != 2) st->g_hist[k & 1] = rr;
  return true;
}

// The state fields one_decide reads, loaded at the start of a launch so that their latency overlaps
// the partial-sum loads (they are set before the launch: no launch writes a field it reads, g_hist
// being read only by a resuming launch, which does not write it)
struct DecideIn {
  int64_t k_special, maxiter;
  int32_t special;
  double g_prev, atol2, rr_replace;
};
__device__ __forceinline__ DecideIn decide_in(const PcgState* st, int64_t k) {
  DecideIn d;
  d.k_special = st->k_special;
  d.special = st->special;
  d.g_prev = st->g_hist[k & 1];
  d.atol2 = st->atol2;
  d.maxiter = st->maxiter;
  d.rr_replace = st->rr_replace;
  return d;
}
// one_decide on prefetched fields (the same decision and the same state writes)
__device__ __forceinline__ bool one_decide_pre(PcgState* st, const DecideIn& din, int64_t k, const double* in,
                                               double& alpha, double& beta, int& special) {
  const double rr = in[0];
  special = k == din.k_special ? din.special : 0;
  alpha = 0.0;
  beta = 0.0;
  if (special == 2) {
    beta = rr / din.g_prev;
  } else if (special == 0) {
    alpha = rr / in[1];
    beta = (rr - 2.0 * alpha * in[2] + alpha * alpha * in[3]) / rr;
  }
  const bool stop_conv = special != 2 && !(rr > din.atol2);
  const bool stop_cap = k >= din.maxiter;
  const bool stop_repl = special == 0 && rr <= din.rr_replace;
  const bool bad = !isfinite(rr) || (!stop_conv && (!isfinite(alpha) || !isfinite(beta)));
  if (stop_conv || stop_cap || stop_repl || bad) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->k_stop = k;
      st->rr_stop = rr;
      st->status = bad ? ST_NONFINITE : stop_conv ? ST_CONVERGED : stop_cap ? ST_MAXITER : ST_REPLACE;
      st->done1 = 1;
      st->done = 1;
    }
    return false;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && special != 2) st->g_hist[k & 1] = rr;
  return true;
}

// ---------------------------------------------------------------------------------------------
// One kernel per iteration on the Jacobi-scaled DIA operator (iteration mode 5).  Launch k reads
// r_{k-1}, q_{k-1} = A p_{k-1}, p_{k-1}, x and the partial sums of launch k-1: gamma_{k-1} = r.r
// (exact), p.q, q.r, q.q.  With alpha = gamma_{k-1} / p.q the next residual norm follows without
// another reduction, gamma_k = gamma_{k-1} - 2 alpha q.r + alpha^2 q.q (from the exact
// gamma_{k-1}, so its rounding does not accumulate over iterations), and beta = gamma_k /
// gamma_{k-1}.  Then, per row: r_k = r_{k-1} - alpha q_{k-1}, x += alpha p_{k-1},
// p_k = r_k + beta p_{k-1} (staged in LDS as above, far neighbours recomputed from r, q and p),
// q_k = A p_k, and the partials of r_k.r_k, p_k.q_k, q_k.r_k, q_k.q_k.  r, q and p each alternate
// between two buffers (launch k reads one and writes the other); x is updated in place.  A stop
// leaves x_{k-1} and r_{k-1} consistent (no deferred update).  After a residual replacement the
// driver reruns launch k-1 with the true residual and alpha = 0, beta = gamma_true / gamma_{k-2} on
// p_{k-2}: the search direction survives the replacement.  One launch per iteration moves 44
// instead of 48 bytes per row and pays one launch and one reduction instead of two.  The first
// batch of loads is issued before the reduction of the partial sums (none of them depends on the
// scalars), and each staged batch's loads are in flight while the previous one is stored.
// (Interleaving r, p, x and q per row, to fetch a far neighbour in one load, measured slower: the
// partially written lines.)
// ---------------------------------------------------------------------------------------------
// x update of a one-kernel launch: every launch (XM_EAGER), or every other launch -- an even launch
// defers its term (XM_DEFER), the odd launch after it adds both (XM_FLUSH); see pcg_one_p
constexpr int XM_EAGER = 0, XM_DEFER = 1, XM_FLUSH = 2;
constexpr int ONE_STAGE_BYTES = 65536;
// FW (far window): the far off-diagonals lie within LDS_FAR_SPAN of each other (the mass-carrying
// 8-diagonal operators of a Kuhn lattice: offsets (nx+1)(ny+1) + {0, 1, nx+1, nx+2}), so p_k of the
// far neighbours is staged as two windows (above and below the tile) of TT + span rows each,
// instead of one range of TT rows per far diagonal and direction: a quarter of the far loads.
constexpr int LDS_FAR_SPAN = 256;

template <int NT, int ND, int NN, typename T, typename V, bool FW = false, int XM = XM_EAGER>
__global__ void __launch_bounds__(NT) pcg_one(PcgArgs<T, V> a) {
  constexpr int NF = ND > 1 ? ND - 1 - NN : 0;  // far off-diagonals
  static_assert(!FW || NF >= 2, "the far window groups two or more far diagonals");
  // rows per tile: pt holds TT + 2 halo rows, rt TT rows, pf two ranges of TT rows per far diagonal
  // (or, with FW, two windows of TT + LDS_FAR_SPAN rows)
  constexpr int TT = FW ? (ONE_STAGE_BYTES / (int)sizeof(T) - 2 * LDS_HALO - 2 * LDS_FAR_SPAN) / 4
                        : (ONE_STAGE_BYTES / (int)sizeof(T) - 2 * LDS_HALO) / (2 + 2 * NF);
  constexpr int NPF = FW ? 2 : NF > 0 ? 2 * NF : 1;
  constexpr int LPF = FW ? TT + LDS_FAR_SPAN : NF > 0 ? TT : 1;
  constexpr int U1 = 4;
  constexpr int U2 = 2;
  constexpr int NDA = ND > 1 ? ND : 2;
  __shared__ double sm[4];
  __shared__ double sm_out[4 * (NT / 64)];
  __shared__ T pt[TT + 2 * LDS_HALO];         // p_k of the tile and its halo
  __shared__ T rt[TT];                        // r_k of the tile
  __shared__ T pf[NPF][LPF];                  // p_k of the far neighbours (i + o_d, i - o_d)
  PcgState* st = a.st;
  const int G = gridDim.x;
  const int64_t k = a.kk;
  const int n = (int)a.n;
  const V* vals = a.vals;
  const int ld = (int)dia_ld(a.n);
  int o[NDA];
#pragma unroll
  for (int d = 1; d < ND; ++d) o[d] = a.off.o[d];
  const int h = NN > 0 ? o[NN] : 0;
  T rv[U1], qv[U1], pv[U1], xv[U1], p2[U1];
  // rows [j0, s1) with the tile's own rows [lo, hi): x (and, flushing, p_{k-2} from the buffer this
  // launch overwrites with p_k, read before the write by the same thread) only for own rows
  auto load1 = [&](int j0, int s1, int lo, int hi) {
#pragma unroll
    for (int u = 0; u < U1; ++u) {
      const int j = j0 + u * NT;
      const bool in = j < s1 && j >= 0 && j < n;
      const bool own = in && j >= lo && j < hi;
      rv[u] = ldc(a.r, j, in);
      qv[u] = ldc(a.q, j, in);
      pv[u] = ldc(a.p_old, j, in);
      if constexpr (XM != XM_DEFER) xv[u] = ldc(a.x, j, own);
      if constexpr (XM == XM_FLUSH) p2[u] = ldc(a.p_new, j, own);
    }
  };
  // Tiles of equal length tl <= TT, nt per workgroup.  When the grid is whole XCD slabs and
  // rows_per_block = nt tl, the slab's tiles are dealt round-robin to the XCD's workgroups, as in
  // pcg_one_q (the tiles in flight form one band that holds their far rows); otherwise each
  // workgroup takes its own contiguous range.
  const int64_t rpb = a.rows_per_block;
  const int ntq = (int)((rpb + TT - 1) / TT);
  int nt, tl, lim;
  int64_t tbase, tstride;
  if (G % N_XCD == 0 && rpb % ntq == 0) {
    nt = ntq;
    tl = (int)(rpb / ntq);
    const int per = G / N_XCD;
    tbase = (int64_t)(blockIdx.x % N_XCD) * per * rpb + (int64_t)(blockIdx.x / N_XCD) * tl;
    tstride = (int64_t)per * tl;
    lim = n;
  } else {
    const int64_t r0l = block_first_row(rpb);
    const int r0 = (int)min(r0l, a.n), r1 = (int)min(a.n, r0l + rpb);
    nt = r1 > r0 ? (r1 - r0 + TT - 1) / TT : 0;
    tl = nt > 0 ? (r1 - r0 + nt - 1) / nt : TT;
    tbase = r0;
    tstride = tl;
    lim = r1;
  }
  auto tile_at = [&](int t, int& t0, int& t1) {
    const int64_t b = tbase + (int64_t)t * tstride;
    t0 = (int)min(b, (int64_t)lim);
    t1 = (int)min(b + tl, (int64_t)lim);
  };
  // the first staged batch of the first tile, loaded before the reduction
  {
    int f0, f1;
    tile_at(0, f0, f1);
    load1(f0 - h + (int)threadIdx.x, f1 + h, f0, f1);
  }
  // the alpha whose x term the previous (deferring) launch left to this one
  const T apd = XM == XM_FLUSH ? (T)st->a_pend[k & 1] : (T)0;
  if (st->done) {  // read after the first loads are issued, as in pcg_one_q
    post_state(st, a.post);
    return;
  }
  double in[4];
  reduce_in<4>(a.part.C + (k & 1) * 4 * (int64_t)G, G, sm, in);
  double alpha, beta;
  int special;
  if (!one_decide(st, k, in, alpha, beta, special)) {
    if constexpr (XM != XM_EAGER) {
      if (blockIdx.x == 0 && threadIdx.x == 0) st->a_pend[(k + 1) & 1] = 0.0;
      // a stopping flushing launch adds the term the launch before it deferred: x = x_{k-1}
      if (XM == XM_FLUSH && special == 0 && apd != (T)0) {
        for (int t = 0; t < nt; ++t) {
          int t0, t1;
          tile_at(t, t0, t1);
          for (int j = t0 + (int)threadIdx.x; j < t1; j += NT) a.x[j] = a.x[j] + apd * a.p_new[j];
        }
      }
    }
    post_state(st, a.post);
    return;
  }
  // alpha = 0 (a first or resumed launch) leaves r and x as they are (q times 0)
  const T al = (T)alpha, be = (T)beta;
  const bool upd = special == 0;
  if constexpr (XM != XM_EAGER) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->a_pend[(k + 1) & 1] = XM == XM_DEFER && upd ? (double)al : 0.0;
  }
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  for (int t = 0; t < nt; ++t) {
    int t0, t1;
    tile_at(t, t0, t1);
    if (t0 >= t1) break;  // tiles past the last row (the same for every thread)
    const int s0 = t0 - h, s1 = t1 + h;
    if (t > 0) load1(s0 + (int)threadIdx.x, s1, t0, t1);
    for (int j0 = s0 + (int)threadIdx.x; j0 < s1; j0 += U1 * NT) {
      T rn[U1], pn[U1], xn[U1];
#pragma unroll
      for (int u = 0; u < U1; ++u) {
        rn[u] = rv[u] - al * qv[u];
        pn[u] = rn[u] + be * pv[u];
        // flushing: the deferred term first, then this launch's, as two eager updates add them
        const T xb = XM == XM_FLUSH && apd != (T)0 ? xv[u] + apd * p2[u] : xv[u];
        xn[u] = XM == XM_DEFER ? (T)0 : xb + al * pv[u];
      }
      load1(j0 + U1 * NT, s1, t0, t1);  // the next batch in flight meanwhile (past the end: row 0, unused)
#pragma unroll
      for (int u = 0; u < U1; ++u) {
        const int j = j0 + u * NT;
        if (j < s1) {
          pt[j - s0] = pn[u];
          if (j >= t0 && j < t1) {
            rt[j - t0] = rn[u];
            sto(&a.r_out[j], rn[u]);
            sto(&a.p_new[j], pn[u]);
            if (XM != XM_DEFER && upd) sto(&a.x[j], xn[u]);
            part[0] += (double)rn[u] * (double)rn[u];
          }
        }
      }
    }
    // p_k of the far neighbour rows (offsets > LDS_HALO), staged in the same pass: their owners
    // read the same rows at about the same time, so these loads hit in L2 (read later, in the
    // SpMV, they had been evicted and came from memory again)
    if constexpr (FW) {
      const int flo = o[NN + 1], span = o[ND - 1] - flo;  // span <= LDS_FAR_SPAN (host-checked)
      const int len = (t1 - t0) + span;
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        const int base = w == 0 ? t0 + flo : t0 - flo - span;
        for (int j0 = (int)threadIdx.x; j0 < len; j0 += U1 * NT) {
          T fr[U1], fq[U1], fp[U1];
#pragma unroll
          for (int u = 0; u < U1; ++u) {
            const int j = base + j0 + u * NT;
            const bool in = j0 + u * NT < len && j >= 0 && j < n;
            fr[u] = ldc(a.r, j, in);
            fq[u] = ldc(a.q, j, in);
            fp[u] = ldc(a.p_old, j, in);
          }
#pragma unroll
          for (int u = 0; u < U1; ++u)
            if (j0 + u * NT < len) pf[w][j0 + u * NT] = (fr[u] - al * fq[u]) + be * fp[u];
        }
      }
    }
#pragma unroll
    for (int f = 0; f < (FW ? 0 : 2 * NF); ++f) {
      const int d = NN + 1 + f / 2;
      const int sh = (f & 1) ? -o[d] : o[d];
      for (int j0 = t0 + (int)threadIdx.x; j0 < t1; j0 += U1 * NT) {
        T fr[U1], fq[U1], fp[U1];
#pragma unroll
        for (int u = 0; u < U1; ++u) {
          const int j = j0 + u * NT + sh;
          const bool in = j0 + u * NT < t1 && j >= 0 && j < n;
          fr[u] = ldc(a.r, j, in);
          fq[u] = ldc(a.q, j, in);
          fp[u] = ldc(a.p_old, j, in);
        }
#pragma unroll
        for (int u = 0; u < U1; ++u)
          if (j0 + u * NT < t1) pf[f][j0 + u * NT - t0] = (fr[u] - al * fq[u]) + be * fp[u];
      }
    }
    __syncthreads();
    for (int i0 = t0 + (int)threadIdx.x; i0 < t1; i0 += U2 * NT) {
      V vu[U2][NDA], vl[U2][NDA];
      T xu[U2][NDA], xl[U2][NDA];
#pragma unroll
      for (int u = 0; u < U2; ++u) {
        const int i = i0 + u * NT;
        const bool ok = i < t1;
#pragma unroll
        for (int d = 1; d < ND; ++d) {
          const V* vd = vals + d * ld;
          vu[u][d] = ldc(vd, i, ok);
          vl[u][d] = ldz(vd, i - o[d], ok && i - o[d] >= 0);
        }
      }
#pragma unroll
      for (int u = 0; u < U2; ++u) {
        const int i = i0 + u * NT;
        if (i < t1) {
          const T pi = pt[i - s0];
          T s = pi;
#pragma unroll
          for (int d = 1; d < ND; ++d) {
            if (d <= NN) {
              xu[u][d] = pt[i + o[d] - s0];
              xl[u][d] = pt[i - o[d] - s0];
            } else if constexpr (FW) {
              const int flo = o[NN + 1], span = o[ND - 1] - flo;
              xu[u][d] = pf[0][i + o[d] - t0 - flo];
              xl[u][d] = pf[1][i - o[d] - t0 + flo + span];
            } else {
              xu[u][d] = pf[2 * (d - NN - 1)][i - t0];
              xl[u][d] = pf[2 * (d - NN - 1) + 1][i - t0];
            }
            s += (T)vu[u][d] * xu[u][d];
            s += (T)vl[u][d] * xl[u][d];
          }
          sto(&a.q_out[i], s);
          part[1] += (double)pi * (double)s;
          part[2] += (double)s * (double)rt[i - t0];
          part[3] += (double)s * (double)s;
        }
      }
    }
    int u0, u1;
    tile_at(t + 1, u0, u1);
    if (t + 1 < nt && u0 < u1) __syncthreads();  // the next tile rewrites pt and rt
  }
  store_out<NT, 4>(part, a.part.C + ((k + 1) & 1) * 4 * (int64_t)G, G, sm_out);
  post_state(st, a.post);
}

// ---------------------------------------------------------------------------------------------
// The one-kernel iteration with register-resident rows (mode 5, k1_vector 1 and 2), for 4-diagonal
// operators.  Same arithmetic and launch contract as pcg_one; the difference is how the memory
// traffic is issued.  A thread owns a quad of 4 consecutive rows (16-byte loads and stores) and a
// workgroup's tile is one quad per thread, TT = 4 NT rows.  Every global load of a tile is issued at
// once: the partial sums of the previous launch first, then the quad's r, q, p, its upper and lower
// off-diagonal values, r, q, p of its far neighbours (offsets above LDS_HALO, recomputed in
// registers), and for the first 2 hq threads one row of r, q, p of the near halo (x follows the
// head: it is needed only at the end).  The reduction of the partials waits only for the loads
// issued before the tile's, so it overlaps the tile's round trip; one round trip then covers the
// whole tile.  Only p_k of the tile and its halo goes through
// LDS (the near neighbours of the SpMV); r_k and the far p_k stay in registers.
// Bounds: the iteration vectors carry PCG_PAD_BYTES of zero pad on both sides (ws_layout,
// zero_pads_k: 32 FP32 rows).  A
// quad that straddles row 0 or row n reads pad rows, which are 0 in every vector and enter only
// through a zero coefficient (the stored upper diagonal is 0 past the last row, the lower one is
// selected to 0 before the first); a quad wholly outside [0, n) loads rows 0..3 instead, finite
// values that only meet zero coefficients or rows past n, whose q is selected to 0.  Pad rows of a
// straddling quad are stored back as the 0 they compute to.
// ---------------------------------------------------------------------------------------------
constexpr int PCG_PAD_BYTES = 128;  // zero pad before and after each iteration vector
// workgroups (partial sums per array) the head reduction takes: every thread loads 4 QRED_G / NT of
// them, so the bound is kept at the default grid (at 1152, the loads past 768 cost 0.7 us per
// launch at 128^3)
constexpr int QRED_G = 768;

// LF (k1_vector 2): the far neighbours' loads (and x) are issued after the head instead, and
// combined into p_k right before the barrier: fewer registers live at the head, for a second, partly
// hidden round trip.  (Issuing the next tile's loads during the current tile's SpMV phase measured
// slower: 211 VGPRs.)
template <int NT, int ND, int NN, bool LF = false>
__global__ void __launch_bounds__(NT) pcg_one_q(PcgArgs<float, float> a) {
  static_assert(ND == 4 && NN >= 2 && NN <= 3, "instantiated for 4-diagonal operators with 2 or 3 near diagonals");
  constexpr int NF = ND - 1 - NN;  // far off-diagonals
  constexpr int NFA = NF > 0 ? NF : 1;
  constexpr int TT = 4 * NT;  // rows per tile
  constexpr int NW = NT / 64;
  constexpr int QRED_F = (4 * QRED_G + NT - 1) / NT;  // partial sums loaded per thread
  constexpr int UH = (2 * LDS_HALO + NT - 1) / NT;    // near halo rows per thread
  using Q = V16<float>;
  __shared__ double sm_w[4 * NW];
  __shared__ double sm_out[4 * NW];
  // p_k of the tile and its near halo, two buffers used by alternate tiles: a tile's writes then
  // need no barrier behind the previous tile's SpMV (the buffer they overwrite was last read two
  // tiles back, before the barrier of the tile between)
  __shared__ __align__(16) float ptb[2][TT + 2 * LDS_HALO];
  PcgState* st = a.st;
  const int G = gridDim.x;
  const int tid = (int)threadIdx.x;
  const int64_t k = a.kk;
  const int n = (int)a.n;
  const float* vals = a.vals;
  const int ld = (int)dia_ld(a.n);
  int o[ND];
#pragma unroll
  for (int d = 1; d < ND; ++d) o[d] = a.off.o[d];
  const int hq = (o[NN] + 3) & ~3;  // near halo rows on either side, whole quads (<= LDS_HALO)
  // Tiles: nt per workgroup, of tl rows (whole quads, at most TT).  When the grid is whole XCD slabs
  // and rows_per_block = nt tl (as the launch sizes it), the slab's tiles are dealt round-robin to
  // the XCD's workgroups: tile t of workgroup l of the XCD starts at slab + (l + t G/8) tl, so the
  // tiles in flight at one time form one contiguous band of the slab, which holds the far rows
  // (offset (n+1)^2) of most of them.  Otherwise each workgroup takes its own contiguous range.
  const int64_t rpb = a.rows_per_block;
  const int ntq = (int)((rpb + TT - 1) / TT);
  const bool ilv = G % N_XCD == 0 && rpb % ntq == 0 && (rpb / ntq) % 4 == 0;
  int nt, tl, lim;
  int64_t tbase, tstride;
  if (ilv) {
    nt = ntq;
    tl = (int)(rpb / ntq);
    const int per = G / N_XCD;
    tbase = (int64_t)(blockIdx.x % N_XCD) * per * rpb + (int64_t)(blockIdx.x / N_XCD) * tl;
    tstride = (int64_t)per * tl;
    lim = n;
  } else {
    const int64_t r0l = block_first_row(rpb);
    const int r0 = (int)min(r0l, a.n), r1 = (int)min(a.n, r0l + rpb);
    nt = r1 > r0 ? (r1 - r0 + TT - 1) / TT : 0;
    tl = nt > 0 ? (((r1 - r0 + nt - 1) / nt + 3) & ~3) : TT;
    tbase = r0;
    tstride = tl;
    lim = r1;
  }
  auto tile_at = [&](int t, int& t0, int& t1) {
    const int64_t b = tbase + (int64_t)t * tstride;
    t0 = (int)min(b, (int64_t)lim);
    t1 = (int)min(b + tl, (int64_t)lim);
  };
  // the head's partial-sum loads go out first: the 4 arrays of G partials as one flat range, entry
  // f = tid + u NT in thread tid
  const double* pin = a.part.C + (k & 1) * 4 * (int64_t)G;
  const int G4 = 4 * G;
  double pv_in[QRED_F];
#pragma unroll
  for (int u = 0; u < QRED_F; ++u) pv_in[u] = pin[min(u * NT + tid, G4 - 1)];
  // the registers one tile loads
  struct Tile {
    Q rv, qv, pv, xv, vu[ND], vl[ND], fur[NFA], fuq[NFA], fup[NFA], fdr[NFA], fdq[NFA], fdp[NFA];
    float hr[UH], hqv[UH], hp[UH];
    int i, jl[ND];
    bool own, hin[UH];
  };
  auto load_far = [&](Tile& c) {
    const int ic = c.own ? c.i : 0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const int od = o[NN + 1 + f];
      const int ju = ic + od < n ? ic + od : 0;
      const int jd = ic - od >= -3 ? ic - od : 0;
      c.fur[f] = ld16(a.r + ju);
      c.fuq[f] = ld16(a.q + ju);
      c.fup[f] = ld16(a.p_old + ju);
      c.fdr[f] = ld16(a.r + jd);
      c.fdq[f] = ld16(a.q + jd);
      c.fdp[f] = ld16(a.p_old + jd);
    }
  };
  auto load_tile = [&](Tile& c, int t0, int t1) {
    c.i = t0 + 4 * tid;
    c.own = c.i < t1;
    const int ic = c.own ? c.i : 0;
    c.rv = ld16(a.r + ic);
    c.qv = ld16(a.q + ic);
    c.pv = ld16(a.p_old + ic);
#pragma unroll
    for (int d = 1; d < ND; ++d) {
      const float* vd = vals + (int64_t)d * ld;
      c.vu[d] = ld16(vd + ic);
      c.jl[d] = ic - o[d];
      c.vl[d] = ld16(vd + (c.jl[d] >= -3 ? c.jl[d] : 0));
    }
    if (!LF) load_far(c);
    // halo rows [t0 - hq, t0) and [t1q, t1q + hq), t1q = the tile end in whole quads
    const int t1q = t0 + ((t1 - t0 + 3) & ~3);
#pragma unroll
    for (int u = 0; u < UH; ++u) {
      const int h = tid + u * NT;
      const int row = h < hq ? t0 - hq + h : t1q + (h - hq);
      c.hin[u] = h < 2 * hq && row >= 0 && row < n;
      const int hc = c.hin[u] ? row : 0;
      c.hr[u] = a.r[hc];
      c.hqv[u] = a.q[hc];
      c.hp[u] = a.p_old[hc];
    }
  };
  auto load_late = [&](Tile& c, bool upd) {
    if (LF) load_far(c);
    if (upd) c.xv = ld16(a.x + (c.own ? c.i : 0));
  };
  Tile c;
  int ft0 = 0, ft1 = 0;
  if (nt > 0) tile_at(0, ft0, ft1);
  // issued unconditionally (a workgroup without rows loads row 0, unused): with the same loads in
  // flight on every path, the head's wait for its partial sums does not also wait for these
  load_tile(c, ft0 < ft1 ? ft0 : 0, ft0 < ft1 ? ft1 : 0);
  // the stop flag is read after the first loads are issued, so its latency overlaps them (a launch
  // past a stop exits with its loads unused: measured 0.35 us less per working launch, 1% less per
  // 128^3 solve, against the extra traffic of the exiting ones)
  if (st->done) {
    post_state(st, a.post);
    return;
  }
  // head: reduce the partials of launch k-1 (fixed order: per thread, per wave, then the waves)
  double in[4];
  {
    const int w = tid >> 6;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < QRED_F; ++u) {
      const int f = u * NT + tid;
#pragma unroll
      for (int m = 0; m < 4; ++m) s[m] += f >= m * G && f < (m + 1) * G ? pv_in[u] : 0.0;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double t = wave_sum(s[m]);
      if ((tid & 63) == 0) sm_w[m * NW + w] = t;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double t = 0.0;
#pragma unroll
      for (int v = 0; v < NW; ++v) t += sm_w[m * NW + v];
      in[m] = t;
    }
  }
  double alpha, beta;
  int special;
  if (!one_decide(st, k, in, alpha, beta, special)) {
    post_state(st, a.post);
    return;
  }
  // alpha = 0 (a first or resumed launch) leaves r and x as they are (q times 0)
  const float al = (float)alpha, be = (float)beta;
  const bool upd = special == 0;
  if (ft0 < ft1) load_late(c, upd);
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  for (int t = 0; t < nt; ++t) {
    int t0, t1;
    tile_at(t, t0, t1);
    if (t0 >= t1) break;  // tiles past the last row (the same for every thread)
    if (t > 0) {
      load_tile(c, t0, t1);
      load_late(c, upd);
    }
    float* pt = ptb[t & 1];
    const int t1q = t0 + ((t1 - t0 + 3) & ~3);
    const int i = c.i;
    // r_k, p_k of the own quad; p_k of the halo rows and of the far neighbours
    Q rn, pn;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      rn.v[e] = c.rv.v[e] - al * c.qv.v[e];
      pn.v[e] = rn.v[e] + be * c.pv.v[e];
    }
    if (c.own) {
      st16(a.r_out + i, rn);
      st16(a.p_new + i, pn);
      *reinterpret_cast<Q*>(&pt[hq + 4 * tid]) = pn;
#pragma unroll
      for (int e = 0; e < 4; ++e) part[0] += (double)rn.v[e] * (double)rn.v[e];
    }
#pragma unroll
    for (int u = 0; u < UH; ++u) {
      const int h = tid + u * NT;
      if (h < 2 * hq)
        pt[h < hq ? h : hq + (t1q - t0) + (h - hq)] = c.hin[u] ? (c.hr[u] - al * c.hqv[u]) + be * c.hp[u] : 0.0f;
    }
    // before the barrier: the far neighbours' p_k and their terms of q = A p_k (sp), and x; the SpMV
    // phase then holds only the near diagonals' values and sp
    Q sp;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float s = pn.v[e];
#pragma unroll
      for (int d = NN + 1; d < ND; ++d) {
        const int f = d - NN - 1;
        const float fu = (c.fur[f].v[e] - al * c.fuq[f].v[e]) + be * c.fup[f].v[e];
        const float fd = (c.fdr[f].v[e] - al * c.fdq[f].v[e]) + be * c.fdp[f].v[e];
        const float cl = c.jl[d] + e >= 0 ? c.vl[d].v[e] : 0.0f;
        s += c.vu[d].v[e] * fu;
        s += cl * fd;
      }
      sp.v[e] = s;
    }
    if (c.own && upd) {
      Q xn;
#pragma unroll
      for (int e = 0; e < 4; ++e) xn.v[e] = c.xv.v[e] + al * c.pv.v[e];
      st16(a.x + i, xn);
    }
    __syncthreads();
    if (c.own) {
      Q qn;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int li = hq + 4 * tid + e;
        float s = sp.v[e];
#pragma unroll
        for (int d = 1; d <= NN; ++d) {
          const float cl = c.jl[d] + e >= 0 ? c.vl[d].v[e] : 0.0f;
          s += c.vu[d].v[e] * pt[li + o[d]];
          s += cl * pt[li - o[d]];
        }
        s = i + e < n ? s : 0.0f;
        qn.v[e] = s;
        part[1] += (double)pn.v[e] * (double)s;
        part[2] += (double)s * (double)rn.v[e];
        part[3] += (double)s * (double)s;
      }
      st16(a.q_out + i, qn);
    }
  }
  store_out<NT, 4>(part, a.part.C + ((k + 1) & 1) * 4 * (int64_t)G, G, sm_out);
  post_state(st, a.post);
}

// ---------------------------------------------------------------------------------------------
// The one-kernel iteration with register-resident rows and software-pipelined tiles (mode 5,
// k1_vector 3), for 4-diagonal operators.  Same arithmetic, bounds handling, tiling and launch
// contract as pcg_one_q; the difference is when each tile's loads are issued.  A tile's loads come
// in three groups: A (r, q, p of the own quad and of the near halo), F (r, q, p of the far
// neighbours and x) and B (the upper and lower off-diagonal values).  Phase 1 of a tile (r_k, p_k,
// x, the far neighbours' p_k) needs A and F; phase 2 (q = A p_k from LDS, the partial sums) needs
// B.  The next tile's A and F are issued right after phase 1 (before the barrier), so they are in
// flight during the barrier and phase 2; the next tile's B is issued after phase 2, in flight during
// the next phase 1.  A workgroup thus always has loads outstanding, instead of alternating a load
// round trip with a compute phase on every tile.  The far-diagonal terms move to phase 2 (their
// p_k, two quads, cross the barrier in registers), so phase 1 holds no diagonal values.
// x is updated every other launch (XM, chosen by the host from the launch parity): a deferring
// launch (XM_DEFER, even k) neither reads nor writes x and records its alpha_{k-1} in
// a_pend[(k+1) & 1]; the flushing launch after it (XM_FLUSH) adds both terms,
// x += a_pend p_{k-2} + alpha_{k-1} p_{k-1}, reading p_{k-2} from the buffer it is about to
// overwrite with p_k (own rows only, read before the write by the same thread).  Per row and pair
// of launches: x read and written once and p_{k-2} read, 12 B instead of 16.  The two products are
// added in the order, and with the same FP32 roundings, as two eager updates, so x is bitwise that
// of updating every launch.  A stop leaves nothing pending: a flushing launch that stops adds the
// deferred term itself (a pass over its rows, x += a_pend p_{k-2}), so x is x_{k-1} at every stop,
// as the residual replacement and the end of the solve need.  Special launches (start, resume)
// have alpha = 0: they never find a term pending (the stop before them flushed it) and defer none.
// ---------------------------------------------------------------------------------------------
template <int NT, int ND, int NN, int XM = XM_EAGER>
__global__ void __launch_bounds__(NT) pcg_one_p(PcgArgs<float, float> a) {
  static_assert(ND == 4 && NN >= 2 && NN <= 3, "instantiated for 4-diagonal operators with 2 or 3 near diagonals");
  constexpr int NF = ND - 1 - NN;  // far off-diagonals
  constexpr int NFA = NF > 0 ? NF : 1;
  constexpr int TT = 4 * NT;  // rows per tile
  constexpr int NW = NT / 64;
  constexpr int QRED_H = (2 * QRED_G + NT - 1) / NT;  // pairs of partial sums loaded per thread
  constexpr int UH = (2 * LDS_HALO + NT - 1) / NT;    // near halo rows per thread
  using Q = V16<float>;
  __shared__ double sm_w[4 * NW];
  __shared__ double sm_out[4 * NW];
  // p_k of the tile and its near halo, two buffers used by alternate tiles (the buffer a tile
  // writes was last read two tiles back, before the barrier of the tile between)
  __shared__ __align__(16) float ptb[2][TT + 2 * LDS_HALO];
  PcgState* st = a.st;
  const int G = gridDim.x;
  const int tid = (int)threadIdx.x;
  const int64_t k = a.kk;
  const int n = (int)a.n;
  const float* vals = a.vals;
  const int ld = (int)dia_ld(a.n);
  int o[ND];
#pragma unroll
  for (int d = 1; d < ND; ++d) o[d] = a.off.o[d];
  const int hq = (o[NN] + 3) & ~3;  // near halo rows on either side, whole quads (<= LDS_HALO)
  // tiles as in pcg_one_q: round-robin within whole XCD slabs when the launch is sized for it
  const int64_t rpb = a.rows_per_block;
  const int ntq = (int)((rpb + TT - 1) / TT);
  const bool ilv = G % N_XCD == 0 && rpb % ntq == 0 && (rpb / ntq) % 4 == 0;
  int nt, tl, lim;
  int64_t tbase, tstride;
  if (ilv) {
    nt = ntq;
    tl = (int)(rpb / ntq);
    const int per = G / N_XCD;
    tbase = (int64_t)(blockIdx.x % N_XCD) * per * rpb + (int64_t)(blockIdx.x / N_XCD) * tl;
    tstride = (int64_t)per * tl;
    lim = n;
  } else {
    const int64_t r0l = block_first_row(rpb);
    const int r0 = (int)min(r0l, a.n), r1 = (int)min(a.n, r0l + rpb);
    nt = r1 > r0 ? (r1 - r0 + TT - 1) / TT : 0;
    tl = nt > 0 ? (((r1 - r0 + nt - 1) / nt + 3) & ~3) : TT;
    tbase = r0;
    tstride = tl;
    lim = r1;
  }
  auto tile_at = [&](int t, int& t0, int& t1) {
    const int64_t b = tbase + (int64_t)t * tstride;
    t0 = (int)min(b, (int64_t)lim);
    t1 = (int)min(b + tl, (int64_t)lim);
  };
  const double* pin = a.part.C + (k & 1) * 4 * (int64_t)G;
  const int G4 = 4 * G;
  // the partial sums as 16-byte pairs (G is even: a pair never spans two arrays), entries
  // 2 (u NT + tid) and the next in thread tid
  double2 pv_in[QRED_H];
#pragma unroll
  for (int u = 0; u < QRED_H; ++u) pv_in[u] = *reinterpret_cast<const double2*>(pin + min(2 * (u * NT + tid), G4 - 2));
  const DecideIn din = decide_in(st, k);
  // the alpha whose x term the previous (deferring) launch left to this one
  const float apd = XM == XM_FLUSH ? (float)st->a_pend[k & 1] : 0.0f;
  // every wave's head loads are queued before any wave of the workgroup issues its tile's: the head
  // waits for them, the tile's loads have time until the head is done
  __syncthreads();
  struct TA {  // group A: the own quad and the near halo rows
    Q rv, qv, pv;
    float hr[UH], hqv[UH], hp[UH];
    int i, t0, t1;
    bool own, hin[UH];
  };
  struct TF {  // group F: the far neighbours, x and (flushing launches) p_{k-2}
    Q fur[NFA], fuq[NFA], fup[NFA], fdr[NFA], fdq[NFA], fdp[NFA], xv, p2;
  };
  struct TB {  // group B: the off-diagonal values
    Q vu[ND], vl[ND];
  };
  auto load_a = [&](TA& c, int t0, int t1) {
    c.t0 = t0;
    c.t1 = t1;
    c.i = t0 + 4 * tid;
    c.own = c.i < t1;
    const int ic = c.own ? c.i : 0;
    c.rv = ld16(a.r + ic);
    c.qv = ld16(a.q + ic);
    c.pv = ld16(a.p_old + ic);
    const int t1q = t0 + ((t1 - t0 + 3) & ~3);
#pragma unroll
    for (int u = 0; u < UH; ++u) {
      const int h = tid + u * NT;
      const int row = h < hq ? t0 - hq + h : t1q + (h - hq);
      c.hin[u] = h < 2 * hq && row >= 0 && row < n;
      const int hc = c.hin[u] ? row : 0;
      c.hr[u] = a.r[hc];
      c.hqv[u] = a.q[hc];
      c.hp[u] = a.p_old[hc];
    }
  };
  auto load_f = [&](TF& c, const TA& ca) {
    const int ic = ca.own ? ca.i : 0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const int od = o[NN + 1 + f];
      const int ju = ic + od < n ? ic + od : 0;
      const int jd = ic - od >= -3 ? ic - od : 0;
      c.fur[f] = ld16(a.r + ju);
      c.fuq[f] = ld16(a.q + ju);
      c.fup[f] = ld16(a.p_old + ju);
      c.fdr[f] = ld16(a.r + jd);
      c.fdq[f] = ld16(a.q + jd);
      c.fdp[f] = ld16(a.p_old + jd);
    }
    // used only by an updating launch (every launch but the special ones)
    if constexpr (XM != XM_DEFER) c.xv = ld16(a.x + ic);
    if constexpr (XM == XM_FLUSH) c.p2 = ld16(a.p_new + ic);  // p_{k-2}, before this launch's p_k
  };
  auto load_b = [&](TB& c, const TA& ca) {
    const int ic = ca.own ? ca.i : 0;
#pragma unroll
    for (int d = 1; d < ND; ++d) {
      const float* vd = vals + (int64_t)d * ld;
      const int jl = ic - o[d];
      c.vu[d] = ld16(vd + ic);
      c.vl[d] = ld16(vd + (jl >= -3 ? jl : 0));
    }
  };
  TA ca;
  TF cf;
  TB cb;
  ca.t0 = ca.t1 = 0;
  int ft0 = 0, ft1 = 0;
  if (nt > 0) tile_at(0, ft0, ft1);
  const bool any = ft0 < ft1;
  // issued unconditionally (a workgroup without rows loads row 0, unused): with the same loads in
  // flight on every path, the head's wait for its partial sums does not also wait for these
  load_a(ca, any ? ft0 : 0, any ? ft1 : 0);
  load_b(cb, ca);
  load_f(cf, ca);
  // the stop flag is read after the first loads are issued, as in pcg_one_q
  if (st->done) {
    post_state(st, a.post);
    return;
  }
  double in[4];
  {
    const int w = tid >> 6;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < QRED_H; ++u) {
      const int f = 2 * (u * NT + tid);
      const double v = pv_in[u].x + pv_in[u].y;
#pragma unroll
      for (int m = 0; m < 4; ++m) s[m] += f >= m * G && f < (m + 1) * G ? v : 0.0;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double t = wave_sum(s[m]);
      if ((tid & 63) == 0) sm_w[m * NW + w] = t;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double t = 0.0;
#pragma unroll
      for (int v = 0; v < NW; ++v) t += sm_w[m * NW + v];
      in[m] = t;
    }
  }
  double alpha, beta;
  int special;
  if (!one_decide_pre(st, din, k, in, alpha, beta, special)) {
    if constexpr (XM != XM_EAGER) {
      if (blockIdx.x == 0 && tid == 0) st->a_pend[(k + 1) & 1] = 0.0;
      // a stopping flushing launch adds the term the launch before it deferred: x = x_{k-1}
      if (XM == XM_FLUSH && special == 0 && apd != 0.0f) {
        for (int t = 0; t < nt; ++t) {
          int t0, t1;
          tile_at(t, t0, t1);
          const int i = t0 + 4 * tid;
          if (i < t1) {
            const Q xv = ld16(a.x + i), p2 = ld16(a.p_new + i);
            Q xn;
#pragma unroll
            for (int e = 0; e < 4; ++e) xn.v[e] = xv.v[e] + apd * p2.v[e];
            st16(a.x + i, xn);
          }
        }
      }
    }
    post_state(st, a.post);
    return;
  }
  const float al = (float)alpha, be = (float)beta;
  const bool upd = special == 0;
  if constexpr (XM != XM_EAGER) {
    if (blockIdx.x == 0 && tid == 0) st->a_pend[(k + 1) & 1] = XM == XM_DEFER && upd ? (double)al : 0.0;
  }
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  // one tile: phase 1, the next tile's A and F, the barrier, phase 2, the next tile's B; returns
  // whether a next tile follows.  The loads are issued unconditionally (the last tile reloads its own
  // rows, unused), so that every path has the same loads in flight and the waits before each use
  // count exactly the loads issued after the one it needs.
  auto step = [&](int t) -> bool {
    const int t0 = ca.t0, t1 = ca.t1;
    float* pt = ptb[t & 1];
    const int t1q = t0 + ((t1 - t0 + 3) & ~3);
    const int i = ca.i;
    const bool own = ca.own;
    // phase 1: r_k, p_k of the own quad, x; p_k of the halo rows and of the far neighbours
    Q rn, pn;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      rn.v[e] = ca.rv.v[e] - al * ca.qv.v[e];
      pn.v[e] = rn.v[e] + be * ca.pv.v[e];
    }
    if (own) {
      st16(a.r_out + i, rn);
      st16(a.p_new + i, pn);
      *reinterpret_cast<Q*>(&pt[hq + 4 * tid]) = pn;
#pragma unroll
      for (int e = 0; e < 4; ++e) part[0] += (double)rn.v[e] * (double)rn.v[e];
    }
#pragma unroll
    for (int u = 0; u < UH; ++u) {
      const int h = tid + u * NT;
      if (h < 2 * hq)
        pt[h < hq ? h : hq + (t1q - t0) + (h - hq)] =
            ca.hin[u] ? (ca.hr[u] - al * ca.hqv[u]) + be * ca.hp[u] : 0.0f;
    }
    Q fu[NFA], fd[NFA];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        fu[f].v[e] = (cf.fur[f].v[e] - al * cf.fuq[f].v[e]) + be * cf.fup[f].v[e];
        fd[f].v[e] = (cf.fdr[f].v[e] - al * cf.fdq[f].v[e]) + be * cf.fdp[f].v[e];
      }
    if (XM != XM_DEFER && own && upd) {
      Q xn;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // flushing: the deferred term first, then this launch's, as two eager updates would add them
        const float xb = XM == XM_FLUSH && apd != 0.0f ? cf.xv.v[e] + apd * cf.p2.v[e] : cf.xv.v[e];
        xn.v[e] = xb + al * ca.pv.v[e];
      }
      st16(a.x + i, xn);
    }
    // the next tile's groups A and F go out now, in flight during the barrier and phase 2
    int n0 = 0, n1 = 0;
    if (t + 1 < nt) tile_at(t + 1, n0, n1);
    const bool more = n0 < n1;
    load_a(ca, more ? n0 : t0, more ? n1 : t1);
    load_f(cf, ca);
    __syncthreads();
    // phase 2: q = A p_k (near diagonals from LDS, far ones from fu / fd), partial sums
    if (own) {
      Q qn;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int li = hq + 4 * tid + e;
        float s = pn.v[e];
#pragma unroll
        for (int d = NN + 1; d < ND; ++d) {
          const int f = d - NN - 1;
          const float cl = i - o[d] + e >= 0 ? cb.vl[d].v[e] : 0.0f;
          s += cb.vu[d].v[e] * fu[f].v[e];
          s += cl * fd[f].v[e];
        }
#pragma unroll
        for (int d = 1; d <= NN; ++d) {
          const float cl = i - o[d] + e >= 0 ? cb.vl[d].v[e] : 0.0f;
          s += cb.vu[d].v[e] * pt[li + o[d]];
          s += cl * pt[li - o[d]];
        }
        s = i + e < n ? s : 0.0f;
        qn.v[e] = s;
        part[1] += (double)pn.v[e] * (double)s;
        part[2] += (double)s * (double)rn.v[e];
        part[3] += (double)s * (double)s;
      }
      st16(a.q_out + i, qn);
    }
    // the next tile's group B, in flight during its phase 1
    load_b(cb, ca);
    return more;
  };
  // the first tiles unrolled (straight-line code keeps the waits exact), any further ones in a loop
  bool go = any;
#pragma unroll
  for (int t = 0; t < 3; ++t)
    if (go) go = step(t);
  for (int t = 3; go; ++t) go = step(t);
  store_out<NT, 4>(part, a.part.C + ((k + 1) & 1) * 4 * (int64_t)G, G, sm_out);
  post_state(st, a.post);
}

// Zero the pads of up to 8 iteration vectors (PCG_PAD_BYTES before and after each of n elements).
struct PadList {
  float* v[8];
  int count;
};

// Zero the vectors of the list whole (with their pads) when bit b of `whole` is set, else only
// their pads: one launch for the set-up of a scaled FP32 solve
__global__ void zero_list_k(int64_t n, PadList l, int whole) {
  constexpr int P = PCG_PAD_BYTES / (int)sizeof(float);
  for (int b = 0; b < l.count; ++b) {
    float* v = l.v[b];
    const int64_t lo = -P, hi = (whole >> b) & 1 ? n + P : P;  // [-P, P) are the two pads, shifted
    for (int64_t j = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < hi; j += (int64_t)gridDim.x * blockDim.x)
      v[(whole >> b) & 1 ? j : (j < 0 ? j : n + j)] = 0.0f;
  }
}

__global__ void zero_pads_k(int64_t n, PadList l) {
  constexpr int P = PCG_PAD_BYTES / (int)sizeof(float);
  const int t = (int)threadIdx.x;  // 2 P threads
  for (int b = 0; b < l.count; ++b) {
    float* v = l.v[b];
    if (t < P)
      v[t - P] = 0.0f;
    else if (t < 2 * P)
      v[n + t - P] = 0.0f;
  }
}

__global__ void set_special_k(PcgState* st, int64_t k, int kind) {
  if (threadIdx.x == 0) {
    st->k_special = k;
    st->special = kind;
  }
}

// ---------------------------------------------------------------------------------------------
// K2: r -= alpha q, partials r.z and r.r
// ---------------------------------------------------------------------------------------------
// Rows [r0, r1) of a workgroup as VN-row tiles: thread t owns tiles t, t + NT, t + 2 NT, ...;
// U of them are loaded together (issued before any use), rows past the last full tile (fewer than
// VN, only at the end of the matrix) go one per thread.
constexpr int TILE_U = 2;

// JAC: r.z with z = dinv r; without it z = r and both partials are r.r
template <int NT, typename T, typename V, bool JAC = true>
__global__ void __launch_bounds__(NT) pcg_k2(PcgArgs<T, V> a) {
  __shared__ double sm[4];                  // broadcast of reduced partial sums
  __shared__ double sm_out[2 * (NT / 64)];  // wave totals of this kernel's partial sums
  PcgState* st = a.st;
  if (st->done1) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
    return;
  }
  const int G = gridDim.x;
  constexpr int VN = V16<T>::N;
  constexpr int U = TILE_U;
  const int64_t r0 = block_first_row(a.rows_per_block);
  const int64_t r1 = min(a.n, r0 + a.rows_per_block);
  const int64_t r1v = r0 < r1 ? r0 + ((r1 - r0) / VN) * VN : r0;
  int64_t t = r0 + (int64_t)threadIdx.x * VN;
  // the first step's loads are issued before the partial-sum reduction, so they overlap it
  V16<T> rv[U], qv[U], dv[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i0 = t + (int64_t)u * VN * NT;
    if (i0 < r1v) {
      rv[u] = ld16(a.r + i0);
      qv[u] = ld16(a.q + i0);
      if (JAC) dv[u] = ld16(a.dinv + i0);
    }
  }
  double pq;
  reduce_in<1>(a.part.A, G, sm, &pq);
  const double gamma = st->gamma_cur;
  const double alpha = gamma / pq;
  const T al = (T)alpha;
  double pg = 0.0, pr = 0.0;
  for (;;) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i0 = t + (int64_t)u * VN * NT;
      if (i0 < r1v) {
#pragma unroll
        for (int j = 0; j < VN; ++j) {
          const T ri = rv[u].v[j] - al * qv[u].v[j];
          rv[u].v[j] = ri;
          const double rr = (double)ri * (double)ri;
          pg += JAC ? (double)ri * (double)(dv[u].v[j] * ri) : rr;
          pr += rr;
        }
        st16(a.r + i0, rv[u]);
      }
    }
    t += (int64_t)U * VN * NT;
    if (t >= r1v) break;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i0 = t + (int64_t)u * VN * NT;
      if (i0 < r1v) {
        rv[u] = ld16(a.r + i0);
        qv[u] = ld16(a.q + i0);
        if (JAC) dv[u] = ld16(a.dinv + i0);
      }
    }
  }
  {
    const int64_t i = r1v + threadIdx.x;
    if (i < r1) {
      const T ri = a.r[i] - al * a.q[i];
      a.r[i] = ri;
      const double rr = (double)ri * (double)ri;
      pg += JAC ? (double)ri * (double)(a.dinv[i] * ri) : rr;
      pr += rr;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alpha;
    st->gamma_prev = gamma;
    st->k = st->k + 1;
  }
  const double outv[2] = {pg, pr};
  store_out<NT, 2>(outv, a.part.B, G, sm_out);
}

// ---------------------------------------------------------------------------------------------
// Split iteration (three kernels): K0 materializes p, KS is a plain SpMV on it, K2 as above.
// Fewer bytes move per row in the SpMV (one neighbour vector instead of r, dinv and p_old), at the
// cost of a third launch and of writing and re-reading p once.
// ---------------------------------------------------------------------------------------------
// JAC: z = dinv r (Jacobi); without it z = r (the Jacobi-scaled operator, identity preconditioner)
template <int NT, typename T, typename V, bool JAC = true>
__global__ void __launch_bounds__(NT) pcg_k0(PcgArgs<T, V> a) {
  __shared__ double sm[4];                  // broadcast of reduced partial sums
  __shared__ double sm_out[2 * (NT / 64)];  // wave totals of this kernel's partial sums
  PcgState* st = a.st;
  if (st->done) return;
  constexpr int VN = V16<T>::N;
  constexpr int U = TILE_U;
  const bool upd_x = st->k > 0;
  const int64_t b0 = block_first_row(a.rows_per_block);
  const int64_t b1 = min(a.n, b0 + a.rows_per_block);
  const int64_t r1v = b0 < b1 ? b0 + ((b1 - b0) / VN) * VN : b0;
  int64_t t = b0 + (int64_t)threadIdx.x * VN;
  // the first step's loads are issued before the partial-sum reduction, so they overlap it
  V16<T> dv[U], rv[U], po[U], xv[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i0 = t + (int64_t)u * VN * NT;
    if (i0 < r1v) {
      if (JAC) dv[u] = ld16(a.dinv + i0);
      rv[u] = ld16(a.r + i0);
      po[u] = ld16(a.p_old + i0);
      if (upd_x) xv[u] = ld16(a.x + i0);
    }
  }
  double gamma;
  T beta, ap;
  int64_t k, r0, r1;
  if (!iter_head<NT, T, V>(a, sm, gamma, beta, ap, k, r0, r1)) return;
  for (;;) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i0 = t + (int64_t)u * VN * NT;
      if (i0 < r1v) {
        V16<T> pi;
#pragma unroll
        for (int j = 0; j < VN; ++j) pi.v[j] = (JAC ? dv[u].v[j] * rv[u].v[j] : rv[u].v[j]) + beta * po[u].v[j];
        st16(a.p_new + i0, pi);
        if (upd_x) {
#pragma unroll
          for (int j = 0; j < VN; ++j) xv[u].v[j] += ap * po[u].v[j];
          st16(a.x + i0, xv[u]);
        }
      }
    }
    t += (int64_t)U * VN * NT;
    if (t >= r1v) break;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i0 = t + (int64_t)u * VN * NT;
      if (i0 < r1v) {
        if (JAC) dv[u] = ld16(a.dinv + i0);
        rv[u] = ld16(a.r + i0);
        po[u] = ld16(a.p_old + i0);
        if (upd_x) xv[u] = ld16(a.x + i0);
      }
    }
  }
  {
    const int64_t i = r1v + threadIdx.x;
    if (i < r1) {
      const T p0 = a.p_old[i];
      a.p_new[i] = (JAC ? a.dinv[i] * a.r[i] : a.r[i]) + beta * p0;
      if (upd_x) a.x[i] += ap * p0;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->gamma_cur = gamma;
}

// UNIT: the main diagonal is 1 and not read (the Jacobi-scaled operator)
// one DIA row of q = A p with bounds checks
template <typename T, typename V, bool UNIT = false>
__device__ __forceinline__ double spmv_dia_row(const PcgArgs<T, V>& a, int64_t i) {
  const int64_t ld = dia_ld(a.n);
  const T* p = a.p_new;
  const T pi = p[i];
  T s = UNIT ? pi : (T)a.vals[i] * pi;
  for (int d = 1; d < a.ndiag; ++d) {
    const int64_t o = a.off.o[d];
    const V* v = a.vals + (int64_t)d * ld;
    if (i + o < a.n) s += (T)v[i] * p[i + o];
    if (i - o >= 0) s += (T)v[i - o] * p[i - o];
  }
  sto(&a.q[i], s);
  return (double)pi * (double)s;
}

// R interior DIA rows i, i + NT, ... of q = A p, all loads in flight before the first use
template <int ND, int R, int NT, typename T, typename V, bool UNIT = false>
__device__ __forceinline__ double spmv_dia_rows(const PcgArgs<T, V>& a, int64_t i) {
  const int64_t ld = dia_ld(a.n);
  const T* p = a.p_new;
  T pv[R], pu[R][ND - 1], pl[R][ND - 1];
  V d0[R], vu[R][ND - 1], vl[R][ND - 1];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t row = i + (int64_t)r * NT;
    pv[r] = p[row];
    if (!UNIT) d0[r] = a.vals[row];
#pragma unroll
    for (int d = 1; d < ND; ++d) {
      const int64_t o = a.off.o[d];
      vu[r][d - 1] = a.vals[(int64_t)d * ld + row];
      vl[r][d - 1] = a.vals[(int64_t)d * ld + row - o];
      pu[r][d - 1] = p[row + o];
      pl[r][d - 1] = p[row - o];
    }
  }
  double part = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    T s = UNIT ? pv[r] : (T)d0[r] * pv[r];
#pragma unroll
    for (int d = 0; d < ND - 1; ++d) {
      s += (T)vu[r][d] * pu[r][d];
      s += (T)vl[r][d] * pl[r][d];
    }
    sto(&a.q[i + (int64_t)r * NT], s);
    part += (double)pv[r] * (double)s;
  }
  return part;
}

// rows [r0, r1) of q = A p, R rows per thread and step (NT apart, coalesced)
template <int ND, int R, int NT, typename T, typename V, bool UNIT = false>
__device__ __forceinline__ double spmv_dia_scalar(const PcgArgs<T, V>& a, int64_t r0, int64_t r1, int64_t omax) {
  double part = 0.0;
  int64_t i = r0 + threadIdx.x;
  for (; i + (int64_t)(R - 1) * NT < r1; i += (int64_t)R * NT) {
    if (i - omax >= 0 && i + (int64_t)(R - 1) * NT + omax < a.n) {
      part += spmv_dia_rows<ND, R, NT, T, V, UNIT>(a, i);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) part += spmv_dia_row<T, V, UNIT>(a, i + (int64_t)r * NT);
    }
  }
  for (; i < r1; i += NT) part += spmv_dia_row<T, V, UNIT>(a, i);
  return part;
}

// VN-row tile [i0, i0 + VN) of q = A p with 16-byte loads (interior tiles only)
template <int ND, typename T, typename V, bool UNIT = false>
__device__ __forceinline__ double spmv_dia_tile(const PcgArgs<T, V>& a, int64_t i0) {
  constexpr int VN = V16<T>::N;
  const int64_t ld = dia_ld(a.n);
  const int nd = ND > 0 ? ND : a.ndiag;
  const T* p = a.p_new;
  const V16<T> pv = ld16(p + i0);
  V16<T> s;
  if (UNIT) {
    s = pv;
  } else {
    const V16<V> d0 = ld16(a.vals + i0);
#pragma unroll
    for (int j = 0; j < VN; ++j) s.v[j] = (T)d0.v[j] * pv.v[j];
  }
#pragma unroll
  for (int d = 1; d < nd; ++d) {
    const int64_t o = a.off.o[d];
    const V* vd = a.vals + (int64_t)d * ld;
    const V16<V> vu = ld16(vd + i0), vl = ld16(vd + i0 - o);
    const V16<T> pu = ld16(p + i0 + o), pl = ld16(p + i0 - o);
#pragma unroll
    for (int j = 0; j < VN; ++j) {
      s.v[j] += (T)vu.v[j] * pu.v[j];
      s.v[j] += (T)vl.v[j] * pl.v[j];
    }
  }
  st16(a.q + i0, s);
  double part = 0.0;
#pragma unroll
  for (int j = 0; j < VN; ++j) part += (double)pv.v[j] * (double)s.v[j];
  return part;
}

// KS: q = A p (p materialized by K0), partial p.q.  With CG set it is KB of the single-reduction
// iteration (launched with p_new = u, q = w): w = A u, partial u.w, and workgroup 0 moves the
// scalars KA staged into the fields KA reads.
template <int NT, int L, bool DIA, int ND, bool CG, typename T, typename V, bool UNIT = false>
__global__ void __launch_bounds__(NT) pcg_ks(PcgArgs<T, V> a) {
  __shared__ double sm[4];                  // broadcast of reduced partial sums
  __shared__ double sm_out[2 * (NT / 64)];  // wave totals of this kernel's partial sums
  PcgState* st = a.st;
  if (st->done1) {
    if (CG && blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
    return;
  }
  if (CG && blockIdx.x == 0 && threadIdx.x == 0) {
    st->k = st->k_next;
    st->gamma_prev = st->gamma_next;
    st->alpha = st->alpha_next;
  }
  const int64_t r0 = block_first_row(a.rows_per_block);
  const int64_t r1 = min(a.n, r0 + a.rows_per_block);
  const T* p = a.p_new;
  double part = 0.0;
  if (DIA) {
    int64_t omax = 0;
    for (int d = 1; d < a.ndiag; ++d) omax = max(omax, (int64_t)a.off.o[d]);
    if (a.vec == 1) {
      constexpr int VN = V16<T>::N;
      for (int64_t i0 = r0 + (int64_t)threadIdx.x * VN; i0 < r1; i0 += (int64_t)VN * NT) {
        if (i0 + VN <= r1 && i0 - omax >= 0 && i0 + VN - 1 + omax < a.n)
          part += spmv_dia_tile<ND, T, V, UNIT>(a, i0);
        else
          for (int64_t i = i0; i < min(r1, i0 + VN); ++i) part += spmv_dia_row<T, V, UNIT>(a, i);
      }
    } else if (ND > 0) {
      part = a.vec == 2 ? spmv_dia_scalar<(ND > 0 ? ND : 1), 4, NT, T, V, UNIT>(a, r0, r1, omax)
                        : spmv_dia_scalar<(ND > 0 ? ND : 1), 2, NT, T, V, UNIT>(a, r0, r1, omax);
    } else {
      for (int64_t i = r0 + threadIdx.x; i < r1; i += NT) part += spmv_dia_row<T, V, UNIT>(a, i);
    }
  } else {
    const int sub = threadIdx.x % L;
    for (int64_t row = r0 + threadIdx.x / L; row < r1; row += NT / L) {
      const int b = a.rowptr[row], e = a.rowptr[row + 1];
      T s = 0;
      for (int j = b + sub; j < e; j += L) s += (T)a.vals[j] * p[a.col[j]];
#pragma unroll
      for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
      if (sub == 0) {
        sto(&a.q[row], s);
        part += (double)p[row] * (double)s;
      }
    }
  }
  store_out<NT, 1>(&part, a.part.A, gridDim.x, sm_out);
}

// KS (or KB with CG set) on a BSR3 operator: the block rows [r0/3, r1/3) of the workgroup (rows
// per workgroup are a multiple of 12), L lanes per block row
template <int NT, int L, bool CG, typename T, typename V>
__global__ void __launch_bounds__(NT) pcg_ks_bsr3(PcgArgs<T, V> a) {
  __shared__ double sm[4];                  // broadcast of reduced partial sums
  __shared__ double sm_out[2 * (NT / 64)];  // wave totals of this kernel's partial sums
  PcgState* st = a.st;
  if (st->done1) {
    if (CG && blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
    return;
  }
  if (CG && blockIdx.x == 0 && threadIdx.x == 0) {
    st->k = st->k_next;
    st->gamma_prev = st->gamma_next;
    st->alpha = st->alpha_next;
  }
  const int64_t r0 = block_first_row(a.rows_per_block);
  const int64_t r1 = min(a.n, r0 + a.rows_per_block);
  const T* p = a.p_new;
  const int sub = threadIdx.x % L;
  double part = 0.0;
  // U block rows per lane group in flight (their loads issued together), NT / L groups apart, and CH
  // blocks per lane of each: for FP32 two rows of one block per lane (C5: 7.6% faster than one row;
  // four rows, or two blocks per lane, cut the occupancy and measured slower), for FP64 one row of two
  // blocks per lane (C5 FP64 solve 594 -> 547 ms against one block: the rows of 42 and 65 blocks
  // leave the serial loop).  The block-row pointers of the next U rows are loaded one batch ahead,
  // so a batch waits on two dependent round trips (block columns and values, then the p gathers)
  // instead of three (C5 mixed solve 379 -> 360 ms).
  constexpr int U = sizeof(T) == 4 ? 2 : 1;
  constexpr int CH = sizeof(T) == 4 ? 1 : 2;
  constexpr int64_t STEP = NT / L;
  const int64_t R1 = r1 / 3;
  int64_t R0 = r0 / 3 + threadIdx.x / L;
  int pa[U], pb[U];
  auto row_ptrs = [&](int64_t Rb) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t R = Rb + u * STEP;
      const int64_t Rc = R < R1 ? R : 0;
      pa[u] = a.rowptr[Rc];
      pb[u] = a.rowptr[Rc + 1];
    }
  };
  row_ptrs(R0);
  for (; R0 < R1; R0 += U * STEP) {
    int64_t Rs[U];
    int bs[U], len[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      Rs[u] = R0 + u * STEP < R1 ? R0 + u * STEP : -1;
      bs[u] = pa[u];
      len[u] = Rs[u] < 0 ? 0 : pb[u] - pa[u];
    }
    row_ptrs(R0 + U * STEP);
    T sr[U][3];
    bsr3_rows_at<L, U, V, T, CH>(a.col, a.vals, p, bs, len, sub, sr);
    if (sub == 0) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (Rs[u] < 0) continue;
        const int64_t R = Rs[u];
        sto(&a.q[3 * R], sr[u][0]);
        sto(&a.q[3 * R + 1], sr[u][1]);
        sto(&a.q[3 * R + 2], sr[u][2]);
        part += (double)p[3 * R] * (double)sr[u][0] + (double)p[3 * R + 1] * (double)sr[u][1] +
                (double)p[3 * R + 2] * (double)sr[u][2];
      }
    }
  }
  store_out<NT, 1>(&part, a.part.A, gridDim.x, sm_out);
}

// ---------------------------------------------------------------------------------------------
// Single-reduction iteration (Chronopoulos & Gear, J. Comput. Appl. Math. 25 (1989) 153-168):
// with u = M r, w = A u, gamma = r.u and delta = u.w reduced together,
//   beta_k = gamma_k / gamma_{k-1},  alpha_k = gamma_k / (delta_k - beta_k gamma_k / alpha_{k-1})
//   p = u + beta p,  s = w + beta s  (s = A p by recursion),  x += alpha p,  r -= alpha s
// so the SpMV acts on u, which is ready as soon as r is updated: two kernels per iteration, KA
// (all vector updates, u, partials r.u and r.r) and KB (w = A u, partial u.w), and one reduction.
// It is the same Krylov iteration as CG in exact arithmetic; the rounding differs, which is why
// it serves the FP32 inner iterations (closed by FP64 residual replacement), not the plain FP64
// solve.
// ---------------------------------------------------------------------------------------------
template <int NT, typename T, typename V>
__global__ void __launch_bounds__(NT) pcg_ka(PcgArgs<T, V> a) {
  __shared__ double sm[4];                  // broadcast of reduced partial sums
  __shared__ double sm_out[2 * (NT / 64)];  // wave totals of this kernel's partial sums
  PcgState* st = a.st;
  if (st->done) return;
  const int G = gridDim.x;
  const int64_t k = st->k;
  // KA both reads and writes r.u / r.r partials: the previous KA's in one buffer, its own in the
  // other, chosen by the parity of k (the same for every workgroup)
  const double* b_in = (k & 1) ? a.part.B2 : a.part.B;
  double* b_out = (k & 1) ? a.part.B : a.part.B2;
  double in2[2], delta;
  reduce_in<2>(b_in, G, sm, in2);
  reduce_in<1>(a.part.A, G, sm + 2, &delta);
  const double gamma = in2[0], rr = in2[1];
  const bool stop_conv = !(rr > st->atol2);
  const bool stop_cap = k >= st->maxiter;
  const bool stop_repl = rr <= st->rr_replace;
  const bool bad_in = !isfinite(gamma) || !isfinite(rr) || !isfinite(delta);
  double beta = 0.0, den = delta;
  if (k > 0) {
    beta = gamma / st->gamma_prev;
    den = delta - beta * gamma / st->alpha;
  }
  const double alpha = gamma / den;
  const bool bad = bad_in || (!stop_conv && !isfinite(alpha));
  if (stop_conv || stop_cap || stop_repl || bad) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done1 = 1;
      st->k_stop = k;
      st->rr_stop = rr;
      st->status = bad ? ST_NONFINITE : stop_conv ? ST_CONVERGED : stop_cap ? ST_MAXITER : ST_REPLACE;
    }
    return;
  }
  const T be = (T)beta, al = (T)alpha;
  const int64_t r0 = block_first_row(a.rows_per_block);
  const int64_t r1 = min(a.n, r0 + a.rows_per_block);
  constexpr int VN = V16<T>::N;
  double pg = 0.0, pr = 0.0;
  for (int64_t i0 = r0 + (int64_t)threadIdx.x * VN; i0 < r1; i0 += (int64_t)VN * NT) {
    if (i0 + VN <= r1) {
      V16<T> rv = ld16(a.r + i0), pv = ld16(a.p_new + i0), sv = ld16(a.s + i0), xv = ld16(a.x + i0);
      const V16<T> dv = ld16(a.dinv + i0), wv = ld16(a.q + i0);
      V16<T> uv;
#pragma unroll
      for (int j = 0; j < VN; ++j) {
        const T u = dv.v[j] * rv.v[j];
        const T p = u + be * pv.v[j];
        const T s = wv.v[j] + be * sv.v[j];
        xv.v[j] += al * p;
        const T r = rv.v[j] - al * s;
        const T un = dv.v[j] * r;
        pv.v[j] = p;
        sv.v[j] = s;
        rv.v[j] = r;
        uv.v[j] = un;
        pg += (double)r * (double)un;
        pr += (double)r * (double)r;
      }
      st16(a.p_new + i0, pv);
      st16(a.s + i0, sv);
      st16(a.x + i0, xv);
      st16(a.r + i0, rv);
      st16(a.u + i0, uv);
    } else {
      for (int64_t i = i0; i < r1; ++i) {
        const T d = a.dinv[i];
        const T p = d * a.r[i] + be * a.p_new[i];
        const T s = a.q[i] + be * a.s[i];
        a.x[i] += al * p;
        const T r = a.r[i] - al * s;
        const T un = d * r;
        a.p_new[i] = p;
        a.s[i] = s;
        a.r[i] = r;
        a.u[i] = un;
        pg += (double)r * (double)un;
        pr += (double)r * (double)r;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->k_next = k + 1;
    st->gamma_next = gamma;
    st->alpha_next = alpha;
  }
  const double outv[2] = {pg, pr};
  store_out<NT, 2>(outv, b_out, G, sm_out);
}

// ---------------------------------------------------------------------------------------------
// residual preparation: r64 = b - Ax (Ax nullable = 0), rs = (T) r64; partials of
// gamma = rs.dinv.rs, rr = rs.rs (into B, for the next K1) and of ||r64||^2, b.b (into P)
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(PCG_NT) residual_prep_k(int64_t n, int64_t rows_per_block,
                                                          const double* __restrict__ b, const double* __restrict__ Ax,
                                                          const T* __restrict__ dinv, T* __restrict__ rs, Partials part,
                                                          int write_rs, T* __restrict__ u_out) {
  __shared__ double sm[PCG_NT / 64];
  const int G = gridDim.x;
  const int64_t r0 = block_first_row(rows_per_block);
  const int64_t r1 = min(n, r0 + rows_per_block);
  double prt = 0, pg = 0, pr = 0, pb = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += PCG_NT) {
    const double bi = b[i];
    const double r = Ax ? bi - Ax[i] : bi;
    prt += r * r;
    pb += bi * bi;
    const T ri = (T)r;
    if (write_rs) rs[i] = ri;
    const T ui = dinv[i] * ri;
    if (u_out) u_out[i] = ui;
    pg += (double)ri * (double)ui;
    pr += (double)ri * (double)ri;
  }
  store_partial(pg, part.B, sm);
  store_partial(pr, part.B + G, sm);
  store_partial(prt, part.P, sm);
  store_partial(pb, part.P + G, sm);
}

// one workgroup: reduce the prep partials into the state and set the tolerances
// check_scale (a scaled FP32 solve's first call): when scale_dia_k flagged the operator, stop before
// the first iteration with ST_UNSCALED instead
__global__ void __launch_bounds__(PCG_NT) finalize_k(PcgState* st, Partials part, int G, double tol, double atol,
                                                     int64_t maxiter, double replace2, int first, int restart,
                                                     int single_reduction, int check_scale = 0) {
  __shared__ double sm[PCG_NT / 64];
  const double rt = reduce_partials(part.P, G, sm);
  const double bb = reduce_partials(part.P + G, G, sm);
  const double rr = reduce_partials(part.B + G, G, sm);
  if (threadIdx.x != 0) return;
  st->rr_true = rt;
  st->bb = bb;
  st->rr0 = rr;
  if (!restart) return;
  st->atol2 = fmax(tol * tol * bb, atol * atol);
  st->maxiter = maxiter;
  st->done = 0;
  st->done1 = 0;
  st->status = ST_RUNNING;
  // two-reduction iterations: any deferred x update was applied at the stop; single-reduction
  // iterations update x at once and need alpha_{k-1} in the next alpha
  if (!single_reduction) st->alpha = 0.0;
  if (first) {
    st->k = 0;
    st->k_next = 0;
    st->gamma_next = 0.0;
    st->alpha_next = 0.0;
    st->outer_done = 0;
    st->a_pend[0] = 0.0;
    st->a_pend[1] = 0.0;
  }
  st->rr_replace = replace2 > 0.0 ? replace2 * rr : -1.0;
  if (check_scale && st->scale_bad) {
    st->done = 1;
    st->done1 = 1;
    st->status = ST_UNSCALED;
    st->k_stop = 0;
    st->outer_done = 1;
  }
}

template <typename T>
__global__ void zero_k(T* __restrict__ a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = T(0);
}

__global__ void accumulate_x_k(int64_t n, double* __restrict__ x, float* __restrict__ xs) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += (double)xs[i];
    xs[i] = 0.0f;
  }
}

__global__ void cast_f32_vec_k(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------------------------
// Jacobi-scaled FP32 inner operator: with S = diag(sqrt(dinv)), Jacobi-PCG on A x = b is plain CG
// on (S A S) y = S b with x = S y, iterate for iterate.  S A S has a unit main diagonal, so the
// inner iterations stream neither dinv nor the main diagonal (52 instead of 64 bytes per row and
// iteration in FP32).
// ---------------------------------------------------------------------------------------------
// sdia[d][i] = (float)(A[i, i+o_d] * s_i * s_{i+o_d}) for the stored upper diagonals d >= 1; *bad is
// set when some (float)(A_ii dinv_i) != 1 (then the unscaled inner iteration is used)
__global__ void scale_dia_k(int64_t n, int ndiag, DiaOffsets off, const double* __restrict__ dia,
                            const double* __restrict__ dinv, float* __restrict__ sdia, int* __restrict__ bad) {
  const int64_t ld = dia_ld(n);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if ((float)(dia[i] * dinv[i]) != 1.0f) atomicOr(bad, 1);
    const double si = sqrt(dinv[i]);
    for (int d = 1; d < ndiag; ++d) {
      const int64_t o = off.o[d];
      sdia[(int64_t)d * ld + i] = i + o < n ? (float)(dia[(int64_t)d * ld + i] * si * sqrt(dinv[i + o])) : 0.0f;
    }
  }
}

// r64 = b - Ax (Ax nullable = 0), rs[i * rs_stride] = (float)(sqrt(dinv) r64); partials of rs.rs
// (both B slots, and c_rr when given: the r.r slot a one-kernel launch reads) and of ||r64||^2, b.b
// (P)
__global__ void __launch_bounds__(PCG_NT) residual_prep_scaled_k(int64_t n, int64_t rows_per_block,
                                                                 const double* __restrict__ b,
                                                                 const double* __restrict__ Ax,
                                                                 const double* __restrict__ dinv,
                                                                 float* __restrict__ rs, int rs_stride,
                                                                 Partials part, double* __restrict__ c_rr) {
  __shared__ double sm[PCG_NT / 64];
  const int G = gridDim.x;
  const int64_t r0 = block_first_row(rows_per_block);
  const int64_t r1 = min(n, r0 + rows_per_block);
  double prt = 0, pr = 0, pb = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += PCG_NT) {
    const double bi = b[i];
    const double r = Ax ? bi - Ax[i] : bi;
    prt += r * r;
    pb += bi * bi;
    const float ri = (float)(r * sqrt(dinv[i]));
    rs[i * rs_stride] = ri;
    pr += (double)ri * (double)ri;
  }
  store_partial(pr, part.B, sm);
  store_partial(pr, part.B + G, sm);
  store_partial(prt, part.P, sm);
  store_partial(pb, part.P + G, sm);
  if (c_rr) store_partial(pr, c_rr, sm);
}

// x64 += sqrt(dinv) xs, xs = 0
__global__ void accumulate_scaled_k(int64_t n, double* __restrict__ x, float* __restrict__ xs,
                                    const double* __restrict__ dinv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += sqrt(dinv[i]) * (double)xs[i];
    xs[i] = 0.0f;
  }
}

// The outer step of the Jacobi-scaled mixed solve, on the device so that no host round trip sits
// between replacement cycles.  After the FP64 residual of the accumulated x was prepared: reduce
// ||r64||^2, b.b and the new FP32 r.r, then either stop (converged on the true residual; the inner
// loop hit the cap or a non-finite value; or it made no progress with an exactly zero FP32
// residual) -- outer_done set, the stop flags left set, so the next inner launches exit at once --
// or restart the inner loop: the next replacement threshold (aimed at the remaining reduction when
// that is smaller than a whole cycle, with margin) and, for one-kernel iterations, the rerun of
// launch kr (special 2; 1 when the stop came at launch 0).  outer_decide runs on thread 0.
__device__ __forceinline__ void outer_decide(PcgState* st, double rt, double bb, double rr, double tol, double atol,
                                             double repl2, int64_t k_start, int64_t kr, int special,
                                             int64_t maxiter) {
  st->rr_true = rt;
  st->bb = bb;
  st->rr0 = rr;
  const int inner = st->status;
  const double atol2 = fmax(tol * tol * bb, atol * atol);
  if (!(rt > atol2)) {
    st->status = isfinite(rt) ? ST_CONVERGED : ST_NONFINITE;
    st->outer_done = 1;
    return;
  }
  if (inner == ST_MAXITER || inner == ST_NONFINITE || !isfinite(rt)) {
    if (!isfinite(rt)) st->status = ST_NONFINITE;
    st->outer_done = 1;
    return;
  }
  if (inner == ST_CONVERGED && st->k_stop == k_start) {
    // the FP32 residual is exactly zero while the FP64 one is above the tolerance
    st->status = ST_MAXITER;
    st->outer_done = 1;
    return;
  }
  const double need = atol2 / rt;  // remaining reduction of r.r
  st->rr_replace = need > repl2 ? 0.5 * need * rr : repl2 * rr;
  st->maxiter = maxiter;
  st->alpha = 0.0;
  if (special) {
    st->k_special = kr;
    st->special = special;
  }
  st->status = ST_RUNNING;
  st->done = 0;
  st->done1 = 0;
}

__global__ void __launch_bounds__(PCG_NT) outer_k(PcgState* st, Partials part, int G, double tol, double atol,
                                                  double repl2, int64_t k_start, int64_t kr, int special,
                                                  int64_t maxiter, PcgState* post) {
  __shared__ double sm[PCG_NT / 64];
  const double rt = reduce_partials(part.P, G, sm);
  const double bb = reduce_partials(part.P + G, G, sm);
  const double rr = reduce_partials(part.B + G, G, sm);
  if (threadIdx.x == 0) outer_decide(st, rt, bb, rr, tol, atol, repl2, k_start, kr, special, maxiter);
  // the decision for the host, which reads it while the next inner loop's first launches run
  post_state(st, post);
}

// ---------------------------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------------------------
static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Ws {
  PcgState* st;
  Partials part;
  void* r;
  void* p0;
  void* p1;
  void* q;
  void* s;       // single-reduction iterations: A p
  void* xs;      // FP32 accumulator (mixed)
  void* q2;      // one-kernel iterations: the second q buffer (scaled FP32 operator only)
  float* dinv32; // mixed
  double* ax;    // FP64 A x
  void* ext;     // the SpMV's own scratch (EBE product buffer; nullable)
  float* sdia;   // mixed on DIA: the Jacobi-scaled FP32 operator (nullable: no room)
};

// sdia_elems: FP32 entries of the Jacobi-scaled operator (0: none); ext_bytes: the SpMV's own
// scratch (an EBE operator's product buffer)
static size_t ws_layout(int64_t n, int mixed, char* base, Ws* w, int64_t sdia_elems = 0, size_t ext_bytes = 0) {
  const size_t vb = mixed ? 4 : 8;
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    void* p = base ? base + off : nullptr;
    off += a256(bytes);
    return p;
  };
  // iteration vectors with PCG_PAD_BYTES of pad before and after (16-byte loads of quads that
  // straddle either end stay inside the allocation)
  auto takev = [&](size_t bytes) -> void* {
    void* p = base ? base + off + PCG_PAD_BYTES : nullptr;
    off += a256(bytes + 2 * PCG_PAD_BYTES);
    return p;
  };
  Ws t;
  t.st = (PcgState*)take(sizeof(PcgState));
  t.part.A = (double*)take(PCG_MAX_BLOCKS * sizeof(double));
  t.part.B = (double*)take(2 * PCG_MAX_BLOCKS * sizeof(double));
  t.part.P = (double*)take(2 * PCG_MAX_BLOCKS * sizeof(double));
  t.part.B2 = (double*)take(2 * PCG_MAX_BLOCKS * sizeof(double));
  t.part.C = (double*)take(8 * PCG_MAX_BLOCKS * sizeof(double));
  t.r = takev(n * vb);
  t.p0 = takev(n * vb);
  t.p1 = takev(n * vb);
  t.q = takev(n * vb);
  t.s = takev(n * vb);
  t.xs = mixed ? takev(n * 4) : nullptr;
  t.dinv32 = mixed ? (float*)take(n * 4) : nullptr;
  t.ax = (double*)take(n * 8);
  t.ext = ext_bytes > 0 ? take(ext_bytes) : nullptr;
  t.sdia = sdia_elems > 0 ? (float*)take(sdia_elems * 4) : nullptr;
  t.q2 = sdia_elems > 0 ? takev(n * 4) : nullptr;
  if (w) *w = t;
  return off;
}

// FP32 entries of the Jacobi-scaled copy of a DIA operator (its stored diagonals)
static int64_t scaled_elems(const jno_op* op) {
  return op && op->format == JNO_FMT_DIA ? (int64_t)op->ndiag * dia_ld(op->n) : 0;
}

// grid of the iteration kernels: a multiple of 8 (XCD slabs), at most PCG_MAX_BLOCKS, and the
// rows per workgroup: a multiple of 4 so every workgroup's first row is 16-byte aligned for FP32
// and FP64, and of 12 for BSR3 (whole 3-row blocks)
static bool quad_ok(const jno_op* op);

// compute units of the current device (cached per device; 256 on MI355X)
static int cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int v = 0;
    cached[dev] = hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0 ? v : 256;
  }
  return cached[dev];
}
// The launch shape an FP32 solve of `op` runs: the register-resident one-kernel shapes (mode 5 with
// k1_vector 1 or 2) apply to the 4-diagonal operators it is built for, at 256 or 512 threads; any
// other operator (BSR3, CSR, 8-diagonal DIA) runs the previous default shape instead.
static Tuning effective_tuning(const Tuning& tu, const jno_op* op) {
  if (tu.split == MODE_ONE && tu.vec >= 1 && !(quad_ok(op) && tu.block <= 512)) return ONE_FALLBACK;
  return tu;
}
static void pcg_grid(const jno_op* op, const Tuning& tu, int* G, int64_t* rpb) {
  const int64_t n = op->n;
  if (op->format == JNO_FMT_EBE && ebe_box_bricks(op->ebe) > 0) {
    // the structured-box SpMV: one workgroup per brick (whole XCD slabs), its partial per workgroup
    const int64_t want = std::min<int64_t>(cdiv(ebe_box_bricks(op->ebe), (int64_t)N_XCD) * N_XCD, PCG_MAX_BLOCKS);
    *G = (int)want;
    *rpb = cdiv(cdiv(n, want), (int64_t)12) * 12;
    return;
  }
  if (tu.split == MODE_ONE && tu.vec >= 1) {
    // pcg_one_q: one tile of 4 block rows per workgroup where the grid cap allows (rounded up to
    // whole XCD slabs), else several equal tiles per workgroup
    // All workgroups are resident at once (at most 4 of 256 threads per CU); a grid of whole
    // multiples of the CU count leaves no CU with an extra workgroup to run after the others (776
    // workgroups at 128^3, 8 more than 3 per CU, measured 24.8 us against 22.8 at 768).
    int64_t want = cdiv(n, (int64_t)tu.block * 4);
    const int64_t ncu = cu_count();
    if (want >= ncu && ncu % N_XCD == 0) want = cdiv(want, ncu) * ncu;
    if (want >= N_XCD) want = cdiv(want, (int64_t)N_XCD) * N_XCD;
    const int64_t cap = std::min<int64_t>(tu.grid_cap, QRED_G);
    if (want > cap) want = std::max<int64_t>(N_XCD, (cap / N_XCD) * N_XCD);
    if (want < 1) want = 1;
    *G = (int)want;
    // rows per workgroup = nt tiles of tl rows exactly (tl a multiple of 4), so that the kernel can
    // deal the tiles of an XCD slab round-robin
    const int64_t r = cdiv(cdiv(n, want), (int64_t)4) * 4;
    const int64_t nt = cdiv(r, (int64_t)tu.block * 4);
    *rpb = nt * (cdiv(cdiv(r, nt), (int64_t)4) * 4);
    return;
  }
  const int64_t lane_rows = op->format == JNO_FMT_DIA || op->format == JNO_FMT_EBE ? n
                            : op->format == JNO_FMT_BSR3                          ? (n / 3) * op->lanes
                                                                                  : n * op->lanes;
  int64_t want = cdiv(lane_rows, (int64_t)tu.block * 2);
  if (want < 1) want = 1;
  if (want > tu.grid_cap) want = tu.grid_cap;
  if (want >= N_XCD) want = (want / N_XCD) * N_XCD;
  *G = (int)want;
  // a multiple of 48 for the one-kernel pcg_one, so that 1 to 4 equal tiles divide it (its tiles
  // are then dealt round-robin within the XCD slabs)
  const int64_t mult = op->format == JNO_FMT_BSR3 || op->format == JNO_FMT_EBE ? 12 : tu.split == MODE_ONE ? 48 : 4;
  *rpb = cdiv(cdiv(n, want), mult) * mult;
}

struct HostPoll {
  PcgState* h = nullptr;  // pinned, 2 slots (filled by copies)
  hipEvent_t ev[2] = {nullptr, nullptr};
  // pinned coherent slots the one-kernel iterations write themselves (post_state: system-scope
  // stores that bypass the GPU caches), and their events, which need no system-scope fence, only
  // the completion of the launch.  Slot 2 holds outer_k's decision (scaled mixed solves).
  PcgState* hk = nullptr;
  hipEvent_t evk[3] = {nullptr, nullptr, nullptr};
  bool ok = false;
  int init() {
    if (ok) return JNO_OK;
    JNO_HIP(hipHostMalloc((void**)&h, 2 * sizeof(PcgState), hipHostMallocDefault));
    JNO_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    JNO_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    JNO_HIP(hipHostMalloc((void**)&hk, 3 * sizeof(PcgState), hipHostMallocCoherent));
    for (int i = 0; i < 3; ++i)
      JNO_HIP(hipEventCreateWithFlags(&evk[i], hipEventDisableTiming | hipEventDisableSystemFence));
    ok = true;
    return JNO_OK;
  }
};
static thread_local HostPoll g_poll;

// How a solve reads the device-side state back.  Pinned (the default): into the thread's pinned
// slots with an event each, one chunk behind the launches, so the queue never drains; one-kernel
// iterations write their slots themselves (kslots), other iterations get a copy per poll.  Sync
// (the allocation-free entry points): into host memory of the solve itself, read synchronously, so
// nothing is allocated or created; the queue drains at each poll.  The control flow (and hence the
// launches and results) is the same in all cases.
struct Poller {
  bool sync = false;
  PcgState local[2];
  PcgState* slots() { return sync ? local : g_poll.h; }
  PcgState* kslots() { return g_poll.hk; }  // written by the launches (pinned poller only)
  int init() {
    if (sync) return JNO_OK;
    JNO_REQUIRE(g_poll.init() == JNO_OK, JNO_EHIP, "pcg: pinned state allocation failed");
    return JNO_OK;
  }
  // by_kernel: the chunk's last launch writes the slot itself (post_state); only the event is needed
  int post(int slot, const PcgState* d, hipStream_t s, bool by_kernel = false) {
    if (by_kernel && !sync) {
      JNO_HIP(hipEventRecord(g_poll.evk[slot], s));
      return JNO_OK;
    }
    JNO_HIP(hipMemcpyAsync(&slots()[slot], d, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    if (sync) {
      JNO_HIP(hipStreamSynchronize(s));
    } else {
      JNO_HIP(hipEventRecord(g_poll.ev[slot], s));
    }
    return JNO_OK;
  }
  int wait(int slot, bool by_kernel = false) {
    if (!sync) JNO_HIP(hipEventSynchronize(by_kernel ? g_poll.evk[slot] : g_poll.ev[slot]));
    return JNO_OK;
  }
  int read(const PcgState* d, hipStream_t s, PcgState* out) {
    JNO_HIP(hipMemcpyAsync(&slots()[0], d, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    JNO_HIP(hipStreamSynchronize(s));
    *out = slots()[0];
    return JNO_OK;
  }
};


// The SpMV-carrying kernel of one iteration: K1 (MODE_FUSED), KS (MODE_SPLIT) or KB
// (MODE_SINGLE), by format, DIA diagonal count (4: the 3-D Kuhn lattice; 3: the 2-D triangle
// lattice) and CSR lanes per row
template <int NT, int MODE, typename T, typename V>
static int launch_spmv_kernel(const jno_op* op, const PcgArgs<T, V>& a, int grid, hipStream_t s) {
  if (op->format == JNO_FMT_EBE) {
    JNO_REQUIRE(MODE != MODE_FUSED, JNO_EINVAL, "pcg: EBE operators run the split or single-reduction iteration");
    PcgStateView sv;
    sv.st = a.st;
    return ebe_spmv_pcg<T>(op, a.p_new, a.q, static_cast<T*>(a.ext), a.part.A, sv, grid, a.rows_per_block, NT,
                           MODE == MODE_SINGLE, s);
  }
  if (op->format == JNO_FMT_BSR3) {
    JNO_REQUIRE(MODE != MODE_FUSED, JNO_EINVAL, "pcg: BSR3 operators run the split or single-reduction iteration");
    constexpr bool CG = MODE == MODE_SINGLE;
    switch (op->lanes) {
      case 1: pcg_ks_bsr3<NT, 1, CG, T, V><<<grid, NT, 0, s>>>(a); break;
      case 2: pcg_ks_bsr3<NT, 2, CG, T, V><<<grid, NT, 0, s>>>(a); break;
      case 4: pcg_ks_bsr3<NT, 4, CG, T, V><<<grid, NT, 0, s>>>(a); break;
      case 8: pcg_ks_bsr3<NT, 8, CG, T, V><<<grid, NT, 0, s>>>(a); break;
      case 16: pcg_ks_bsr3<NT, 16, CG, T, V><<<grid, NT, 0, s>>>(a); break;
      case 32: pcg_ks_bsr3<NT, 32, CG, T, V><<<grid, NT, 0, s>>>(a); break;
      case 64: pcg_ks_bsr3<NT, 64, CG, T, V><<<grid, NT, 0, s>>>(a); break;
      default: JNO_REQUIRE(false, JNO_EINVAL, "pcg: lanes per block row must be a power of two in 1..64");
    }
    JNO_LAUNCH_CHECK();
    return JNO_OK;
  }
#define JNO_K(L_, DIA_, ND_)                                                          \
  do {                                                                                \
    if (MODE == MODE_FUSED)                                                           \
      pcg_k1<NT, L_, DIA_, ND_, T, V><<<grid, NT, 0, s>>>(a);                         \
    else                                                                              \
      pcg_ks<NT, L_, DIA_, ND_, MODE == MODE_SINGLE, T, V><<<grid, NT, 0, s>>>(a);    \
  } while (0)
  if (op->format == JNO_FMT_DIA) {
    switch (op->ndiag) {
      case 4: JNO_K(1, true, 4); break;
      case 3: JNO_K(1, true, 3); break;
      default: JNO_K(1, true, 0); break;
    }
  } else {
    switch (op->lanes) {
      case 1: JNO_K(1, false, 0); break;
      case 2: JNO_K(2, false, 0); break;
      case 4: JNO_K(4, false, 0); break;
      case 8: JNO_K(8, false, 0); break;
      case 16: JNO_K(16, false, 0); break;
      case 32: JNO_K(32, false, 0); break;
      case 64: JNO_K(64, false, 0); break;
      default: JNO_REQUIRE(false, JNO_EINVAL, "pcg: lanes per row must be a power of two in 1..64");
    }
  }
#undef JNO_K
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

// KB of the single-reduction iteration: w (in q) = A u, partial u.w (the SpMV kernel reads p_new)
template <typename T, typename V>
static PcgArgs<T, V> kb_args(const PcgArgs<T, V>& a) {
  PcgArgs<T, V> b = a;
  b.p_new = a.u;
  return b;
}

template <int NT, typename T, typename V>
static int launch_kb_nt(const jno_op* op, const PcgArgs<T, V>& a, int grid, hipStream_t s) {
  return launch_spmv_kernel<NT, MODE_SINGLE, T, V>(op, kb_args(a), grid, s);
}

template <typename T, typename V>
static int launch_kb(const jno_op* op, const PcgArgs<T, V>& a, int grid, hipStream_t s) {
  switch (a.block) {
    case 256: return launch_kb_nt<256, T, V>(op, a, grid, s);
    case 512: return launch_kb_nt<512, T, V>(op, a, grid, s);
    default: return launch_kb_nt<1024, T, V>(op, a, grid, s);
  }
}

// One iteration: K1 K2 | K0 KS K2 | KA KB.  ev (optional) receives three events: after K0 or KA
// (at the start for K1 K2), after the SpMV kernel, after K2 (right after KB for KA KB).
// Near off-diagonals of a DIA operator for the LDS-staged kernels: the count of stored offsets
// (after the main diagonal) up to LDS_HALO, or -1 when the offsets do not ascend.
static int near_count(const jno_op* op) {
  int nn = 0;
  for (int d = 1; d < op->ndiag; ++d) {
    if (op->offsets[d] <= op->offsets[d - 1]) return -1;
    if (op->offsets[d] <= LDS_HALO) nn = d;
  }
  return nn;
}

static bool lds_ok(const jno_op* op) {
  if (op->format != JNO_FMT_DIA || op->ndiag < 1 || op->ndiag > 8) return false;
  const int nn = near_count(op);
  // instantiated diagonal sets: up to 4 diagonals (Poisson stiffness on Kuhn lattices and their
  // sub-sets), and 8 diagonals with 3 near ones inside one far window (mass-carrying operators on
  // the same lattices, e.g. the transient step operator, nx <= 254) or with 1 near one
  if (nn < 0) return false;
  if (op->ndiag <= 4) return true;
  if (op->ndiag != 8) return false;
  return nn == 1 || (nn == 3 && op->offsets[7] - op->offsets[4] <= LDS_FAR_SPAN);
}

// the register-resident one-kernel iteration: 4 ascending diagonals with 2 or 3 near ones (Kuhn
// lattices up to 254 cells along the first axis)
static bool quad_ok(const jno_op* op) {
  if (op->format != JNO_FMT_DIA || op->ndiag != 4) return false;
  const int nn = near_count(op);
  return nn == 2 || nn == 3;
}

// pcg_one (one = true) or pcg_k1_lds by diagonal count and near count (FP32 iterations only)
template <int NT, typename T, typename V>
static int launch_lds(bool one, const jno_op* op, const PcgArgs<T, V>& a, int grid, hipStream_t s) {
  JNO_REQUIRE(lds_ok(op), JNO_EINVAL, "pcg: no LDS-staged kernel for this DIA diagonal set");
  if constexpr (std::is_same<T, float>::value) {
    const int key = op->ndiag * 8 + near_count(op);
#define JNO_LDS_CASE(ND_, NN_, FW_)                                \
  case ND_ * 8 + NN_:                                              \
    if (one)                                                       \
      pcg_one<NT, ND_, NN_, T, V, FW_><<<grid, NT, 0, s>>>(a);     \
    else                                                           \
      pcg_k1_lds<NT, ND_, NN_, T, V><<<grid, NT, 0, s>>>(a);       \
    break;
    switch (key) {
      JNO_LDS_CASE(1, 0, false)
      JNO_LDS_CASE(2, 0, false)
      JNO_LDS_CASE(2, 1, false)
      JNO_LDS_CASE(3, 0, false)
      JNO_LDS_CASE(3, 1, false)
      JNO_LDS_CASE(3, 2, false)
      JNO_LDS_CASE(4, 0, false)
      JNO_LDS_CASE(4, 1, false)
      JNO_LDS_CASE(4, 2, false)
      JNO_LDS_CASE(4, 3, false)
      JNO_LDS_CASE(8, 1, false)
      // the transient step operator M + theta dt A: x every other launch, as in pcg_one_p
      case 8 * 8 + 3:
        if (!one)
          pcg_k1_lds<NT, 8, 3, T, V><<<grid, NT, 0, s>>>(a);
        else if (a.kk & 1)
          pcg_one<NT, 8, 3, T, V, true, XM_FLUSH><<<grid, NT, 0, s>>>(a);
        else
          pcg_one<NT, 8, 3, T, V, true, XM_DEFER><<<grid, NT, 0, s>>>(a);
        break;
      default: JNO_REQUIRE(false, JNO_EINVAL, "pcg: no LDS-staged kernel for this diagonal set");
    }
#undef JNO_LDS_CASE
    JNO_LAUNCH_CHECK();
    return JNO_OK;
  } else {
    JNO_REQUIRE(false, JNO_EINVAL, "pcg: the LDS-staged kernels run FP32 iterations");
  }
}

// pcg_one_q by near count (4-diagonal operators, FP32 iterations)
template <int NT, typename T, typename V>
static int launch_quad(const jno_op* op, const PcgArgs<T, V>& a, int grid, hipStream_t s) {
  JNO_REQUIRE(quad_ok(op), JNO_EINVAL, "pcg: no register-resident kernel for this DIA diagonal set");
  JNO_REQUIRE(grid <= QRED_G, JNO_EINVAL, "pcg: %d workgroups exceed the head reduction of %d", grid, QRED_G);
  if constexpr (std::is_same<T, float>::value && std::is_same<V, float>::value && NT <= 512) {
    const int key = near_count(op) * 4 + a.vec;
    switch (key) {
      case 9: pcg_one_q<NT, 4, 2><<<grid, NT, 0, s>>>(a); break;
      case 10: pcg_one_q<NT, 4, 2, true><<<grid, NT, 0, s>>>(a); break;
      // the pipelined kernel reads the partial sums as pairs (an even grid); an odd one runs LF
      // x every other launch: even launches defer their x term, odd ones flush both
      case 11:
        if (grid % 2) pcg_one_q<NT, 4, 2, true><<<grid, NT, 0, s>>>(a);
        else if (a.kk & 1) pcg_one_p<NT, 4, 2, XM_FLUSH><<<grid, NT, 0, s>>>(a);
        else pcg_one_p<NT, 4, 2, XM_DEFER><<<grid, NT, 0, s>>>(a);
        break;
      case 13: pcg_one_q<NT, 4, 3><<<grid, NT, 0, s>>>(a); break;
      case 14: pcg_one_q<NT, 4, 3, true><<<grid, NT, 0, s>>>(a); break;
      default:
        if (grid % 2) pcg_one_q<NT, 4, 3, true><<<grid, NT, 0, s>>>(a);
        else if (a.kk & 1) pcg_one_p<NT, 4, 3, XM_FLUSH><<<grid, NT, 0, s>>>(a);
        else pcg_one_p<NT, 4, 3, XM_DEFER><<<grid, NT, 0, s>>>(a);
        break;
    }
    JNO_LAUNCH_CHECK();
    return JNO_OK;
  } else {
    JNO_REQUIRE(false, JNO_EINVAL, "pcg: the register-resident kernel runs FP32 iterations, 256 or 512 threads");
  }
}

template <int NT, typename T, typename V>
static int launch_iter_nt(const jno_op* op, PcgArgs<T, V>& a, int grid, hipStream_t s, hipEvent_t* ev) {
  int rc;
  if (a.unit && a.split == MODE_ONE) {
    if (ev) JNO_HIP(hipEventRecord(ev[0], s));
    if (a.vec >= 1 && NT <= 512 && quad_ok(op))
      rc = launch_quad<NT, T, V>(op, a, grid, s);
    else
      rc = launch_lds<NT, T, V>(true, op, a, grid, s);
    if (rc) return rc;
    if (ev) JNO_HIP(hipEventRecord(ev[1], s));
    if (ev) JNO_HIP(hipEventRecord(ev[2], s));
    return JNO_OK;
  }
  if (a.unit && a.split == MODE_FUSED) {
    // Jacobi-scaled DIA operator, two kernels: K1 and K2 without dinv, unit main diagonal
    if (ev) JNO_HIP(hipEventRecord(ev[0], s));
    if (a.vec == 3 && lds_ok(op)) {
      rc = launch_lds<NT, T, V>(false, op, a, grid, s);
      if (rc) return rc;
    } else {
      switch (op->ndiag) {
        case 4: pcg_k1<NT, 1, true, 4, T, V, false><<<grid, NT, 0, s>>>(a); break;
        case 3: pcg_k1<NT, 1, true, 3, T, V, false><<<grid, NT, 0, s>>>(a); break;
        default: pcg_k1<NT, 1, true, 0, T, V, false><<<grid, NT, 0, s>>>(a); break;
      }
    }
    JNO_LAUNCH_CHECK();
    if (ev) JNO_HIP(hipEventRecord(ev[1], s));
    pcg_k2<NT, T, V, false><<<grid, NT, 0, s>>>(a);
    JNO_LAUNCH_CHECK();
    if (ev) JNO_HIP(hipEventRecord(ev[2], s));
    return JNO_OK;
  }
  if (a.unit) {
    // Jacobi-scaled DIA operator, split iteration: K0 and K2 without dinv, KS with the unit main
    // diagonal
    pcg_k0<NT, T, V, false><<<grid, NT, 0, s>>>(a);
    JNO_LAUNCH_CHECK();
    if (ev) JNO_HIP(hipEventRecord(ev[0], s));
    switch (op->ndiag) {
      case 4: pcg_ks<NT, 1, true, 4, false, T, V, true><<<grid, NT, 0, s>>>(a); break;
      case 3: pcg_ks<NT, 1, true, 3, false, T, V, true><<<grid, NT, 0, s>>>(a); break;
      default: pcg_ks<NT, 1, true, 0, false, T, V, true><<<grid, NT, 0, s>>>(a); break;
    }
    JNO_LAUNCH_CHECK();
    if (ev) JNO_HIP(hipEventRecord(ev[1], s));
    pcg_k2<NT, T, V, false><<<grid, NT, 0, s>>>(a);
    JNO_LAUNCH_CHECK();
    if (ev) JNO_HIP(hipEventRecord(ev[2], s));
    return JNO_OK;
  }
  if (a.split == MODE_SINGLE) {
    pcg_ka<NT, T, V><<<grid, NT, 0, s>>>(a);
    JNO_LAUNCH_CHECK();
    if (ev) JNO_HIP(hipEventRecord(ev[0], s));
    rc = launch_kb_nt<NT, T, V>(op, a, grid, s);
    if (rc) return rc;
    if (ev) JNO_HIP(hipEventRecord(ev[1], s));
    if (ev) JNO_HIP(hipEventRecord(ev[2], s));
    return JNO_OK;
  }
  if (a.split == MODE_SPLIT) {
    pcg_k0<NT, T, V><<<grid, NT, 0, s>>>(a);
    JNO_LAUNCH_CHECK();
    if (ev) JNO_HIP(hipEventRecord(ev[0], s));
    rc = launch_spmv_kernel<NT, MODE_SPLIT, T, V>(op, a, grid, s);
  } else {
    if (ev) JNO_HIP(hipEventRecord(ev[0], s));
    rc = launch_spmv_kernel<NT, MODE_FUSED, T, V>(op, a, grid, s);
  }
  if (rc) return rc;
  if (ev) JNO_HIP(hipEventRecord(ev[1], s));
  pcg_k2<NT, T, V><<<grid, NT, 0, s>>>(a);
  JNO_LAUNCH_CHECK();
  if (ev) JNO_HIP(hipEventRecord(ev[2], s));
  return JNO_OK;
}

template <typename T, typename V>
static int launch_iter(const jno_op* op, PcgArgs<T, V>& a, int grid, hipStream_t s, hipEvent_t* ev) {
  switch (a.block) {
    case 256: return launch_iter_nt<256, T, V>(op, a, grid, s, ev);
    case 512: return launch_iter_nt<512, T, V>(op, a, grid, s, ev);
    default: return launch_iter_nt<1024, T, V>(op, a, grid, s, ev);
  }
}

// Optional live kernel timing: on every stride-th iteration, HIP events are recorded on the
// solve's stream around its K1 and K2 (E0 K1 E1 K2 E2); the sampled iterations that did work are
// summed.  One-kernel iterations are timed as windows of PROF_WIN consecutive launches between two
// events (the launches run back to back, so the window's length over PROF_WIN is the average
// launch duration, with the events' own cost spread over the window).
constexpr int PROF_WIN = 4;
struct KernelProf {
  bool on = false;
  int stride = 8;  // time every stride-th iteration (events perturb the queue very little then)
  std::vector<hipEvent_t> ev;
  double ms[3] = {0, 0, 0};  // K0 (split iterations only), SpMV kernel (K1 or KS), K2
  int64_t n = 0;             // sampled iterations
  int ensure(size_t n) {
    while (ev.size() < n) {
      hipEvent_t e;
      // timing events only: without the system-scope fence a record does not flush the caches
      // between the kernels it brackets
      JNO_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
      ev.push_back(e);
    }
    return JNO_OK;
  }
};
static thread_local KernelProf g_prof;  // per thread: times the solves of the thread that enabled it

// Runs K1/K2 chunks until the device reports a stop; returns the final state in *out.
// Pacing of an inner loop's launches (one-kernel iterations).  Launches enqueued after the stop
// exit at once but still cost a launch each (about 3 us), and the host learns of the stop one poll
// late, so a fixed chunk wastes up to two chunks per stop.  The first two chunks are small (an
// early stop, e.g. the outer step's, is seen after few launches), and later chunk sizes follow the
// observed convergence rate of r.r: the launches still needed to reach the stop threshold, less
// those already queued, plus a small margin.  log_rate carries the rate from one inner loop to the
// next.
struct InnerPace {
  double log_rate = 0.0;  // ln(r.r) per iteration, 0 = unknown
};
constexpr int PACE_FIRST = 4;
// launches enqueued beyond the predicted stop (1 against 3: 128^3 solve 14.37 -> 14.34 ms, heat
// step 2.112 -> 2.095 ms)
constexpr int PACE_MARGIN = 1;

template <typename T, typename V>
static int run_inner(Poller& poll, const jno_op* op, PcgArgs<T, V> a, T* pbuf[2], int64_t k0, int chunk, int grid,
                     hipStream_t s, PcgState* out, int64_t maxiter, T* rbuf[2] = nullptr, T* qbuf[2] = nullptr,
                     InnerPace* pace = nullptr, bool outer_posted = false) {
  int rc0 = poll.init();
  if (rc0) return rc0;
  const bool prof = g_prof.on && !poll.sync;  // the allocation-free path creates no events
  int64_t k = k0;
  int slot = 0, pending = 0, round = 0;
  int64_t upto[2] = {k0, k0};  // launches enqueued before each outstanding poll
  int64_t k_ref = -1;          // first sampled launch of this inner loop and its r.r
  double rr_ref = 0.0;
  int64_t k_last = -1;  // the previous sample
  double rr_last = 0.0;
  size_t nev = 0;
  std::vector<int64_t> sampled;  // iteration index of each (start, K0, SpMV, K2) event quadruple
                                 // (one-kernel iterations: first launch of each event pair's window)
  int64_t win = -1;              // first launch of the open timing window (one-kernel iterations)
  int n = pace ? std::min(chunk, PACE_FIRST) : chunk;
  bool polled = false;
  const bool kpost = rbuf && !poll.sync;  // one-kernel launches write the pinned poll slots
  for (;;) {
    for (int i = 0; i < n; ++i, ++k) {
      a.p_old = pbuf[k & 1];
      a.p_new = pbuf[(k + 1) & 1];
      if (rbuf) {  // one-kernel iterations: launch index and the r and q halves
        a.kk = k;
        a.r = rbuf[k & 1];
        a.r_out = rbuf[(k + 1) & 1];
        a.q = qbuf[k & 1];
        a.q_out = qbuf[(k + 1) & 1];
        // the chunk's last launch writes the poll slot itself (pinned poller)
        a.post = kpost && i == n - 1 ? &poll.kslots()[slot] : nullptr;
      }
      if (prof && rbuf) {
        if (k % g_prof.stride == 0) {
          if (g_prof.ensure(nev + 2) != JNO_OK) return JNO_EHIP;
          JNO_HIP(hipEventRecord(g_prof.ev[nev], s));
          win = k;
        }
        int rc = launch_iter<T, V>(op, a, grid, s, nullptr);
        if (rc) return rc;
        if (win >= 0 && k == win + PROF_WIN - 1) {
          JNO_HIP(hipEventRecord(g_prof.ev[nev + 1], s));
          sampled.push_back(win);
          nev += 2;
          win = -1;
        }
      } else if (prof && k % g_prof.stride == 0) {
        if (g_prof.ensure(nev + 4) != JNO_OK) return JNO_EHIP;
        JNO_HIP(hipEventRecord(g_prof.ev[nev], s));
        int rc = launch_iter<T, V>(op, a, grid, s, &g_prof.ev[nev + 1]);
        nev += 4;
        sampled.push_back(k);
        if (rc) return rc;
      } else {
        int rc = launch_iter<T, V>(op, a, grid, s, nullptr);
        if (rc) return rc;
      }
    }
    int rcp = poll.post(slot, a.st, s, kpost);
    if (rcp) return rcp;
    // the device stops an inner loop within maxiter launches; a host that has not seen the stop
    // two chunks past that reads a state the device did not write (never expected)
    JNO_REQUIRE(k - k0 - 2 * (int64_t)chunk - 8 <= maxiter, JNO_EHIP,
                "pcg: no stop reported by the device after %lld launches", (long long)(k - k0));
    upto[slot] = k;
    ++round;
    n = chunk;
    if (outer_posted && round == 1) {
      // the outer step before this loop posted its decision (slot 2): read it while the first
      // chunk runs, so that a stopped solve enqueues no second chunk
      rcp = poll.wait(2, true);
      if (rcp) return rcp;
      if (poll.kslots()[2].outer_done) {
        *out = poll.kslots()[2];
        polled = true;
        break;
      }
    }
    if (pending) {
      const int prev = slot ^ 1;
      rcp = poll.wait(prev, kpost);
      if (rcp) return rcp;
      const PcgState& h = kpost ? poll.kslots()[prev] : poll.slots()[prev];
      if (h.done1) {
        // one-kernel iterations: the launches after the stop exit before writing any state, so
        // the polled state is final and the queue need not drain before the caller's next commands
        if (rbuf && !poll.sync) {
          *out = h;
          polled = true;
        }
        break;
      }
      if (pace && rbuf) {
        // r.r read by the last launch before that poll (a resuming launch does not record it, but
        // the first chunk holds more than one launch)
        const int64_t kl = upto[prev] - 1;
        const double rr = h.g_hist[kl & 1];
        if (k_ref < 0) {
          k_ref = kl;
          rr_ref = rr;
        } else if (kl > k_ref && rr > 0.0 && rr_ref > 0.0) {
          pace->log_rate = std::log(rr / rr_ref) / (double)(kl - k_ref);
        }
        // CG converges faster as a cycle proceeds: the rate since the previous poll, when it is
        // the faster one, predicts the stop better than the cycle's average
        double rate = pace->log_rate;
        if (k_last >= 0 && kl > k_last && rr > 0.0 && rr_last > 0.0)
          rate = std::min(rate, std::log(rr / rr_last) / (double)(kl - k_last));
        k_last = kl;
        rr_last = rr;
        const double target = std::max(h.atol2, h.rr_replace);
        if (rate < -1e-3 && rr > 0.0 && target > 0.0) {
          const double need = rr > target ? std::log(target / rr) / rate : 0.0;
          const int64_t queued = k - upto[prev];
          const double want = std::ceil(need) - (double)queued + PACE_MARGIN;
          n = (int)std::max(1.0, std::min((double)chunk, want));
        }
      }
    } else if (pace) {
      n = std::min(chunk, 2 * PACE_FIRST);  // the second chunk, enqueued before the first poll
    }
    pending = 1;
    slot ^= 1;
  }
  if (!polled) {
    int rcr = poll.read(a.st, s, out);
    if (rcr) return rcr;
  }
  if (prof && rbuf) {
    for (size_t t = 0; t < sampled.size(); ++t) {
      if (sampled[t] + PROF_WIN - 1 >= out->k_stop) break;  // every launch of the window did work
      float ms = 0;
      JNO_HIP(hipEventElapsedTime(&ms, g_prof.ev[2 * t], g_prof.ev[2 * t + 1]));
      g_prof.ms[1] += ms;
      g_prof.n += PROF_WIN;
    }
  } else if (prof) {
    for (size_t t = 0; t < sampled.size(); ++t) {
      if (sampled[t] >= out->k_stop) break;  // launches after the stop did no work
      for (int j = 0; j < 3; ++j) {
        float ms = 0;
        JNO_HIP(hipEventElapsedTime(&ms, g_prof.ev[4 * t + j], g_prof.ev[4 * t + j + 1]));
        g_prof.ms[j] += ms;
      }
      g_prof.n += 1;
    }
  }
  return JNO_OK;
}

static int read_state(Poller& poll, PcgState* d, PcgState* out, hipStream_t s) {
  int rc = poll.init();
  if (rc) return rc;
  return poll.read(d, s, out);
}

static int finish_status(const PcgState& f, jno_pcg_stats* stats, int64_t iters) {
  if (f.status == ST_NONFINITE) {
    set_error("pcg: non-finite value in the iteration after %lld iterations", (long long)iters);
    return JNO_ENONFINITE;
  }
  if (f.status == ST_MAXITER) {
    set_error("pcg: no convergence in %lld iterations (relative residual %.3e)", (long long)iters,
              stats ? stats->relres : -1.0);
    return JNO_ENOTCONV;
  }
  return JNO_OK;
}

static int check_op(const jno_op* op, int need32) {
  JNO_REQUIRE(op, JNO_EINVAL, "pcg: null operator");
  JNO_REQUIRE(op->format == JNO_FMT_CSR || op->format == JNO_FMT_DIA || op->format == JNO_FMT_BSR3 ||
                  op->format == JNO_FMT_EBE,
              JNO_EINVAL, "pcg: unknown operator format");
  JNO_REQUIRE(op->vals64, JNO_EINVAL, "pcg: operator has no FP64 values");
  JNO_REQUIRE(!need32 || op->vals32, JNO_EINVAL, "pcg: mixed precision needs the operator's FP32 values");
  if (op->format == JNO_FMT_EBE) JNO_REQUIRE(op->ebe && op->n == 3 * op->ebe->n_nodes, JNO_EINVAL, "pcg: bad EBE operator");
  if (op->format == JNO_FMT_CSR || op->format == JNO_FMT_BSR3)
    JNO_REQUIRE(op->rowptr && op->col, JNO_EINVAL, "pcg: CSR operator without indices");
  if (op->format == JNO_FMT_BSR3) JNO_REQUIRE(op->n % 3 == 0, JNO_EINVAL, "pcg: a BSR3 operator has 3k rows");
  if (op->format == JNO_FMT_DIA)
    JNO_REQUIRE(op->ndiag >= 1 && op->ndiag <= JNO_DIA_MAX && op->offsets[0] == 0, JNO_EINVAL,
                "pcg: bad DIA operator");
  return JNO_OK;
}

template <typename T, typename V>
static PcgArgs<T, V> make_args(const jno_op* op, const V* vals, const T* dinv, T* r, T* q, T* x, const Ws& w,
                               int64_t rpb, const Tuning& tu) {
  PcgArgs<T, V> a;
  a.n = op->n;
  a.rows_per_block = rpb;
  a.rowptr = op->rowptr;
  a.col = op->col;
  a.vals = vals;
  a.ndiag = op->ndiag;
  a.vec = tu.vec;
  a.block = tu.block;
  // the fused K1 recomputes p for gathered neighbours; BSR3 only has the plain SpMV kernel.  The
  // driver sets `unit` when it runs the scaled operator; otherwise MODE_SCALED is the split mode.
  int mode = tu.split == MODE_SCALED                                 ? MODE_SPLIT
             : (tu.split == MODE_SCALED_FUSED || tu.split == MODE_ONE) ? MODE_FUSED
                                                                       : tu.split;
  if ((op->format == JNO_FMT_BSR3 || op->format == JNO_FMT_EBE) && mode == MODE_FUSED) mode = MODE_SPLIT;
  a.split = mode;
  for (int d = 0; d < JNO_DIA_MAX; ++d) a.off.o[d] = op->offsets[d];
  a.dinv = dinv;
  a.r = r;
  a.q = q;
  a.x = x;
  a.st = w.st;
  a.part = w.part;
  a.post = nullptr;
  a.p_old = nullptr;
  a.p_new = nullptr;
  a.unit = 0;
  a.s = (T*)w.s;
  a.u = (T*)w.p1;  // single-reduction iterations keep p in place in p0; p1 holds u
  a.ext = w.ext;
  a.kk = 0;
  a.r_out = nullptr;
  a.q_out = nullptr;
  return a;
}

// A launch shape is valid when it is one jno_pcg_set_tuning accepts (the message names the field).
static int check_shape(const jno_launch_shape& t, const char* who) {
  JNO_REQUIRE(t.grid_cap >= 8 && t.grid_cap <= PCG_MAX_BLOCKS && t.grid_cap % 8 == 0, JNO_EINVAL,
              "%s: grid_cap must be a multiple of 8 in 8..%d", who, PCG_MAX_BLOCKS);
  JNO_REQUIRE(t.block == 256 || t.block == 512 || t.block == 1024, JNO_EINVAL, "%s: block must be 256, 512 or 1024",
              who);
  JNO_REQUIRE(t.split >= MODE_FUSED && t.split <= MODE_ONE, JNO_EINVAL,
              "%s: split must be 0 (K1 K2), 1 (K0 KS K2), 2 (single reduction, KA KB), 3 (K0 KS K2 on the "
              "Jacobi-scaled operator), 4 (K1 K2 on the Jacobi-scaled operator) or 5 (one kernel per iteration on "
              "the Jacobi-scaled operator)",
              who);
  JNO_REQUIRE(t.k1_vector >= 0 && t.k1_vector <= 3, JNO_EINVAL,
              "%s: k1_vector must be 0 (2 rows per thread), 1 (16-byte tiles), 2 (4 rows) or 3 (p staged in LDS, "
              "iteration mode 4); with mode 5: 0 the LDS-staged kernel, 1 and 2 the register-resident one (far "
              "loads before / after the head), 3 the register-resident one with pipelined tiles",
              who);
  return JNO_OK;
}

static int check_tuning(const jno_tuning* t, const char* who) {
  for (int p = 0; p < 2; ++p) {
    int rc = check_shape(t->iter[p], who);
    if (rc) return rc;
  }
  JNO_REQUIRE(t->chunk >= 1 && t->chunk <= 1024, JNO_EINVAL, "%s: chunk must be in 1..1024", who);
  JNO_REQUIRE(t->pacing == 0 || t->pacing == 1, JNO_EINVAL, "%s: pacing must be 0 or 1", who);
  return JNO_OK;
}

// The tuning of one solve on `op` (its own, validated, or a copy of the defaults).
static int call_tuning(const jno_op* op, CallTuning* out, const char* who) {
  if (op->tuning) {
    int rc = check_tuning(op->tuning, who);
    if (rc) return rc;
    for (int p = 0; p < 2; ++p) {
      const jno_launch_shape& t = op->tuning->iter[p];
      out->it[p] = Tuning{t.grid_cap, t.k1_vector, t.block, t.split};
    }
    out->chunk = op->tuning->chunk;
    out->pace = op->tuning->pacing;
    return JNO_OK;
  }
  std::lock_guard<std::mutex> lk(g_def_mu);
  *out = g_def;
  return JNO_OK;
}

}  // namespace jno

using namespace jno;

extern "C" size_t jno_pcg_workspace_size(int64_t n, int mixed) { return ws_layout(n, mixed ? 1 : 0, nullptr, nullptr); }

extern "C" size_t jno_pcg_workspace_size_op(const jno_op* h_op, int mixed) {
  if (!h_op) return 0;
  return ws_layout(h_op->n, mixed ? 1 : 0, nullptr, nullptr, mixed ? scaled_elems(h_op) : 0, ebe_scratch_bytes(h_op));
}

extern "C" int jno_pcg_profile(int enable) {
  g_prof.on = enable != 0;
  g_prof.ms[0] = g_prof.ms[1] = g_prof.ms[2] = 0;
  g_prof.n = 0;
  return JNO_OK;
}

extern "C" int jno_pcg_profile_read(double* h_ms, int64_t* h_count) {
  if (h_ms)
    for (int j = 0; j < 3; ++j) h_ms[j] = g_prof.ms[j];
  if (h_count) *h_count = g_prof.n;
  return JNO_OK;
}

extern "C" int jno_pcg_set_tuning(int precision, int grid_cap, int k1_vector, int block, int split) {
  JNO_REQUIRE(precision == 0 || precision == 1, JNO_EINVAL,
              "jno_pcg_set_tuning: precision must be 0 (FP64 iterations) or 1 (FP32 iterations)");
  const jno_launch_shape t = {grid_cap, k1_vector, block, split};
  int rc = check_shape(t, "jno_pcg_set_tuning");
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_def_mu);
  g_def.it[precision] = Tuning{grid_cap, k1_vector, block, split};
  return JNO_OK;
}

extern "C" int jno_pcg_get_tuning(int precision, int* grid_cap, int* k1_vector, int* block, int* split) {
  JNO_REQUIRE(precision == 0 || precision == 1, JNO_EINVAL,
              "jno_pcg_get_tuning: precision must be 0 (FP64 iterations) or 1 (FP32 iterations)");
  std::lock_guard<std::mutex> lk(g_def_mu);
  const Tuning& t = g_def.it[precision];
  if (grid_cap) *grid_cap = t.grid_cap;
  if (k1_vector) *k1_vector = t.vec;
  if (block) *block = t.block;
  if (split) *split = t.split;
  return JNO_OK;
}

extern "C" int jno_pcg_set_chunk(int chunk) {
  JNO_REQUIRE(chunk >= 1 && chunk <= 1024, JNO_EINVAL, "jno_pcg_set_chunk: chunk must be in 1..1024");
  std::lock_guard<std::mutex> lk(g_def_mu);
  g_def.chunk = chunk;
  return JNO_OK;
}

extern "C" int jno_pcg_set_pacing(int enable) {
  std::lock_guard<std::mutex> lk(g_def_mu);
  g_def.pace = enable ? 1 : 0;
  return JNO_OK;
}

extern "C" int jno_pcg_default_tuning(jno_tuning* h_out) {
  JNO_REQUIRE(h_out, JNO_EINVAL, "jno_pcg_default_tuning: null output");
  std::lock_guard<std::mutex> lk(g_def_mu);
  for (int p = 0; p < 2; ++p) {
    const Tuning& t = g_def.it[p];
    h_out->iter[p] = jno_launch_shape{t.grid_cap, t.vec, t.block, t.split};
  }
  h_out->chunk = g_def.chunk;
  h_out->pacing = g_def.pace;
  return JNO_OK;
}

extern "C" int jno_pcg_check_tuning(const jno_tuning* h_tuning) {
  JNO_REQUIRE(h_tuning, JNO_EINVAL, "jno_pcg_check_tuning: null tuning");
  return check_tuning(h_tuning, "jno_pcg_check_tuning");
}

extern "C" int jno_pcg_f64(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x, int use_x0,
                           double tol, double atol, int64_t maxiter, jno_pcg_stats* h_stats, void* d_ws,
                           size_t ws_bytes, void* stream) {
  int rc = check_op(h_op, 0);
  if (rc) return rc;
  const int64_t n = h_op->n;
  JNO_REQUIRE(d_dinv && d_b && d_x, JNO_EINVAL, "jno_pcg_f64: null vector");
  const size_t ext = ebe_scratch_bytes(h_op);
  JNO_REQUIRE(ws_bytes >= ws_layout(n, 0, nullptr, nullptr, 0, ext), JNO_EWORKSPACE, "jno_pcg_f64: workspace too small");
  JNO_REQUIRE(maxiter >= 0, JNO_EINVAL, "jno_pcg_f64: maxiter must be >= 0");
  CallTuning ct;
  rc = call_tuning(h_op, &ct, "jno_pcg_f64");
  if (rc) return rc;
  Poller poll;
  hipStream_t s = as_stream(stream);
  Ws w;
  ws_layout(n, 0, static_cast<char*>(d_ws), &w, 0, ext);
  const auto t0 = std::chrono::steady_clock::now();
  int G;
  int64_t rpb;
  const Tuning tu = ct.it[0];
  pcg_grid(h_op, tu, &G, &rpb);
  const int gz = grid_for(n, PCG_NT, 2048);
  double* r = (double*)w.r;
  const bool single = tu.split == MODE_SINGLE;
  double* pb[2] = {(double*)w.p0, single ? (double*)w.p0 : (double*)w.p1};
  zero_k<double><<<gz, PCG_NT, 0, s>>>((double*)w.p0, n);
  zero_k<double><<<gz, PCG_NT, 0, s>>>((double*)w.p1, n);
  zero_k<double><<<gz, PCG_NT, 0, s>>>((double*)w.s, n);
  if (use_x0) {
    rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s, w.ext);
    if (rc) return rc;
  } else {
    zero_k<double><<<gz, PCG_NT, 0, s>>>(d_x, n);
  }
  residual_prep_k<double><<<G, PCG_NT, 0, s>>>(n, rpb, d_b, use_x0 ? w.ax : nullptr, d_dinv, r, w.part, 1,
                                                single ? (double*)w.p1 : nullptr);
  finalize_k<<<1, PCG_NT, 0, s>>>(w.st, w.part, G, tol, atol, maxiter, 0.0, 1, 1, single);
  JNO_LAUNCH_CHECK();
  PcgArgs<double, double> a =
      make_args<double, double>(h_op, h_op->vals64, d_dinv, r, (double*)w.q, d_x, w, rpb, tu);
  if (single) {
    rc = launch_kb<double, double>(h_op, a, G, s);  // w = A u, u.w for the first KA
    if (rc) return rc;
  }
  PcgState f;
  rc = run_inner<double, double>(poll, h_op, a, pb, 0, ct.chunk, G, s, &f, maxiter);
  if (rc) return rc;
  // true residual ||b - A x|| / ||b||  (the eager guard of _maybe_residual_check)
  rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s, w.ext);
  if (rc) return rc;
  residual_prep_k<double><<<G, PCG_NT, 0, s>>>(n, rpb, d_b, w.ax, d_dinv, r, w.part, 0, nullptr);
  finalize_k<<<1, PCG_NT, 0, s>>>(w.st, w.part, G, tol, atol, maxiter, 0.0, 0, 0, single);
  JNO_LAUNCH_CHECK();
  PcgState e;
  rc = read_state(poll, w.st, &e, s);
  if (rc) return rc;
  const auto t1 = std::chrono::steady_clock::now();
  const double bn = sqrt(e.bb);
  if (h_stats) {
    h_stats->iterations = f.k_stop;
    h_stats->refinements = 0;
    h_stats->relres = bn > 0 ? sqrt(e.rr_true) / bn : sqrt(e.rr_true);
    h_stats->relres_recursive = bn > 0 ? sqrt(f.rr_stop) / bn : sqrt(f.rr_stop);
    h_stats->seconds = std::chrono::duration<double>(t1 - t0).count();
  }
  return finish_status(f, h_stats, f.k_stop);
}

namespace jno {

// Argument checks of jno_pcg_mixed and its allocation-free variant.
int pcg_mixed_check(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x, int64_t maxiter,
                    double replace_factor, const char* who) {
  int rc = check_op(h_op, 1);
  if (rc) return rc;
  JNO_REQUIRE(d_dinv && d_b && d_x, JNO_EINVAL, "%s: null vector", who);
  JNO_REQUIRE(replace_factor > 0.0 && replace_factor < 1.0, JNO_EINVAL, "%s: replace_factor must be in (0, 1)", who);
  JNO_REQUIRE(maxiter >= 0, JNO_EINVAL, "%s: maxiter must be >= 0", who);
  CallTuning ct;
  return call_tuning(h_op, &ct, who);
}

// FP32 inner iterations on the Jacobi-scaled DIA operator (MODE_SCALED).  Returns JNO_OK with
// *used = false, having changed nothing the caller relies on, when the operator's scaled main
// diagonal is not exactly 1 in FP32 (a zero-diagonal guard row): the caller then runs the
// unscaled inner iteration.  The inner loop stops only on residual replacement or the cap; the
// FP64 true residual decides convergence, and when the remaining reduction is smaller than a
// replacement cycle the next cycle is aimed at it (with margin), so the last cycle does not
// overshoot the tolerance by a full cycle.
static int mixed_scaled(Poller& poll, const CallTuning& ct, const jno_op* h_op, const double* d_dinv,
                        const double* d_b, double* d_x, int use_x0, double tol, double atol, int64_t maxiter,
                        double repl2, jno_pcg_stats* h_stats, const Ws& w, int G, int64_t rpb, const Tuning& tu,
                        hipStream_t s, bool* used) {
  *used = false;
  const int64_t n = h_op->n;
  const int gz = grid_for(n, PCG_NT, 2048);
  const auto t0 = std::chrono::steady_clock::now();
  DiaOffsets off;
  for (int d = 0; d < JNO_DIA_MAX; ++d) off.o[d] = h_op->offsets[d];
  // the check of the scaled main diagonal stays on the device: finalize_k stops the solve before its
  // first iteration when scale_dia_k flagged it, and the host learns that at its first poll
  int* bad = &w.st->scale_bad;
  JNO_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  scale_dia_k<<<gz, PCG_NT, 0, s>>>(n, h_op->ndiag, off, h_op->vals64, d_dinv, w.sdia, bad);
  JNO_LAUNCH_CHECK();
  // one kernel per iteration: r, q (and p) alternate between two buffers by launch parity
  const bool one = tu.split == MODE_ONE && lds_ok(h_op) && w.q2;
  float* rs = (float*)w.r;
  float* pb[2] = {(float*)w.p0, (float*)w.p1};
  float* rb[2] = {rs, (float*)w.s};
  float* qb[2] = {(float*)w.q, (float*)w.q2};
  float* xs = (float*)w.xs;
  if (one) {  // p, x, the second r buffer and both q buffers whole, the first r buffer's pads
    PadList pl = {{pb[0], pb[1], xs, rb[1], qb[0], qb[1], rb[0], nullptr}, 7};
    zero_list_k<<<gz, PCG_NT, 0, s>>>(n, pl, 0x3F);
  } else {
    zero_k<float><<<gz, PCG_NT, 0, s>>>(pb[0], n);
    zero_k<float><<<gz, PCG_NT, 0, s>>>(pb[1], n);
    zero_k<float><<<gz, PCG_NT, 0, s>>>(xs, n);
  }
  int rc;
  if (use_x0) {
    rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s, w.ext);
    if (rc) return rc;
  } else {
    zero_k<double><<<gz, PCG_NT, 0, s>>>(d_x, n);
  }
  residual_prep_scaled_k<<<G, PCG_NT, 0, s>>>(n, rpb, d_b, use_x0 ? w.ax : nullptr, d_dinv, rs, 1, w.part, one ? w.part.C : nullptr);
  // inner tolerance 0: the FP64 residual below decides convergence
  finalize_k<<<1, PCG_NT, 0, s>>>(w.st, w.part, G, 0.0, 0.0, maxiter, repl2, 1, 1, 0, 1);
  JNO_LAUNCH_CHECK();
  if (one) {
    set_special_k<<<1, 64, 0, s>>>(w.st, 0, 1);
    JNO_LAUNCH_CHECK();
  }
  PcgArgs<float, float> a = make_args<float, float>(h_op, w.sdia, nullptr, rs, (float*)w.q, xs, w, rpb, tu);
  a.unit = 1;
  a.split = one ? MODE_ONE : tu.split == MODE_SCALED_FUSED || tu.split == MODE_ONE ? MODE_FUSED : MODE_SPLIT;
  // Each cycle: the inner loop runs to its stop; x is accumulated, the FP64 residual prepared and
  // the outer step decides on the device whether to stop or restart, with no host round trip: the
  // host learns an outer stop from outer_k's own post (one-kernel iterations, pinned poller), read
  // once the next inner loop's first chunk is queued, or else from that loop's first poll (its
  // launches exit at once).
  int64_t k = 0, refinements = 0, resumes = 0;
  bool last_resume = false;
  PcgState f;
  InnerPace pace;
  // one-kernel iterations with the pinned poller: outer_k posts its decision to the third slot
  rc = poll.init();
  if (rc) return rc;
  const bool opost = one && !poll.sync;
  for (;;) {
    const int64_t k_start = k;
    rc = run_inner<float, float>(poll, h_op, a, pb, k, ct.chunk, G, s, &f, maxiter, one ? rb : nullptr, one ? qb : nullptr,
                                 one && ct.pace ? &pace : nullptr, opost && refinements > 0);
    if (rc) return rc;
    if (f.status == ST_UNSCALED) return JNO_OK;  // *used stays false: nothing the caller relies on changed
    *used = true;
    if (f.outer_done) break;
    k = f.k_stop;
    // one-kernel iterations resume by rerunning launch k-1 (from p_{k-2}, which launch k did not
    // overwrite): its r input half takes the true residual, its r.r slot the new partials; a stop
    // at launch 0 (nothing to resume from) starts over
    const int64_t kr = one && k > 0 ? k - 1 : k;
    accumulate_scaled_k<<<gz, PCG_NT, 0, s>>>(n, d_x, xs, d_dinv);
    JNO_LAUNCH_CHECK();
    rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s, w.ext);
    if (rc) return rc;
    residual_prep_scaled_k<<<G, PCG_NT, 0, s>>>(n, rpb, d_b, w.ax, d_dinv, one ? rb[kr & 1] : rs, 1,
                                                 w.part, one ? w.part.C + (kr & 1) * 4 * (int64_t)G : nullptr);
    ++refinements;
    const bool resume = one && k > 0;
    outer_k<<<1, PCG_NT, 0, s>>>(w.st, w.part, G, tol, atol, repl2, k_start, kr, one ? (resume ? 2 : 1) : 0,
                                 one ? maxiter - resumes - (resume ? 1 : 0) : maxiter,
                                 opost ? &poll.kslots()[2] : nullptr);
    JNO_LAUNCH_CHECK();
    if (opost) JNO_HIP(hipEventRecord(g_poll.evk[2], s));
    if (resume) ++resumes;
    last_resume = resume;
    k = kr;
  }
  if (last_resume) --resumes;  // the last outer step stopped instead of resuming
  k = f.k_stop + resumes;       // launches that did work
  const double rr_true = f.rr_true, bb = f.bb;
  const auto t1 = std::chrono::steady_clock::now();
  const double bn = sqrt(bb);
  if (h_stats) {
    h_stats->iterations = k;
    h_stats->refinements = refinements;
    h_stats->relres = bn > 0 ? sqrt(rr_true) / bn : sqrt(rr_true);
    h_stats->relres_recursive = h_stats->relres;  // the last replacement set r to the true residual
    h_stats->seconds = std::chrono::duration<double>(t1 - t0).count();
  }
  return finish_status(f, h_stats, k);
}

// The FP32-inner / FP64-refinement solve behind jno_pcg_mixed and its allocation-free variant:
// `sync_poll` reads the device state back synchronously (no pinned memory, no events).  The
// arguments are checked by the callers.
int pcg_mixed_impl(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x, int use_x0, double tol,
                   double atol, int64_t maxiter, double replace_factor, jno_pcg_stats* h_stats, void* d_ws,
                   size_t ws_bytes, hipStream_t s, bool sync_poll) {
  const int64_t n = h_op->n;
  CallTuning ct;
  int rc = call_tuning(h_op, &ct, "jno_pcg_mixed");
  if (rc) return rc;
  Poller poll;
  poll.sync = sync_poll;
  const Tuning tu = effective_tuning(ct.it[1], h_op);
  const int64_t se = scaled_elems(h_op);
  const size_t ext = ebe_scratch_bytes(h_op);
  JNO_REQUIRE(ws_bytes >= ws_layout(n, 1, nullptr, nullptr, 0, ext), JNO_EWORKSPACE, "jno_pcg_mixed: workspace too small");
  const bool room = se > 0 && ws_bytes >= ws_layout(n, 1, nullptr, nullptr, se, ext);
  Ws w;
  ws_layout(n, 1, static_cast<char*>(d_ws), &w, room ? se : 0, ext);
  int G;
  int64_t rpb;
  pcg_grid(h_op, tu, &G, &rpb);
  if (room && (tu.split == MODE_SCALED || tu.split == MODE_SCALED_FUSED || tu.split == MODE_ONE)) {
    bool used = false;
    rc = mixed_scaled(poll, ct, h_op, d_dinv, d_b, d_x, use_x0, tol, atol, maxiter, replace_factor * replace_factor,
                      h_stats, w, G, rpb, tu, s, &used);
    if (rc || used) return rc;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const int gz = grid_for(n, PCG_NT, 2048);
  float* rs = (float*)w.r;
  const bool single = tu.split == MODE_SINGLE;
  float* pb[2] = {(float*)w.p0, single ? (float*)w.p0 : (float*)w.p1};
  float* xs = (float*)w.xs;
  float* u_out = single ? (float*)w.p1 : nullptr;
  zero_k<float><<<gz, PCG_NT, 0, s>>>((float*)w.p0, n);
  zero_k<float><<<gz, PCG_NT, 0, s>>>((float*)w.p1, n);
  zero_k<float><<<gz, PCG_NT, 0, s>>>((float*)w.s, n);
  zero_k<float><<<gz, PCG_NT, 0, s>>>(xs, n);
  cast_f32_vec_k<<<gz, PCG_NT, 0, s>>>(n, d_dinv, w.dinv32);
  if (use_x0) {
    rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s, w.ext);
    if (rc) return rc;
  } else {
    zero_k<double><<<gz, PCG_NT, 0, s>>>(d_x, n);
  }
  const double repl2 = replace_factor * replace_factor;
  residual_prep_k<float><<<G, PCG_NT, 0, s>>>(n, rpb, d_b, use_x0 ? w.ax : nullptr, w.dinv32, rs, w.part, 1,
                                               u_out);
  finalize_k<<<1, PCG_NT, 0, s>>>(w.st, w.part, G, tol, atol, maxiter, repl2, 1, 1, single);
  JNO_LAUNCH_CHECK();
  PcgArgs<float, float> a = make_args<float, float>(h_op, h_op->vals32, w.dinv32, rs, (float*)w.q, xs, w, rpb, tu);
  if (single) {
    rc = launch_kb<float, float>(h_op, a, G, s);  // w = A u, u.w for the first KA
    if (rc) return rc;
  }
  int64_t k = 0, refinements = 0;
  PcgState f, e;
  double rr_true = 0.0, bb = 0.0;
  for (;;) {
    rc = run_inner<float, float>(poll, h_op, a, pb, k, ct.chunk, G, s, &f, maxiter);
    if (rc) return rc;
    k = f.k_stop;
    // FP64 refinement: x += xs, r = b - A x, rs = (float) r
    accumulate_x_k<<<gz, PCG_NT, 0, s>>>(n, d_x, xs);
    JNO_LAUNCH_CHECK();
    rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s, w.ext);
    if (rc) return rc;
    // the single-reduction KA of iteration k reads the r.u / r.r partials of buffer (k odd ? B2 : B)
    Partials pp = w.part;
    if (single && (k & 1)) pp.B = w.part.B2;
    residual_prep_k<float><<<G, PCG_NT, 0, s>>>(n, rpb, d_b, w.ax, w.dinv32, rs, pp, 1, u_out);
    // sets the state for a restart too (two-reduction iterations: alpha = 0, the deferred x update
    // of the next K1 is void, it was applied at the stop; p_{k-1} and gamma_{k-1} are kept; the
    // single-reduction iteration keeps p, s = A p, gamma_{k-1}, alpha_{k-1} and needs w = A u)
    finalize_k<<<1, PCG_NT, 0, s>>>(w.st, pp, G, tol, atol, maxiter, repl2, 0, 1, single);
    JNO_LAUNCH_CHECK();
    if (single) {
      rc = launch_kb<float, float>(h_op, a, G, s);
      if (rc) return rc;
    }
    rc = read_state(poll, w.st, &e, s);
    if (rc) return rc;
    ++refinements;
    rr_true = e.rr_true;
    bb = e.bb;
    const double atol2 = fmax(tol * tol * bb, atol * atol);
    if (!(rr_true > atol2)) {
      f.status = ST_CONVERGED;
      break;
    }
    if (f.status == ST_MAXITER || f.status == ST_NONFINITE || !std::isfinite(rr_true)) {
      if (!std::isfinite(rr_true)) f.status = ST_NONFINITE;
      break;
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  const double bn = sqrt(bb);
  if (h_stats) {
    h_stats->iterations = k;
    h_stats->refinements = refinements;
    h_stats->relres = bn > 0 ? sqrt(rr_true) / bn : sqrt(rr_true);
    h_stats->relres_recursive = bn > 0 ? sqrt(e.rr0) / bn : sqrt(e.rr0);
    h_stats->seconds = std::chrono::duration<double>(t1 - t0).count();
  }
  return finish_status(f, h_stats, k);
}

}  // namespace jno

extern "C" int jno_pcg_mixed(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x, int use_x0,
                             double tol, double atol, int64_t maxiter, double replace_factor,
                             jno_pcg_stats* h_stats, void* d_ws, size_t ws_bytes, void* stream) {
  int rc = jno::pcg_mixed_check(h_op, d_dinv, d_b, d_x, maxiter, replace_factor, "jno_pcg_mixed");
  if (rc) return rc;
  JNO_REQUIRE(ws_bytes >= jno_pcg_workspace_size(h_op->n, 1), JNO_EWORKSPACE, "jno_pcg_mixed: workspace too small");
  return jno::pcg_mixed_impl(h_op, d_dinv, d_b, d_x, use_x0, tol, atol, maxiter, replace_factor, h_stats, d_ws,
                             ws_bytes, as_stream(stream), false);
}
