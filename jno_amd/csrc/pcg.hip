// Jacobi-preconditioned CG: plain FP64 (jax.scipy.sparse.linalg.cg semantics as called by
// jno.solve.cg, jno/solve.py:191-206) and FP32 inner iterations with FP64 residual replacement.
//
// One iteration is two kernels:
//   K1  p_k = z_k + beta p_{k-1}  (z = dinv r, recomputed for every gathered neighbour, so p is
//       never a separate pass),  x += alpha_{k-1} p_{k-1} (the x update is deferred one iteration
//       and fused here),  q = A p_k,  partial p.q
//   K2  alpha_k = gamma_k / p.q,  r -= alpha_k q,  partials r.z and r.r
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
#include <chrono>
#include <cmath>
#include <vector>

#include "common.h"
#include "sparse.h"

namespace jno {

constexpr int PCG_NT = 256;
constexpr int PCG_MAX_BLOCKS = 4096;  // workspace capacity for per-workgroup partials

// Launch shape of the iteration kernels, one per iteration precision (set by the autotuner through
// jno_pcg_set_tuning).  The defaults are the fastest shapes of a sweep over block x grid cap x DIA
// tile variant on a 2.1M-row 4-diagonal operator: FP32 iterations 1024 threads x 256 workgroups
// with one row per thread, FP64 iterations 512 x 1024 with 16-byte vector tiles.
struct Tuning {
  int grid_cap;  // workgroups of the iteration kernels (a multiple of 8)
  int vec;       // DIA K1: VN-row vector tiles (1) or one row per thread (0)
  int block;     // threads per workgroup (256, 512, 1024)
};
static Tuning g_tune[2] = {{1024, 1, 512}, {256, 0, 1024}};  // [0] FP64 iterations, [1] FP32 iterations
constexpr int N_XCD = 8;

enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_MAXITER = 2, ST_NONFINITE = 3, ST_REPLACE = 4 };

struct PcgState {
  int32_t done;    // written by K2 (block 0), read by K1
  int32_t done1;   // written by K1 (block 0), read by K2
  int32_t status;  // why K1 stopped
  int32_t pad;
  int64_t k;       // iterations done (written by K2 block 0)
  int64_t k_stop;  // iteration at which K1 stopped
  double rr_stop;     // recursive r.r at the stop
  double gamma_cur;   // gamma_k for K2 (K1 block 0)
  double gamma_prev;  // gamma_{k-1}       (K2 block 0)
  double alpha;       // alpha_{k-1}       (K2 block 0)
  double atol2;       // stop when r.r <= atol2
  double rr_replace;  // FP32 inner: replace the residual when r.r <= rr_replace
  double rr_true;     // ||b - A x||^2 (finalize)
  double bb;          // b . b (finalize)
  double rr0;         // r.r after the last residual preparation (finalize)
  int64_t maxiter;
};

// partial-sum areas (per workgroup): A = K1's p.q; B = K2's (r.z, r.r); P = prep's (rr_true, b.b)
struct Partials {
  double* A;  // [G]
  double* B;  // [2][G]
  double* P;  // [2][G]
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
  DiaOffsets off;
  const T* dinv;
  T* r;
  const T* p_old;
  T* p_new;
  T* q;
  T* x;
  PcgState* st;
  Partials part;
};

// contiguous row range of this workgroup, XCD-aware (gridDim.x is a multiple of 8 or < 8)
__device__ __forceinline__ int64_t block_first_row(int64_t rows_per_block) {
  const int G = gridDim.x, b = blockIdx.x;
  const int logical = (G % N_XCD == 0) ? (b % N_XCD) * (G / N_XCD) + b / N_XCD : b;
  return (int64_t)logical * rows_per_block;
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

template <typename T>
__device__ __forceinline__ void st16(T* p, const V16<T>& r) {
  __builtin_memcpy(p, &r, 16);
}

// DIA row tile [i0, i0+VN): p_new = z + beta p_old on the tile, q = A p_new, deferred x update;
// returns the tile's p.q.  `interior` tiles need no bounds checks on neighbour rows.
template <typename T, typename V>
__device__ __forceinline__ double dia_tile(const PcgArgs<T, V>& a, int64_t i0, T beta, T ap, bool upd_x) {
  constexpr int VN = V16<T>::N;
  const int64_t ld = dia_ld(a.n);
  const V16<T> dv = ld16(a.dinv + i0), rv = ld16(a.r + i0), po = ld16(a.p_old + i0);
  V16<T> pi, s;
#pragma unroll
  for (int j = 0; j < VN; ++j) pi.v[j] = dv.v[j] * rv.v[j] + beta * po.v[j];
  {
    const V16<V> d0 = ld16(a.vals + i0);
#pragma unroll
    for (int j = 0; j < VN; ++j) s.v[j] = (T)d0.v[j] * pi.v[j];
  }
  for (int d = 1; d < a.ndiag; ++d) {
    const int64_t o = a.off.o[d];
    const V* vd = a.vals + (int64_t)d * ld;
    const V16<V> vu = ld16(vd + i0), vl = ld16(vd + i0 - o);
    const V16<T> du = ld16(a.dinv + i0 + o), ru = ld16(a.r + i0 + o), pu = ld16(a.p_old + i0 + o);
    const V16<T> dl = ld16(a.dinv + i0 - o), rl = ld16(a.r + i0 - o), pl = ld16(a.p_old + i0 - o);
#pragma unroll
    for (int j = 0; j < VN; ++j) {
      s.v[j] += (T)vu.v[j] * (du.v[j] * ru.v[j] + beta * pu.v[j]);
      s.v[j] += (T)vl.v[j] * (dl.v[j] * rl.v[j] + beta * pl.v[j]);
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
template <typename T, typename V>
__device__ __forceinline__ double dia_row(const PcgArgs<T, V>& a, int64_t i, T beta, T ap, bool upd_x) {
  const int64_t ld = dia_ld(a.n);
  const T po = a.p_old[i];
  const T pi = a.dinv[i] * a.r[i] + beta * po;
  T s = (T)a.vals[i] * pi;
  for (int d = 1; d < a.ndiag; ++d) {
    const int64_t o = a.off.o[d];
    const V* v = a.vals + (int64_t)d * ld;
    if (i + o < a.n) s += (T)v[i] * pnew_at(a.dinv, a.r, a.p_old, beta, i + o);
    if (i - o >= 0) s += (T)v[i - o] * pnew_at(a.dinv, a.r, a.p_old, beta, i - o);
  }
  a.p_new[i] = pi;
  a.q[i] = s;
  if (upd_x) a.x[i] += ap * po;
  return (double)pi * (double)s;
}

template <int NT, int L, bool DIA, typename T, typename V>
__global__ void __launch_bounds__(NT) pcg_k1(PcgArgs<T, V> a) {
  __shared__ double sm[NT / 64];
  PcgState* st = a.st;
  if (st->done) return;
  const int G = gridDim.x;
  const double gamma = reduce_partials<NT>(a.part.B, G, sm);
  const double rr = reduce_partials<NT>(a.part.B + G, G, sm);
  const int64_t k = st->k;
  const double alpha_prev = st->alpha;
  const bool stop_conv = !(rr > st->atol2);
  const bool stop_cap = k >= st->maxiter;
  const bool stop_repl = rr <= st->rr_replace;
  const bool bad = !isfinite(gamma) || !isfinite(rr);
  const T ap = (T)alpha_prev;
  const int64_t r0 = block_first_row(a.rows_per_block);
  const int64_t r1 = min(a.n, r0 + a.rows_per_block);
  if (stop_conv || stop_cap || stop_repl || bad) {
    // finish the deferred x update and stop
    if (k > 0 && alpha_prev != 0.0)
      for (int64_t i = r0 + threadIdx.x; i < r1; i += NT) a.x[i] += ap * a.p_old[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done1 = 1;
      st->k_stop = k;
      st->rr_stop = rr;
      st->status = bad ? ST_NONFINITE : stop_conv ? ST_CONVERGED : stop_cap ? ST_MAXITER : ST_REPLACE;
    }
    return;
  }
  const double beta_d = (k == 0) ? 0.0 : gamma / st->gamma_prev;
  const T beta = (T)beta_d;
  double part = 0.0;
  if (DIA && a.vec) {
    // each thread owns VN consecutive rows; the block's first row is VN-aligned
    constexpr int VN = V16<T>::N;
    int64_t omax = 0;
    for (int d = 1; d < a.ndiag; ++d) omax = max(omax, (int64_t)a.off.o[d]);
    for (int64_t i0 = r0 + (int64_t)threadIdx.x * VN; i0 < r1; i0 += (int64_t)VN * NT) {
      if (i0 + VN <= r1 && i0 - omax >= 0 && i0 + VN - 1 + omax < a.n) {
        part += dia_tile<T, V>(a, i0, beta, ap, k > 0);
      } else {
        for (int64_t i = i0; i < min(r1, i0 + VN); ++i) part += dia_row<T, V>(a, i, beta, ap, k > 0);
      }
    }
  } else if (DIA) {
    for (int64_t i = r0 + threadIdx.x; i < r1; i += NT) part += dia_row<T, V>(a, i, beta, ap, k > 0);
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
        a.p_new[row] = pi;
        a.q[row] = s;
        if (k > 0) a.x[row] += ap * po;
        part += (double)pi * (double)s;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->gamma_cur = gamma;
  store_partial<NT>(part, a.part.A, sm);
}

// ---------------------------------------------------------------------------------------------
// K2: r -= alpha q, partials r.z and r.r
// ---------------------------------------------------------------------------------------------
template <int NT, typename T, typename V>
__global__ void __launch_bounds__(NT) pcg_k2(PcgArgs<T, V> a) {
  __shared__ double sm[NT / 64];
  PcgState* st = a.st;
  if (st->done1) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
    return;
  }
  const int G = gridDim.x;
  const int64_t r0 = block_first_row(a.rows_per_block);
  const int64_t r1 = min(a.n, r0 + a.rows_per_block);
  constexpr int VN = V16<T>::N;
  // the first tile's loads are issued before the partial-sum reduction, so they overlap it
  const int64_t f0 = r0 + (int64_t)threadIdx.x * VN;
  const bool pre = f0 + VN <= r1;
  V16<T> rv0, qv0, dv0;
  if (pre) {
    rv0 = ld16(a.r + f0);
    qv0 = ld16(a.q + f0);
    dv0 = ld16(a.dinv + f0);
  }
  const double pq = reduce_partials<NT>(a.part.A, G, sm);
  const double gamma = st->gamma_cur;
  const double alpha = gamma / pq;
  const T al = (T)alpha;
  double pg = 0.0, pr = 0.0;
  for (int64_t i0 = f0; i0 < r1; i0 += (int64_t)VN * NT) {
    if (i0 + VN <= r1) {
      const bool first = pre && i0 == f0;
      V16<T> rv = first ? rv0 : ld16(a.r + i0);
      const V16<T> qv = first ? qv0 : ld16(a.q + i0), dv = first ? dv0 : ld16(a.dinv + i0);
#pragma unroll
      for (int j = 0; j < VN; ++j) {
        const T ri = rv.v[j] - al * qv.v[j];
        rv.v[j] = ri;
        pg += (double)ri * (double)(dv.v[j] * ri);
        pr += (double)ri * (double)ri;
      }
      st16(a.r + i0, rv);
    } else {
      for (int64_t i = i0; i < r1; ++i) {
        const T ri = a.r[i] - al * a.q[i];
        a.r[i] = ri;
        pg += (double)ri * (double)(a.dinv[i] * ri);
        pr += (double)ri * (double)ri;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alpha;
    st->gamma_prev = gamma;
    st->k = st->k + 1;
  }
  store_partial<NT>(pg, a.part.B, sm);
  store_partial<NT>(pr, a.part.B + G, sm);
}

// ---------------------------------------------------------------------------------------------
// residual preparation: r64 = b - Ax (Ax nullable = 0), rs = (T) r64; partials of
// gamma = rs.dinv.rs, rr = rs.rs (into B, for the next K1) and of ||r64||^2, b.b (into P)
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(PCG_NT) residual_prep_k(int64_t n, int64_t rows_per_block,
                                                          const double* __restrict__ b, const double* __restrict__ Ax,
                                                          const T* __restrict__ dinv, T* __restrict__ rs, Partials part,
                                                          int write_rs) {
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
    pg += (double)ri * (double)(dinv[i] * ri);
    pr += (double)ri * (double)ri;
  }
  store_partial(pg, part.B, sm);
  store_partial(pr, part.B + G, sm);
  store_partial(prt, part.P, sm);
  store_partial(pb, part.P + G, sm);
}

// one workgroup: reduce the prep partials into the state and set the tolerances
__global__ void __launch_bounds__(PCG_NT) finalize_k(PcgState* st, Partials part, int G, double tol, double atol,
                                                     int64_t maxiter, double replace2, int first, int restart) {
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
  st->alpha = 0.0;  // any deferred x update was applied at the stop
  if (first) st->k = 0;
  st->rr_replace = replace2 > 0.0 ? replace2 * rr : -1.0;
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
  void* xs;      // FP32 accumulator (mixed)
  float* dinv32; // mixed
  double* ax;    // FP64 A x
};

static size_t ws_layout(int64_t n, int mixed, char* base, Ws* w) {
  const size_t vb = mixed ? 4 : 8;
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    void* p = base ? base + off : nullptr;
    off += a256(bytes);
    return p;
  };
  Ws t;
  t.st = (PcgState*)take(sizeof(PcgState));
  t.part.A = (double*)take(PCG_MAX_BLOCKS * sizeof(double));
  t.part.B = (double*)take(2 * PCG_MAX_BLOCKS * sizeof(double));
  t.part.P = (double*)take(2 * PCG_MAX_BLOCKS * sizeof(double));
  t.r = take(n * vb);
  t.p0 = take(n * vb);
  t.p1 = take(n * vb);
  t.q = take(n * vb);
  t.xs = mixed ? take(n * 4) : nullptr;
  t.dinv32 = mixed ? (float*)take(n * 4) : nullptr;
  t.ax = (double*)take(n * 8);
  if (w) *w = t;
  return off;
}

// grid of the iteration kernels: a multiple of 8 (XCD slabs), at most PCG_MAX_BLOCKS, and the
// rows per workgroup
static void pcg_grid(int64_t n, int lanes, const Tuning& tu, int* G, int64_t* rpb) {
  int64_t want = cdiv(n * lanes, (int64_t)tu.block * 2);
  if (want < 1) want = 1;
  if (want > tu.grid_cap) want = tu.grid_cap;
  if (want >= N_XCD) want = (want / N_XCD) * N_XCD;
  *G = (int)want;
  // a multiple of 4 rows so every workgroup's first row is 16-byte aligned for FP32 and FP64
  *rpb = cdiv(cdiv(n, want), 4) * 4;
}

struct HostPoll {
  PcgState* h = nullptr;  // pinned, 2 slots
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ok = false;
  int init() {
    if (ok) return JNO_OK;
    JNO_HIP(hipHostMalloc((void**)&h, 2 * sizeof(PcgState), hipHostMallocDefault));
    JNO_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    JNO_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    ok = true;
    return JNO_OK;
  }
};
static thread_local HostPoll g_poll;

template <int NT, typename T, typename V>
static int launch_iter_nt(const jno_op* op, PcgArgs<T, V>& a, int grid, hipStream_t s, hipEvent_t after_k1,
                          hipEvent_t after_k2) {
  if (op->format == JNO_FMT_DIA) {
    pcg_k1<NT, 1, true, T, V><<<grid, NT, 0, s>>>(a);
  } else {
    switch (op->lanes) {
      case 1: pcg_k1<NT, 1, false, T, V><<<grid, NT, 0, s>>>(a); break;
      case 2: pcg_k1<NT, 2, false, T, V><<<grid, NT, 0, s>>>(a); break;
      case 4: pcg_k1<NT, 4, false, T, V><<<grid, NT, 0, s>>>(a); break;
      case 8: pcg_k1<NT, 8, false, T, V><<<grid, NT, 0, s>>>(a); break;
      case 16: pcg_k1<NT, 16, false, T, V><<<grid, NT, 0, s>>>(a); break;
      case 32: pcg_k1<NT, 32, false, T, V><<<grid, NT, 0, s>>>(a); break;
      case 64: pcg_k1<NT, 64, false, T, V><<<grid, NT, 0, s>>>(a); break;
      default: JNO_REQUIRE(false, JNO_EINVAL, "pcg: lanes per row must be a power of two in 1..64");
    }
  }
  JNO_LAUNCH_CHECK();
  if (after_k1) JNO_HIP(hipEventRecord(after_k1, s));
  pcg_k2<NT, T, V><<<grid, NT, 0, s>>>(a);
  JNO_LAUNCH_CHECK();
  if (after_k2) JNO_HIP(hipEventRecord(after_k2, s));
  return JNO_OK;
}

template <typename T, typename V>
static int launch_iter(const jno_op* op, PcgArgs<T, V>& a, int grid, hipStream_t s, hipEvent_t after_k1,
                       hipEvent_t after_k2) {
  switch (a.block) {
    case 256: return launch_iter_nt<256, T, V>(op, a, grid, s, after_k1, after_k2);
    case 512: return launch_iter_nt<512, T, V>(op, a, grid, s, after_k1, after_k2);
    default: return launch_iter_nt<1024, T, V>(op, a, grid, s, after_k1, after_k2);
  }
}

// Optional live kernel timing: on every stride-th iteration, HIP events are recorded on the
// solve's stream around its K1 and K2 (E0 K1 E1 K2 E2); the sampled iterations that did work are
// summed.
struct KernelProf {
  bool on = false;
  int stride = 8;  // time every stride-th iteration (events perturb the queue very little then)
  std::vector<hipEvent_t> ev;
  double k1_ms = 0, k2_ms = 0;
  int64_t k1_n = 0, k2_n = 0;
  int ensure(size_t n) {
    while (ev.size() < n) {
      hipEvent_t e;
      JNO_HIP(hipEventCreate(&e));
      ev.push_back(e);
    }
    return JNO_OK;
  }
};
static KernelProf g_prof;

// Runs K1/K2 chunks until the device reports a stop; returns the final state in *out.
template <typename T, typename V>
static int run_inner(const jno_op* op, PcgArgs<T, V> a, T* pbuf[2], int64_t k0, int chunk, int grid, hipStream_t s,
                     PcgState* out) {
  JNO_REQUIRE(g_poll.init() == JNO_OK, JNO_EHIP, "pcg: pinned state allocation failed");
  int64_t k = k0;
  int slot = 0, pending = 0;
  size_t nev = 0;
  std::vector<int64_t> sampled;  // iteration index of each (E0, E1, E2) event triple
  for (;;) {
    for (int i = 0; i < chunk; ++i, ++k) {
      a.p_old = pbuf[k & 1];
      a.p_new = pbuf[(k + 1) & 1];
      if (g_prof.on && k % g_prof.stride == 0) {
        if (g_prof.ensure(nev + 3) != JNO_OK) return JNO_EHIP;
        JNO_HIP(hipEventRecord(g_prof.ev[nev], s));
        int rc = launch_iter<T, V>(op, a, grid, s, g_prof.ev[nev + 1], g_prof.ev[nev + 2]);
        nev += 3;
        sampled.push_back(k);
        if (rc) return rc;
      } else {
        int rc = launch_iter<T, V>(op, a, grid, s, nullptr, nullptr);
        if (rc) return rc;
      }
    }
    JNO_HIP(hipMemcpyAsync(&g_poll.h[slot], a.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    JNO_HIP(hipEventRecord(g_poll.ev[slot], s));
    if (pending) {
      const int prev = slot ^ 1;
      JNO_HIP(hipEventSynchronize(g_poll.ev[prev]));
      if (g_poll.h[prev].done1) break;
    }
    pending = 1;
    slot ^= 1;
  }
  JNO_HIP(hipMemcpyAsync(&g_poll.h[0], a.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  *out = g_poll.h[0];
  if (g_prof.on) {
    for (size_t t = 0; t < sampled.size(); ++t) {
      if (sampled[t] >= out->k_stop) break;  // launches after the stop did no work
      float t1 = 0, t2 = 0;
      JNO_HIP(hipEventElapsedTime(&t1, g_prof.ev[3 * t], g_prof.ev[3 * t + 1]));
      JNO_HIP(hipEventElapsedTime(&t2, g_prof.ev[3 * t + 1], g_prof.ev[3 * t + 2]));
      g_prof.k1_ms += t1;
      g_prof.k2_ms += t2;
      g_prof.k1_n += 1;
      g_prof.k2_n += 1;
    }
  }
  return JNO_OK;
}

static int read_state(PcgState* d, PcgState* out, hipStream_t s) {
  JNO_REQUIRE(g_poll.init() == JNO_OK, JNO_EHIP, "pcg: pinned state allocation failed");
  JNO_HIP(hipMemcpyAsync(&g_poll.h[0], d, sizeof(PcgState), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  *out = g_poll.h[0];
  return JNO_OK;
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
  JNO_REQUIRE(op->format == JNO_FMT_CSR || op->format == JNO_FMT_DIA, JNO_EINVAL, "pcg: unknown operator format");
  JNO_REQUIRE(op->vals64, JNO_EINVAL, "pcg: operator has no FP64 values");
  JNO_REQUIRE(!need32 || op->vals32, JNO_EINVAL, "pcg: mixed precision needs the operator's FP32 values");
  if (op->format == JNO_FMT_CSR) JNO_REQUIRE(op->rowptr && op->col, JNO_EINVAL, "pcg: CSR operator without indices");
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
  for (int d = 0; d < JNO_DIA_MAX; ++d) a.off.o[d] = op->offsets[d];
  a.dinv = dinv;
  a.r = r;
  a.q = q;
  a.x = x;
  a.st = w.st;
  a.part = w.part;
  a.p_old = nullptr;
  a.p_new = nullptr;
  return a;
}

static int g_chunk = 8;

}  // namespace jno

using namespace jno;

extern "C" size_t jno_pcg_workspace_size(int64_t n, int mixed) { return ws_layout(n, mixed ? 1 : 0, nullptr, nullptr); }

extern "C" int jno_pcg_profile(int enable) {
  g_prof.on = enable != 0;
  g_prof.k1_ms = g_prof.k2_ms = 0;
  g_prof.k1_n = g_prof.k2_n = 0;
  return JNO_OK;
}

extern "C" int jno_pcg_profile_read(double* h_k1_ms, int64_t* h_k1_n, double* h_k2_ms, int64_t* h_k2_n) {
  if (h_k1_ms) *h_k1_ms = g_prof.k1_ms;
  if (h_k1_n) *h_k1_n = g_prof.k1_n;
  if (h_k2_ms) *h_k2_ms = g_prof.k2_ms;
  if (h_k2_n) *h_k2_n = g_prof.k2_n;
  return JNO_OK;
}

extern "C" int jno_pcg_set_tuning(int precision, int grid_cap, int k1_vector, int block) {
  JNO_REQUIRE(precision == 0 || precision == 1, JNO_EINVAL,
              "jno_pcg_set_tuning: precision must be 0 (FP64 iterations) or 1 (FP32 iterations)");
  JNO_REQUIRE(grid_cap >= 8 && grid_cap <= PCG_MAX_BLOCKS && grid_cap % 8 == 0, JNO_EINVAL,
              "jno_pcg_set_tuning: grid_cap must be a multiple of 8 in 8..%d", PCG_MAX_BLOCKS);
  JNO_REQUIRE(block == 256 || block == 512 || block == 1024, JNO_EINVAL,
              "jno_pcg_set_tuning: block must be 256, 512 or 1024");
  g_tune[precision] = Tuning{grid_cap, k1_vector ? 1 : 0, block};
  return JNO_OK;
}

extern "C" int jno_pcg_get_tuning(int precision, int* grid_cap, int* k1_vector, int* block) {
  JNO_REQUIRE(precision == 0 || precision == 1, JNO_EINVAL,
              "jno_pcg_get_tuning: precision must be 0 (FP64 iterations) or 1 (FP32 iterations)");
  if (grid_cap) *grid_cap = g_tune[precision].grid_cap;
  if (k1_vector) *k1_vector = g_tune[precision].vec;
  if (block) *block = g_tune[precision].block;
  return JNO_OK;
}

extern "C" int jno_pcg_set_chunk(int chunk) {
  JNO_REQUIRE(chunk >= 1 && chunk <= 1024, JNO_EINVAL, "jno_pcg_set_chunk: chunk must be in 1..1024");
  g_chunk = chunk;
  return JNO_OK;
}

extern "C" int jno_pcg_f64(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x, int use_x0,
                           double tol, double atol, int64_t maxiter, jno_pcg_stats* h_stats, void* d_ws,
                           size_t ws_bytes, void* stream) {
  int rc = check_op(h_op, 0);
  if (rc) return rc;
  const int64_t n = h_op->n;
  JNO_REQUIRE(d_dinv && d_b && d_x, JNO_EINVAL, "jno_pcg_f64: null vector");
  JNO_REQUIRE(ws_bytes >= jno_pcg_workspace_size(n, 0), JNO_EWORKSPACE, "jno_pcg_f64: workspace too small");
  JNO_REQUIRE(maxiter >= 0, JNO_EINVAL, "jno_pcg_f64: maxiter must be >= 0");
  hipStream_t s = as_stream(stream);
  Ws w;
  ws_layout(n, 0, static_cast<char*>(d_ws), &w);
  const auto t0 = std::chrono::steady_clock::now();
  int G;
  int64_t rpb;
  const Tuning tu = g_tune[0];
  pcg_grid(n, h_op->format == JNO_FMT_DIA ? 1 : h_op->lanes, tu, &G, &rpb);
  const int gz = grid_for(n, PCG_NT, 2048);
  double* r = (double*)w.r;
  double* pb[2] = {(double*)w.p0, (double*)w.p1};
  zero_k<double><<<gz, PCG_NT, 0, s>>>(pb[0], n);
  zero_k<double><<<gz, PCG_NT, 0, s>>>(pb[1], n);
  if (use_x0) {
    rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s);
    if (rc) return rc;
  } else {
    zero_k<double><<<gz, PCG_NT, 0, s>>>(d_x, n);
  }
  residual_prep_k<double><<<G, PCG_NT, 0, s>>>(n, rpb, d_b, use_x0 ? w.ax : nullptr, d_dinv, r, w.part, 1);
  finalize_k<<<1, PCG_NT, 0, s>>>(w.st, w.part, G, tol, atol, maxiter, 0.0, 1, 1);
  JNO_LAUNCH_CHECK();
  PcgArgs<double, double> a =
      make_args<double, double>(h_op, h_op->vals64, d_dinv, r, (double*)w.q, d_x, w, rpb, tu);
  PcgState f;
  rc = run_inner<double, double>(h_op, a, pb, 0, g_chunk, G, s, &f);
  if (rc) return rc;
  // true residual ||b - A x|| / ||b||  (the eager guard of _maybe_residual_check)
  rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s);
  if (rc) return rc;
  residual_prep_k<double><<<G, PCG_NT, 0, s>>>(n, rpb, d_b, w.ax, d_dinv, r, w.part, 0);
  finalize_k<<<1, PCG_NT, 0, s>>>(w.st, w.part, G, tol, atol, maxiter, 0.0, 0, 0);
  JNO_LAUNCH_CHECK();
  PcgState e;
  rc = read_state(w.st, &e, s);
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

extern "C" int jno_pcg_mixed(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x, int use_x0,
                             double tol, double atol, int64_t maxiter, double replace_factor,
                             jno_pcg_stats* h_stats, void* d_ws, size_t ws_bytes, void* stream) {
  int rc = check_op(h_op, 1);
  if (rc) return rc;
  const int64_t n = h_op->n;
  JNO_REQUIRE(d_dinv && d_b && d_x, JNO_EINVAL, "jno_pcg_mixed: null vector");
  JNO_REQUIRE(ws_bytes >= jno_pcg_workspace_size(n, 1), JNO_EWORKSPACE, "jno_pcg_mixed: workspace too small");
  JNO_REQUIRE(replace_factor > 0.0 && replace_factor < 1.0, JNO_EINVAL,
              "jno_pcg_mixed: replace_factor must be in (0, 1)");
  JNO_REQUIRE(maxiter >= 0, JNO_EINVAL, "jno_pcg_mixed: maxiter must be >= 0");
  hipStream_t s = as_stream(stream);
  Ws w;
  ws_layout(n, 1, static_cast<char*>(d_ws), &w);
  const auto t0 = std::chrono::steady_clock::now();
  int G;
  int64_t rpb;
  const Tuning tu = g_tune[1];
  pcg_grid(n, h_op->format == JNO_FMT_DIA ? 1 : h_op->lanes, tu, &G, &rpb);
  const int gz = grid_for(n, PCG_NT, 2048);
  float* rs = (float*)w.r;
  float* pb[2] = {(float*)w.p0, (float*)w.p1};
  float* xs = (float*)w.xs;
  zero_k<float><<<gz, PCG_NT, 0, s>>>(pb[0], n);
  zero_k<float><<<gz, PCG_NT, 0, s>>>(pb[1], n);
  zero_k<float><<<gz, PCG_NT, 0, s>>>(xs, n);
  cast_f32_vec_k<<<gz, PCG_NT, 0, s>>>(n, d_dinv, w.dinv32);
  if (use_x0) {
    rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s);
    if (rc) return rc;
  } else {
    zero_k<double><<<gz, PCG_NT, 0, s>>>(d_x, n);
  }
  const double repl2 = replace_factor * replace_factor;
  residual_prep_k<float><<<G, PCG_NT, 0, s>>>(n, rpb, d_b, use_x0 ? w.ax : nullptr, w.dinv32, rs, w.part, 1);
  finalize_k<<<1, PCG_NT, 0, s>>>(w.st, w.part, G, tol, atol, maxiter, repl2, 1, 1);
  JNO_LAUNCH_CHECK();
  PcgArgs<float, float> a = make_args<float, float>(h_op, h_op->vals32, w.dinv32, rs, (float*)w.q, xs, w, rpb, tu);
  int64_t k = 0, refinements = 0;
  PcgState f, e;
  double rr_true = 0.0, bb = 0.0;
  for (;;) {
    rc = run_inner<float, float>(h_op, a, pb, k, g_chunk, G, s, &f);
    if (rc) return rc;
    k = f.k_stop;
    // FP64 refinement: x += xs, r = b - A x, rs = (float) r
    accumulate_x_k<<<gz, PCG_NT, 0, s>>>(n, d_x, xs);
    JNO_LAUNCH_CHECK();
    rc = op_spmv<double, double>(h_op, h_op->vals64, d_x, w.ax, s);
    if (rc) return rc;
    residual_prep_k<float><<<G, PCG_NT, 0, s>>>(n, rpb, d_b, w.ax, w.dinv32, rs, w.part, 1);
    // sets the state for a restart too (alpha = 0: the deferred x update of the next K1 is void,
    // it was applied at the stop; p_{k-1} and gamma_{k-1} are kept)
    finalize_k<<<1, PCG_NT, 0, s>>>(w.st, w.part, G, tol, atol, maxiter, repl2, 0, 1);
    JNO_LAUNCH_CHECK();
    rc = read_state(w.st, &e, s);
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
