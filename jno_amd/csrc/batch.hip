// Batched Jacobi-PCG: many independent systems (uneven sizes, each with its own operator, or one
// shared operator with many right-hand sides) solved together with FP32 inner iterations and FP64
// residual replacement.
//
// Every system ("item") runs its own recurrence and its own small state machine on the device.
// A step is the fixed kernel triple (KA, KS, KB) over one flattened table of (item, row block) entries:
//   KA  reduces the item's partial sums left by the previous KB and decides the item's step:
//         ITER     p = z + beta p_old, deferred x update
//         REPLACE  (the FP32 residual dropped by replace_factor^2, met the tolerance or the cap)
//                  x64 += xs + alpha p_old, xs = 0
//         DONE     (after a REPLACE whose true FP64 residual met the tolerance, or at the cap)
//   KS  ITER     q = A32 p, partial p.q
//   KB  ITER     r -= alpha q (alpha from the reduced p.q), partials r.z, r.r
//       REPLACE  r64 = b - A64 x64, rs = (float) r64, partials ||r64||^2, r.z, r.r, b.b
// DIA items run on their Jacobi-scaled operator S A S (S = diag(sqrt(dinv)), unit main diagonal, as
// the single-system mixed solver does): p = r + beta p and r.z = r.r stream no dinv and the SpMV no
// main diagonal (48 instead of 60 bytes per row and step); x = S y and r = S (b - A x) are formed
// in FP64 at the replacements, and the inner loop stops only to replace (the last cycle aimed at
// the remaining reduction), so the FP64 true residual alone decides convergence.  An item whose
// (float)(A_ii dinv_i) is not exactly 1 somewhere, and every CSR item, runs unscaled.
// Each entry reduces its item's partials itself in a fixed order, so no value crosses workgroups
// inside a kernel (no atomics or cache-flushing fences on the step path) and all entries of an
// item take the same decision.  Per-item fields are written by the item's first entry only, into
// fields that kernel does not read.  The step arguments are constant, so a chunk of steps is
// captured once in a hipGraph and replayed; items do not wait for each other.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "common.h"
#include "sparse.h"

namespace jno {

constexpr int BT_NT = 512;
constexpr int BT_RPB = 2048;   // rows per table entry (every entry reduces its item's partials: fewer,
                               // larger entries cut that; 1024 rows measured 9% slower at C4, 4096 4%)
constexpr int BT_GRID = 8192;  // grid-stride workgroups per launch (2048: 8% slower at C4)

enum { PH_ITER = 0, PH_REPLACE = 1, PH_DONE = 2 };
enum { BS_RUNNING = 0, BS_CONVERGED = 1, BS_MAXITER = 2, BS_NONFINITE = 3 };

struct BItem {
  int64_t n, off;
  int32_t format, ndiag;
  int32_t offs[JNO_DIA_MAX];
  const int32_t* rowptr;
  const int32_t* col;
  const float* v32;
  const double* v64;
  const double* b;
  double* x;
  int32_t e0, ne;  // entry range in the table
  const float* s32;      // scaled operator: the ndiag - 1 stored upper diagonals of S A S
  const double* dinv64;  // the item's FP64 Jacobi inverse diagonal
  int32_t scaled, pad;
};

struct BState {
  // written by KA's first entry, read by KB
  int32_t phase_b;
  int32_t done;       // written by KB's first entry (once KA decided DONE), read by KA
  double gamma_cur;   // gamma_k
  // written by KA's first entry when it closes a REPLACE, read by later KAs
  double atol2, rr_replace, rr_true, bb, rr_last;
  int64_t refinements;
  int32_t status;
  // written by KB's first entry, read by KA
  int32_t last_kb;    // kind of the previous KB (its partials)
  int64_t k;
  double alpha, gamma_prev;
};

struct BArgs {
  const BItem* items;
  BState* st;
  const int2* table;  // (item, local block)
  int32_t n_entries;
  float* rs;
  float* p[2];
  float* q;
  float* xs;
  float* dinv;
  double* PA;         // [n_entries]       KA: p.q
  double* PB;         // [4][n_entries]    KB: ITER (r.z, r.r) / REPLACE (rr_true, r.z, r.r, b.b)
  int32_t* active;
  double tol2, atol2_abs, repl2;
  int64_t maxiter;
};

// The scalar fields of an item, held in registers for an entry's whole pass (the kernels store to
// global memory, so fields read through a reference would be reloaded after every store); the
// diagonal offsets stay in memory (a register copy of the array would be indexed dynamically and
// spill to scratch)
struct ItemView {
  int64_t n, off;
  int32_t format, ndiag;
  const int32_t* offs;
  const int32_t* rowptr;
  const int32_t* col;
  const float* v32;
  const double* v64;
  const double* b;
  double* x;
  int32_t e0, ne;
  const float* s32;
  const double* dinv64;
  bool scaled;
};

__device__ __forceinline__ ItemView item_view(const BItem* items, int i) {
  const BItem& s = items[i];
  ItemView v;
  v.n = s.n;
  v.off = s.off;
  v.format = s.format;
  v.ndiag = s.ndiag;
  v.offs = s.offs;
  v.rowptr = s.rowptr;
  v.col = s.col;
  v.v32 = s.v32;
  v.v64 = s.v64;
  v.b = s.b;
  v.x = s.x;
  v.e0 = s.e0;
  v.ne = s.ne;
  v.s32 = s.s32;
  v.dinv64 = s.dinv64;
  v.scaled = s.scaled != 0;
  return v;
}

// row i of the scaled operator S A S (unit main diagonal; the stored upper diagonals of s32, the
// lower ones through symmetry)
__device__ __forceinline__ float scaled_row(const ItemView& it, int64_t i, const float* x) {
  const int64_t ld = dia_ld(it.n);
  float s = x[i];
  for (int d = 1; d < it.ndiag; ++d) {
    const int64_t o = it.offs[d];
    const float* v = it.s32 + (int64_t)(d - 1) * ld;
    if (i + o < it.n) s += v[i] * x[i + o];
    if (i - o >= 0) s += v[i - o] * x[i - o];
  }
  return s;
}

template <typename T, typename V>
__device__ __forceinline__ T mat_row(const ItemView& it, int64_t i, const V* vals, const T* x, bool dia) {
  if (dia) {
    T s = (T)vals[i] * x[i];
    for (int d = 1; d < it.ndiag; ++d) {
      const int64_t o = it.offs[d];
      const V* v = vals + (int64_t)d * dia_ld(it.n);
      if (i + o < it.n) s += (T)v[i] * x[i + o];
      if (i - o >= 0) s += (T)v[i - o] * x[i - o];
    }
    return s;
  }
  T s = 0;
  for (int j = it.rowptr[i]; j < it.rowptr[i + 1]; ++j) s += (T)vals[j] * x[it.col[j]];
  return s;
}

// item total of one partial array, fixed order; every thread gets it
__device__ __forceinline__ double item_sum(const double* __restrict__ P, const ItemView& it, double* sm) {
  double s = 0.0;
  for (int e = threadIdx.x; e < it.ne; e += BT_NT) s += P[it.e0 + e];
  return block_sum<BT_NT>(s, sm);
}

__device__ __forceinline__ void put(double v, double* __restrict__ P, int entry, double* sm) {
  v = block_sum<BT_NT>(v, sm);
  if (threadIdx.x == 0) P[entry] = v;
}

template <bool DIA>
__global__ void __launch_bounds__(BT_NT) batch_ka(BArgs a) {
  __shared__ double sm[BT_NT / 64];
  const int64_t NE = a.n_entries;
  for (int entry = blockIdx.x; entry < a.n_entries; entry += gridDim.x) {
    const int2 te = a.table[entry];
    const ItemView it = item_view(a.items, te.x);
    if ((it.format == JNO_FMT_DIA) != DIA) continue;
    BState* S = a.st + te.x;
    if (S->done) continue;  // entry-uniform
    const bool first = te.y == 0;
    const int64_t k = S->k;
    // ---- decide this step from the previous KB's partials
    int phase;
    double gamma, rr;
    if (S->last_kb == PH_ITER) {
      gamma = item_sum(a.PB, it, sm);
      rr = item_sum(a.PB + NE, it, sm);
      const bool bad = !isfinite(gamma) || !isfinite(rr);
      phase = (bad || rr <= S->rr_replace || !(rr > S->atol2) || k >= a.maxiter) ? PH_REPLACE : PH_ITER;
      if (bad && first && threadIdx.x == 0) S->status = BS_NONFINITE;
    } else {  // the previous KB formed the true residual
      const double rt = item_sum(a.PB, it, sm);
      gamma = item_sum(a.PB + NE, it, sm);
      rr = item_sum(a.PB + 2 * NE, it, sm);
      const double bb = item_sum(a.PB + 3 * NE, it, sm);
      const double atol2 = fmax(a.tol2 * bb, a.atol2_abs);
      // a non-finite true residual (e.g. from a NaN in b) stops the item as non-finite: !(rt > atol2)
      // alone would take a NaN for convergence
      const bool bad = !isfinite(rt) || S->status == BS_NONFINITE;
      phase = (bad || !(rt > atol2) || k >= a.maxiter) ? PH_DONE : PH_ITER;
      if (first && threadIdx.x == 0) {
        // scaled items: the inner loop stops only to replace, the last cycle aimed at the
        // remaining reduction (with margin) when that is smaller than a whole cycle
        const double need = atol2 / rt;
        S->atol2 = it.scaled ? 0.0 : atol2;
        S->rr_replace = (it.scaled && need > a.repl2) ? 0.5 * need * rr : a.repl2 * rr;
        S->rr_true = rt;
        S->bb = bb;
        S->rr_last = rr;
        S->refinements = S->refinements + 1;
        if (phase == PH_DONE) {
          S->status = bad ? BS_NONFINITE : !(rt > atol2) ? BS_CONVERGED : BS_MAXITER;
          atomicSub(a.active, 1);
        }
      }
    }
    if (first && threadIdx.x == 0) {
      S->phase_b = phase;
      S->gamma_cur = gamma;
    }
    if (phase == PH_DONE) continue;
    const int64_t r0 = (int64_t)te.y * BT_RPB, r1 = min(it.n, r0 + BT_RPB);
    const float* pold = a.p[k & 1] + it.off;
    float* xs = a.xs + it.off;
    const double alpha_prev = S->alpha;
    const float ap = (float)alpha_prev;
    if (phase == PH_REPLACE) {
      for (int64_t i = r0 + threadIdx.x; i < r1; i += BT_NT) {
        float d = xs[i];
        if (k > 0 && alpha_prev != 0.0) d += ap * pold[i];
        it.x[i] += it.scaled ? sqrt(it.dinv64[i]) * (double)d : (double)d;  // x = S y when scaled
        xs[i] = 0.0f;
      }
      continue;
    }
    float* pnew = a.p[(k + 1) & 1] + it.off;
    const float beta = (k == 0) ? 0.0f : (float)(gamma / S->gamma_prev);
    const float* r = a.rs + it.off;
    const float* dv = a.dinv + it.off;
    // p = z + beta p_old and the deferred x update, one 4-row tile per thread (16-byte accesses)
    static_assert(BT_RPB == 4 * BT_NT, "one 4-row tile per thread");
    const int64_t t0 = r0 + 4 * (int64_t)threadIdx.x;
    if (t0 + 4 <= r1) {
      const float4 dq = it.scaled ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : *reinterpret_cast<const float4*>(dv + t0);
      const float4 rq = *reinterpret_cast<const float4*>(r + t0);
      const float4 po = *reinterpret_cast<const float4*>(pold + t0);
      float4 pi;
      if (it.scaled) {  // z = r
        pi.x = rq.x + beta * po.x;
        pi.y = rq.y + beta * po.y;
        pi.z = rq.z + beta * po.z;
        pi.w = rq.w + beta * po.w;
      } else {
        pi.x = dq.x * rq.x + beta * po.x;
        pi.y = dq.y * rq.y + beta * po.y;
        pi.z = dq.z * rq.z + beta * po.z;
        pi.w = dq.w * rq.w + beta * po.w;
      }
      *reinterpret_cast<float4*>(pnew + t0) = pi;
      if (k > 0) {
        float4 xq = *reinterpret_cast<const float4*>(xs + t0);
        xq.x += ap * po.x;
        xq.y += ap * po.y;
        xq.z += ap * po.z;
        xq.w += ap * po.w;
        *reinterpret_cast<float4*>(xs + t0) = xq;
      }
    } else {
      for (int64_t i = t0; i < min(r1, t0 + 4); ++i) {
        const float po = pold[i];
        pnew[i] = (it.scaled ? r[i] : dv[i] * r[i]) + beta * po;
        if (k > 0) xs[i] += ap * po;
      }
    }
  }
}

// KS: q = A32 p for the items KA set to ITER, partial p.q
template <bool DIA>
__global__ void __launch_bounds__(BT_NT) batch_ks(BArgs a) {
  __shared__ double sm[BT_NT / 64];
  for (int entry = blockIdx.x; entry < a.n_entries; entry += gridDim.x) {
    const int2 te = a.table[entry];
    const ItemView it = item_view(a.items, te.x);
    if ((it.format == JNO_FMT_DIA) != DIA) continue;
    const BState* S = a.st + te.x;
    if (S->phase_b != PH_ITER) continue;  // entry-uniform
    const int64_t k = S->k;
    const float* p = a.p[(k + 1) & 1] + it.off;
    float* q = a.q + it.off;
    const int64_t r0 = (int64_t)te.y * BT_RPB, r1 = min(it.n, r0 + BT_RPB);
    double part = 0.0;
    int64_t i0 = r0 + threadIdx.x;
    if (DIA && it.scaled && it.ndiag == 4) {
      // the 3-D Kuhn lattice on the scaled operator: unit diagonal, 3 stored upper diagonals
      constexpr int BR = 2;
      const int64_t oo[3] = {it.offs[1], it.offs[2], it.offs[3]};
      const int64_t omax = max(max(oo[0], oo[1]), oo[2]);
      const int64_t ld = dia_ld(it.n);
      for (; i0 + (BR - 1) * BT_NT < r1 && i0 - omax >= 0 && i0 + (BR - 1) * BT_NT + omax < it.n; i0 += BR * BT_NT) {
        float pv[BR], vu[BR][3], vl[BR][3], pu[BR][3], pl[BR][3];
#pragma unroll
        for (int t = 0; t < BR; ++t) {
          const int64_t i = i0 + t * BT_NT;
          pv[t] = p[i];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const int64_t o = oo[d];
            const float* v = it.s32 + (int64_t)d * ld;
            vu[t][d] = v[i];
            vl[t][d] = v[i - o];
            pu[t][d] = p[i + o];
            pl[t][d] = p[i - o];
          }
        }
#pragma unroll
        for (int t = 0; t < BR; ++t) {
          float sacc = pv[t];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            sacc += vu[t][d] * pu[t][d];
            sacc += vl[t][d] * pl[t][d];
          }
          q[i0 + t * BT_NT] = sacc;
          part += (double)pv[t] * (double)sacc;
        }
      }
    } else if (DIA && it.ndiag == 4) {
      // the 3-D Kuhn lattice: BR rows per thread with all 14 BR loads in flight before the first use
      constexpr int BR = 2;
      const int64_t oo[3] = {it.offs[1], it.offs[2], it.offs[3]};
      const int64_t omax = max(max(oo[0], oo[1]), oo[2]);
      const int64_t ld = dia_ld(it.n);
      for (; i0 + (BR - 1) * BT_NT < r1 && i0 - omax >= 0 && i0 + (BR - 1) * BT_NT + omax < it.n; i0 += BR * BT_NT) {
        float pv[BR], d0[BR], vu[BR][3], vl[BR][3], pu[BR][3], pl[BR][3];
#pragma unroll
        for (int t = 0; t < BR; ++t) {
          const int64_t i = i0 + t * BT_NT;
          pv[t] = p[i];
          d0[t] = it.v32[i];
#pragma unroll
          for (int d = 1; d < 4; ++d) {
            const int64_t o = oo[d - 1];
            const float* v = it.v32 + (int64_t)d * ld;
            vu[t][d - 1] = v[i];
            vl[t][d - 1] = v[i - o];
            pu[t][d - 1] = p[i + o];
            pl[t][d - 1] = p[i - o];
          }
        }
#pragma unroll
        for (int t = 0; t < BR; ++t) {
          float sacc = d0[t] * pv[t];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            sacc += vu[t][d] * pu[t][d];
            sacc += vl[t][d] * pl[t][d];
          }
          q[i0 + t * BT_NT] = sacc;
          part += (double)pv[t] * (double)sacc;
        }
      }
    }
    for (int64_t i = i0; i < r1; i += BT_NT) {
      const float s = (DIA && it.scaled) ? scaled_row(it, i, p) : mat_row<float, float>(it, i, it.v32, p, DIA);
      q[i] = s;
      part += (double)p[i] * (double)s;
    }
    put(part, a.PA, entry, sm);
  }
}

template <bool DIA>
__global__ void __launch_bounds__(BT_NT) batch_kb(BArgs a) {
  __shared__ double sm[BT_NT / 64];
  const int64_t NE = a.n_entries;
  for (int entry = blockIdx.x; entry < a.n_entries; entry += gridDim.x) {
    const int2 te = a.table[entry];
    const ItemView it = item_view(a.items, te.x);
    if ((it.format == JNO_FMT_DIA) != DIA) continue;
    BState* S = a.st + te.x;
    const int phase = S->phase_b;
    const bool first = te.y == 0;
    if (phase == PH_DONE) {
      // KA never writes `done`, so its fast exit reads a value that is stable during KA
      if (first && threadIdx.x == 0) S->done = 1;
      continue;
    }
    const int64_t r0 = (int64_t)te.y * BT_RPB, r1 = min(it.n, r0 + BT_RPB);
    float* r = a.rs + it.off;
    const float* dv = a.dinv + it.off;
    if (phase == PH_ITER) {
      const double pq = item_sum(a.PA, it, sm);
      const double gamma = S->gamma_cur;
      const double alpha = gamma / pq;
      const float al = (float)alpha;
      const float* q = a.q + it.off;
      double pg = 0, pr = 0;
      // an entry is at most BT_RPB = 4 * BT_NT rows: one 4-row tile per thread, 16-byte accesses
      static_assert(BT_RPB == 4 * BT_NT, "one 4-row tile per thread");
      const int64_t t0 = r0 + 4 * (int64_t)threadIdx.x;
      if (t0 + 4 <= r1) {
        float4 rv = *reinterpret_cast<const float4*>(r + t0);
        const float4 qv = *reinterpret_cast<const float4*>(q + t0);
        // z = r on the scaled operator: r.z = r.r
        const float4 dq = it.scaled ? make_float4(1.0f, 1.0f, 1.0f, 1.0f) : *reinterpret_cast<const float4*>(dv + t0);
        rv.x = rv.x - al * qv.x;
        rv.y = rv.y - al * qv.y;
        rv.z = rv.z - al * qv.z;
        rv.w = rv.w - al * qv.w;
        *reinterpret_cast<float4*>(r + t0) = rv;
        pg += (double)rv.x * (double)(dq.x * rv.x);
        pg += (double)rv.y * (double)(dq.y * rv.y);
        pg += (double)rv.z * (double)(dq.z * rv.z);
        pg += (double)rv.w * (double)(dq.w * rv.w);
        pr += (double)rv.x * (double)rv.x;
        pr += (double)rv.y * (double)rv.y;
        pr += (double)rv.z * (double)rv.z;
        pr += (double)rv.w * (double)rv.w;
      } else {
        for (int64_t i = t0; i < min(r1, t0 + 4); ++i) {
          const float ri = r[i] - al * q[i];
          r[i] = ri;
          pg += (double)ri * (double)(it.scaled ? ri : dv[i] * ri);
          pr += (double)ri * (double)ri;
        }
      }
      put(pg, a.PB, entry, sm);
      put(pr, a.PB + NE, entry, sm);
      if (first && threadIdx.x == 0) {
        S->alpha = alpha;
        S->gamma_prev = gamma;
        S->k = S->k + 1;
        S->last_kb = PH_ITER;
      }
    } else {  // PH_REPLACE: true residual in FP64
      double prt = 0, pg = 0, pr = 0, pb = 0;
      for (int64_t i = r0 + threadIdx.x; i < r1; i += BT_NT) {
        const double bi = it.b[i];
        const double ri = bi - mat_row<double, double>(it, i, it.v64, it.x, DIA);
        const float rf = it.scaled ? (float)(ri * sqrt(it.dinv64[i])) : (float)ri;  // r = S r64 when scaled
        r[i] = rf;
        prt += ri * ri;
        pb += bi * bi;
        pg += (double)rf * (double)(it.scaled ? rf : dv[i] * rf);
        pr += (double)rf * (double)rf;
      }
      put(prt, a.PB, entry, sm);
      put(pg, a.PB + NE, entry, sm);
      put(pr, a.PB + 2 * NE, entry, sm);
      put(pb, a.PB + 3 * NE, entry, sm);
      if (first && threadIdx.x == 0) {
        S->alpha = 0.0;  // the deferred x update was applied by KA
        S->last_kb = PH_REPLACE;
      }
    }
  }
}

__global__ void batch_init_k(BArgs a, int n_items) {
  // per entry: zero xs, p, x
  for (int entry = blockIdx.x; entry < a.n_entries; entry += gridDim.x) {
    const int2 te = a.table[entry];
    const BItem& it = a.items[te.x];
    const int64_t r0 = (int64_t)te.y * BT_RPB, r1 = min(it.n, r0 + BT_RPB);
    for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
      it.x[i] = 0.0;
      a.xs[it.off + i] = 0.0f;
      a.p[0][it.off + i] = 0.0f;
      a.p[1][it.off + i] = 0.0f;
    }
  }
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n_items) {
    BState s = {};
    s.phase_b = PH_REPLACE;  // the first KB forms r = b, the norms and the tolerances
    s.last_kb = PH_REPLACE;
    s.status = BS_RUNNING;
    s.refinements = -1;
    a.st[g] = s;
  }
}

__global__ void batch_dinv_k(const BItem* items, const double* const* dinvs, const int2* table, int n_entries,
                             float* dinv32) {
  for (int entry = blockIdx.x; entry < n_entries; entry += gridDim.x) {
    const int2 te = table[entry];
    const BItem& it = items[te.x];
    const double* d = dinvs[te.x];
    const int64_t r0 = (int64_t)te.y * BT_RPB, r1 = min(it.n, r0 + BT_RPB);
    for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) dinv32[it.off + i] = (float)d[i];
  }
}

// Scaled operator of the DIA items, per entry's rows: s32[d-1][i] = (float)(A[i, i+o_d] s_i s_{i+o_d})
// (s = sqrt(dinv)) for the stored upper diagonals d >= 1; bad[item] is set where (float)(A_ii dinv_i)
// is not exactly 1 (that item then runs unscaled)
__global__ void batch_scale_k(const BItem* items, const double* const* dinvs, const int2* table, int n_entries,
                              int* bad) {
  for (int entry = blockIdx.x; entry < n_entries; entry += gridDim.x) {
    const int2 te = table[entry];
    const BItem& it = items[te.x];
    if (it.format != JNO_FMT_DIA || !it.s32) continue;
    const double* dv = dinvs[te.x];
    const int64_t ld = dia_ld(it.n);
    float* s32 = const_cast<float*>(it.s32);
    const int64_t r0 = (int64_t)te.y * BT_RPB, r1 = min(it.n, r0 + BT_RPB);
    bool ok = true;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
      ok = ok && (float)(it.v64[i] * dv[i]) == 1.0f;
      const double si = sqrt(dv[i]);
      for (int d = 1; d < it.ndiag; ++d) {
        const int64_t o = it.offs[d];
        s32[(int64_t)(d - 1) * ld + i] = i + o < it.n ? (float)(it.v64[(int64_t)d * ld + i] * si * sqrt(dv[i + o])) : 0.0f;
      }
    }
    if (!ok) atomicOr(&bad[te.x], 1);
  }
}

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

struct BLayout {
  BItem* items;
  BState* st;
  int2* table;
  const double** dinvs;
  double* PA;
  double* PB;
  int32_t* active;
  int32_t* bad;  // [n_items]: the item's operator cannot be Jacobi-scaled exactly
  float *rs, *p0, *p1, *q, *xs, *dinv;
  float* sdia;   // scaled upper diagonals of the DIA items, one block per item
  size_t bytes;
};

static BLayout blayout(int n_items, int64_t n_entries, int64_t ntot, int64_t nsd, char* base) {
  BLayout L{};
  size_t off = 0;
  auto take = [&](size_t b) -> char* {
    char* p = base ? base + off : nullptr;
    off += a256(b);
    return p;
  };
  L.items = (BItem*)take(sizeof(BItem) * n_items);
  L.st = (BState*)take(sizeof(BState) * n_items);
  L.table = (int2*)take(sizeof(int2) * n_entries);
  L.dinvs = (const double**)take(sizeof(double*) * n_items);
  L.PA = (double*)take(sizeof(double) * n_entries);
  L.PB = (double*)take(sizeof(double) * 4 * n_entries);
  L.active = (int32_t*)take(256);
  L.rs = (float*)take(4 * ntot);
  L.p0 = (float*)take(4 * ntot);
  L.p1 = (float*)take(4 * ntot);
  L.q = (float*)take(4 * ntot);
  L.xs = (float*)take(4 * ntot);
  L.dinv = (float*)take(4 * ntot);
  L.bad = (int32_t*)take(sizeof(int32_t) * n_items);
  L.sdia = (float*)take(4 * nsd);
  L.bytes = off;
  return L;
}

// FP32 entries of the scaled copies of the DIA items' stored upper diagonals
static int64_t scaled_rows(const jno_op& op) { return op.format == JNO_FMT_DIA ? (op.ndiag - 1) * dia_ld(op.n) : 0; }

// each item's rows start 64-aligned in the concatenated FP32 vectors (16-byte vector tiles)
static int64_t padded_rows(int64_t n) { return (n + 63) & ~(int64_t)63; }

static int64_t count_entries(int n, const jno_batch_item* items, int64_t* ntot, int64_t* nsd) {
  int64_t ne = 0, nt = 0, ns = 0;
  for (int i = 0; i < n; ++i) {
    ne += cdiv(items[i].op.n, BT_RPB);
    nt += padded_rows(items[i].op.n);
    ns += scaled_rows(items[i].op);
  }
  if (ntot) *ntot = nt;
  if (nsd) *nsd = ns;
  return ne;
}

// Pinned slots and events of the host's polls of the active-item count: created once per thread
// and reused by every batched solve of that thread.
struct BatchPoll {
  int32_t* h = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ok = false;
  int init() {
    if (ok) return JNO_OK;
    JNO_HIP(hipHostMalloc((void**)&h, 2 * sizeof(int32_t), hipHostMallocDefault));
    JNO_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    JNO_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    ok = true;
    return JNO_OK;
  }
};
static thread_local BatchPoll g_bpoll;

// The per-call graph of a batched solve and its capture stream, released on every return path
// (after the solve's stream has drained, so no queued replay still refers to them).
struct BatchGraph {
  hipStream_t s = nullptr;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~BatchGraph() {
    if (exec || graph || cap) (void)hipStreamSynchronize(s);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

}  // namespace jno

using namespace jno;

extern "C" size_t jno_batch_workspace_size(int n_items, const jno_batch_item* h_items) {
  if (n_items <= 0 || !h_items) return 0;
  int64_t ntot, nsd;
  const int64_t ne = count_entries(n_items, h_items, &ntot, &nsd);
  return blayout(n_items, ne, ntot, nsd, nullptr).bytes;
}

extern "C" int jno_batch_pcg_mixed(int n_items, const jno_batch_item* h_items, double tol, double atol, int64_t maxiter,
                                   double replace_factor, int steps_per_graph, jno_pcg_stats* h_stats, void* d_ws,
                                   size_t ws_bytes, void* stream) {
  JNO_REQUIRE(n_items > 0 && h_items && h_stats, JNO_EINVAL, "jno_batch_pcg_mixed: empty batch");
  JNO_REQUIRE(replace_factor > 0.0 && replace_factor < 1.0, JNO_EINVAL, "jno_batch_pcg_mixed: replace_factor in (0,1)");
  JNO_REQUIRE(maxiter >= 0 && steps_per_graph >= 0, JNO_EINVAL, "jno_batch_pcg_mixed: bad maxiter/steps");
  int64_t ntot, nsd;
  const int64_t ne = count_entries(n_items, h_items, &ntot, &nsd);
  JNO_REQUIRE(ne < INT32_MAX, JNO_EINVAL, "jno_batch_pcg_mixed: batch too large");
  JNO_REQUIRE(ws_bytes >= jno_batch_workspace_size(n_items, h_items), JNO_EWORKSPACE,
              "jno_batch_pcg_mixed: workspace too small");
  hipStream_t s = as_stream(stream);
  BLayout L = blayout(n_items, ne, ntot, nsd, static_cast<char*>(d_ws));
  // host staging of the item descriptors and the entry table
  std::vector<BItem> items(n_items);
  std::vector<int2> table(ne);
  std::vector<const double*> dinvs(n_items);
  int64_t off = 0, e = 0, soff = 0;
  bool any_dia = false, any_csr = false;
  for (int i = 0; i < n_items; ++i) {
    const jno_batch_item& h = h_items[i];
    JNO_REQUIRE(h.op.vals32 && h.op.vals64 && h.dinv && h.b && h.x, JNO_EINVAL,
                "jno_batch_pcg_mixed: item %d lacks FP32/FP64 values or vectors", i);
    JNO_REQUIRE(h.op.format == JNO_FMT_DIA || (h.op.format == JNO_FMT_CSR && h.op.rowptr && h.op.col), JNO_EINVAL,
                "jno_batch_pcg_mixed: item %d has a bad operator", i);
    JNO_REQUIRE(h.op.n > 0, JNO_EINVAL, "jno_batch_pcg_mixed: item %d is empty", i);
    BItem& it = items[i];
    it.n = h.op.n;
    it.off = off;
    it.format = h.op.format;
    it.ndiag = h.op.ndiag;
    for (int d = 0; d < JNO_DIA_MAX; ++d) it.offs[d] = h.op.offsets[d];
    it.rowptr = h.op.rowptr;
    it.col = h.op.col;
    it.v32 = h.op.vals32;
    it.v64 = h.op.vals64;
    it.b = h.b;
    it.x = h.x;
    it.e0 = (int32_t)e;
    it.ne = (int32_t)cdiv(it.n, BT_RPB);
    for (int j = 0; j < it.ne; ++j) table[e++] = make_int2(i, j);
    dinvs[i] = h.dinv;
    it.dinv64 = h.dinv;
    it.s32 = it.format == JNO_FMT_DIA ? L.sdia + soff : nullptr;
    it.scaled = 0;
    soff += scaled_rows(h.op);
    off += padded_rows(it.n);
    any_dia |= it.format == JNO_FMT_DIA;
    any_csr |= it.format == JNO_FMT_CSR;
  }
  JNO_HIP(hipMemcpyAsync(L.items, items.data(), sizeof(BItem) * n_items, hipMemcpyHostToDevice, s));
  JNO_HIP(hipMemcpyAsync(L.table, table.data(), sizeof(int2) * ne, hipMemcpyHostToDevice, s));
  JNO_HIP(hipMemcpyAsync(L.dinvs, dinvs.data(), sizeof(double*) * n_items, hipMemcpyHostToDevice, s));
  const int32_t h_active = n_items;
  JNO_HIP(hipMemcpyAsync(L.active, &h_active, sizeof(int32_t), hipMemcpyHostToDevice, s));
  BArgs a;
  a.items = L.items;
  a.st = L.st;
  a.table = L.table;
  a.n_entries = (int32_t)ne;
  a.rs = L.rs;
  a.p[0] = L.p0;
  a.p[1] = L.p1;
  a.q = L.q;
  a.xs = L.xs;
  a.dinv = L.dinv;
  a.PA = L.PA;
  a.PB = L.PB;
  a.active = L.active;
  a.tol2 = tol * tol;
  a.atol2_abs = atol * atol;
  a.repl2 = replace_factor * replace_factor;
  a.maxiter = maxiter;
  const int grid = (int)std::min<int64_t>(ne, BT_GRID);
  batch_init_k<<<(unsigned)std::max<int64_t>(grid, cdiv(n_items, BT_NT)), BT_NT, 0, s>>>(a, n_items);
  JNO_LAUNCH_CHECK();
  batch_dinv_k<<<grid, BT_NT, 0, s>>>(L.items, L.dinvs, L.table, (int)ne, L.dinv);
  JNO_LAUNCH_CHECK();
  if (any_dia) {
    // the DIA items' scaled operators; an item whose Jacobi scaling is not exact runs unscaled
    JNO_HIP(hipMemsetAsync(L.bad, 0, sizeof(int32_t) * n_items, s));
    batch_scale_k<<<grid, BT_NT, 0, s>>>(L.items, L.dinvs, L.table, (int)ne, L.bad);
    JNO_LAUNCH_CHECK();
    std::vector<int32_t> bad(n_items);
    JNO_HIP(hipMemcpyAsync(bad.data(), L.bad, sizeof(int32_t) * n_items, hipMemcpyDeviceToHost, s));
    JNO_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n_items; ++i) items[i].scaled = items[i].format == JNO_FMT_DIA && !bad[i];
    JNO_HIP(hipMemcpyAsync(L.items, items.data(), sizeof(BItem) * n_items, hipMemcpyHostToDevice, s));
  }
  JNO_HIP(hipStreamSynchronize(s));  // host staging vectors go out of scope after this call

  auto launch_kb = [&](hipStream_t st) -> int {
    if (any_dia) batch_kb<true><<<grid, BT_NT, 0, st>>>(a);
    if (any_csr) batch_kb<false><<<grid, BT_NT, 0, st>>>(a);
    JNO_LAUNCH_CHECK();
    return JNO_OK;
  };
  auto launch_step = [&](hipStream_t st) -> int {
    if (any_dia) batch_ka<true><<<grid, BT_NT, 0, st>>>(a);
    if (any_csr) batch_ka<false><<<grid, BT_NT, 0, st>>>(a);
    if (any_dia) batch_ks<true><<<grid, BT_NT, 0, st>>>(a);
    if (any_csr) batch_ks<false><<<grid, BT_NT, 0, st>>>(a);
    JNO_LAUNCH_CHECK();
    return launch_kb(st);
  };

  int rc = g_bpoll.init();
  if (rc) return rc;
  int32_t* h_act = g_bpoll.h;
  hipEvent_t* ev = g_bpoll.ev;
  rc = launch_kb(s);  // the first true residual r = b
  if (rc) return rc;
  BatchGraph bg;
  bg.s = s;
  const int chunk = steps_per_graph > 0 ? steps_per_graph : 8;
  if (steps_per_graph > 0) {
    // capture one chunk of steps on a private stream (default priority), replay it on `s`
    JNO_HIP(hipStreamCreateWithFlags(&bg.cap, hipStreamNonBlocking));
    JNO_HIP(hipStreamBeginCapture(bg.cap, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk && rc == JNO_OK; ++i) rc = launch_step(bg.cap);
    hipError_t ce = hipStreamEndCapture(bg.cap, &bg.graph);
    if (rc) return rc;
    JNO_HIP(ce);
    JNO_HIP(hipGraphInstantiate(&bg.exec, bg.graph, nullptr, nullptr, 0));
  }
  hipGraphExec_t exec = bg.exec;
  const auto t0 = std::chrono::steady_clock::now();
  int slot = 0, pending = 0;
  int64_t launched = 0;
  // every item is DONE after at most maxiter iterations plus one replacement per iteration
  const int64_t max_steps = 2 * (maxiter + 2) + 4 * (int64_t)chunk;
  for (;;) {
    if (exec) {
      JNO_HIP(hipGraphLaunch(exec, s));
    } else {
      for (int i = 0; i < chunk; ++i) {
        rc = launch_step(s);
        if (rc) return rc;
      }
    }
    launched += chunk;
    JNO_HIP(hipMemcpyAsync(&h_act[slot], L.active, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    JNO_HIP(hipEventRecord(ev[slot], s));
    if (pending) {
      JNO_HIP(hipEventSynchronize(ev[slot ^ 1]));
      if (h_act[slot ^ 1] <= 0) break;
    }
    if (launched > max_steps) break;
    pending = 1;
    slot ^= 1;
  }
  JNO_HIP(hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::vector<BState> st(n_items);
  JNO_HIP(hipMemcpyAsync(st.data(), L.st, sizeof(BState) * n_items, hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  int worst = JNO_OK;
  for (int i = 0; i < n_items; ++i) {
    const double bn = sqrt(st[i].bb);
    h_stats[i].iterations = st[i].k;
    h_stats[i].refinements = st[i].refinements;
    h_stats[i].relres = bn > 0 ? sqrt(st[i].rr_true) / bn : sqrt(st[i].rr_true);
    h_stats[i].relres_recursive = bn > 0 ? sqrt(st[i].rr_last) / bn : sqrt(st[i].rr_last);
    h_stats[i].seconds = secs;
    if (st[i].status == BS_NONFINITE) worst = JNO_ENONFINITE;
    else if (st[i].status != BS_CONVERGED && worst == JNO_OK) worst = JNO_ENOTCONV;
  }
  if (worst == JNO_ENONFINITE) set_error("jno_batch_pcg_mixed: a non-finite value in at least one system");
  if (worst == JNO_ENOTCONV) set_error("jno_batch_pcg_mixed: at least one system did not converge in %lld iterations",
                                       (long long)maxiter);
  return worst;
}
