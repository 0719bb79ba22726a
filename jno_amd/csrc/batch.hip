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
#include "pcg_common.h"
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

// ---------------------------------------------------------------------------------------------
// One kernel per step (batch_one_k): the single-system one-kernel iteration carried over to the
// batch, for items on a Jacobi-scaled 4-diagonal DIA operator (the 3-D Kuhn lattice; near offsets
// o1 < o2 <= B1_HALO, far o3).  Launch k of an item reads r_{k-1}, q_{k-1} = A p_{k-1}, p_{k-1} and
// the partials of its previous launch (r.r, p.q, q.r, q.q); alpha = gamma_{k-1} / p.q,
// gamma_k = gamma_{k-1} - 2 alpha q.r + alpha^2 q.q, beta = gamma_k / gamma_{k-1}; then per row
// r_k = r_{k-1} - alpha q_{k-1}, xs += alpha p_{k-1} (applied every other launch: b1_body_q),
// p_k = r_k + beta p_{k-1} (near neighbours
// through LDS, the far ones recomputed from r, q, p), q_k = A p_k and the partials of r_k.r_k,
// p_k.q_k, q_k.r_k, q_k.q_k.  r, q and p alternate between two buffers by the parity of k.
// An item's phases, one launch each:
//   ITER  as above; when r.r has fallen to the replacement threshold (or at the cap, or on a
//         non-finite scalar) the launch instead folds xs into x64 (x += S xs, xs = 0): phase R2
//   R2    r64 = b - A64 x64 in FP64, r = S r64 into the buffer the resumed launch reads; partials
//         r.r (scaled), ||r64||^2, b.b
//   RES   decides on the true residual (converged, cap, non-finite: DONE); otherwise runs launch
//         kr = k_stop - 1 again with alpha = 0 and beta = gamma_true / gamma_{kr-1} on p_{kr-1},
//         which the stopped launch left intact, so the search direction survives the replacement
//         (kr = 0: the start, beta = 0)
//   DONE  nothing (the state is carried to the other buffer)
// The first launch is R2 from x64 = 0.  Per-item state and partial sums are double-buffered by the
// launch parity (a kernel argument: launches alternate 0, 1, 0, ...), so a launch reads only what
// the previous launch wrote; every entry of an item derives the same decision from the same
// values, and the item's first entry stores the next state.  No value crosses workgroups inside a
// launch (no atomics on the step path apart from the active-item count, no fences).
// ---------------------------------------------------------------------------------------------
constexpr int B1_NT = 256;
constexpr int B1_HALO = 128;           // near halo rows staged in LDS (o2 of a Kuhn lattice up to 127 cells)
constexpr int B1_RPE_UNIT = 2048;      // rows per table entry: a multiple of this (whole tiles of 4 B1_NT rows)
constexpr int B1_RPE = 4096;           // default rows per table entry
constexpr int B1_GRID = 4096;          // default workgroups per launch (C4: 106.7 ms against 111.8 at 2048)

enum { P1_R2 = 0, P1_RES = 1, P1_ITER = 2, P1_DONE = 3 };

struct B1State {
  int32_t phase, status, bad;
  float a_pend;         // the alpha whose xs term the item's previous (deferring) launch left pending
  int64_t k;            // ITER: the index of this launch
  int64_t kr;           // R2 / RES: the launch a resume runs (iterations done)
  int64_t refinements;
  double g_hist[2];     // g_hist[k & 1] = gamma_{k-1}, written by the first run of launch k
  double rr_replace, rr_true, bb, rr_last;
};

struct B1Args {
  const BItem* items;
  B1State* st[2];     // by launch parity
  const int2* table;  // (item, local entry)
  int32_t n_entries;
  int32_t rpe;        // rows per entry (a multiple of B1_RPE_UNIT)
  float* r[2];
  float* q[2];
  float* p[2];
  float* xs;
  double* P[2];       // [4][n_entries] by launch parity
  int32_t* active;
  double tol2, atol2_abs, repl2;
  int64_t maxiter;
};

// Values uniform across the workgroup moved to scalar registers.  The compiler loads the table
// entry, the item descriptor and the state with vector loads (the kernel stores to global memory,
// so it cannot prove them unclobbered), which would hold some 50 vector registers per lane.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
  const int32_t lo = __builtin_amdgcn_readfirstlane((int32_t)(v & 0xffffffffLL));
  const int32_t hi = __builtin_amdgcn_readfirstlane((int32_t)(v >> 32));
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}
__device__ __forceinline__ double uni(double v) { return __longlong_as_double(uni((int64_t)__double_as_longlong(v))); }
template <typename T>
__device__ __forceinline__ T* uni(T* p) {
  return reinterpret_cast<T*>(uni((int64_t) reinterpret_cast<intptr_t>(p)));
}

__device__ __forceinline__ ItemView item_view_uni(const BItem* items, int i) {
  ItemView v = item_view(items, i);
  v.n = uni(v.n);
  v.off = uni(v.off);
  v.format = uni(v.format);
  v.ndiag = uni(v.ndiag);
  v.offs = uni(v.offs);
  v.rowptr = uni(v.rowptr);
  v.col = uni(v.col);
  v.v32 = uni(v.v32);
  v.v64 = uni(v.v64);
  v.b = uni(v.b);
  v.x = uni(v.x);
  v.e0 = uni(v.e0);
  v.ne = uni(v.ne);
  v.s32 = uni(v.s32);
  v.dinv64 = uni(v.dinv64);
  v.scaled = uni((int32_t)v.scaled) != 0;
  return v;
}

__device__ __forceinline__ B1State state_uni(const B1State* S) {
  B1State s = *S;
  s.phase = uni(s.phase);
  s.status = uni(s.status);
  s.bad = uni(s.bad);
  s.a_pend = __int_as_float(uni(__float_as_int(s.a_pend)));
  s.k = uni(s.k);
  s.kr = uni(s.kr);
  s.refinements = uni(s.refinements);
  s.g_hist[0] = uni(s.g_hist[0]);
  s.g_hist[1] = uni(s.g_hist[1]);
  s.rr_replace = uni(s.rr_replace);
  s.rr_true = uni(s.rr_true);
  s.bb = uni(s.bb);
  s.rr_last = uni(s.rr_last);
  return s;
}

// item totals of the 4 partial arrays P[m * NE + e0 + j], j < ne, in a fixed order, identical in
// every thread: array m is summed by the threads m NT/4 .. (m+1) NT/4 - 1, then their waves' sums in
// wave order
template <int NT>
__device__ __forceinline__ void b1_item_totals(const double* __restrict__ P, int64_t NE, int e0, int ne,
                                               double* sm, double* out) {
  constexpr int GS = NT / 4, WPG = GS / 64;
  const int m = threadIdx.x / GS, l = threadIdx.x % GS;
  const double* src = P + (int64_t)m * NE + e0;
  double s = 0.0;
  for (int j = l; j < ne; j += GS) s += src[j];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < WPG; ++w) t += sm[mm * WPG + w];
    out[mm] = t;
  }
  __syncthreads();  // sm is reused
}

// the entry's 4 partials into P[m * NE + entry]
template <int NT>
__device__ __forceinline__ void b1_store_partials(const double* v, double* __restrict__ P, int64_t NE, int entry,
                                                  double* sm) {
  constexpr int NW = NT / 64;
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const double s = wave_sum(v[m]);
    if ((threadIdx.x & 63) == 0) sm[m * NW + w] = s;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) t += sm[threadIdx.x * NW + ww];
    P[(int64_t)threadIdx.x * NE + entry] = t;
  }
  __syncthreads();
}

// a row of an item's vector through a 32-bit byte offset from the item's (uniform) base: the loads
// then take the scalar-base addressing form, one 32-bit offset register each instead of a 64-bit
// address (item rows are below 2^29, checked on the host)
__device__ __forceinline__ float ldr(const float* __restrict__ base, uint32_t row) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (row << 2));
}
__device__ __forceinline__ void str(float* __restrict__ base, uint32_t row, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + (row << 2)) = v;
}

// ldr / str for 16-byte quads of rows: unaligned quads (the neighbours at odd offsets) are served
// by gfx950's vector loads at 4-byte alignment; the base is the item's vector shifted 4 rows down,
// so quads reaching up to 3 rows before the item's first row keep an unsigned offset
typedef float F4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ F4 ldq(const float* __restrict__ base4, uint32_t row4) {
  F4 r;
  __builtin_memcpy(&r, reinterpret_cast<const char*>(base4) + (row4 << 2), 16);
  return r;
}
__device__ __forceinline__ void stq(float* __restrict__ base4, uint32_t row4, F4 v) {
  __builtin_memcpy(reinterpret_cast<char*>(base4) + (row4 << 2), &v, 16);
}

// The same launch with register-resident quads: a thread owns 4 consecutive rows (16-byte loads and
// stores) and a tile is one quad per thread, 4 NT rows.  Items sit back to back in the vectors
// (64-row aligned), so a quad at an item's first or last rows reads rows of the neighbouring item
// or of the padding, which may hold anything, NaN included: every term of such a row is selected
// after its product (a coefficient selected to 0 times a NaN would still be NaN).  Per row the sum
// runs p, then the upper and lower terms of each off-diagonal in ascending offset order.  Rows of
// the last quad past the item's end are stored as 0.  Every load
// stays in the workspace: at most 3 rows before an item (the buffer before the first) or 3 past its
// padded end (the buffer after the last).
// xs is updated every other launch of an item: a launch with even kk defers its term alpha p_{k-1}
// (no xs traffic) and the next one (odd kk, `flush`) adds both, xs += apd p_{k-2} + alpha p_{k-1},
// reading p_{k-2} from the buffer it is about to overwrite with p_k (own rows, read before the
// write by the same thread), in the order and with the roundings of two eager updates.
template <int NT>
__device__ __forceinline__ void b1_body_q(const B1Args& a, const ItemView& it, int64_t r0, int64_t r1, int64_t kk,
                                          float al, float be, float apd, float* pt, double* part) {
  constexpr int TT = 4 * NT;
  const int n = (int)it.n;
  const int o1 = it.offs[1], o2 = it.offs[2], o3 = it.offs[3];
  const int hq = (o2 + 3) & ~3;  // near halo rows on either side, whole quads (<= B1_HALO)
  const int64_t ld = dia_ld(it.n);
  const int cur = (int)(kk & 1), nxt = cur ^ 1;
  // bases 4 rows down: row j is at offset j + 4 >= 1 for every j >= -3
  const float* __restrict__ R = a.r[cur] + it.off - 4;
  const float* __restrict__ Q = a.q[cur] + it.off - 4;
  const float* __restrict__ Pp = a.p[cur] + it.off - 4;
  float* __restrict__ Rn = a.r[nxt] + it.off - 4;
  float* __restrict__ Qn = a.q[nxt] + it.off - 4;
  float* __restrict__ Pn = a.p[nxt] + it.off - 4;
  float* __restrict__ XS = a.xs + it.off - 4;
  const float* __restrict__ v1 = it.s32 - 4;
  const float* __restrict__ v2 = it.s32 + ld - 4;
  const float* __restrict__ v3 = it.s32 + 2 * ld - 4;
  const bool upd = al != 0.0f;
  const bool flush = (kk & 1) != 0;  // item-uniform: this launch writes xs
  const int tid = threadIdx.x;
  for (int t0 = (int)r0; t0 < (int)r1; t0 += TT) {
    const int t1 = min((int)r1, t0 + TT);
    const int i = t0 + 4 * tid;
    const bool own = i < t1;
    const uint32_t ic4 = (uint32_t)((own ? i : 0) + 4);
    const F4 rv = ldq(R, ic4), qv = ldq(Q, ic4), pv = ldq(Pp, ic4);
    F4 xv = F4{0.0f, 0.0f, 0.0f, 0.0f}, p2 = F4{0.0f, 0.0f, 0.0f, 0.0f};
    if (flush) {
      xv = ldq(XS, ic4);
      p2 = ldq(Pn, ic4);  // p_{k-2}, before this launch's p_k
    }
    // one near-halo row per thread below 2 hq (rows outside the item staged as 0)
    const bool hal = tid < 2 * hq;
    const int jh = tid < hq ? t0 - hq + tid : t1 + (tid - hq);
    const bool hok = hal && jh >= 0 && jh < n;
    const uint32_t jh4 = (uint32_t)((hok ? jh : 0) + 4);
    const float rh = ldr(R, jh4), qh = ldr(Q, jh4), ph = ldr(Pp, jh4);
    bool vr[4];
    float rn[4], pn[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      vr[u] = own && i + u < t1;
      const float r_ = upd ? rv[u] - al * qv[u] : rv[u];
      const float p_ = r_ + be * pv[u];
      rn[u] = vr[u] ? r_ : 0.0f;
      pn[u] = vr[u] ? p_ : 0.0f;
    }
    if (own) {
      const int li = i - t0 + hq;
#pragma unroll
      for (int u = 0; u < 4; ++u) pt[li + u] = pn[u];
      const uint32_t i4 = (uint32_t)(i + 4);
      stq(Rn, i4, F4{rn[0], rn[1], rn[2], rn[3]});
      stq(Pn, i4, F4{pn[0], pn[1], pn[2], pn[3]});
      if (flush && upd) {
        F4 xn;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float xb = apd != 0.0f ? xv[u] + apd * p2[u] : xv[u];
          xn[u] = vr[u] ? xb + al * pv[u] : 0.0f;
        }
        stq(XS, i4, xn);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) part[0] += (double)rn[u] * (double)rn[u];
    }
    if (hal) {
      const float rr = upd ? rh - al * qh : rh;
      pt[jh - t0 + hq] = hok ? rr + be * ph : 0.0f;
    }
    // far neighbours' p_k and the coefficients (upper at the own rows, lower at i - o), in one batch
    // of loads; indices that do not reach a valid row are clamped to row 0 (their terms are selected)
    const int ib = own ? i : 0;
    const uint32_t ju4 = (uint32_t)((ib + o3 < n ? ib + o3 : 0) + 4);
    const uint32_t jd4 = (uint32_t)((ib - o3 >= -3 ? ib - o3 : 0) + 4);
    const F4 fur = ldq(R, ju4), fuq = ldq(Q, ju4), fup = ldq(Pp, ju4);
    const F4 fdr = ldq(R, jd4), fdq = ldq(Q, jd4), fdp = ldq(Pp, jd4);
    const F4 cu1 = ldq(v1, ic4), cu2 = ldq(v2, ic4), cu3 = ldq(v3, ic4);
    const F4 cl1 = ldq(v1, (uint32_t)((ib - o1 >= -3 ? ib - o1 : 0) + 4));
    const F4 cl2 = ldq(v2, (uint32_t)((ib - o2 >= -3 ? ib - o2 : 0) + 4));
    const F4 cl3 = ldq(v3, jd4);
    __syncthreads();
    if (own) {
      const int li = i - t0 + hq;
      float qo[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int iu = i + u;
        const float fu = (upd ? fur[u] - al * fuq[u] : fur[u]) + be * fup[u];
        const float fl = (upd ? fdr[u] - al * fdq[u] : fdr[u]) + be * fdp[u];
        const float t_u1 = cu1[u] * pt[li + u + o1], t_l1 = cl1[u] * pt[li + u - o1];
        const float t_u2 = cu2[u] * pt[li + u + o2], t_l2 = cl2[u] * pt[li + u - o2];
        const float t_u3 = cu3[u] * fu, t_l3 = cl3[u] * fl;
        float s_ = pn[u];
        s_ += iu + o1 < n ? t_u1 : 0.0f;
        s_ += iu - o1 >= 0 ? t_l1 : 0.0f;
        s_ += iu + o2 < n ? t_u2 : 0.0f;
        s_ += iu - o2 >= 0 ? t_l2 : 0.0f;
        s_ += iu + o3 < n ? t_u3 : 0.0f;
        s_ += iu - o3 >= 0 ? t_l3 : 0.0f;
        qo[u] = vr[u] ? s_ : 0.0f;
        part[1] += (double)pn[u] * (double)qo[u];
        part[2] += (double)qo[u] * (double)rn[u];
        part[3] += (double)qo[u] * (double)qo[u];
      }
      stq(Qn, (uint32_t)(i + 4), F4{qo[0], qo[1], qo[2], qo[3]});
    }
    __syncthreads();  // the next tile rewrites pt
  }
}

// At least 3 waves per SIMD (at most 168 VGPRs; the kernel takes 157): three 256-thread
// workgroups per CU.  C4 (256 systems, 29.1M rows): 103.2 ms per solve (106.6 with xs updated every
// launch), where rows one per thread (512 threads, 4 rows per thread, LDS-staged, 127 VGPRs) took
// 112.9 ms; issuing the first tile's r, q, p before the reduction of the partial sums pushed the
// kernel to 168 VGPRs with spills and measured 112.0 ms.
template <int NT>
__global__ void __launch_bounds__(NT, 3) batch_one_k(B1Args a, int par) {
  constexpr int TT = 4 * NT;
  __shared__ __align__(16) float pt[TT + 2 * B1_HALO];
  __shared__ double sm[4 * (NT / 64)];
  const int64_t NE = a.n_entries;
  // XCD-aware: the entries split into 8 contiguous slabs, one per XCD (workgroups are dealt to the
  // XCDs round-robin), so an item's neighbouring entries run on one XCD at about the same time and
  // share their far rows through its L2
  const int G = gridDim.x;
  int64_t e_lo = 0, e_hi = NE, e_first = blockIdx.x, e_step = G;
  if (G % N_XCD == 0) {
    const int x = blockIdx.x % N_XCD, per = G / N_XCD;
    e_lo = NE * x / N_XCD;
    e_hi = NE * (x + 1) / N_XCD;
    e_first = e_lo + blockIdx.x / N_XCD;
    e_step = per;
  }
  const B1State* __restrict__ Sin = a.st[par];
  B1State* __restrict__ Sout = a.st[par ^ 1];
  const double* __restrict__ Pin = a.P[par];
  double* __restrict__ Pout = a.P[par ^ 1];
  for (int64_t entry = e_first; entry < e_hi; entry += e_step) {
    (void)e_lo;
    int2 te = a.table[entry];
    te.x = uni(te.x);
    te.y = uni(te.y);
    const ItemView it = item_view_uni(a.items, te.x);
    B1State S = state_uni(Sin + te.x);
    const bool first = te.y == 0;
    const bool lead = first && threadIdx.x == 0;
    if (S.phase == P1_DONE) {
      if (lead) Sout[te.x] = S;
      continue;
    }
    const int64_t r0 = (int64_t)te.y * a.rpe, r1 = min(it.n, r0 + a.rpe);
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    if (S.phase == P1_R2) {
      // true residual in FP64 on the unscaled operator, r = S r64 into the resumed launch's buffer
      float* __restrict__ Rw = a.r[S.kr & 1] + it.off;
      const int64_t ld = dia_ld(it.n);
      for (int64_t i = r0 + threadIdx.x; i < r1; i += NT) {
        const double bi = it.b[i];
        double ax = it.v64[i] * it.x[i];
#pragma unroll
        for (int d = 1; d < 4; ++d) {
          const int64_t o = it.offs[d];
          const double* vd = it.v64 + (int64_t)d * ld;
          if (i + o < it.n) ax += vd[i] * it.x[i + o];
          if (i - o >= 0) ax += vd[i - o] * it.x[i - o];
        }
        const double ri = bi - ax;
        const float rf = (float)(ri * sqrt(it.dinv64[i]));
        Rw[i] = rf;
        part[0] += (double)rf * (double)rf;
        part[1] += ri * ri;
        part[2] += bi * bi;
      }
      S.phase = P1_RES;
      b1_store_partials<NT>(part, Pout, NE, (int)entry, sm);
      if (lead) Sout[te.x] = S;
      continue;
    }
    // the quad form issues its first tile's loads (which need only the buffer parity) before the
    // reduction of the partial sums, so their round trip overlaps it
    double in[4];
    b1_item_totals<NT>(Pin, NE, it.e0, it.ne, sm, in);
    int64_t kk;
    double alpha, beta;
    if (S.phase == P1_RES) {
      const double rr = in[0], rt = in[1], bb = in[2];
      const double atol2 = fmax(a.tol2 * bb, a.atol2_abs);
      const bool bad = !isfinite(rt) || S.bad;
      S.refinements += 1;
      S.rr_true = rt;
      S.bb = bb;
      S.rr_last = rr;
      if (bad || !(rt > atol2) || S.kr >= a.maxiter) {
        S.phase = P1_DONE;
        S.status = bad ? BS_NONFINITE : !(rt > atol2) ? BS_CONVERGED : BS_MAXITER;
        if (lead) {
          Sout[te.x] = S;
          atomicSub(a.active, 1);
        }
        continue;
      }
      // the inner loop stops only to replace; the last cycle aims at the remaining reduction
      const double need = atol2 / rt;
      S.rr_replace = need > a.repl2 ? 0.5 * need * rr : a.repl2 * rr;
      kk = S.kr;
      alpha = 0.0;
      // selects, not a runtime index (which would place S in scratch memory)
      beta = kk == 0 ? 0.0 : rr / ((kk & 1) ? S.g_hist[1] : S.g_hist[0]);
      S.phase = P1_ITER;
      S.k = kk + 1;
      S.a_pend = 0.0f;  // the stop's fold left nothing pending, and this launch (alpha = 0) defers nothing
    } else {  // P1_ITER
      kk = S.k;
      const double rr = in[0];
      alpha = rr / in[1];
      beta = (rr - 2.0 * alpha * in[2] + alpha * alpha * in[3]) / rr;
      const bool stop_repl = rr <= S.rr_replace;
      const bool bad = !isfinite(rr) || !isfinite(alpha) || !isfinite(beta);
      if (stop_repl || bad || kk > a.maxiter) {
        // x64 += S xs, xs = 0 on the entry's rows, with the term the previous launch deferred
        // (an odd launch: xs + a_pend p_{k-2}, p_{k-2} in the buffer this launch would have
        // written); the true residual follows in the next launch
        float* __restrict__ XS = a.xs + it.off;
        const float apd = (kk & 1) ? S.a_pend : 0.0f;
        const float* __restrict__ P2 = a.p[(kk & 1) ^ 1] + it.off;
        for (int64_t i = r0 + threadIdx.x; i < r1; i += NT) {
          const float xsi = apd != 0.0f ? XS[i] + apd * P2[i] : XS[i];
          it.x[i] += sqrt(it.dinv64[i]) * (double)xsi;
          XS[i] = 0.0f;
        }
        S.a_pend = 0.0f;
        S.phase = P1_R2;
        S.kr = kk - 1;
        S.bad = S.bad || (bad && !stop_repl);
        if (lead) Sout[te.x] = S;
        continue;
      }
      if (kk & 1)
        S.g_hist[1] = rr;
      else
        S.g_hist[0] = rr;
      S.k = kk + 1;
    }
    // even launches defer their xs term to the next launch, odd ones add it (and their own)
    const float apd_in = S.a_pend;
    S.a_pend = (kk & 1) ? 0.0f : (float)alpha;
    b1_body_q<NT>(a, it, r0, r1, kk, (float)alpha, (float)beta, apd_in, pt, part);
    b1_store_partials<NT>(part, Pout, NE, (int)entry, sm);
    if (lead) Sout[te.x] = S;
  }
}

__global__ void batch_one_init_k(B1Args a, int n_items) {
  // per entry: zero x64, xs, p and q (a start reads p_{-1} and q_{-1} times zero)
  for (int entry = blockIdx.x; entry < a.n_entries; entry += gridDim.x) {
    const int2 te = a.table[entry];
    const BItem& it = a.items[te.x];
    const int64_t r0 = (int64_t)te.y * a.rpe, r1 = min(it.n, r0 + a.rpe);
    for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
      it.x[i] = 0.0;
      a.xs[it.off + i] = 0.0f;
      a.p[0][it.off + i] = 0.0f;
      a.p[1][it.off + i] = 0.0f;
      a.q[0][it.off + i] = 0.0f;
      a.q[1][it.off + i] = 0.0f;
    }
  }
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n_items) {
    B1State s = {};
    s.phase = P1_R2;  // the first launch forms r = b and the norms
    s.status = BS_RUNNING;
    s.refinements = -1;
    a.st[0][g] = s;
    a.st[1][g] = s;
  }
}

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

struct BLayout {
  BItem* items;
  BState* st;
  B1State* st1[2];
  int2* table;
  const double** dinvs;
  double* PA;
  double* PB;        // [4][n_entries]; the one-kernel step uses it as P[0]
  double* PB1;       // [4][n_entries]: the one-kernel step's P[1]
  int32_t* active;
  int32_t* bad;  // [n_items]: the item's operator cannot be Jacobi-scaled exactly
  float *rs, *p0, *p1, *q, *xs, *dinv, *r1, *q1;
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
  L.st1[0] = (B1State*)take(sizeof(B1State) * n_items);
  L.st1[1] = (B1State*)take(sizeof(B1State) * n_items);
  L.table = (int2*)take(sizeof(int2) * n_entries);
  L.dinvs = (const double**)take(sizeof(double*) * n_items);
  L.PA = (double*)take(sizeof(double) * n_entries);
  L.PB = (double*)take(sizeof(double) * 4 * n_entries);
  L.PB1 = (double*)take(sizeof(double) * 4 * n_entries);
  L.active = (int32_t*)take(256);
  L.rs = (float*)take(4 * ntot);
  L.p0 = (float*)take(4 * ntot);
  L.p1 = (float*)take(4 * ntot);
  L.q = (float*)take(4 * ntot);
  L.xs = (float*)take(4 * ntot);
  L.dinv = (float*)take(4 * ntot);
  L.r1 = (float*)take(4 * ntot);
  L.q1 = (float*)take(4 * ntot);
  L.bad = (int32_t*)take(sizeof(int32_t) * n_items);
  L.sdia = (float*)take(4 * nsd);
  L.bytes = off;
  return L;
}

// FP32 entries of the scaled copies of the DIA items' stored upper diagonals
static int64_t scaled_rows(const jno_op& op) { return op.format == JNO_FMT_DIA ? (op.ndiag - 1) * dia_ld(op.n) : 0; }

// each item's rows start 64-aligned in the concatenated FP32 vectors (16-byte vector tiles)
static int64_t padded_rows(int64_t n) { return (n + 63) & ~(int64_t)63; }

// entries of the three-kernel table (BT_RPB rows each): the most any table of the call has
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

// the item table and (item, local entry) list for entries of rpe rows
static int64_t build_table(std::vector<BItem>& items, std::vector<int2>& table, int64_t rpe) {
  int64_t e = 0;
  table.clear();
  for (size_t i = 0; i < items.size(); ++i) {
    BItem& it = items[i];
    it.e0 = (int32_t)e;
    it.ne = (int32_t)cdiv(it.n, rpe);
    for (int j = 0; j < it.ne; ++j) table.push_back(make_int2((int)i, j));
    e += it.ne;
  }
  return e;
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

static int check_batch_tuning(const jno_batch_tuning* t) {
  if (!t) return JNO_OK;
  JNO_REQUIRE(t->one_kernel == 0 || t->one_kernel == 1, JNO_EINVAL, "jno_batch_tuning: one_kernel is 0 or 1");
  JNO_REQUIRE(t->rows_per_entry == 0 || (t->rows_per_entry > 0 && t->rows_per_entry % B1_RPE_UNIT == 0 &&
                                         t->rows_per_entry <= 65536),
              JNO_EINVAL, "jno_batch_tuning: rows_per_entry is 0 or a multiple of %d up to 65536", B1_RPE_UNIT);
  JNO_REQUIRE(t->grid == 0 || (t->grid > 0 && t->grid % N_XCD == 0 && t->grid <= 65536), JNO_EINVAL,
              "jno_batch_tuning: grid is 0 or a multiple of %d up to 65536", N_XCD);
  return JNO_OK;
}

// every item allows the one-kernel step: a Jacobi-scaled 4-diagonal DIA operator with ascending
// offsets whose second one fits the LDS halo, and fewer than 2^29 rows (32-bit byte offsets)
static bool one_kernel_ok(const std::vector<BItem>& items) {
  for (const BItem& it : items) {
    if (it.format != JNO_FMT_DIA || it.ndiag != 4 || !it.scaled || it.n >= ((int64_t)1 << 29)) return false;
    if (!(0 < it.offs[1] && it.offs[1] < it.offs[2] && it.offs[2] <= B1_HALO && it.offs[2] < it.offs[3])) return false;
  }
  return true;
}

// Host loop shared by both step forms: enqueue chunks of steps (replayed from a hipGraph when
// steps_per_graph > 0) and poll the active-item count one chunk behind, so the queue never drains.
template <typename Step>
static int run_steps(Step&& launch_step, int steps_per_graph, int chunk, int64_t max_steps, int32_t* d_active,
                     hipStream_t s, int64_t* launched_out) {
  int rc = g_bpoll.init();
  if (rc) return rc;
  int32_t* h_act = g_bpoll.h;
  hipEvent_t* ev = g_bpoll.ev;
  BatchGraph bg;
  bg.s = s;
  if (steps_per_graph > 0) {
    // capture one chunk of steps on a private stream (default priority), replay it on `s`
    JNO_HIP(hipStreamCreateWithFlags(&bg.cap, hipStreamNonBlocking));
    JNO_HIP(hipStreamBeginCapture(bg.cap, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk && rc == JNO_OK; ++i) rc = launch_step(bg.cap, i);
    hipError_t ce = hipStreamEndCapture(bg.cap, &bg.graph);
    if (rc) return rc;
    JNO_HIP(ce);
    JNO_HIP(hipGraphInstantiate(&bg.exec, bg.graph, nullptr, nullptr, 0));
  }
  int slot = 0, pending = 0;
  int64_t launched = 0;
  for (;;) {
    if (bg.exec) {
      JNO_HIP(hipGraphLaunch(bg.exec, s));
    } else {
      for (int i = 0; i < chunk; ++i) {
        rc = launch_step(s, i);
        if (rc) return rc;
      }
    }
    launched += chunk;
    JNO_HIP(hipMemcpyAsync(&h_act[slot], d_active, sizeof(int32_t), hipMemcpyDeviceToHost, s));
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
  *launched_out = launched;
  return JNO_OK;
}

static int batch_impl(int n_items, const jno_batch_item* h_items, double tol, double atol, int64_t maxiter,
                      double replace_factor, int steps_per_graph, jno_batch_tuning* h_tuning,
                      jno_pcg_stats* h_stats, void* d_ws, size_t ws_bytes, void* stream) {
  JNO_REQUIRE(n_items > 0 && h_items && h_stats, JNO_EINVAL, "jno_batch_pcg_mixed: empty batch");
  JNO_REQUIRE(replace_factor > 0.0 && replace_factor < 1.0, JNO_EINVAL, "jno_batch_pcg_mixed: replace_factor in (0,1)");
  JNO_REQUIRE(maxiter >= 0 && steps_per_graph >= 0, JNO_EINVAL, "jno_batch_pcg_mixed: bad maxiter/steps");
  int rc = check_batch_tuning(h_tuning);
  if (rc) return rc;
  const bool want_one = !h_tuning || h_tuning->one_kernel;
  const int64_t rpe = h_tuning && h_tuning->rows_per_entry ? h_tuning->rows_per_entry : B1_RPE;
  int64_t ntot, nsd;
  const int64_t ne = count_entries(n_items, h_items, &ntot, &nsd);
  JNO_REQUIRE(ne < INT32_MAX, JNO_EINVAL, "jno_batch_pcg_mixed: batch too large");
  JNO_REQUIRE(ws_bytes >= jno_batch_workspace_size(n_items, h_items), JNO_EWORKSPACE,
              "jno_batch_pcg_mixed: workspace too small");
  hipStream_t s = as_stream(stream);
  BLayout L = blayout(n_items, ne, ntot, nsd, static_cast<char*>(d_ws));
  // host staging of the item descriptors and the entry table
  std::vector<BItem> items(n_items);
  std::vector<int2> table;
  std::vector<const double*> dinvs(n_items);
  int64_t off = 0, soff = 0;
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
    dinvs[i] = h.dinv;
    it.dinv64 = h.dinv;
    it.s32 = it.format == JNO_FMT_DIA ? L.sdia + soff : nullptr;
    it.scaled = 0;
    soff += scaled_rows(h.op);
    off += padded_rows(it.n);
    any_dia |= it.format == JNO_FMT_DIA;
    any_csr |= it.format == JNO_FMT_CSR;
  }
  build_table(items, table, BT_RPB);
  JNO_HIP(hipMemcpyAsync(L.items, items.data(), sizeof(BItem) * n_items, hipMemcpyHostToDevice, s));
  JNO_HIP(hipMemcpyAsync(L.table, table.data(), sizeof(int2) * ne, hipMemcpyHostToDevice, s));
  JNO_HIP(hipMemcpyAsync(L.dinvs, dinvs.data(), sizeof(double*) * n_items, hipMemcpyHostToDevice, s));
  const int32_t h_active = n_items;
  JNO_HIP(hipMemcpyAsync(L.active, &h_active, sizeof(int32_t), hipMemcpyHostToDevice, s));
  const int grid = (int)std::min<int64_t>(ne, BT_GRID);
  if (any_dia) {
    // the DIA items' scaled operators; an item whose Jacobi scaling is not exact runs unscaled
    JNO_HIP(hipMemsetAsync(L.bad, 0, sizeof(int32_t) * n_items, s));
    batch_scale_k<<<grid, BT_NT, 0, s>>>(L.items, L.dinvs, L.table, (int)ne, L.bad);
    JNO_LAUNCH_CHECK();
    std::vector<int32_t> bad(n_items);
    JNO_HIP(hipMemcpyAsync(bad.data(), L.bad, sizeof(int32_t) * n_items, hipMemcpyDeviceToHost, s));
    JNO_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n_items; ++i) items[i].scaled = items[i].format == JNO_FMT_DIA && !bad[i];
  }
  const bool one = want_one && one_kernel_ok(items);
  if (h_tuning) h_tuning->ran_one_kernel = one ? 1 : 0;
  int64_t ne_run = ne;
  if (one) ne_run = build_table(items, table, rpe);
  JNO_HIP(hipMemcpyAsync(L.items, items.data(), sizeof(BItem) * n_items, hipMemcpyHostToDevice, s));
  if (one) JNO_HIP(hipMemcpyAsync(L.table, table.data(), sizeof(int2) * ne_run, hipMemcpyHostToDevice, s));
  JNO_HIP(hipStreamSynchronize(s));  // host staging vectors go out of scope after this call

  const auto t0 = std::chrono::steady_clock::now();
  int64_t launched = 0;
  int worst = JNO_OK;
  std::vector<jno_pcg_stats> out(n_items);
  if (one) {
    B1Args a;
    a.items = L.items;
    a.st[0] = L.st1[0];
    a.st[1] = L.st1[1];
    a.table = L.table;
    a.n_entries = (int32_t)ne_run;
    a.rpe = (int32_t)rpe;
    a.r[0] = L.rs;
    a.r[1] = L.r1;
    a.q[0] = L.q;
    a.q[1] = L.q1;
    a.p[0] = L.p0;
    a.p[1] = L.p1;
    a.xs = L.xs;
    a.P[0] = L.PB;
    a.P[1] = L.PB1;
    a.active = L.active;
    a.tol2 = tol * tol;
    a.atol2_abs = atol * atol;
    a.repl2 = replace_factor * replace_factor;
    a.maxiter = maxiter;
    const int g1 = h_tuning && h_tuning->grid ? h_tuning->grid : B1_GRID;
    batch_one_init_k<<<(unsigned)std::max<int64_t>(std::min<int64_t>(ne_run, 4096), cdiv(n_items, B1_NT)), B1_NT, 0,
                       s>>>(a, n_items);
    JNO_LAUNCH_CHECK();
    // launches alternate the state parity, so a chunk holds an even number of them
    int chunk = steps_per_graph > 0 ? steps_per_graph : 8;
    chunk += chunk & 1;
    auto step = [&](hipStream_t st, int i) -> int {
      batch_one_k<B1_NT><<<g1, B1_NT, 0, st>>>(a, i & 1);
      JNO_LAUNCH_CHECK();
      return JNO_OK;
    };
    // every replacement costs three launches (the stop, R2 and the resumed launch)
    rc = run_steps(step, steps_per_graph > 0 ? chunk : 0, chunk, 4 * (maxiter + 2) + 4 * (int64_t)chunk, L.active, s,
                   &launched);
    if (rc) return rc;
    std::vector<B1State> st(n_items);
    JNO_HIP(hipMemcpyAsync(st.data(), L.st1[launched & 1], sizeof(B1State) * n_items, hipMemcpyDeviceToHost, s));
    JNO_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n_items; ++i) {
      const double bn = sqrt(st[i].bb);
      out[i].iterations = st[i].kr;
      out[i].refinements = st[i].refinements;
      out[i].relres = bn > 0 ? sqrt(st[i].rr_true) / bn : sqrt(st[i].rr_true);
      out[i].relres_recursive = bn > 0 ? sqrt(st[i].rr_last) / bn : sqrt(st[i].rr_last);
      const int status = st[i].phase == P1_DONE ? st[i].status : BS_RUNNING;
      if (status == BS_NONFINITE) worst = JNO_ENONFINITE;
      else if (status != BS_CONVERGED && worst == JNO_OK) worst = JNO_ENOTCONV;
    }
  } else {
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
    batch_init_k<<<(unsigned)std::max<int64_t>(grid, cdiv(n_items, BT_NT)), BT_NT, 0, s>>>(a, n_items);
    JNO_LAUNCH_CHECK();
    batch_dinv_k<<<grid, BT_NT, 0, s>>>(L.items, L.dinvs, L.table, (int)ne, L.dinv);
    JNO_LAUNCH_CHECK();
    auto launch_kb = [&](hipStream_t st) -> int {
      if (any_dia) batch_kb<true><<<grid, BT_NT, 0, st>>>(a);
      if (any_csr) batch_kb<false><<<grid, BT_NT, 0, st>>>(a);
      JNO_LAUNCH_CHECK();
      return JNO_OK;
    };
    auto step = [&](hipStream_t st, int) -> int {
      if (any_dia) batch_ka<true><<<grid, BT_NT, 0, st>>>(a);
      if (any_csr) batch_ka<false><<<grid, BT_NT, 0, st>>>(a);
      if (any_dia) batch_ks<true><<<grid, BT_NT, 0, st>>>(a);
      if (any_csr) batch_ks<false><<<grid, BT_NT, 0, st>>>(a);
      JNO_LAUNCH_CHECK();
      return launch_kb(st);
    };
    rc = launch_kb(s);  // the first true residual r = b
    if (rc) return rc;
    const int chunk = steps_per_graph > 0 ? steps_per_graph : 8;
    // every item is DONE after at most maxiter iterations plus one replacement per iteration
    rc = run_steps(step, steps_per_graph, chunk, 2 * (maxiter + 2) + 4 * (int64_t)chunk, L.active, s, &launched);
    if (rc) return rc;
    std::vector<BState> st(n_items);
    JNO_HIP(hipMemcpyAsync(st.data(), L.st, sizeof(BState) * n_items, hipMemcpyDeviceToHost, s));
    JNO_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n_items; ++i) {
      const double bn = sqrt(st[i].bb);
      out[i].iterations = st[i].k;
      out[i].refinements = st[i].refinements;
      out[i].relres = bn > 0 ? sqrt(st[i].rr_true) / bn : sqrt(st[i].rr_true);
      out[i].relres_recursive = bn > 0 ? sqrt(st[i].rr_last) / bn : sqrt(st[i].rr_last);
      if (st[i].status == BS_NONFINITE) worst = JNO_ENONFINITE;
      else if (st[i].status != BS_CONVERGED && worst == JNO_OK) worst = JNO_ENOTCONV;
    }
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < n_items; ++i) {
    out[i].seconds = secs;
    h_stats[i] = out[i];
  }
  if (worst == JNO_ENONFINITE) set_error("jno_batch_pcg_mixed: a non-finite value in at least one system");
  if (worst == JNO_ENOTCONV) set_error("jno_batch_pcg_mixed: at least one system did not converge in %lld iterations",
                                       (long long)maxiter);
  return worst;
}

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
  return batch_impl(n_items, h_items, tol, atol, maxiter, replace_factor, steps_per_graph, nullptr, h_stats, d_ws,
                    ws_bytes, stream);
}

extern "C" int jno_batch_pcg_mixed_tuned(int n_items, const jno_batch_item* h_items, double tol, double atol,
                                         int64_t maxiter, double replace_factor, int steps_per_graph,
                                         jno_batch_tuning* h_tuning, jno_pcg_stats* h_stats, void* d_ws,
                                         size_t ws_bytes, void* stream) {
  return batch_impl(n_items, h_items, tol, atol, maxiter, replace_factor, steps_per_graph, h_tuning, h_stats, d_ws,
                    ws_bytes, stream);
}
