// PCG state and helpers shared by the PCG engine (pcg.hip) and the SpMV kernels it launches from
// other files (the element-by-element operator, ebe.hip).
#pragma once
#include "common.h"

namespace jno {

constexpr int N_XCD = 8;

// ST_UNSCALED: a scaled FP32 solve found its operator's scaled main diagonal not exactly 1 and
// stopped before its first iteration (the unscaled inner iteration runs instead)
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_MAXITER = 2, ST_NONFINITE = 3, ST_REPLACE = 4, ST_UNSCALED = 5 };

struct PcgState {
  int32_t done;    // written by K2 (block 0), read by K1
  int32_t done1;   // written by K1 (block 0), read by K2
  int32_t status;  // why K1 stopped
  int32_t scale_bad;  // scaled FP32 solves: set when the scaled main diagonal is not 1 (scale_dia_k)
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
  // single-reduction iterations: KA stages k+1, gamma_k, alpha_k here, KB copies them to k,
  // gamma_prev, alpha (KA reads those, so it cannot write them itself)
  int64_t k_next;
  double gamma_next;
  double alpha_next;
  // one-kernel iterations: the launch with index k_special starts (special = 1) or resumes after a
  // residual replacement (2) with alpha = 0; g_hist[k & 1] = the exact r.r read by launch k
  int64_t k_special;
  int32_t special;
  int32_t outer_done;  // scaled mixed solves: the outer step decided to stop (convergence, cap, ...)
  double g_hist[2];
  // one-kernel iterations with x updated every other launch: a_pend[k & 1] = the (FP32) alpha whose
  // x term launch k-1 deferred to launch k (0 when nothing is pending)
  double a_pend[2];
};

// A kernel-side view of the PCG state for SpMV kernels launched by the PCG engine; empty (st =
// nullptr) for a plain y = A x.  begin_spmv: true when the iteration has stopped (the launch exits;
// the single-reduction KB also raises `done`); otherwise KB's workgroup 0 moves the scalars KA
// staged into the fields KA reads.
struct PcgStateView {
  PcgState* st = nullptr;
  template <bool CG>
  __device__ __forceinline__ bool begin_spmv() const {
    if (!st) return false;
    if (st->done1) {
      if (CG && blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
      return true;
    }
    if (CG && blockIdx.x == 0 && threadIdx.x == 0) {
      st->k = st->k_next;
      st->gamma_prev = st->gamma_next;
      st->alpha = st->alpha_next;
    }
    return false;
  }
  // host side: the device address of the stop flag (nullptr without a state)
  const int32_t* stop_flag() const { return st ? &st->done1 : nullptr; }
};

// contiguous row range of this workgroup, XCD-aware (gridDim.x is a multiple of 8 or < 8)
__device__ __forceinline__ int64_t block_first_row(int64_t rows_per_block) {
  const int G = gridDim.x, b = blockIdx.x;
  const int logical = (G % N_XCD == 0) ? (b % N_XCD) * (G / N_XCD) + b / N_XCD : b;
  return (int64_t)logical * rows_per_block;
}

// Workgroup totals of M per-thread values into dst[m*G + blockIdx.x]: each wave reduces, the wave
// totals go through sm_out[M][NT/64], and wave 0 adds them (a fixed butterfly).
template <int NT, int M>
__device__ __forceinline__ void store_out(const double* v, double* __restrict__ dst, int G, double* sm_out) {
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const double s = wave_sum(v[m]);
    if ((threadIdx.x & 63) == 0) sm_out[m * (NT / 64) + w] = s;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const double t = wave_sum(threadIdx.x < NT / 64 ? sm_out[m * (NT / 64) + threadIdx.x] : 0.0);
      if (threadIdx.x == 0) dst[(int64_t)m * G + blockIdx.x] = t;
    }
  }
}

// The element-by-element operator's SpMV as the PCG's KS (or single-reduction KB with cg): q = A p
// with the partial p.q per workgroup, XCD-aware rows (rpb a multiple of 3); yloc = the product
// buffer in the PCG workspace (ebe_scratch_bytes)
template <typename T>
int ebe_spmv_pcg(const jno_op* op, const T* p, T* q, T* yloc, double* part, PcgStateView st, int grid, int64_t rpb,
                 int nt, bool cg, hipStream_t s);
size_t ebe_scratch_bytes(const jno_op* op);
// the structured-box SpMV of an element-by-element operator (ebe_box.hip, FP32): bricks of its grid
// (0: not a box), and the launch (grid <= 0: one workgroup per brick)
int64_t ebe_box_bricks(const jno_ebe* e);
template <typename T>
int ebe_box_launch(const jno_op* op, const T* x, T* y, double* part, PcgStateView st, int grid, bool cg,
                   hipStream_t s);

template <int NT>
__device__ __forceinline__ void store_out_1(double v, double* __restrict__ dst, double* sm_out) {
  store_out<NT, 1>(&v, dst, gridDim.x, sm_out);
}

}  // namespace jno
