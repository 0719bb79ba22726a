// Internal declarations shared by the sparse kernels and the PCG engine.
#pragma once
#include "common.h"

namespace jno {

struct DiaOffsets {
  int32_t o[JNO_DIA_MAX];
};

// DIA leading dimension: every stored diagonal starts on a 256-byte boundary, so 16-byte vector
// loads of a diagonal at a 4-aligned row are aligned
__host__ __device__ __forceinline__ int64_t dia_ld(int64_t n) { return (n + 63) & ~(int64_t)63; }

template <typename V, typename X>
int op_spmv(const jno_op* op, const V* vals, const X* x, X* y, hipStream_t s);

// One block row R of y = A x in the BSR3 layout (3x3 blocks; the 9 entries of block row R stored
// entry-major: entry e = 3i+j of block b at 9*browptr[R] + e*len + b, len = blocks in the row), on
// L lanes: lane `sub` takes blocks sub, sub + L, ... so each entry read is coalesced across the
// lanes; the three row sums are reduced over the L lanes and every lane gets them.
template <int L, typename V, typename X>
__device__ __forceinline__ void bsr3_row(const int32_t* __restrict__ browptr, const int32_t* __restrict__ bcol,
                                         const V* __restrict__ bvals, const X* __restrict__ x, int64_t R, int sub,
                                         X& s0, X& s1, X& s2) {
  const int bs = browptr[R], len = browptr[R + 1] - bs;
  const V* v = bvals + 9 * (int64_t)bs;
  s0 = s1 = s2 = X(0);
  for (int b = sub; b < len; b += L) {
    const int64_t c = 3 * (int64_t)bcol[bs + b];
    const X x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
    s0 += (X)v[b] * x0 + (X)v[len + b] * x1 + (X)v[2 * len + b] * x2;
    s1 += (X)v[3 * len + b] * x0 + (X)v[4 * len + b] * x1 + (X)v[5 * len + b] * x2;
    s2 += (X)v[6 * len + b] * x0 + (X)v[7 * len + b] * x1 + (X)v[8 * len + b] * x2;
  }
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o, L);
    s1 += __shfl_xor(s1, o, L);
    s2 += __shfl_xor(s2, o, L);
  }
}

}  // namespace jno
