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

// U block rows R[0..U) of y = A x in the BSR3 layout (3x3 blocks; the 9 entries of block row R
// stored entry-major: entry e = 3i+j of block b at 9*browptr[R] + e*len + b, len = blocks in the
// row), on L lanes: lane `sub` takes blocks sub, sub + L, ... so each entry read is coalesced across
// the lanes, and the three row sums are reduced over the L lanes (every lane gets them).  Every
// load of the U rows' first block per lane is issued before the first use, so the dependent chain
// block-row pointer -> block column -> x is paid once per U rows.  Blocks past the first L of a row
// go one at a time.  Rows with R[u] < 0 are skipped (sums 0).
template <int L, int U, typename V, typename X>
__device__ __forceinline__ void bsr3_rows(const int32_t* __restrict__ browptr, const int32_t* __restrict__ bcol,
                                          const V* __restrict__ bvals, const X* __restrict__ x, const int64_t* R,
                                          int sub, X (*s)[3]) {
  int bs[U], len[U], c[U];
  V v[U][9];
  X x0[U], x1[U], x2[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t Ru = R[u] < 0 ? 0 : R[u];
    bs[u] = browptr[Ru];
    len[u] = R[u] < 0 ? 0 : browptr[Ru + 1] - bs[u];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    // unconditional loads from clamped indices (no branch, so all of them stay in flight)
    const bool ok = sub < len[u];
    c[u] = 3 * bcol[ok ? bs[u] + sub : 0];
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const V w = bvals[ok ? 9 * (int64_t)bs[u] + e * len[u] + sub : 0];
      v[u][e] = ok ? w : V(0);
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    x0[u] = x[c[u]];
    x1[u] = x[c[u] + 1];
    x2[u] = x[c[u] + 2];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    s[u][0] = (X)v[u][0] * x0[u] + (X)v[u][1] * x1[u] + (X)v[u][2] * x2[u];
    s[u][1] = (X)v[u][3] * x0[u] + (X)v[u][4] * x1[u] + (X)v[u][5] * x2[u];
    s[u][2] = (X)v[u][6] * x0[u] + (X)v[u][7] * x1[u] + (X)v[u][8] * x2[u];
    const V* vb = bvals + 9 * (int64_t)bs[u];
    for (int b = sub + L; b < len[u]; b += L) {
      const int64_t cb = 3 * (int64_t)bcol[bs[u] + b];
      const X y0 = x[cb], y1 = x[cb + 1], y2 = x[cb + 2];
      s[u][0] += (X)vb[b] * y0 + (X)vb[len[u] + b] * y1 + (X)vb[2 * len[u] + b] * y2;
      s[u][1] += (X)vb[3 * len[u] + b] * y0 + (X)vb[4 * len[u] + b] * y1 + (X)vb[5 * len[u] + b] * y2;
      s[u][2] += (X)vb[6 * len[u] + b] * y0 + (X)vb[7 * len[u] + b] * y1 + (X)vb[8 * len[u] + b] * y2;
    }
  }
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1)
#pragma unroll
    for (int u = 0; u < U; ++u) {
      s[u][0] += __shfl_xor(s[u][0], o, L);
      s[u][1] += __shfl_xor(s[u][1], o, L);
      s[u][2] += __shfl_xor(s[u][2], o, L);
    }
}

}  // namespace jno
