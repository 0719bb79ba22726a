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

// y = A x with a solver operator; `ext` is the product buffer an EBE operator needs (nullable: its
// ebe->scratch)
template <typename V, typename X>
int op_spmv(const jno_op* op, const V* vals, const X* x, X* y, hipStream_t s, void* ext = nullptr);
// y = A x with an element-by-element operator (ebe.hip)
template <typename T>
int ebe_spmv(const jno_op* op, const T* x, T* y, void* yloc, hipStream_t s);

// U block rows of y = A x in the BSR3 layout (3x3 blocks; the 9 entries of block row R stored
// entry-major: entry e = 3i+j of block b at 9*browptr[R] + e*len + b, len = blocks in the row), given
// each row's first block bs[u] and block count len[u] (0 for a row to skip: sums 0), on L lanes:
// lane `sub` takes blocks sub, sub + L, ... so each entry read is coalesced across the lanes, and the
// three row sums are reduced over the L lanes (every lane gets them).  The loads of the first CH
// blocks per lane of all U rows are issued before the first use, so the dependent chain block
// column -> x is paid once per U rows for rows of up to CH L blocks; blocks past those go one at a
// time.  (P2 vector rows on a Kuhn lattice have 19, 27, 42 or 65 blocks: with L = 32 and CH = 2 only
// the last block of a vertex row is left to the serial loop.)
template <int L, int U, typename V, typename X, int CH = 1>
__device__ __forceinline__ void bsr3_rows_at(const int32_t* __restrict__ bcol, const V* __restrict__ bvals,
                                             const X* __restrict__ x, const int* bs, const int* len, int sub,
                                             X (*s)[3]) {
  int c[U][CH];
  V v[U][CH][9];
  X xv[U][CH][3];
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int h = 0; h < CH; ++h) {
      // unconditional loads from clamped indices (no branch, so all of them stay in flight)
      const int b = sub + h * L;
      const bool ok = b < len[u];
      c[u][h] = 3 * bcol[ok ? bs[u] + b : 0];
#pragma unroll
      for (int e = 0; e < 9; ++e) {
        const V w = bvals[ok ? 9 * (int64_t)bs[u] + e * len[u] + b : 0];
        v[u][h][e] = ok ? w : V(0);
      }
    }
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int h = 0; h < CH; ++h)
#pragma unroll
      for (int j = 0; j < 3; ++j) xv[u][h][j] = x[c[u][h] + j];
#pragma unroll
  for (int u = 0; u < U; ++u) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      X t = (X)v[u][0][3 * i] * xv[u][0][0] + (X)v[u][0][3 * i + 1] * xv[u][0][1] + (X)v[u][0][3 * i + 2] * xv[u][0][2];
#pragma unroll
      for (int h = 1; h < CH; ++h)
        t += (X)v[u][h][3 * i] * xv[u][h][0] + (X)v[u][h][3 * i + 1] * xv[u][h][1] + (X)v[u][h][3 * i + 2] * xv[u][h][2];
      s[u][i] = t;
    }
    const V* vb = bvals + 9 * (int64_t)bs[u];
    for (int b = sub + CH * L; b < len[u]; b += L) {
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

// The same for block rows R[0..U) (R[u] < 0: skipped), reading their block-row pointers first.
template <int L, int U, typename V, typename X, int CH = 1>
__device__ __forceinline__ void bsr3_rows(const int32_t* __restrict__ browptr, const int32_t* __restrict__ bcol,
                                          const V* __restrict__ bvals, const X* __restrict__ x, const int64_t* R,
                                          int sub, X (*s)[3]) {
  int bs[U], len[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t Ru = R[u] < 0 ? 0 : R[u];
    bs[u] = browptr[Ru];
    len[u] = R[u] < 0 ? 0 : browptr[Ru + 1] - bs[u];
  }
  bsr3_rows_at<L, U, V, X, CH>(bcol, bvals, x, bs, len, sub, s);
}

// position of entry (r, k) of a 30 x 30 class block (pitch 32) in the element-by-element operator's
// class table (the MFMA operand order of its element pass)
__host__ __device__ constexpr int ebe_table_index(int r, int k) {
  return ((r / 16) * 8 + k / 4) * 64 + (k % 4) * 16 + r % 16;
}

}  // namespace jno
