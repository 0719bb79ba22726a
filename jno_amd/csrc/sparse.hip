// CSR utilities: SpMV (lanes-per-row CSR and symmetric DIA), Jacobi diagonal, Dirichlet
// elimination on an assembled CSR, exact-zero compaction, and the operator analysis that decides
// whether the index-free symmetric diagonal layout applies.
#include <type_traits>

#include "common.h"
#include "sparse.h"

namespace jno {

// ---------------------------------------------------------------------------------------------
// SpMV
// ---------------------------------------------------------------------------------------------
template <int L, typename V, typename X>
__global__ void __launch_bounds__(256) csr_spmv_k(int64_t n, const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ col, const V* __restrict__ vals,
                                                  const X* __restrict__ x, X* __restrict__ y) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int sub = threadIdx.x % L;
  for (int64_t row = gid / L; row < n; row += (int64_t)gridDim.x * blockDim.x / L) {
    const int b = rowptr[row], e = rowptr[row + 1];
    X s = 0;
    for (int j = b + sub; j < e; j += L) s += (X)vals[j] * x[col[j]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
    if (sub == 0) y[row] = s;
  }
}

template <int L, typename V, typename X>
__global__ void __launch_bounds__(256) bsr3_spmv_k(int64_t nb, const int32_t* __restrict__ browptr,
                                                   const int32_t* __restrict__ bcol, const V* __restrict__ bvals,
                                                   const X* __restrict__ x, X* __restrict__ y) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int sub = threadIdx.x % L;
  // U block rows per lane group in flight and CH blocks per lane, as in the PCG's BSR3 kernel (two
  // rows of one block for FP32, one row of two blocks for FP64)
  constexpr int U = sizeof(X) == 4 ? 2 : 1;
  constexpr int CH = sizeof(X) == 4 ? 1 : 2;
  const int64_t step = (int64_t)gridDim.x * blockDim.x / L;
  for (int64_t R0 = gid / L; R0 < nb; R0 += U * step) {
    int64_t Rs[U];
#pragma unroll
    for (int u = 0; u < U; ++u) Rs[u] = R0 + u * step < nb ? R0 + u * step : -1;
    X sr[U][3];
    bsr3_rows<L, U, V, X, CH>(browptr, bcol, bvals, x, Rs, sub, sr);
    if (sub == 0) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (Rs[u] < 0) continue;
        y[3 * Rs[u]] = sr[u][0];
        y[3 * Rs[u] + 1] = sr[u][1];
        y[3 * Rs[u] + 2] = sr[u][2];
      }
    }
  }
}

// BSR3 applicability: every block row's three rows have the same length 3*len, and row 3R+i holds
// columns 3c_b + j (j = 0, 1, 2) for its blocks b in one ascending order
__global__ void bsr3_check_k(int64_t nb, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                             int* __restrict__ bad) {
  for (int64_t R = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; R < nb; R += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = 3 * R;
    const int b0 = rowptr[r], len3 = rowptr[r + 1] - b0;
    bool ok = len3 % 3 == 0 && b0 % 9 == 0;
    for (int i = 1; i < 3 && ok; ++i) ok = rowptr[r + i + 1] - rowptr[r + i] == len3;
    for (int i = 0; i < 3 && ok; ++i) {
      const int ri = rowptr[r + i];
      for (int m = 0; m < len3 && ok; ++m) {
        const int cb = col[b0 + (m / 3) * 3];
        ok = cb % 3 == 0 && col[ri + m] == cb + m % 3;
      }
    }
    if (!ok) atomicOr(bad, 1);
  }
}

__global__ void bsr3_fill_k(int64_t nb, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                            const double* __restrict__ vals, int32_t* __restrict__ browptr,
                            int32_t* __restrict__ bcol, double* __restrict__ v64, float* __restrict__ v32) {
  for (int64_t R = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; R < nb; R += (int64_t)gridDim.x * blockDim.x) {
    const int b0 = rowptr[3 * R], len = (rowptr[3 * R + 1] - b0) / 3, bs = b0 / 9;
    browptr[R] = bs;
    if (R == nb - 1) browptr[nb] = rowptr[3 * nb] / 9;
    for (int b = 0; b < len; ++b) {
      bcol[bs + b] = col[b0 + 3 * b] / 3;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const double v = vals[rowptr[3 * R + i] + 3 * b + j];
          const int64_t at = 9 * (int64_t)bs + (int64_t)(3 * i + j) * len + b;
          if (v64) v64[at] = v;
          if (v32) v32[at] = (float)v;
        }
    }
  }
}

template <typename V, typename X>
__global__ void __launch_bounds__(256) dia_spmv_k(int64_t n, int ndiag, DiaOffsets off, const V* __restrict__ dia,
                                                  const X* __restrict__ x, X* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    X s = (X)dia[i] * x[i];
    for (int d = 1; d < ndiag; ++d) {
      const int64_t o = off.o[d];
      const V* v = dia + (int64_t)d * dia_ld(n);
      if (i + o < n) s += (X)v[i] * x[i + o];
      if (i - o >= 0) s += (X)v[i - o] * x[i - o];
    }
    y[i] = s;
  }
}

template <typename V, typename X>
int op_spmv(const jno_op* op, const V* vals, const X* x, X* y, hipStream_t s, void* ext) {
  JNO_REQUIRE(op && vals && x && y, JNO_EINVAL, "spmv: null pointer (is the FP32 copy built?)");
  const int64_t n = op->n;
  if (n == 0) return JNO_OK;
  if (op->format == JNO_FMT_EBE) {
    if constexpr (std::is_same<V, X>::value) {
      return ebe_spmv<X>(op, x, y, ext, s);
    } else {
      JNO_REQUIRE(false, JNO_EINVAL, "spmv: an EBE operator computes in its table's precision");
    }
  }
  if (op->format == JNO_FMT_DIA) {
    DiaOffsets off;
    for (int d = 0; d < JNO_DIA_MAX; ++d) off.o[d] = op->offsets[d];
    dia_spmv_k<V, X><<<grid_for(n, 256, 16384), 256, 0, s>>>(n, op->ndiag, off, vals, x, y);
  } else if (op->format == JNO_FMT_BSR3) {
    JNO_REQUIRE(n % 3 == 0, JNO_EINVAL, "spmv: a BSR3 operator has a multiple of 3 rows");
    const int64_t nb = n / 3;
    const int L = op->lanes;
    const int g = grid_for(nb * L, 256, 16384);
    switch (L) {
      case 1: bsr3_spmv_k<1, V, X><<<g, 256, 0, s>>>(nb, op->rowptr, op->col, vals, x, y); break;
      case 2: bsr3_spmv_k<2, V, X><<<g, 256, 0, s>>>(nb, op->rowptr, op->col, vals, x, y); break;
      case 4: bsr3_spmv_k<4, V, X><<<g, 256, 0, s>>>(nb, op->rowptr, op->col, vals, x, y); break;
      case 8: bsr3_spmv_k<8, V, X><<<g, 256, 0, s>>>(nb, op->rowptr, op->col, vals, x, y); break;
      case 16: bsr3_spmv_k<16, V, X><<<g, 256, 0, s>>>(nb, op->rowptr, op->col, vals, x, y); break;
      case 32: bsr3_spmv_k<32, V, X><<<g, 256, 0, s>>>(nb, op->rowptr, op->col, vals, x, y); break;
      case 64: bsr3_spmv_k<64, V, X><<<g, 256, 0, s>>>(nb, op->rowptr, op->col, vals, x, y); break;
      default: JNO_REQUIRE(false, JNO_EINVAL, "spmv: lanes per block row must be a power of two in 1..64");
    }
  } else {
    const int L = op->lanes;
    const int g = grid_for(n * L, 256, 16384);
    switch (L) {
      case 1: csr_spmv_k<1, V, X><<<g, 256, 0, s>>>(n, op->rowptr, op->col, vals, x, y); break;
      case 2: csr_spmv_k<2, V, X><<<g, 256, 0, s>>>(n, op->rowptr, op->col, vals, x, y); break;
      case 4: csr_spmv_k<4, V, X><<<g, 256, 0, s>>>(n, op->rowptr, op->col, vals, x, y); break;
      case 8: csr_spmv_k<8, V, X><<<g, 256, 0, s>>>(n, op->rowptr, op->col, vals, x, y); break;
      case 16: csr_spmv_k<16, V, X><<<g, 256, 0, s>>>(n, op->rowptr, op->col, vals, x, y); break;
      case 32: csr_spmv_k<32, V, X><<<g, 256, 0, s>>>(n, op->rowptr, op->col, vals, x, y); break;
      case 64: csr_spmv_k<64, V, X><<<g, 256, 0, s>>>(n, op->rowptr, op->col, vals, x, y); break;
      default: JNO_REQUIRE(false, JNO_EINVAL, "spmv: lanes per row must be a power of two in 1..64");
    }
  }
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

template int op_spmv<double, double>(const jno_op*, const double*, const double*, double*, hipStream_t, void*);
template int op_spmv<float, float>(const jno_op*, const float*, const float*, float*, hipStream_t, void*);

// ---------------------------------------------------------------------------------------------
// Jacobi, Dirichlet, cast
// ---------------------------------------------------------------------------------------------
__global__ void jacobi_k(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                         const double* __restrict__ vals, double* __restrict__ dinv) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double d = 0.0;  // duplicates on the diagonal are summed (matrix_diagonal, linear.py:88-94)
    for (int j = rowptr[r]; j < rowptr[r + 1]; ++j)
      if (col[j] == (int)r) d += vals[j];
    dinv[r] = fabs(d) > 1e-30 ? 1.0 / d : 1.0;
  }
}

__global__ void dirichlet_k(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                            double* __restrict__ vals, double* __restrict__ b, const uint8_t* __restrict__ dmask,
                            const double* __restrict__ g) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const int rb = rowptr[r], re = rowptr[r + 1];
    if (dmask[r]) {
      bool seen = false;  // a duplicated diagonal would sum to 1 only if one copy carries it
      for (int j = rb; j < re; ++j) {
        const bool diag = col[j] == (int)r && !seen;
        seen |= diag;
        vals[j] = diag ? 1.0 : 0.0;
      }
      if (b) b[r] = g ? g[r] : 0.0;
    } else {
      double lift = 0.0;
      for (int j = rb; j < re; ++j) {
        const int c = col[j];
        if (dmask[c]) {
          if (g) lift += vals[j] * g[c];
          vals[j] = 0.0;
        }
      }
      if (b) b[r] -= lift;
    }
  }
}

__global__ void cast_f32_k(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------------------------
// exact-zero compaction (compress_eager, fem_utils.py:4063-4091)
// ---------------------------------------------------------------------------------------------
__global__ void nz_count_k(int64_t n, const int32_t* __restrict__ rowptr, const double* __restrict__ vals,
                           int32_t* __restrict__ cnt) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int c = 0;
    for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) c += fabs(vals[j]) > 0.0 ? 1 : 0;
    cnt[r] = c;
  }
}

__global__ void nz_fill_k(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                          const double* __restrict__ vals, const int32_t* __restrict__ rowptr_o,
                          int32_t* __restrict__ col_o, double* __restrict__ vals_o) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int o = rowptr_o[r];
    for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      const double v = vals[j];
      if (fabs(v) > 0.0) {
        col_o[o] = col[j];
        vals_o[o] = v;
        ++o;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// operator analysis: distinct diagonal offsets (LDS hash per block, merged into a global table)
// and numerical symmetry
// ---------------------------------------------------------------------------------------------
constexpr int OFF_SLOTS = 128;
constexpr int32_t OFF_EMPTY = INT32_MIN;

__device__ __forceinline__ bool hash_insert(int32_t* table, int32_t key) {
  uint32_t h = ((uint32_t)key * 2654435761u) & (OFF_SLOTS - 1);
  for (int p = 0; p < OFF_SLOTS; ++p) {
    int32_t cur = table[h];
    if (cur == key) return true;
    if (cur == OFF_EMPTY) {
      int32_t old = atomicCAS(&table[h], OFF_EMPTY, key);
      if (old == OFF_EMPTY || old == key) return true;
    }
    h = (h + 1) & (OFF_SLOTS - 1);
  }
  return false;
}

// Entries whose value is exactly 0.0 (when vals is given) are skipped: they are the couplings
// compress_eager drops, and a DIA layout does not need a diagonal for them.
__global__ void __launch_bounds__(256) offsets_k(int64_t n, const int32_t* __restrict__ rowptr,
                                                 const int32_t* __restrict__ col, const double* __restrict__ vals,
                                                 int32_t* __restrict__ gtable, int32_t* __restrict__ overflow) {
  __shared__ int32_t table[OFF_SLOTS];
  __shared__ int s_over;
  for (int i = threadIdx.x; i < OFF_SLOTS; i += blockDim.x) table[i] = OFF_EMPTY;
  if (threadIdx.x == 0) s_over = 0;
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    if (s_over) break;
    for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      if (vals && vals[j] == 0.0) continue;
      if (!hash_insert(table, col[j] - (int32_t)r)) {
        s_over = 1;
        break;
      }
    }
  }
  __syncthreads();
  if (s_over) {
    if (threadIdx.x == 0) atomicExch(overflow, 1);
    return;
  }
  for (int i = threadIdx.x; i < OFF_SLOTS; i += blockDim.x) {
    const int32_t k = table[i];
    if (k != OFF_EMPTY && !hash_insert(gtable, k)) atomicExch(overflow, 1);
  }
}

__global__ void symmetry_k(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                           const double* __restrict__ vals, int32_t* __restrict__ bad) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      const int c = col[j];
      if (c == (int)r || (vals && vals[j] == 0.0)) continue;
      // binary search (c, r)
      int lo = rowptr[c], hi = rowptr[c + 1] - 1, hit = -1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int cm = col[mid];
        if (cm == (int)r) { hit = mid; break; }
        if (cm < (int)r) lo = mid + 1; else hi = mid - 1;
      }
      const bool ok = hit >= 0 ? (vals ? vals[hit] == vals[j] : true) : (vals ? vals[j] == 0.0 : false);
      if (!ok) atomicAdd(bad, 1);
    }
  }
}

__global__ void dia_fill_k(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                           const double* __restrict__ vals, int ndiag, DiaOffsets off, double* __restrict__ d64,
                           float* __restrict__ d32) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double acc[JNO_DIA_MAX];
#pragma unroll
    for (int d = 0; d < JNO_DIA_MAX; ++d) acc[d] = 0.0;
    for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      const int o = col[j] - (int)r;
      const double v = vals[j];
#pragma unroll
      for (int d = 0; d < JNO_DIA_MAX; ++d)
        if (d < ndiag && off.o[d] == o) acc[d] += v;
    }
#pragma unroll
    for (int d = 0; d < JNO_DIA_MAX; ++d) {
      if (d < ndiag) {
        if (d64) d64[(int64_t)d * dia_ld(n) + r] = acc[d];
        if (d32) d32[(int64_t)d * dia_ld(n) + r] = (float)acc[d];
      }
    }
  }
}

// After dia_fill_k: every stored entry off the stored diagonal set must be exactly 0, every lower
// entry (r, c), c < r, must equal bitwise the stored upper value A[c, r] = dia[d][c], and every
// nonzero upper entry (r, c), c > r, must have its lower twin (c, r) in the pattern (the layout reads
// A[c, r] through symmetry; rows hold ascending columns).  Sets *bad otherwise (the layout would
// not reproduce A).
__global__ void dia_check_k(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                            const double* __restrict__ vals, int ndiag, DiaOffsets off,
                            const double* __restrict__ d64, int* __restrict__ bad) {
  const int64_t ld = dia_ld(n);
  bool ok = true;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      const int c = col[j];
      const int o = c - (int)r;
      const int a = o < 0 ? -o : o;
      int dd = -1;
#pragma unroll
      for (int d = 0; d < JNO_DIA_MAX; ++d)
        if (d < ndiag && off.o[d] == a) dd = d;
      const double v = vals[j];
      if (dd < 0) {
        ok = ok && v == 0.0;
      } else if (o < 0) {
        ok = ok && __double_as_longlong(v) == __double_as_longlong(d64[(int64_t)dd * ld + c]);
      } else if (o > 0 && v != 0.0) {
        int lo = rowptr[c], hi = rowptr[c + 1];
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (col[mid] < (int)r) lo = mid + 1;
          else hi = mid;
        }
        ok = ok && lo < rowptr[c + 1] && col[lo] == (int)r;
      }
    }
  }
  if (!ok) atomicOr(bad, 1);
}

}  // namespace jno

using namespace jno;

extern "C" int jno_jacobi(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
                          double* d_dinv, void* stream) {
  if (n_rows == 0) return JNO_OK;
  jacobi_k<<<grid_for(n_rows, 256), 256, 0, as_stream(stream)>>>(n_rows, d_rowptr, d_col, d_vals, d_dinv);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_dirichlet_eliminate(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                                       double* d_vals, double* d_b, const uint8_t* d_dmask, const double* d_g,
                                       void* stream) {
  JNO_REQUIRE(d_dmask, JNO_EINVAL, "jno_dirichlet_eliminate: mask is required");
  if (n_rows == 0) return JNO_OK;
  dirichlet_k<<<grid_for(n_rows, 256), 256, 0, as_stream(stream)>>>(n_rows, d_rowptr, d_col, d_vals, d_b, d_dmask,
                                                                    d_g);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_spmv_f64(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
                            const double* d_x, double* d_y, void* stream) {
  jno_op op{};
  op.format = JNO_FMT_CSR;
  op.lanes = 8;
  op.n = n_rows;
  op.rowptr = d_rowptr;
  op.col = d_col;
  op.vals64 = d_vals;
  return op_spmv<double, double>(&op, d_vals, d_x, d_y, as_stream(stream));
}

extern "C" int jno_op_spmv_f64(const jno_op* h_op, const double* d_x, double* d_y, void* stream) {
  JNO_REQUIRE(h_op, JNO_EINVAL, "jno_op_spmv_f64: null operator");
  return op_spmv<double, double>(h_op, h_op->vals64, d_x, d_y, as_stream(stream));
}

extern "C" int jno_op_spmv_f32(const jno_op* h_op, const float* d_x, float* d_y, void* stream) {
  JNO_REQUIRE(h_op, JNO_EINVAL, "jno_op_spmv_f32: null operator");
  return op_spmv<float, float>(h_op, h_op->vals32, d_x, d_y, as_stream(stream));
}

extern "C" int jno_cast_f32(int64_t n, const double* d_in, float* d_out, void* stream) {
  if (n == 0) return JNO_OK;
  cast_f32_k<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(n, d_in, d_out);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_drop_zeros_count(int64_t n_rows, const int32_t* d_rowptr, const double* d_vals,
                                    int32_t* d_rowptr_out, int64_t* h_nnz_out, void* d_ws, size_t ws_bytes,
                                    void* stream) {
  hipStream_t s = as_stream(stream);
  const size_t need = (size_t)(n_rows + 1) * 4 + 256 + exclusive_scan_ws(n_rows);
  JNO_REQUIRE(ws_bytes >= need, JNO_EWORKSPACE, "jno_drop_zeros_count: workspace %zu < %zu", ws_bytes, need);
  int32_t* cnt = static_cast<int32_t*>(d_ws);
  char* w = static_cast<char*>(d_ws) + (((size_t)(n_rows + 1) * 4 + 255) & ~(size_t)255);
  int64_t* total = reinterpret_cast<int64_t*>(w);
  w += 256;
  nz_count_k<<<grid_for(n_rows, 256), 256, 0, s>>>(n_rows, d_rowptr, d_vals, cnt);
  JNO_LAUNCH_CHECK();
  int rc = exclusive_scan_i32(cnt, d_rowptr_out, n_rows, total, w, exclusive_scan_ws(n_rows), s);
  if (rc) return rc;
  int64_t h_total = 0;
  JNO_HIP(hipMemcpyAsync(&h_total, total, 8, hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  if (h_nnz_out) *h_nnz_out = h_total;
  return JNO_OK;
}

extern "C" int jno_drop_zeros_fill(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                                   const double* d_vals, const int32_t* d_rowptr_out, int32_t* d_col_out,
                                   double* d_vals_out, void* stream) {
  if (n_rows == 0) return JNO_OK;
  nz_fill_k<<<grid_for(n_rows, 256), 256, 0, as_stream(stream)>>>(n_rows, d_rowptr, d_col, d_vals, d_rowptr_out,
                                                                   d_col_out, d_vals_out);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" size_t jno_op_analyze_workspace_size(void) { return 1024; }

extern "C" int jno_op_analyze(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
                              int32_t* h_ndiag, int32_t* h_offsets, void* d_ws, size_t ws_bytes, void* stream) {
  JNO_REQUIRE(ws_bytes >= jno_op_analyze_workspace_size(), JNO_EWORKSPACE, "jno_op_analyze: workspace too small");
  JNO_REQUIRE(h_ndiag && h_offsets, JNO_EINVAL, "jno_op_analyze: null output");
  hipStream_t s = as_stream(stream);
  int32_t* table = static_cast<int32_t*>(d_ws);
  int32_t* flags = table + OFF_SLOTS;  // [0] overflow, [1] asymmetric count
  {
    int32_t h_init[OFF_SLOTS + 2];
    for (int i = 0; i < OFF_SLOTS; ++i) h_init[i] = OFF_EMPTY;
    h_init[OFF_SLOTS] = 0;
    h_init[OFF_SLOTS + 1] = 0;
    JNO_HIP(hipMemcpyAsync(table, h_init, sizeof(h_init), hipMemcpyHostToDevice, s));
    JNO_HIP(hipStreamSynchronize(s));
  }
  offsets_k<<<grid_for(n_rows, 256, 2048), 256, 0, s>>>(n_rows, d_rowptr, d_col, d_vals, table, flags);
  JNO_LAUNCH_CHECK();
  symmetry_k<<<grid_for(n_rows, 256, 8192), 256, 0, s>>>(n_rows, d_rowptr, d_col, d_vals, flags + 1);
  JNO_LAUNCH_CHECK();
  int32_t h[OFF_SLOTS + 2];
  JNO_HIP(hipMemcpyAsync(h, table, sizeof(h), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  if (h[OFF_SLOTS] || h[OFF_SLOTS + 1]) {
    *h_ndiag = -1;
    return JNO_OK;
  }
  int32_t offs[OFF_SLOTS];
  int cnt = 0;
  for (int i = 0; i < OFF_SLOTS; ++i)
    if (h[i] != OFF_EMPTY) offs[cnt++] = h[i];
  // ascending insertion sort
  for (int i = 1; i < cnt; ++i)
    for (int j = i; j > 0 && offs[j - 1] > offs[j]; --j) {
      int32_t t = offs[j];
      offs[j] = offs[j - 1];
      offs[j - 1] = t;
    }
  int nd = 0;
  bool has0 = false;
  for (int i = 0; i < cnt; ++i) {
    if (offs[i] == 0) has0 = true;
    if (offs[i] >= 0) {
      if (nd >= JNO_DIA_MAX) {
        *h_ndiag = -1;
        return JNO_OK;
      }
      nd++;
    }
  }
  int k = 0;
  if (!has0) {
    if (nd >= JNO_DIA_MAX) {
      *h_ndiag = -1;
      return JNO_OK;
    }
    h_offsets[k++] = 0;
  }
  for (int i = 0; i < cnt; ++i)
    if (offs[i] >= 0) h_offsets[k++] = offs[i];
  *h_ndiag = k;
  return JNO_OK;
}

extern "C" int jno_op_fill_dia(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
                               int ndiag, const int32_t* h_offsets, double* d_dia64, float* d_dia32, void* stream) {
  JNO_REQUIRE(ndiag >= 1 && ndiag <= JNO_DIA_MAX && h_offsets, JNO_EINVAL, "jno_op_fill_dia: bad offsets");
  JNO_REQUIRE(h_offsets[0] == 0, JNO_EINVAL, "jno_op_fill_dia: first offset must be the main diagonal");
  DiaOffsets off;
  for (int d = 0; d < JNO_DIA_MAX; ++d) off.o[d] = d < ndiag ? h_offsets[d] : 0;
  if (n_rows == 0) return JNO_OK;
  dia_fill_k<<<grid_for(n_rows, 256), 256, 0, as_stream(stream)>>>(n_rows, d_rowptr, d_col, d_vals, ndiag, off,
                                                                    d_dia64, d_dia32);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_op_fill_dia_checked(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                                       const double* d_vals, int ndiag, const int32_t* h_offsets, double* d_dia64,
                                       float* d_dia32, int* h_ok, void* d_ws, size_t ws_bytes, void* stream) {
  JNO_REQUIRE(h_ok && d_dia64 && d_ws, JNO_EINVAL, "jno_op_fill_dia_checked: null pointer");
  JNO_REQUIRE(ws_bytes >= 256, JNO_EWORKSPACE, "jno_op_fill_dia_checked: workspace too small (256 bytes)");
  *h_ok = 0;
  int rc = jno_op_fill_dia(n_rows, d_rowptr, d_col, d_vals, ndiag, h_offsets, d_dia64, d_dia32, stream);
  if (rc) return rc;
  if (n_rows == 0) {
    *h_ok = 1;
    return JNO_OK;
  }
  hipStream_t s = as_stream(stream);
  DiaOffsets off;
  for (int d = 0; d < JNO_DIA_MAX; ++d) off.o[d] = d < ndiag ? h_offsets[d] : 0;
  int* bad = static_cast<int*>(d_ws);
  JNO_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  dia_check_k<<<grid_for(n_rows, 256), 256, 0, s>>>(n_rows, d_rowptr, d_col, d_vals, ndiag, off, d_dia64, bad);
  JNO_LAUNCH_CHECK();
  int h_bad = 1;
  JNO_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  *h_ok = h_bad ? 0 : 1;
  return JNO_OK;
}

extern "C" size_t jno_op_bsr3_workspace_size(void) { return 256; }

extern "C" int jno_op_bsr3_count(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, int64_t* h_nnzb,
                                 void* d_ws, size_t ws_bytes, void* stream) {
  JNO_REQUIRE(h_nnzb, JNO_EINVAL, "jno_op_bsr3_count: null h_nnzb");
  *h_nnzb = -1;
  if (n_rows <= 0 || n_rows % 3 != 0) return JNO_OK;
  JNO_REQUIRE(d_rowptr && d_col && d_ws, JNO_EINVAL, "jno_op_bsr3_count: null pointer");
  JNO_REQUIRE(ws_bytes >= jno_op_bsr3_workspace_size(), JNO_EWORKSPACE, "jno_op_bsr3_count: workspace too small");
  hipStream_t s = as_stream(stream);
  int* bad = static_cast<int*>(d_ws);
  JNO_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  const int64_t nb = n_rows / 3;
  bsr3_check_k<<<grid_for(nb, 256), 256, 0, s>>>(nb, d_rowptr, d_col, bad);
  JNO_LAUNCH_CHECK();
  int h[2] = {0, 0};
  JNO_HIP(hipMemcpyAsync(&h[0], bad, sizeof(int), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipMemcpyAsync(&h[1], d_rowptr + n_rows, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  if (!h[0] && h[1] % 9 == 0) *h_nnzb = h[1] / 9;
  return JNO_OK;
}

extern "C" int jno_op_fill_bsr3(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
                                int32_t* d_browptr, int32_t* d_bcol, double* d_bvals64, float* d_bvals32,
                                void* stream) {
  JNO_REQUIRE(n_rows > 0 && n_rows % 3 == 0, JNO_EINVAL, "jno_op_fill_bsr3: rows must be a positive multiple of 3");
  JNO_REQUIRE(d_rowptr && d_col && d_vals && d_browptr && d_bcol, JNO_EINVAL, "jno_op_fill_bsr3: null pointer");
  const int64_t nb = n_rows / 3;
  bsr3_fill_k<<<grid_for(nb, 256), 256, 0, as_stream(stream)>>>(nb, d_rowptr, d_col, d_vals, d_browptr, d_bcol,
                                                                 d_bvals64, d_bvals32);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
