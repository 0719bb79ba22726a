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

}  // namespace jno
