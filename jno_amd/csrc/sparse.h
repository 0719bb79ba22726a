// Internal declarations shared by the sparse kernels and the PCG engine.
#pragma once
#include "common.h"

namespace jno {

struct DiaOffsets {
  int32_t o[JNO_DIA_MAX];
};

template <typename V, typename X>
int op_spmv(const jno_op* op, const V* vals, const X* x, X* y, hipStream_t s);

}  // namespace jno
