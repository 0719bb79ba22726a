// Shared helpers of the jno_fem HIP library (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/jno_fem.h"

namespace jno {

void set_error(const char* fmt, ...);

#define JNO_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t _e = (call);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      ::jno::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(_e));   \
      return JNO_EHIP;                                                                       \
    }                                                                                        \
  } while (0)

#define JNO_REQUIRE(cond, code, ...)   \
  do {                                 \
    if (!(cond)) {                     \
      ::jno::set_error(__VA_ARGS__);   \
      return (code);                   \
    }                                  \
  } while (0)

#define JNO_LAUNCH_CHECK() JNO_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int WAVE = 64;

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Number of workgroups for a grid-stride kernel: enough to fill all 256 CUs several times over.
inline int grid_for(int64_t work, int block, int max_blocks = 4096) {
  int64_t g = cdiv(work, block);
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return static_cast<int>(g);
}

// ---- wavefront reductions (wave64; every lane of the wave active) --------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// v moved across lanes by the DPP control CTRL (rows outside ROW_MASK read 0), as two 32-bit moves
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The FP64 wave sum by DPP row operations (VALU lane moves, no LDS round trips): pairs and quads
// (quad_perm [1,0,3,2], [2,3,0,1]), half rows and rows (row_half_mirror, row_mirror), then rows 0+1
// and 2+3 (row_bcast:15 into rows 1 and 3) and all four (row_bcast:31 into rows 2 and 3); lane 63
// ends with the sum, which is broadcast.  A fixed order, so the result is deterministic.
template <>
__device__ __forceinline__ double wave_sum<double>(double v) {
  v += dpp_move<0xB1, 0xF>(v);
  v += dpp_move<0x4E, 0xF>(v);
  v += dpp_move<0x141, 0xF>(v);
  v += dpp_move<0x140, 0xF>(v);
  v += dpp_move<0x142, 0xA>(v);
  v += dpp_move<0x143, 0xC>(v);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum for blockDim.x = NT (multiple of 64).  All threads get the result.
template <int NT, typename T>
__device__ __forceinline__ T block_sum(T v, T* smem /* >= NT/64 */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  T t = (threadIdx.x < NT / 64) ? smem[threadIdx.x] : T(0);
  if (wid == 0) t = wave_sum(t);
  if (threadIdx.x == 0) smem[0] = t;
  __syncthreads();
  T r = smem[0];
  __syncthreads();
  return r;
}

// a x b with every product and difference rounded separately (no FMA contraction)
__device__ __forceinline__ void cross_rn(const double* a, const double* b, double* c) {
  c[0] = __dsub_rn(__dmul_rn(a[1], b[2]), __dmul_rn(a[2], b[1]));
  c[1] = __dsub_rn(__dmul_rn(a[2], b[0]), __dmul_rn(a[0], b[2]));
  c[2] = __dsub_rn(__dmul_rn(a[0], b[1]), __dmul_rn(a[1], b[0]));
}

// Exclusive scan of int32 counts into int32 offsets (n+1 entries, last = total).  Three-phase,
// workspace: cdiv(n, 1024) + 1 int64.  Writes the total to *d_total (int64, device).
int exclusive_scan_i32(const int32_t* d_in, int32_t* d_out, int64_t n, int64_t* d_total, void* d_ws,
                       size_t ws_bytes, hipStream_t stream);
size_t exclusive_scan_ws(int64_t n);

// numpy.linspace(lo, hi, n+1)[i]: i*step + lo with step = (hi-lo)/n, last entry exactly hi.
// The products and sums are rounded separately (no FMA contraction) so the value is the same
// double numpy produces.
__device__ __forceinline__ double linspace_at(double lo, double hi, int n, int i) {
  if (i == n) return hi;
  const double step = __ddiv_rn(__dsub_rn(hi, lo), (double)n);
  return __dadd_rn(__dmul_rn((double)i, step), lo);
}

}  // namespace jno
