// Error reporting and the device-wide exclusive scan shared by the pattern and compaction passes.
#include <stdarg.h>
#include <string.h>

#include "common.h"

namespace jno {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

// ---- exclusive scan ------------------------------------------------------------------------
constexpr int SCAN_NT = 256;
constexpr int SCAN_PER = 4;
constexpr int SCAN_TILE = SCAN_NT * SCAN_PER;

size_t exclusive_scan_ws(int64_t n) { return (size_t)(cdiv(n, SCAN_TILE) + 1) * sizeof(int64_t); }

__global__ void __launch_bounds__(SCAN_NT) scan_tile_sums(const int32_t* __restrict__ in, int64_t n,
                                                           int64_t* __restrict__ tile_sums) {
  __shared__ int64_t sm[SCAN_NT / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_PER; ++i) {
    int64_t idx = base + (int64_t)i * SCAN_NT + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  s = block_sum<SCAN_NT>(s, sm);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

// One block scans the tile sums in place (exclusive) and writes the grand total.
__global__ void __launch_bounds__(1024) scan_tile_prefix(int64_t* __restrict__ tile_sums, int64_t ntiles,
                                                         int64_t* __restrict__ total) {
  __shared__ int64_t sm[1024];
  int64_t carry = 0;
  for (int64_t base = 0; base < ntiles; base += 1024) {
    int64_t idx = base + threadIdx.x;
    int64_t v = idx < ntiles ? tile_sums[idx] : 0;
    sm[threadIdx.x] = v;
    __syncthreads();
    // Hillis-Steele inclusive scan in LDS
    for (int o = 1; o < 1024; o <<= 1) {
      int64_t t = threadIdx.x >= o ? sm[threadIdx.x - o] : 0;
      __syncthreads();
      sm[threadIdx.x] += t;
      __syncthreads();
    }
    int64_t incl = sm[threadIdx.x];
    if (idx < ntiles) tile_sums[idx] = carry + incl - v;
    int64_t blk_total = sm[1023];
    __syncthreads();
    carry += blk_total;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(SCAN_NT) scan_tile_apply(const int32_t* __restrict__ in, int64_t n,
                                                            const int64_t* __restrict__ tile_sums,
                                                            const int64_t* __restrict__ total,
                                                            int32_t* __restrict__ out) {
  __shared__ int64_t sm[SCAN_TILE];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  // load tile (coalesced), scan in LDS per thread chunk
  for (int i = 0; i < SCAN_PER; ++i) {
    int64_t idx = base + (int64_t)i * SCAN_NT + threadIdx.x;
    sm[i * SCAN_NT + threadIdx.x] = idx < n ? in[idx] : 0;
  }
  __syncthreads();
  // each thread sums its contiguous SCAN_PER run
  int64_t run[SCAN_PER];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_PER; ++i) {
    run[i] = s;
    s += sm[threadIdx.x * SCAN_PER + i];
  }
  __shared__ int64_t ts[SCAN_NT];
  ts[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < SCAN_NT; o <<= 1) {
    int64_t t = threadIdx.x >= o ? ts[threadIdx.x - o] : 0;
    __syncthreads();
    ts[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t excl = ts[threadIdx.x] - s + tile_sums[blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < SCAN_PER; ++i) sm[threadIdx.x * SCAN_PER + i] = excl + run[i];
  __syncthreads();
  for (int i = 0; i < SCAN_PER; ++i) {
    int64_t idx = base + (int64_t)i * SCAN_NT + threadIdx.x;
    if (idx < n) out[idx] = (int32_t)sm[i * SCAN_NT + threadIdx.x];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = (int32_t)*total;
}

int exclusive_scan_i32(const int32_t* d_in, int32_t* d_out, int64_t n, int64_t* d_total, void* d_ws,
                       size_t ws_bytes, hipStream_t stream) {
  JNO_REQUIRE(ws_bytes >= exclusive_scan_ws(n), JNO_EWORKSPACE, "scan workspace too small");
  int64_t ntiles = cdiv(n, SCAN_TILE);
  if (ntiles == 0) ntiles = 1;
  int64_t* tile_sums = static_cast<int64_t*>(d_ws);
  scan_tile_sums<<<(unsigned)ntiles, SCAN_NT, 0, stream>>>(d_in, n, tile_sums);
  JNO_LAUNCH_CHECK();
  scan_tile_prefix<<<1, 1024, 0, stream>>>(tile_sums, ntiles, d_total);
  JNO_LAUNCH_CHECK();
  scan_tile_apply<<<(unsigned)ntiles, SCAN_NT, 0, stream>>>(d_in, n, tile_sums, d_total, d_out);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

}  // namespace jno

extern "C" int jno_last_error(char* buf, size_t len) {
  const std::string& e = jno::g_last_error;
  if (buf && len) {
    size_t n = e.size() < len - 1 ? e.size() : len - 1;
    memcpy(buf, e.data(), n);
    buf[n] = 0;
  }
  return (int)e.size();
}

extern "C" const char* jno_version(void) { return "jno_fem 0.1 (gfx950)"; }
