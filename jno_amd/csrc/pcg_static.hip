// The allocation-free mixed-precision PCG entry point (jno_pcg_mixed_noalloc).
//
// The solve is jno_pcg_mixed's (same kernels, launch shapes and results); its workspace is the
// static device array `scratch` below instead of a caller's buffer, and the device state is read
// back synchronously into host memory of the call itself, so the call allocates nothing: no device
// or pinned memory, no events, no streams.  The array lives in this file's own code object: it is
// placed on a GPU (JNO_PCG_SCRATCH_BYTES there, until the process ends) when this entry point is
// first called on that GPU.  Every call on one GPU shares that GPU's array and leaves its contents
// behind; the solve clears what must start at zero.  Calls that overlap on one GPU (two threads or
// two streams at once) race on it.
#include <mutex>

#include "common.h"

namespace jno {
int pcg_mixed_check(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x, int64_t maxiter,
                    double replace_factor, const char* who);
int pcg_mixed_impl(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x, int use_x0, double tol,
                   double atol, int64_t maxiter, double replace_factor, jno_pcg_stats* h_stats, void* d_ws,
                   size_t ws_bytes, hipStream_t s, bool sync_poll);
}  // namespace jno

using namespace jno;

static __device__ __align__(256) unsigned char scratch[JNO_PCG_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[MAX_DEVICES] = {nullptr};  // the array's device address, per GPU

// The current GPU's copy of `scratch`, looked up once per GPU with that GPU current.  Host code
// uses this address (never the array's name, which on the host is only a stand-in).
static int scratch_ptr(void** out) {
  int dev = -1;
  JNO_HIP(hipGetDevice(&dev));
  JNO_REQUIRE(dev >= 0 && dev < MAX_DEVICES, JNO_EINVAL, "jno_pcg_mixed_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    JNO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    JNO_REQUIRE(p != nullptr, JNO_EHIP, "jno_pcg_mixed_noalloc: no device address for the static scratch");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return JNO_OK;
}

extern "C" int jno_pcg_mixed_noalloc(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x,
                                     int use_x0, double tol, double atol, int64_t maxiter, double replace_factor,
                                     jno_pcg_stats* h_stats, void* stream) {
  int rc = pcg_mixed_check(h_op, d_dinv, d_b, d_x, maxiter, replace_factor, "jno_pcg_mixed_noalloc");
  if (rc) return rc;
  // the operator's full mixed workspace (with room for the Jacobi-scaled FP32 copy of a DIA
  // operator) must fit; nothing is queued otherwise
  const size_t need = jno_pcg_workspace_size_op(h_op, 1);
  JNO_REQUIRE(need <= JNO_PCG_SCRATCH_BYTES, JNO_ECAPACITY,
              "jno_pcg_mixed_noalloc: a %lld-row operator needs %zu bytes of workspace, more than the %zu-byte "
              "static scratch (use jno_pcg_mixed with a caller workspace)",
              (long long)h_op->n, need, (size_t)JNO_PCG_SCRATCH_BYTES);
  void* ws = nullptr;
  rc = scratch_ptr(&ws);
  if (rc) return rc;
  return pcg_mixed_impl(h_op, d_dinv, d_b, d_x, use_x0, tol, atol, maxiter, replace_factor, h_stats, ws,
                        JNO_PCG_SCRATCH_BYTES, as_stream(stream), true);
}
