// A7 CSR sparsity pattern build (the bit-exact index structure) and the node->element incidence
// that the row-pull assembly reuses.
//
//   count:  node degree histogram -> scan -> incidence fill -> per-node (one wavefront each):
//           rank-sort the incidence list (deterministic assembly order), gather every node of every
//           incident element into LDS, count distinct ones -> dof-level row counts -> scan = rowptr
//   fill:   per node again: distinct neighbours in ascending order, expanded to dofs n*vec+c.
//
// The result equals np.unique over the (row, col) keys of all element pairs, which is what
// compress_plan (jno/utils/solver/fem_utils.py:4009-4060) returns.
#include "common.h"

namespace jno {

constexpr int PAT_NT = 256;               // 4 wavefronts per block, one node per wavefront
constexpr int PAT_MAX_DEG = 512;          // incident (element, local) entries per node
constexpr int PAT_MAX_CAND = 2048;        // candidate neighbour nodes per node (deg * nloc)

__global__ void degree_count(const int32_t* __restrict__ cells, int64_t m, int32_t* __restrict__ deg) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&deg[cells[i]], 1);
}

__global__ void incidence_fill(const int32_t* __restrict__ cells, int64_t m, int32_t* __restrict__ cursor,
                               int32_t* __restrict__ inc) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    int pos = atomicAdd(&cursor[cells[i]], 1);
    inc[pos] = (int32_t)i;
  }
}

__global__ void fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void copy_i32(const int32_t* __restrict__ a, int32_t* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// Gather the candidate neighbour nodes of `node` into LDS `cand`; returns how many.
__device__ __forceinline__ int gather_candidates(const int32_t* __restrict__ cells, int nloc,
                                                 const int32_t* __restrict__ inc, int beg, int deg,
                                                 int32_t* cand, int lane) {
  const int m = deg * nloc;
  for (int i = lane; i < m; i += WAVE) {
    int ent = inc[beg + i / nloc];
    int e = ent / nloc;
    cand[i] = cells[(int64_t)e * nloc + (i % nloc)];
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  return m;
}

// first[i] = 1 when no earlier candidate has the same value.  Writes flags into `flag`.
__device__ __forceinline__ int mark_first(const int32_t* cand, int m, uint8_t* flag, int lane) {
  int cnt = 0;
  for (int i = lane; i < m; i += WAVE) {
    const int v = cand[i];
    bool first = true;
    for (int j = 0; j < i; ++j) first &= (cand[j] != v);
    flag[i] = first ? 1 : 0;
    cnt += first ? 1 : 0;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  return wave_sum_int(cnt);
}

__global__ void __launch_bounds__(PAT_NT) pattern_count_k(const int32_t* __restrict__ cells, int nloc, int vec,
                                                          int64_t n_nodes, const int32_t* __restrict__ inc_ptr,
                                                          int32_t* __restrict__ inc, int32_t* __restrict__ rowcnt,
                                                          int32_t* __restrict__ status) {
  __shared__ int32_t s_cand[PAT_NT / WAVE][PAT_MAX_CAND];
  __shared__ int32_t s_inc[PAT_NT / WAVE][PAT_MAX_DEG];
  __shared__ uint8_t s_flag[PAT_NT / WAVE][PAT_MAX_CAND];
  const int wid = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  int32_t* cand = s_cand[wid];
  int32_t* sinc = s_inc[wid];
  uint8_t* flag = s_flag[wid];
  const int64_t nwaves = (int64_t)gridDim.x * (PAT_NT / WAVE);
  for (int64_t node = blockIdx.x * (int64_t)(PAT_NT / WAVE) + wid; node < n_nodes; node += nwaves) {
    const int beg = inc_ptr[node], deg = inc_ptr[node + 1] - beg;
    if (deg > PAT_MAX_DEG || deg * nloc > PAT_MAX_CAND) {
      if (lane == 0) atomicMax(status, deg);
      continue;
    }
    // rank-sort the incidence entries (all distinct) so assembly sums in ascending element order
    for (int i = lane; i < deg; i += WAVE) sinc[i] = inc[beg + i];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (int i = lane; i < deg; i += WAVE) {
      const int v = sinc[i];
      int r = 0;
      for (int j = 0; j < deg; ++j) r += (sinc[j] < v) ? 1 : 0;
      inc[beg + r] = v;
    }
    const int m = gather_candidates(cells, nloc, sinc, 0, deg, cand, lane);
    const int uniq = mark_first(cand, m, flag, lane);
    if (lane < vec) rowcnt[node * vec + lane] = uniq * vec;
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ void __launch_bounds__(PAT_NT) pattern_fill_k(const int32_t* __restrict__ cells, int nloc, int vec,
                                                         int64_t n_nodes, const int32_t* __restrict__ inc_ptr,
                                                         const int32_t* __restrict__ inc,
                                                         const int32_t* __restrict__ rowptr,
                                                         int32_t* __restrict__ col) {
  __shared__ int32_t s_cand[PAT_NT / WAVE][PAT_MAX_CAND];
  __shared__ uint8_t s_flag[PAT_NT / WAVE][PAT_MAX_CAND];
  const int wid = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  int32_t* cand = s_cand[wid];
  uint8_t* flag = s_flag[wid];
  const int64_t nwaves = (int64_t)gridDim.x * (PAT_NT / WAVE);
  for (int64_t node = blockIdx.x * (int64_t)(PAT_NT / WAVE) + wid; node < n_nodes; node += nwaves) {
    const int beg = inc_ptr[node], deg = inc_ptr[node + 1] - beg;
    const int m = gather_candidates(cells, nloc, inc, beg, deg, cand, lane);
    const int uniq = mark_first(cand, m, flag, lane);
    // position of each first occurrence among the distinct values
    for (int i = lane; i < m; i += WAVE) {
      if (!flag[i]) continue;
      const int v = cand[i];
      int pos = 0;
      for (int j = 0; j < m; ++j) pos += (flag[j] && cand[j] < v) ? 1 : 0;
      for (int c = 0; c < vec; ++c) {
        const int64_t row = node * vec + c;
        int32_t* out = col + rowptr[row] + (int64_t)pos * vec;
        for (int c2 = 0; c2 < vec; ++c2) out[c2] = v * vec + c2;
      }
    }
    (void)uniq;
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ void max_row_k(const int32_t* __restrict__ rowcnt, int64_t n, int32_t* __restrict__ out) {
  int v = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v = max(v, rowcnt[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, v);
}

// workspace layout: [deg/cursor i32 n_nodes+1][rowcnt i32 ndof+1][status i32 4][total i64 2][scan ws]
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace jno

using namespace jno;

extern "C" size_t jno_pattern_workspace_size(int64_t n_cells, int nloc, int64_t n_nodes) {
  (void)n_cells;
  (void)nloc;
  const int64_t vec_max = 3;
  const int64_t ndof = n_nodes * vec_max;
  return align256((n_nodes + 1) * 4) + align256((ndof + 1) * 4) + 256 + 256 +
         align256(exclusive_scan_ws(ndof > n_nodes ? ndof : n_nodes));
}

extern "C" int jno_pattern_count(const int32_t* d_cells, int64_t n_cells, int nloc, int vec, int64_t n_nodes,
                                 int32_t* d_inc_ptr, int32_t* d_inc, int32_t* d_rowptr, int64_t* h_nnz,
                                 int32_t* h_max_row, void* d_ws, size_t ws_bytes, void* stream) {
  JNO_REQUIRE(nloc >= 1 && nloc <= 10, JNO_EINVAL, "jno_pattern_count: nloc must be in 1..10");
  JNO_REQUIRE(vec >= 1 && vec <= 3, JNO_EINVAL, "jno_pattern_count: vec must be in 1..3");
  JNO_REQUIRE(n_cells * nloc < INT32_MAX && n_nodes * vec < INT32_MAX, JNO_EINVAL,
              "jno_pattern_count: mesh exceeds int32 indexing");
  JNO_REQUIRE(ws_bytes >= jno_pattern_workspace_size(n_cells, nloc, n_nodes), JNO_EWORKSPACE,
              "jno_pattern_count: workspace too small");
  hipStream_t s = as_stream(stream);
  const int64_t m = n_cells * nloc, ndof = n_nodes * vec;
  char* w = static_cast<char*>(d_ws);
  int32_t* deg = reinterpret_cast<int32_t*>(w);
  w += align256((n_nodes + 1) * 4);
  int32_t* rowcnt = reinterpret_cast<int32_t*>(w);
  w += align256((ndof + 1) * 4);
  int32_t* status = reinterpret_cast<int32_t*>(w);
  w += 256;
  int64_t* total = reinterpret_cast<int64_t*>(w);
  w += 256;
  void* scan_ws = w;
  size_t scan_bytes = ws_bytes - (size_t)(w - static_cast<char*>(d_ws));

  fill_i32<<<grid_for(n_nodes + 1, 256), 256, 0, s>>>(deg, n_nodes + 1, 0);
  fill_i32<<<1, 64, 0, s>>>(status, 4, 0);
  JNO_LAUNCH_CHECK();
  if (m) degree_count<<<grid_for(m, 256), 256, 0, s>>>(d_cells, m, deg);
  JNO_LAUNCH_CHECK();
  int rc = exclusive_scan_i32(deg, d_inc_ptr, n_nodes, total, scan_ws, scan_bytes, s);
  if (rc) return rc;
  copy_i32<<<grid_for(n_nodes, 256), 256, 0, s>>>(d_inc_ptr, deg, n_nodes);
  JNO_LAUNCH_CHECK();
  if (m) incidence_fill<<<grid_for(m, 256), 256, 0, s>>>(d_cells, m, deg, d_inc);
  JNO_LAUNCH_CHECK();
  fill_i32<<<grid_for(ndof + 1, 256), 256, 0, s>>>(rowcnt, ndof + 1, 0);
  pattern_count_k<<<grid_for(n_nodes, PAT_NT / WAVE, 16384), PAT_NT, 0, s>>>(d_cells, nloc, vec, n_nodes,
                                                                             d_inc_ptr, d_inc, rowcnt, status);
  JNO_LAUNCH_CHECK();
  max_row_k<<<grid_for(ndof, 256, 1024), 256, 0, s>>>(rowcnt, ndof, status + 1);
  JNO_LAUNCH_CHECK();
  rc = exclusive_scan_i32(rowcnt, d_rowptr, ndof, total, scan_ws, scan_bytes, s);
  if (rc) return rc;
  int32_t h_status[2];
  int64_t h_total;
  JNO_HIP(hipMemcpyAsync(h_status, status, sizeof(h_status), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipMemcpyAsync(&h_total, total, sizeof(h_total), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  JNO_REQUIRE(h_status[0] == 0, JNO_ECAPACITY,
              "jno_pattern_count: a node has %d incident cells (capacity %d entries, %d candidates)", h_status[0],
              PAT_MAX_DEG, PAT_MAX_CAND);
  JNO_REQUIRE(h_total < INT32_MAX, JNO_ECAPACITY, "jno_pattern_count: nnz %lld exceeds int32",
              (long long)h_total);
  if (h_nnz) *h_nnz = h_total;
  if (h_max_row) *h_max_row = h_status[1];
  return JNO_OK;
}

extern "C" int jno_pattern_fill(const int32_t* d_cells, int64_t n_cells, int nloc, int vec, int64_t n_nodes,
                                const int32_t* d_inc_ptr, const int32_t* d_inc, const int32_t* d_rowptr,
                                int32_t* d_col, void* stream) {
  (void)n_cells;
  JNO_REQUIRE(nloc >= 1 && nloc <= 10 && vec >= 1 && vec <= 3, JNO_EINVAL, "jno_pattern_fill: bad nloc/vec");
  pattern_fill_k<<<grid_for(n_nodes, PAT_NT / WAVE, 16384), PAT_NT, 0, as_stream(stream)>>>(
      d_cells, nloc, vec, n_nodes, d_inc_ptr, d_inc, d_rowptr, d_col);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
