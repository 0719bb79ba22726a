// A7 CSR sparsity pattern build (the bit-exact index structure) and the node->element incidence
// that the row-pull assembly reuses.
//
//   count:  node degree histogram -> scan -> incidence fill -> per-node (one wavefront each): sort
//           the incidence list (deterministic assembly order), gather every node of every incident
//           element, sort them and count the distinct ones -> dof-level row counts -> scan = rowptr
//   fill:   per node again: distinct neighbours in ascending order, expanded to dofs n*vec+c.
// The sorts are bitonic networks in registers across the wavefront for nodes of up to 64 incident
// entries and 256 candidates; larger nodes (unstructured fans) compare their candidates pairwise in
// LDS.
//
// The result equals np.unique over the (row, col) keys of all element pairs, which is what
// compress_plan (jno/utils/solver/fem_utils.py:4009-4060) returns.
#include "common.h"
#include "kuhn.h"

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

// ---------------------------------------------------------------------------------------------
// The sort-based path (nodes with at most 64 incident entries and 256 candidates: every node of a
// P1 or P2 Kuhn lattice): one wavefront per node sorts in registers instead of comparing all pairs.
// Values are held in a blocked layout, element i = K lane + r in register r of lane `lane`; a
// bitonic network over 64 K elements (Batcher) sorts them ascending, INT_MAX padding last.
// ---------------------------------------------------------------------------------------------
constexpr int PAT_PAD = 0x7fffffff;

template <int K>
__device__ __forceinline__ void wave_bitonic(int (&v)[K], int lane) {
  constexpr int M = 64 * K;
#pragma unroll
  for (int s = 2; s <= M; s <<= 1) {
#pragma unroll
    for (int d = s >> 1; d > 0; d >>= 1) {
      if (d >= K) {
        // the partner of element i is i ^ d, in lane lane ^ (d / K), same register
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const int i = lane * K + r;
          const int pv = __shfl_xor(v[r], d / K, 64);
          const bool up = (i & s) == 0, lower = (i & d) == 0;
          v[r] = (up == lower) ? min(v[r], pv) : max(v[r], pv);
        }
      } else {
#pragma unroll
        for (int r = 0; r < K; ++r) {
          if (r & d) continue;
          const int i = lane * K + r;
          const bool up = (i & s) == 0;
          const int a = v[r], b = v[r | d];
          v[r] = up ? min(a, b) : max(a, b);
          v[r | d] = up ? max(a, b) : min(a, b);
        }
      }
    }
  }
}

// The node's incidence entries, sorted ascending (assembly sums in ascending element order): lane l
// holds entry l (PAT_PAD past deg); written back when `store`
__device__ __forceinline__ int sorted_incidence(int32_t* __restrict__ inc, int beg, int deg, int lane, bool store) {
  int v[1] = {lane < deg ? inc[beg + lane] : PAT_PAD};
  if (store) {
    wave_bitonic<1>(v, lane);
    if (lane < deg) inc[beg + lane] = v[0];
  }
  return v[0];
}

// Sorted candidate neighbour nodes of a node (every node of every incident element), K per lane;
// returns the count of distinct ones and sets first[r] for the first occurrence of each
template <int K>
__device__ __forceinline__ int sorted_candidates(const int32_t* __restrict__ cells, int nloc, int ent, int deg,
                                                 int lane, int (&c)[K], bool (&first)[K]) {
  const int m = deg * nloc;
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const int i = lane * K + r;
    const bool in = i < m;
    const int q = in ? i / nloc : 0;  // q < deg <= 64 for the candidates; element 0 otherwise, unused
    const int e = __shfl(ent, q, 64) / nloc;
    const int32_t cv = cells[(int64_t)(in ? e : 0) * nloc + (in ? i - q * nloc : 0)];
    c[r] = in ? cv : PAT_PAD;
  }
  wave_bitonic<K>(c, lane);
  const int prev_last = __shfl_up(c[K - 1], 1, 64);
  int cnt = 0;
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const int prev = r > 0 ? c[r - 1] : (lane > 0 ? prev_last : -1);
    first[r] = c[r] != PAT_PAD && c[r] != prev;
    cnt += first[r] ? 1 : 0;
  }
  return cnt;
}

// exclusive prefix sum over the 64 lanes (every lane gets its own prefix), and the total
__device__ __forceinline__ int wave_exclusive(int v, int lane, int* total) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

template <int K>
__device__ __forceinline__ void count_sorted(const int32_t* __restrict__ cells, int nloc, int vec, int64_t node, int ent,
                                             int deg, int lane, int32_t* __restrict__ rowcnt) {
  int c[K];
  bool first[K];
  const int cnt = sorted_candidates<K>(cells, nloc, ent, deg, lane, c, first);
  const int uniq = wave_sum_int(cnt);
  if (lane < vec) rowcnt[node * vec + lane] = uniq * vec;
}

template <int K>
__device__ __forceinline__ void fill_sorted(const int32_t* __restrict__ cells, int nloc, int vec, int64_t node, int ent,
                                            int deg, int lane, const int32_t* __restrict__ rowptr,
                                            int32_t* __restrict__ col) {
  int c[K];
  bool first[K];
  const int cnt = sorted_candidates<K>(cells, nloc, ent, deg, lane, c, first);
  int total;
  int pos = wave_exclusive(cnt, lane, &total);
#pragma unroll
  for (int r = 0; r < K; ++r) {
    if (!first[r]) continue;
    for (int cc = 0; cc < vec; ++cc) {
      int32_t* out = col + rowptr[node * vec + cc] + (int64_t)pos * vec;
      for (int c2 = 0; c2 < vec; ++c2) out[c2] = c[r] * vec + c2;
    }
    ++pos;
  }
}

// the candidate registers per lane for m candidates (0: too many for the sort-based path)
__device__ __forceinline__ int sort_width(int deg, int m) {
  return deg > 64 ? 0 : m <= 64 ? 1 : m <= 128 ? 2 : m <= 256 ? 4 : 0;
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
    const int kw = sort_width(deg, deg * nloc);  // wavefront-uniform
    if (kw) {
      const int ent = sorted_incidence(inc, beg, deg, lane, true);
      if (kw == 1) count_sorted<1>(cells, nloc, vec, node, ent, deg, lane, rowcnt);
      else if (kw == 2) count_sorted<2>(cells, nloc, vec, node, ent, deg, lane, rowcnt);
      else count_sorted<4>(cells, nloc, vec, node, ent, deg, lane, rowcnt);
      continue;
    }
    if (deg > PAT_MAX_DEG || deg * nloc > PAT_MAX_CAND) {
      if (lane == 0) atomicMax(status, deg);
      continue;
    }
    // larger nodes: rank-sort the incidence entries (all distinct) so assembly sums in ascending
    // element order, and compare the candidates pairwise
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
    const int kw = sort_width(deg, deg * nloc);  // wavefront-uniform
    if (kw) {
      // the incidence was sorted by the count phase
      const int ent = lane < deg ? inc[beg + lane] : PAT_PAD;
      if (kw == 1) fill_sorted<1>(cells, nloc, vec, node, ent, deg, lane, rowptr, col);
      else if (kw == 2) fill_sorted<2>(cells, nloc, vec, node, ent, deg, lane, rowptr, col);
      else fill_sorted<4>(cells, nloc, vec, node, ent, deg, lane, rowptr, col);
      continue;
    }
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

// ---------------------------------------------------------------------------------------------
// Content keys of index arrays (the plan cache keys a pattern on its cells' content, as compress_plan
// memoizes on a digest of its index array, jno/utils/solver/fem_utils.py:4027-4060): two
// independent 64-bit sums of a position-dependent mix of every entry (splitmix64's finalizer of
// (i << 32 | v) under two seeds).  Integer sums are exact, so the key does not depend on the order
// of the reduction; the equality check guards a hit against a collision.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) hash_i32_k(const int32_t* __restrict__ a, int64_t n,
                                                   unsigned long long* __restrict__ out) {
  __shared__ unsigned long long sm[2][4];
  unsigned long long h0 = 0, h1 = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t x = ((uint64_t)i << 32) | (uint32_t)a[i];
    h0 += mix64(x + 0x9e3779b97f4a7c15ull);
    h1 += mix64(x ^ 0xd1b54a32d192ed03ull);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    h0 += __shfl_xor(h0, o, 64);
    h1 += __shfl_xor(h1, o, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sm[0][w] = h0;
    sm[1][w] = h1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], sm[0][0] + sm[0][1] + sm[0][2] + sm[0][3]);
    atomicAdd(&out[1], sm[1][0] + sm[1][1] + sm[1][2] + sm[1][3]);
  }
}

__global__ void equal_i32_k(const int32_t* __restrict__ a, const int32_t* __restrict__ b, int64_t n,
                            int* __restrict__ diff) {
  bool d = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d = d || a[i] != b[i];
  if (d) atomicOr(diff, 1);
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

extern "C" int jno_hash_i32(const int32_t* d_data, int64_t n, uint64_t* h_hash, void* d_ws, size_t ws_bytes,
                            void* stream) {
  JNO_REQUIRE(h_hash && d_ws && (d_data || n == 0) && n >= 0, JNO_EINVAL, "jno_hash_i32: null pointer");
  JNO_REQUIRE(ws_bytes >= 256, JNO_EWORKSPACE, "jno_hash_i32: workspace too small (256 bytes)");
  hipStream_t s = as_stream(stream);
  unsigned long long* acc = static_cast<unsigned long long*>(d_ws);
  fill_i32<<<1, 64, 0, s>>>(reinterpret_cast<int32_t*>(acc), 4, 0);
  JNO_LAUNCH_CHECK();
  if (n > 0) hash_i32_k<<<grid_for(n, 256, 2048), 256, 0, s>>>(d_data, n, acc);
  JNO_LAUNCH_CHECK();
  unsigned long long h[2];
  JNO_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  h_hash[0] = h[0];
  h_hash[1] = h[1];
  return JNO_OK;
}

extern "C" int jno_equal_i32(const int32_t* d_a, const int32_t* d_b, int64_t n, int* h_equal, void* d_ws,
                             size_t ws_bytes, void* stream) {
  JNO_REQUIRE(h_equal && d_ws && ((d_a && d_b) || n == 0) && n >= 0, JNO_EINVAL, "jno_equal_i32: null pointer");
  JNO_REQUIRE(ws_bytes >= 256, JNO_EWORKSPACE, "jno_equal_i32: workspace too small (256 bytes)");
  hipStream_t s = as_stream(stream);
  int* diff = static_cast<int*>(d_ws);
  fill_i32<<<1, 64, 0, s>>>(diff, 1, 0);
  JNO_LAUNCH_CHECK();
  if (n > 0) equal_i32_k<<<grid_for(n, 256, 4096), 256, 0, s>>>(d_a, d_b, n, diff);
  JNO_LAUNCH_CHECK();
  int h = 1;
  JNO_HIP(hipMemcpyAsync(&h, diff, sizeof(int), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  *h_equal = h ? 0 : 1;
  return JNO_OK;
}

// ---------------------------------------------------------------------------------------------
// The plan of a structured Kuhn box in closed form: the same rowptr / col / inc_ptr / inc the
// sort-based builder produces for that mesh, with no sort and no atomics.  Node (i, j, k) couples
// to the 15 stencil neighbours that lie in the box (every such pair shares a tet: corners c1 and
// c2 = c1 + d of a voxel lie on a common monotone path), in ascending id = slot order; its incident
// (element, local vertex) entries are the box_inc order, ascending by element.  A row's offsets are
// counts of the earlier rows: the rows where a given slot (or voxel corner) exists form a sub-box,
// and the lexicographically earlier points of a sub-box have a closed form.
// ---------------------------------------------------------------------------------------------
namespace jno {

// points (a, b, c) of the sub-box [a0, a1] x [b0, b1] x [c0, c1] that precede (i, j, k) in
// lexicographic order (empty ranges count nothing)
__device__ __forceinline__ int64_t lex_before(int i, int j, int k, int a0, int a1, int b0, int b1, int c0, int c1) {
  const int64_t na = max(0, a1 - a0 + 1), nb = max(0, b1 - b0 + 1), nc = max(0, c1 - c0 + 1);
  int64_t cnt = (int64_t)min(max(i - a0, 0), (int)na) * nb * nc;
  if (i >= a0 && i <= a1) {
    cnt += (int64_t)min(max(j - b0, 0), (int)nb) * nc;
    if (j >= b0 && j <= b1) cnt += min(max(k - c0, 0), (int)nc);
  }
  return cnt;
}

__global__ void pattern_box_k(int nx, int ny, int nz, int32_t* __restrict__ rowptr, int32_t* __restrict__ col,
                              int32_t* __restrict__ inc_ptr, int32_t* __restrict__ inc) {
  const int64_t S2 = nz + 1, S1 = (int64_t)(ny + 1) * (nz + 1);
  const int64_t N = S1 * (nx + 1);
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r <= N; r += (int64_t)gridDim.x * blockDim.x) {
    // r = N: the point after the last row ((nx+1, 0, 0)), whose counts are the totals
    const int k = r < N ? (int)(r % S2) : 0;
    const int j = r < N ? (int)((r / S2) % (ny + 1)) : 0;
    const int i = r < N ? (int)(r / S1) : nx + 1;
    int64_t rp = 0;
    static_for<0, 15>([&](auto S) {
      using O = SlotOffset<S>;
      rp += lex_before(i, j, k, max(0, -O::di), nx - max(0, O::di), max(0, -O::dj), ny - max(0, O::dj),
                       max(0, -O::dk), nz - max(0, O::dk));
    });
    int64_t ip = 0;
    static_for<0, 8>([&](auto C) {
      constexpr int ci = (C >> 2) & 1, cj = (C >> 1) & 1, ck = C & 1;
      constexpr int mult = C == 0 || C == 7 ? 6 : 2;  // the types holding corner C
      ip += mult * lex_before(i, j, k, ci, nx - 1 + ci, cj, ny - 1 + cj, ck, nz - 1 + ck);
    });
    rowptr[r] = (int32_t)rp;
    inc_ptr[r] = (int32_t)ip;
    if (r == N) continue;
    static_for<0, 15>([&](auto S) {
      using O = SlotOffset<S>;
      if (i + O::di >= 0 && i + O::di <= nx && j + O::dj >= 0 && j + O::dj <= ny && k + O::dk >= 0 && k + O::dk <= nz)
        col[rp++] = (int32_t)(r + O::di * S1 + O::dj * S2 + O::dk);
    });
    static_for<0, 24>([&](auto Q) {
      constexpr BoxInc L = box_inc(Q);
      const int vi = i - ((L.c >> 2) & 1), vj = j - ((L.c >> 1) & 1), vk = k - (L.c & 1);
      if (vi >= 0 && vi < nx && vj >= 0 && vj < ny && vk >= 0 && vk < nz)
        inc[ip++] = (int32_t)(4 * (6 * (((int64_t)vi * ny + vj) * nz + vk) + L.t) + L.a);
    });
  }
}

// the nodes on the box surface (every node of a facet used by one cell, for the Kuhn box)
__global__ void mask_box_k(int nx, int ny, int nz, uint8_t* __restrict__ mask) {
  const int64_t S2 = nz + 1, S1 = (int64_t)(ny + 1) * (nz + 1);
  const int64_t N = S1 * (nx + 1);
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(r % S2), j = (int)((r / S2) % (ny + 1)), i = (int)(r / S1);
    mask[r] = (i == 0 || i == nx || j == 0 || j == ny || k == 0 || k == nz) ? 1 : 0;
  }
}

}  // namespace jno

// nnz of the structured box's pattern: per stencil slot, the nodes whose neighbour lies in the box
static int64_t box_nnz(int nx, int ny, int nz) {
  int64_t nnz = 0;
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
      for (int c = -1; c <= 1; ++c) {
        const bool pos = a >= 0 && b >= 0 && c >= 0, neg = a <= 0 && b <= 0 && c <= 0;
        if (pos || neg) nnz += (int64_t)(nx + 1 - (a != 0)) * (ny + 1 - (b != 0)) * (nz + 1 - (c != 0));
      }
  return nnz;
}

extern "C" int jno_pattern_kuhn_box_nnz(int nx, int ny, int nz, int64_t* h_nnz, int32_t* h_max_row) {
  JNO_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2 && h_nnz && h_max_row, JNO_EINVAL,
              "jno_pattern_kuhn_box_nnz: the box needs at least 2 cells per axis");
  const int64_t nnz = box_nnz(nx, ny, nz);
  JNO_REQUIRE(nnz < INT32_MAX && (int64_t)24 * nx * ny * nz < INT32_MAX, JNO_ECAPACITY,
              "jno_pattern_kuhn_box_nnz: the pattern exceeds 32-bit indices");
  *h_nnz = nnz;
  *h_max_row = 15;  // an interior node couples to all 15 stencil neighbours
  return JNO_OK;
}

extern "C" int jno_pattern_kuhn_box(int nx, int ny, int nz, int32_t* d_rowptr, int32_t* d_col, int32_t* d_inc_ptr,
                                    int32_t* d_inc, void* stream) {
  int64_t nnz;
  int32_t mr;
  int rc = jno_pattern_kuhn_box_nnz(nx, ny, nz, &nnz, &mr);
  if (rc) return rc;
  JNO_REQUIRE(d_rowptr && d_col && d_inc_ptr && d_inc, JNO_EINVAL, "jno_pattern_kuhn_box: null pointer");
  const int64_t N = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  jno::pattern_box_k<<<jno::grid_for(N + 1, 256, 16384), 256, 0, jno::as_stream(stream)>>>(nx, ny, nz, d_rowptr, d_col,
                                                                                          d_inc_ptr, d_inc);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_mask_kuhn_box(int nx, int ny, int nz, uint8_t* d_mask, void* stream) {
  JNO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && d_mask, JNO_EINVAL, "jno_mask_kuhn_box: bad argument");
  const int64_t N = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  jno::mask_box_k<<<jno::grid_for(N, 256, 8192), 256, 0, jno::as_stream(stream)>>>(nx, ny, nz, d_mask);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
