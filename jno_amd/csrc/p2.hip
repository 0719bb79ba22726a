// P2 (10-node tet) path: node promotion, element coloring, 3-D linear-elasticity element matrices
// on MFMA with an atomic-free colored scatter into the CSR plan, the traction load and the
// compliance gradient.
#include "common.h"

namespace jno {

// basix P2 edge order (jno/utils/solver/fem_topology.py:44): local dof 4+j sits on edge j
__constant__ int8_t c_edge[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};

__device__ __forceinline__ int edge_index(int la, int lb) {
  const int lo = la < lb ? la : lb, hi = la < lb ? lb : la;
  // (2,3)->0 (1,3)->1 (1,2)->2 (0,3)->3 (0,2)->4 (0,1)->5
  if (lo == 2) return 0;
  if (lo == 1) return hi == 3 ? 1 : 2;
  return hi == 3 ? 3 : (hi == 2 ? 4 : 5);
}

// ---------------------------------------------------------------------------------------------
// A2 promotion: new node ids in cell-major / local-dof-minor first-appearance order
// (_promote_to_degree, jno/utils/solver/fem_utils.py:1979-2044); an edge is identified by its two
// vertices, which is the coordinate-hash identity on a conforming mesh.
// ---------------------------------------------------------------------------------------------
__global__ void p2_first_k(const int4* __restrict__ cells, int64_t T, const int32_t* __restrict__ inc_ptr,
                           const int32_t* __restrict__ inc, int32_t* __restrict__ minpos, int32_t* __restrict__ flag) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < 6 * T; o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = o / 6;
    const int j = (int)(o % 6);
    const int4 c = cells[e];
    const int v[4] = {c.x, c.y, c.z, c.w};
    const int a = v[c_edge[j][0]], b = v[c_edge[j][1]];
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    int best = (int)o;
    for (int q = inc_ptr[lo]; q < inc_ptr[lo + 1]; ++q) {
      const int e2 = inc[q] >> 2, la = inc[q] & 3;
      if ((int64_t)e2 * 6 >= best) break;  // incidence is ascending: no earlier cell follows
      const int4 c2 = cells[e2];
      const int w[4] = {c2.x, c2.y, c2.z, c2.w};
      int lb = -1;
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (w[t] == hi) lb = t;
      if (lb >= 0) {
        const int p = e2 * 6 + edge_index(la, lb);
        best = p < best ? p : best;
      }
    }
    minpos[o] = best;
    flag[o] = (best == (int)o) ? 1 : 0;
  }
}

// P2 reference points as barycentric weights of the 4 vertices (basix DOF order)
__constant__ double c_p2w[10][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1},
                                    {0, 0, 0.5, 0.5}, {0, 0.5, 0, 0.5}, {0, 0.5, 0.5, 0},
                                    {0.5, 0, 0, 0.5}, {0.5, 0, 0.5, 0}, {0.5, 0.5, 0, 0}};

__global__ void p2_fill_k(const double* __restrict__ pts, const int4* __restrict__ cells, int64_t T, int64_t N,
                          const int32_t* __restrict__ minpos, const int32_t* __restrict__ rank,
                          int32_t* __restrict__ cells10, double* __restrict__ pts2) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < 6 * T; o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = o / 6;
    const int j = (int)(o % 6);
    const int4 c = cells[e];
    const int v[4] = {c.x, c.y, c.z, c.w};
    const int64_t id = N + rank[minpos[o]];
    cells10[e * 10 + 4 + j] = (int32_t)id;
    if (j == 0) {
#pragma unroll
      for (int a = 0; a < 4; ++a) cells10[e * 10 + a] = v[a];
    }
    if (minpos[o] == (int32_t)o) {
      // same arithmetic as the einsum over the four weighted vertices, without FMA contraction
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        double s = __dmul_rn(c_p2w[4 + j][0], pts[3 * (int64_t)v[0] + g]);
#pragma unroll
        for (int a = 1; a < 4; ++a) s = __dadd_rn(s, __dmul_rn(c_p2w[4 + j][a], pts[3 * (int64_t)v[a] + g]));
        pts2[3 * id + g] = s;
      }
    }
  }
}

__global__ void copy_f64(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// ---------------------------------------------------------------------------------------------
// element coloring: Jones-Plassmann rounds with first-fit colors; colors of the previous round are
// read from one buffer and written to another, so the result does not depend on timing
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

constexpr int MAX_COLORS = 128;

__global__ void color_round_k(const int32_t* __restrict__ cells, int64_t T, int nloc,
                              const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                              const int32_t* __restrict__ cold, int32_t* __restrict__ cnew,
                              int32_t* __restrict__ remaining, int32_t* __restrict__ overflow) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < T; e += (int64_t)gridDim.x * blockDim.x) {
    if (cold[e] >= 0) {
      cnew[e] = cold[e];
      continue;
    }
    const uint32_t pe = hash32((uint32_t)e);
    bool top = true;
    uint64_t used0 = 0, used1 = 0;
    for (int a = 0; a < 4 && top; ++a) {
      const int v = cells[e * nloc + a];
      for (int q = inc_ptr[v]; q < inc_ptr[v + 1]; ++q) {
        const int64_t e2 = inc[q] / nloc;
        if (e2 == e) continue;
        const int c2 = cold[e2];
        if (c2 < 0) {
          const uint32_t p2 = hash32((uint32_t)e2);
          if (p2 > pe || (p2 == pe && e2 > e)) {
            top = false;
            break;
          }
        } else if (c2 < 64) {
          used0 |= 1ull << c2;
        } else {
          used1 |= 1ull << (c2 - 64);
        }
      }
    }
    if (top) {
      int c;
      if (~used0) c = __ffsll((long long)~used0) - 1;
      else if (~used1) c = 64 + __ffsll((long long)~used1) - 1;
      else {
        c = -1;
        atomicExch(overflow, 1);
      }
      cnew[e] = c;
      if (c < 0) atomicAdd(remaining, 1);
    } else {
      cnew[e] = -1;
      atomicAdd(remaining, 1);
    }
  }
}

__global__ void color_count_k(const int32_t* __restrict__ color, int64_t T, int32_t* __restrict__ cnt) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < T; e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[color[e]], 1);
}

__global__ void color_fill_k(const int32_t* __restrict__ color, int64_t T, int32_t* __restrict__ cursor,
                             int32_t* __restrict__ order) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < T; e += (int64_t)gridDim.x * blockDim.x)
    order[atomicAdd(&cursor[color[e]], 1)] = (int32_t)e;
}

__global__ void fill_i32_k(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void fill_f64_k(double* __restrict__ p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------------------------
// P2 geometry and reference gradients
// ---------------------------------------------------------------------------------------------
// Degree-2 exact 4-point rule on the reference tet, weights 1/24.
__constant__ double c_qpt[4][3] = {{0.1381966011250105, 0.1381966011250105, 0.1381966011250105},
                                   {0.5854101966249685, 0.1381966011250105, 0.1381966011250105},
                                   {0.1381966011250105, 0.5854101966249685, 0.1381966011250105},
                                   {0.1381966011250105, 0.1381966011250105, 0.5854101966249685}};
constexpr double QW = 1.0 / 24.0;

// d phi_a / d xi_k of the P2 basis at quad point q (basix order)
__device__ __forceinline__ double p2_gref(int q, int a, int k) {
  const double xi0 = c_qpt[q][0], xi1 = c_qpt[q][1], xi2 = c_qpt[q][2];
  const double l[4] = {1.0 - xi0 - xi1 - xi2, xi0, xi1, xi2};
  // d lambda_v / d xi_k: lambda_0 -> -1, lambda_{k+1} -> 1
  auto dl = [&](int v) -> double { return v == 0 ? -1.0 : (v == k + 1 ? 1.0 : 0.0); };
  if (a < 4) return (4.0 * l[a] - 1.0) * dl(a);
  const int u = c_edge[a - 4][0], w = c_edge[a - 4][1];
  return 4.0 * (l[u] * dl(w) + l[w] * dl(u));
}

struct Geo {
  double Ji[3][3];  // Ji[k][i] = d xi_k / d x_i
  double adet;
};

__device__ __forceinline__ Geo geo_of(const double* __restrict__ pts, const int32_t* __restrict__ c) {
  const double* p0 = pts + 3 * (int64_t)c[0];
  double e[3][3];
#pragma unroll
  for (int v = 0; v < 3; ++v)
#pragma unroll
    for (int i = 0; i < 3; ++i) e[v][i] = pts[3 * (int64_t)c[v + 1] + i] - p0[i];
  double c23[3], c31[3], c12[3];
  cross_rn(e[1], e[2], c23);
  cross_rn(e[2], e[0], c31);
  cross_rn(e[0], e[1], c12);
  const double det =
      __dadd_rn(__dadd_rn(__dmul_rn(e[0][0], c23[0]), __dmul_rn(e[0][1], c23[1])), __dmul_rn(e[0][2], c23[2]));
  Geo g;
  const double inv = 1.0 / det;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    g.Ji[0][i] = c23[i] * inv;
    g.Ji[1][i] = c31[i] * inv;
    g.Ji[2][i] = c12[i] * inv;
  }
  g.adet = fabs(det);
  return g;
}

// physical gradient component i of basis a at quad point q
__device__ __forceinline__ double gphys(const Geo& g, int q, int a, int i) {
  return p2_gref(q, a, 0) * g.Ji[0][i] + p2_gref(q, a, 1) * g.Ji[1][i] + p2_gref(q, a, 2) * g.Ji[2][i];
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------------------
// Elasticity element matrices on MFMA (v_mfma_f64_16x16x4f64), one wavefront per element.
//   K[(a,i),(b,j)] = lam X^T Xc [(a,i),(b,j)] + mu X^T Xc [(a,j),(b,i)] + mu delta_ij (H^T Hc)[a,b]
// with X[q,(a,i)] = G_q[a,i], Xc = c_q X, H[(q,k),a] = G_q[a,k], c_q = w_q |det J| E_e.
// MFMA operand layout (16x16x4 f64): A[r][k] in lane r + 16k, B[k][c] in lane c + 16k, and
// D[4t + lane/16][lane%16] in element t of the lane's double4 (one row per output group, four
// groups interleaved by the 16-lane input block).
// K1 = X^T X is formed without c_q and scaled afterwards, so it is bitwise symmetric; the final
// combination is rounded term by term, so K itself is bitwise symmetric like the reference's.
// ---------------------------------------------------------------------------------------------
constexpr int EL_WAVES = 4;

// The MFMA tiles of one element, computed by one wavefront: sK = X^T X (30x30, padded 32x32) and
// sS = H^T H (10x10, padded 16x16); the caller's wavefront reads them after a wave barrier.
__device__ __forceinline__ void p2_elem_tiles(const Geo& g, int lane, double (*sK)[33], double (*sS)[17]) {
  const int r16 = lane % 16, k4 = lane / 16;
  // ---- K1 = X^T X: four 16x16 blocks, K = 4 quad points
  double xa[2];
#pragma unroll
  for (int blk = 0; blk < 2; ++blk) {
    const int rc = blk * 16 + r16;  // (a,i) index of A rows / B cols
    xa[blk] = rc < 30 ? gphys(g, k4, rc / 3, rc % 3) : 0.0;
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      dbl4 acc = {0.0, 0.0, 0.0, 0.0};
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[bi], xa[bj], acc, 0, 0, 0);
#pragma unroll
      for (int t = 0; t < 4; ++t) sK[bi * 16 + 4 * t + k4][bj * 16 + r16] = acc[t];
    }
  // ---- S = H^T H: 10x10 (padded 16x16), K = 12 = (q, k) pairs in three steps of 4
  {
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const int m = 4 * s + k4, q = m / 3, kk = m % 3;
      const double gv = r16 < 10 ? gphys(g, q, r16, kk) : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(gv, gv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) sS[4 * t + k4][r16] = acc[t];
  }
  // the LDS tiles written above are read by other lanes of this wavefront
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// entry ((a,i),(b,j)) of the element matrix before the factor c_q, rounded term by term (bitwise
// symmetric)
__device__ __forceinline__ double p2_kv(const double (*sK)[33], const double (*sS)[17], double lam, double mu, int a,
                                        int i, int b, int j) {
  double kv = __dadd_rn(__dmul_rn(lam, sK[3 * a + i][3 * b + j]), __dmul_rn(mu, sK[3 * a + j][3 * b + i]));
  if (i == j) kv = __dadd_rn(kv, __dmul_rn(mu, sS[a][b]));
  return kv;
}

__global__ void __launch_bounds__(64 * EL_WAVES) elast_p2_k(
    const double* __restrict__ pts, const int32_t* __restrict__ cells10, const int32_t* __restrict__ order,
    int32_t n_el, const double* __restrict__ E, double lam, double mu, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ col, const uint8_t* __restrict__ rank, double* __restrict__ vals) {
  __shared__ double sK[EL_WAVES][32][33];
  __shared__ double sS[EL_WAVES][16][17];
  const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int idx = blockIdx.x * EL_WAVES + wid;
  if (idx >= n_el) return;  // wave-uniform
  const int e = order ? order[idx] : idx;
  const int32_t* c = cells10 + (int64_t)e * 10;
  const Geo g = geo_of(pts, c);
  // all four quad weights are 1/24, so c_q = w |det J| E_e is one scalar per element
  const double cq = QW * g.adet * E[e];
  p2_elem_tiles(g, lane, sK[wid], sS[wid]);
  // ---- colored scatter: 100 node pairs over 64 lanes, 3x3 block each
  for (int pr = lane; pr < 100; pr += 64) {
    const int a = pr / 10, b = pr % 10;
    const int na = c[a];
    int lo;
    if (rank) {  // the pair's place in node a's row, found once per plan (jno_p2_pair_rank)
      lo = rank[(int64_t)e * 100 + pr];
    } else {
      const int nb = c[b];
      const int rb = rowptr[3 * na];
      const int len = (rowptr[3 * na + 1] - rb) / 3;
      lo = 0;
      int hi = len - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[rb + 3 * mid] < 3 * nb) lo = mid + 1;
        else hi = mid;
      }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double* out = vals + rowptr[3 * na + i] + 3 * lo;
#pragma unroll
      for (int j = 0; j < 3; ++j) out[j] += __dmul_rn(cq, p2_kv(sK[wid], sS[wid], lam, mu, a, i, b, j));
    }
  }
}

// Dense element matrices (E = 1) of the listed cells, one wavefront per cell, as 32 x 32 values (zero
// past 30) in the MFMA operand order of the element-by-element operator's SpMV: entry (r, k) at
// ((r / 16) * 8 + k / 4) * 64 + (k % 4) * 16 + r % 16, so that each of the SpMV's 16 operand loads is
// one contiguous 64-value run across the wavefront.  Same arithmetic as elast_p2_k's values
// (c_q = w |det J| with E = 1).
__global__ void __launch_bounds__(64 * EL_WAVES) elast_p2_block_k(const double* __restrict__ pts,
                                                                  const int32_t* __restrict__ cells10,
                                                                  const int32_t* __restrict__ list, int32_t n_list,
                                                                  double lam, double mu, double* __restrict__ t64,
                                                                  float* __restrict__ t32) {
  __shared__ double sK[EL_WAVES][32][33];
  __shared__ double sS[EL_WAVES][16][17];
  const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int idx = blockIdx.x * EL_WAVES + wid;
  if (idx >= n_list) return;  // wave-uniform
  const int32_t* c = cells10 + (int64_t)list[idx] * 10;
  const Geo g = geo_of(pts, c);
  const double cq = QW * g.adet;
  p2_elem_tiles(g, lane, sK[wid], sS[wid]);
  for (int w = lane; w < 32 * 32; w += 64) {
    const int r = w / 32, k = w % 32;
    double v = 0.0;
    if (r < 30 && k < 30) v = __dmul_rn(cq, p2_kv(sK[wid], sS[wid], lam, mu, r / 3, r % 3, k / 3, k % 3));
    const int o = ((r / 16) * 8 + k / 4) * 64 + (k % 4) * 16 + r % 16;
    t64[(int64_t)idx * 1024 + o] = v;
    if (t32) t32[(int64_t)idx * 1024 + o] = (float)v;
  }
}

// vertex differences x_{v+1} - x_0 of each cell (the only geometry the element matrices depend on)
__global__ void cell_jacobians_k(const double* __restrict__ pts, const int32_t* __restrict__ cells10, int64_t T,
                                 double* __restrict__ jac) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < T; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* c = cells10 + e * 10;
    const double* p0 = pts + 3 * (int64_t)c[0];
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
      for (int i = 0; i < 3; ++i) jac[e * 9 + 3 * v + i] = pts[3 * (int64_t)c[v + 1] + i] - p0[i];
  }
}

// ---------------------------------------------------------------------------------------------
// traction load on boundary facets of a plane: int_F t phi = |F|/3 t on the three edge nodes
// ---------------------------------------------------------------------------------------------
__constant__ int8_t c_faces2[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
__constant__ int8_t c_face_edges2[4][3] = {{6, 8, 9}, {5, 7, 9}, {4, 7, 8}, {4, 5, 6}};

__global__ void traction_p2_k(const double* __restrict__ pts, const int32_t* __restrict__ cells10, int64_t T,
                              const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc, int axis,
                              double value, double t0, double t1, double t2, double* __restrict__ b) {
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < 4 * T; w += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = w >> 2;
    const int f = (int)(w & 3);
    const int32_t* ce = cells10 + e * 10;
    const int v0 = ce[c_faces2[f][0]], v1 = ce[c_faces2[f][1]], v2 = ce[c_faces2[f][2]];
    if (pts[3 * (int64_t)v0 + axis] != value || pts[3 * (int64_t)v1 + axis] != value ||
        pts[3 * (int64_t)v2 + axis] != value)
      continue;
    bool shared = false;  // a facet on the plane counts only if it is a boundary facet
    for (int s = inc_ptr[v0]; s < inc_ptr[v0 + 1] && !shared; ++s) {
      const int64_t e2 = inc[s] / 10;
      if (e2 == e) continue;
      const int32_t* c2 = cells10 + e2 * 10;
      bool h1 = false, h2 = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        h1 |= c2[q] == v1;
        h2 |= c2[q] == v2;
      }
      shared = h1 && h2;
    }
    if (shared) continue;
    double d1[3], d2[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      d1[i] = pts[3 * (int64_t)v1 + i] - pts[3 * (int64_t)v0 + i];
      d2[i] = pts[3 * (int64_t)v2 + i] - pts[3 * (int64_t)v0 + i];
    }
    const double cx = d1[1] * d2[2] - d1[2] * d2[1], cy = d1[2] * d2[0] - d1[0] * d2[2],
                 cz = d1[0] * d2[1] - d1[1] * d2[0];
    const double w3 = 0.5 * sqrt(cx * cx + cy * cy + cz * cz) / 3.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int64_t node = ce[c_face_edges2[f][q]];
      // an edge node on a surface receives at most two contributions, so the sum is order-free
      atomicAdd(&b[3 * node + 0], t0 * w3);
      atomicAdd(&b[3 * node + 1], t1 * w3);
      atomicAdd(&b[3 * node + 2], t2 * w3);
    }
  }
}

// dC/dE_e = -u_e^T (K_e/E_e) u_e = -sum_q w_q |det| (lam (div u)^2 + 2 mu eps:eps)
__global__ void compliance_grad_p2_k(const double* __restrict__ pts, const int32_t* __restrict__ cells10, int64_t T,
                                     double lam, double mu, const double* __restrict__ u, double* __restrict__ grad) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < T; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* c = cells10 + e * 10;
    const Geo g = geo_of(pts, c);
    double ue[10][3];
#pragma unroll
    for (int a = 0; a < 10; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) ue[a][i] = u[3 * (int64_t)c[a] + i];
    double acc = 0.0;
    for (int q = 0; q < 4; ++q) {
      double gu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // gu[i][k] = d u_i / d x_k
      for (int a = 0; a < 10; ++a) {
        double gp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) gp[k] = gphys(g, q, a, k);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k) gu[i][k] += ue[a][i] * gp[k];
      }
      const double div = gu[0][0] + gu[1][1] + gu[2][2];
      double ee = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double s = 0.5 * (gu[i][k] + gu[k][i]);
          ee += s * s;
        }
      acc += QW * g.adet * (lam * div * div + 2.0 * mu * ee);
    }
    grad[e] = -acc;
  }
}

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace jno

using namespace jno;

extern "C" size_t jno_promote_p2_workspace_size(int64_t n_cells) {
  return 3 * a256((size_t)(6 * n_cells + 1) * 4) + 256 + exclusive_scan_ws(6 * n_cells);
}

extern "C" int jno_promote_p2(const double* d_points, const int32_t* d_cells, int64_t n_cells, int64_t n_nodes,
                              const int32_t* d_inc_ptr, const int32_t* d_inc, int32_t* d_cells10, double* d_points2,
                              int64_t* h_n_nodes2, void* d_ws, size_t ws_bytes, void* stream) {
  JNO_REQUIRE(ws_bytes >= jno_promote_p2_workspace_size(n_cells), JNO_EWORKSPACE, "jno_promote_p2: workspace too small");
  JNO_REQUIRE(6 * n_cells < INT32_MAX && n_nodes + 6 * n_cells < INT32_MAX, JNO_EINVAL,
              "jno_promote_p2: mesh exceeds int32 ids");
  JNO_REQUIRE(d_points && d_cells && d_cells10 && d_points2 && h_n_nodes2, JNO_EINVAL, "jno_promote_p2: null pointer");
  hipStream_t s = as_stream(stream);
  const int64_t m = 6 * n_cells;
  char* w = static_cast<char*>(d_ws);
  int32_t* minpos = reinterpret_cast<int32_t*>(w);
  w += a256((m + 1) * 4);
  int32_t* flag = reinterpret_cast<int32_t*>(w);
  w += a256((m + 1) * 4);
  int32_t* rank = reinterpret_cast<int32_t*>(w);
  w += a256((m + 1) * 4);
  int64_t* total = reinterpret_cast<int64_t*>(w);
  w += 256;
  const size_t scan_bytes = ws_bytes - (size_t)(w - static_cast<char*>(d_ws));
  const int4* cells = reinterpret_cast<const int4*>(d_cells);
  if (m) {
    p2_first_k<<<grid_for(m, 256), 256, 0, s>>>(cells, n_cells, d_inc_ptr, d_inc, minpos, flag);
    JNO_LAUNCH_CHECK();
  }
  int rc = exclusive_scan_i32(flag, rank, m, total, w, scan_bytes, s);
  if (rc) return rc;
  copy_f64<<<grid_for(3 * n_nodes, 256), 256, 0, s>>>(d_points, d_points2, 3 * n_nodes);
  JNO_LAUNCH_CHECK();
  if (m) {
    p2_fill_k<<<grid_for(m, 256), 256, 0, s>>>(d_points, cells, n_cells, n_nodes, minpos, rank, d_cells10, d_points2);
    JNO_LAUNCH_CHECK();
  }
  int64_t h_total = 0;
  JNO_HIP(hipMemcpyAsync(&h_total, total, 8, hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  *h_n_nodes2 = n_nodes + h_total;
  return JNO_OK;
}

extern "C" size_t jno_color_workspace_size(int64_t n_cells) { return 2 * a256((size_t)n_cells * 4) + 1024; }

extern "C" int jno_color_cells(const int32_t* d_cells, int64_t n_cells, int nloc, const int32_t* d_inc_ptr,
                               const int32_t* d_inc, int32_t* d_order, int32_t* h_color_ptr, int32_t max_colors,
                               int32_t* h_ncolors, void* d_ws, size_t ws_bytes, void* stream) {
  JNO_REQUIRE(ws_bytes >= jno_color_workspace_size(n_cells), JNO_EWORKSPACE, "jno_color_cells: workspace too small");
  JNO_REQUIRE(nloc >= 4 && d_cells && d_order && h_color_ptr && h_ncolors, JNO_EINVAL, "jno_color_cells: bad args");
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(d_ws);
  int32_t* ca = reinterpret_cast<int32_t*>(w);
  w += a256(n_cells * 4);
  int32_t* cb = reinterpret_cast<int32_t*>(w);
  w += a256(n_cells * 4);
  int32_t* flags = reinterpret_cast<int32_t*>(w);  // [0] remaining [1] overflow, [2..130) counts/cursors
  fill_i32_k<<<grid_for(n_cells, 256), 256, 0, s>>>(ca, n_cells, -1);
  JNO_LAUNCH_CHECK();
  int32_t h_flags[2];
  int rounds = 0;
  for (;;) {
    fill_i32_k<<<1, 64, 0, s>>>(flags, 2, 0);
    color_round_k<<<grid_for(n_cells, 256), 256, 0, s>>>(d_cells, n_cells, nloc, d_inc_ptr, d_inc, ca, cb, flags,
                                                         flags + 1);
    JNO_LAUNCH_CHECK();
    int32_t* t = ca;
    ca = cb;
    cb = t;
    JNO_HIP(hipMemcpyAsync(h_flags, flags, sizeof(h_flags), hipMemcpyDeviceToHost, s));
    JNO_HIP(hipStreamSynchronize(s));
    JNO_REQUIRE(!h_flags[1], JNO_ECAPACITY, "jno_color_cells: more than %d colors needed", MAX_COLORS);
    if (h_flags[0] == 0) break;
    JNO_REQUIRE(++rounds < 100000, JNO_ECAPACITY, "jno_color_cells: coloring did not finish");
  }
  // group cells by color
  int32_t* cnt = flags + 2;
  fill_i32_k<<<1, 256, 0, s>>>(cnt, MAX_COLORS, 0);
  color_count_k<<<grid_for(n_cells, 256), 256, 0, s>>>(ca, n_cells, cnt);
  JNO_LAUNCH_CHECK();
  int32_t h_cnt[MAX_COLORS];
  JNO_HIP(hipMemcpyAsync(h_cnt, cnt, sizeof(h_cnt), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  int nc = 0;
  for (int c = 0; c < MAX_COLORS; ++c)
    if (h_cnt[c]) nc = c + 1;
  JNO_REQUIRE(nc <= max_colors, JNO_ECAPACITY, "jno_color_cells: %d colors exceed the caller's %d", nc, max_colors);
  h_color_ptr[0] = 0;
  for (int c = 0; c < nc; ++c) h_color_ptr[c + 1] = h_color_ptr[c] + h_cnt[c];
  JNO_HIP(hipMemcpyAsync(cnt, h_color_ptr, sizeof(int32_t) * nc, hipMemcpyHostToDevice, s));
  color_fill_k<<<grid_for(n_cells, 256), 256, 0, s>>>(ca, n_cells, cnt, d_order);
  JNO_LAUNCH_CHECK();
  JNO_HIP(hipStreamSynchronize(s));
  *h_ncolors = nc;
  return JNO_OK;
}

// the place of every (cell, local node a, local node b) pair in node a's row of the vec = 3 pattern
// (in nodes: the row's 3 (place) + d columns), by binary search; a place past 254 flags the table
// unusable
__global__ void pair_rank_k(const int32_t* __restrict__ cells10, int64_t T, const int32_t* __restrict__ rowptr,
                            const int32_t* __restrict__ col, uint8_t* __restrict__ rank, int32_t* __restrict__ bad) {
  int f = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < 100 * T; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = q / 100;
    const int pr = (int)(q % 100), a = pr / 10, b = pr % 10;
    const int na = cells10[e * 10 + a], nb = cells10[e * 10 + b];
    const int rb = rowptr[3 * na];
    const int len = (rowptr[3 * na + 1] - rb) / 3;
    int lo = 0, hi = len - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (col[rb + 3 * mid] < 3 * nb) lo = mid + 1;
      else hi = mid;
    }
    if (lo > 254 || col[rb + 3 * lo] != 3 * nb) f = 1;
    rank[q] = (uint8_t)(lo > 254 ? 255 : lo);
  }
  if (f) atomicOr(bad, 1);
}

extern "C" int jno_p2_pair_rank(const int32_t* d_cells10, int64_t n_cells, const int32_t* d_rowptr,
                                const int32_t* d_col, uint8_t* d_rank, int32_t* h_ok, void* d_ws, size_t ws_bytes,
                                void* stream) {
  JNO_REQUIRE(d_cells10 && d_rowptr && d_col && d_rank && h_ok && d_ws && ws_bytes >= sizeof(int32_t), JNO_EINVAL,
              "jno_p2_pair_rank: null pointer or workspace too small");
  *h_ok = 0;
  hipStream_t s = as_stream(stream);
  int32_t* bad = static_cast<int32_t*>(d_ws);
  JNO_HIP(hipMemsetAsync(bad, 0, sizeof(int32_t), s));
  if (n_cells > 0) {
    pair_rank_k<<<grid_for(100 * n_cells, 256), 256, 0, s>>>(d_cells10, n_cells, d_rowptr, d_col, d_rank, bad);
    JNO_LAUNCH_CHECK();
  }
  int32_t hb = 1;
  JNO_HIP(hipMemcpyAsync(&hb, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  *h_ok = hb == 0 ? 1 : 0;
  return JNO_OK;
}

extern "C" int jno_assemble_p2_elasticity(const double* d_points2, const int32_t* d_cells10, int64_t n_cells,
                                          const int32_t* d_rowptr, const int32_t* d_col, int64_t nnz, const double* d_E,
                                          double nu, const int32_t* d_order, const int32_t* h_color_ptr, int ncolors,
                                          const uint8_t* d_rank, double* d_vals, void* stream) {
  JNO_REQUIRE(nu > -1.0 && nu < 0.5, JNO_EINVAL, "jno_assemble_p2_elasticity: Poisson ratio must be in (-1, 0.5)");
  JNO_REQUIRE(d_points2 && d_cells10 && d_rowptr && d_col && d_E && d_order && h_color_ptr && d_vals, JNO_EINVAL,
              "jno_assemble_p2_elasticity: null pointer");
  JNO_REQUIRE(h_color_ptr[ncolors] == n_cells, JNO_EINVAL, "jno_assemble_p2_elasticity: coloring covers %d of %lld cells",
              h_color_ptr[ncolors], (long long)n_cells);
  hipStream_t s = as_stream(stream);
  const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 1.0 / (2.0 * (1.0 + nu));
  fill_f64_k<<<grid_for(nnz, 256), 256, 0, s>>>(d_vals, nnz, 0.0);
  JNO_LAUNCH_CHECK();
  for (int c = 0; c < ncolors; ++c) {
    const int32_t n_el = h_color_ptr[c + 1] - h_color_ptr[c];
    if (!n_el) continue;
    elast_p2_k<<<(unsigned)cdiv(n_el, EL_WAVES), 64 * EL_WAVES, 0, s>>>(
        d_points2, d_cells10, d_order + h_color_ptr[c], n_el, d_E, lam, mu, d_rowptr, d_col, d_rank, d_vals);
    JNO_LAUNCH_CHECK();
  }
  return JNO_OK;
}

extern "C" int jno_traction_p2(const double* d_points2, const int32_t* d_cells10, int64_t n_cells,
                               const int32_t* d_inc_ptr, const int32_t* d_inc, int axis, double value,
                               const double* h_t, double* d_b, void* stream) {
  JNO_REQUIRE(axis >= 0 && axis < 3 && h_t, JNO_EINVAL, "jno_traction_p2: bad axis or traction");
  if (n_cells == 0) return JNO_OK;
  traction_p2_k<<<grid_for(4 * n_cells, 256), 256, 0, as_stream(stream)>>>(
      d_points2, d_cells10, n_cells, d_inc_ptr, d_inc, axis, value, h_t[0], h_t[1], h_t[2], d_b);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_compliance_grad_p2(const double* d_points2, const int32_t* d_cells10, int64_t n_cells, double nu,
                                      const double* d_u, double* d_grad, void* stream) {
  JNO_REQUIRE(nu > -1.0 && nu < 0.5, JNO_EINVAL, "jno_compliance_grad_p2: Poisson ratio must be in (-1, 0.5)");
  if (n_cells == 0) return JNO_OK;
  const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 1.0 / (2.0 * (1.0 + nu));
  compliance_grad_p2_k<<<grid_for(n_cells, 256, 8192), 256, 0, as_stream(stream)>>>(d_points2, d_cells10, n_cells, lam,
                                                                                    mu, d_u, d_grad);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_p2_cell_jacobians(const double* d_points2, const int32_t* d_cells10, int64_t n_cells, double* d_jac,
                                     void* stream) {
  JNO_REQUIRE(d_points2 && d_cells10 && d_jac, JNO_EINVAL, "jno_p2_cell_jacobians: null pointer");
  if (n_cells == 0) return JNO_OK;
  cell_jacobians_k<<<grid_for(n_cells, 256), 256, 0, as_stream(stream)>>>(d_points2, d_cells10, n_cells, d_jac);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_ebe_class_blocks(const double* d_points2, const int32_t* d_cells10, int n_classes,
                                    const int32_t* d_rep, double nu, double* d_table64, float* d_table32,
                                    void* stream) {
  JNO_REQUIRE(nu > -1.0 && nu < 0.5, JNO_EINVAL, "jno_ebe_class_blocks: Poisson ratio must be in (-1, 0.5)");
  JNO_REQUIRE(d_points2 && d_cells10 && d_rep && d_table64 && n_classes >= 0, JNO_EINVAL,
              "jno_ebe_class_blocks: null pointer");
  if (n_classes == 0) return JNO_OK;
  const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 1.0 / (2.0 * (1.0 + nu));
  elast_p2_block_k<<<(unsigned)cdiv(n_classes, EL_WAVES), 64 * EL_WAVES, 0, as_stream(stream)>>>(
      d_points2, d_cells10, d_rep, n_classes, lam, mu, d_table64, d_table32);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
