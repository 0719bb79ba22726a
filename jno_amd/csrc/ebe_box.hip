// The element-by-element P2 elasticity SpMV on a structured Kuhn box, without a product buffer.
//
// The P2 nodes of the Kuhn box of nx x ny x nz voxels are exactly the points of the half lattice
// (2nx+1) x (2ny+1) x (2nz+1): a vertex sits at twice its lattice coordinates, an edge node at the
// sum of its two vertices' ones.  A node's incident cells and the other nodes of each are therefore
// fixed offsets that depend only on the node's parity (I%2, J%2, K%2), and every cell of Kuhn type
// t has the element block E_e K_{class(t)}.  The operator is the same as the general
// element-by-element one (ebe.hip, the reference's ElementOperator.mv,
// jno/utils/solver/element_operator.py:89-167); only the order of the floating-point sums differs.
//
// ebe_box_k: one workgroup per brick of 4 x 4 x 4 voxels (the lanes of a wavefront), eleven
// wavefronts, one per node kind:
//   kinds 0-3   the vertex nodes (parity 000), 6 of their 24 incidences each;
//   kinds 4-6   axis-edge nodes (parity 100, 010, 001), 6 incidences each;
//   kinds 7-9   face-diagonal nodes (parity 110, 101, 011), 4 each;
//   kind 10     body-diagonal nodes (parity 111), 6.
// The brick first stages x of the 11^3 half-lattice points its cells touch (constrained dofs read
// as 0: P x) and E of its 5^3 x 6 cells (one voxel of halo below) in LDS; then each lane forms, per
// incidence (cell, local node a), the three rows 3a..3a+2 of E_e K_t times the cell's 30 staged
// values (explicit FMAs; K from the class table through uniform scalar loads) and sums them; the
// vertex kinds add their four partial sums in a fixed order.  No atomics, no product buffer: the
// results are bit-reproducible.  The bricks tile the virtual voxels [0, nx] x [0, ny] x [0, nz]: a
// node (I, J, K) belongs to voxel (I/2, J/2, K/2), so the nodes on the upper faces belong to a
// layer of voxels that holds no cells (their cells are masked by the domain bounds).
#include <type_traits>

#include "common.h"
#include "kuhn.h"
#include "pcg_common.h"
#include "sparse.h"

namespace jno {

// A brick is PBX x PBY x PBZ voxels; a lane owns the nodes of voxels (li, lj, lk + 4u), u < NPL, so
// that each coefficient read from a scalar register feeds NPL fused multiply-adds.
constexpr int NPL = 1;
constexpr int PBX = 4, PBY = 4, PBZ = 4 * NPL;
// staged half-lattice points per axis (one voxel of halo below): 2 P + 3
constexpr int SX = 2 * PBX + 3, SY = 2 * PBY + 3, SZ = 2 * PBZ + 3;
constexpr int PB_NX = SX * SY * SZ;
// x is staged as the eight parity sub-lattices of the brick's points, HX x HY x HZ entries each:
// point (X, Y, Z) at SUB (4 (X%2) + 2 (Y%2) + Z%2) + (X/2 HY + Y/2) HZ + Z/2, so that the lanes of a
// wavefront, whose nodes step by 2 along each axis, read consecutive 16-byte entries
constexpr int HX = PBX + 2, HY = PBY + 2, HZ = PBZ + 2;
constexpr int SUB = HX * HY * HZ;
constexpr int PB_NS = 8 * SUB;
__host__ __device__ constexpr int box_slot(int X, int Y, int Z) {
  return (4 * (X & 1) + 2 * (Y & 1) + (Z & 1)) * SUB + ((X >> 1) * HY + (Y >> 1)) * HZ + (Z >> 1);
}
__host__ __device__ constexpr int floor2(int v) { return v >= 0 ? v / 2 : -((1 - v) / 2); }
// staged voxels per axis (one of halo below) and cells
constexpr int VX = PBX + 1, VY = PBY + 1, VZ = PBZ + 1;
constexpr int PB_NE = VX * VY * VZ * 6;
constexpr int PB_NT = 11 * 64;    // one wavefront per node kind

// half-lattice offset (0..2) along axis d (0: i, 1: j, 2: k) of P2 local node a of Kuhn type t
// inside its voxel: twice the corner bit for a vertex, the sum of the two corner bits for the node of
// basix edge a - 4 ((2,3),(1,3),(1,2),(0,3),(0,2),(0,1); jno/utils/solver/fem_topology.py:44)
__host__ __device__ constexpr int corner_bit(int c, int d) { return (c >> (2 - d)) & 1; }
__host__ __device__ constexpr int p2_hoff(int t, int a, int d) {
  constexpr int E[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
  return a < 4 ? 2 * corner_bit(kuhn_corner(t, a), d)
               : corner_bit(kuhn_corner(t, E[a - 4][0]), d) + corner_bit(kuhn_corner(t, E[a - 4][1]), d);
}
__host__ __device__ constexpr int p2_parity(int t, int a) {
  return (p2_hoff(t, a, 0) % 2) * 4 + (p2_hoff(t, a, 1) % 2) * 2 + p2_hoff(t, a, 2) % 2;
}

// node kind -> its parity and its range of incidences (in the order t ascending, a ascending)
__host__ __device__ constexpr int kind_parity(int kind) {
  constexpr int P[11] = {0, 0, 0, 0, 4, 2, 1, 6, 5, 3, 7};
  return P[kind];
}
__host__ __device__ constexpr int kind_first(int kind) { return kind < 4 ? 6 * kind : 0; }
__host__ __device__ constexpr int parity_count(int p) {
  int n = 0;
  for (int t = 0; t < 6; ++t)
    for (int a = 0; a < 10; ++a)
      if (p2_parity(t, a) == p) ++n;
  return n;
}
__host__ __device__ constexpr int kind_count(int kind) { return kind < 4 ? 6 : parity_count(kind_parity(kind)); }
// the q-th incidence (t * 10 + a) of parity p
__host__ __device__ constexpr int parity_inc(int p, int q) {
  int n = 0;
  for (int t = 0; t < 6; ++t)
    for (int a = 0; a < 10; ++a)
      if (p2_parity(t, a) == p) {
        if (n == q) return t * 10 + a;
        ++n;
      }
  return -1;
}
static_assert(parity_count(0) == 24 && parity_count(4) == 6 && parity_count(6) == 4 && parity_count(7) == 6,
              "Kuhn P2 incidence counts");

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec4<double> {
  typedef double type __attribute__((ext_vector_type(4)));
};

template <typename T>
struct EbeBox {
  int nx, ny, nz;          // voxels
  int bx, by, bz;          // bricks per axis (of the virtual voxels 0..n)
  int64_t n_bricks;
  const int32_t* perm;     // node of each half-lattice point
  const uint8_t* mlat;     // its constrained-dof bits
  const T* E;              // per cell, cell order
  const T* rows;           // the coefficient table (BOX_ROWS values)
  const T* x;
  T* y;
  double* part;            // nullable: per-workgroup partial x.y
  PcgStateView st;
};

// FP32 values of the coefficient table: for each Kuhn type t, local node a and other local node b,
// the 3 x 3 block K_t[3a + r][3b + c] at ((10 t + a) 10 + b) 16 + 3 r + c
constexpr int BOX_ROWS = 6 * 10 * 10 * 16;
struct jno_i6 {
  int v[6];
};

// One lane's NPL nodes of kind KIND: for each, the sum over its incidences of rows 3a..3a+2 of
// E_e K_t x_e.  (li, lj, lk): the lane's first voxel in the brick; (gi, gj, gk): the brick's first voxel.
template <int KIND, typename T>
__device__ __forceinline__ void box_kind(const EbeBox<T>& a, const typename Vec4<T>::type* __restrict__ xs,
                                         const T* __restrict__ es, int li, int lj, int lk, int gi, int gj, int gk,
                                         T acc[NPL][3]) {
  constexpr int P = kind_parity(KIND);
  constexpr int PI = P >> 2, PJ = (P >> 1) & 1, PK = P & 1;
  // the lane's part of every staged address: its node u is point (2 (li + 1) + PI, ..., 2 (lk + 4u + 1) + PK)
  const int lbase = (li * HY + lj) * HZ + lk;
  static_for<0, kind_count(KIND)>([&](auto qc) {
    constexpr int ta = parity_inc(P, kind_first(KIND) + decltype(qc)::value);
    constexpr int t = ta / 10, la = ta % 10;
    constexpr int hi = p2_hoff(t, la, 0), hj = p2_hoff(t, la, 1), hk = p2_hoff(t, la, 2);
    // the cell's voxel, staged (origin one voxel below the brick) and global
    const int vi = li + 1 - hi / 2, vj = lj + 1 - hj / 2;
    const int ci = gi + vi - 1, cj = gj + vj - 1;
    const bool valid_ij = ci >= 0 && ci < a.nx && cj >= 0 && cj < a.ny;
    T e[NPL];
    bool valid[NPL];
    static_for<0, NPL>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int vk = lk + 4 * u + 1 - hk / 2, ck = gk + vk - 1;
      valid[u] = valid_ij && ck >= 0 && ck < a.nz;
      e[u] = es[((vi * VY + vj) * VZ + vk) * 6 + t];
    });
    // the coefficients are the same for every lane: read through the constant address space, so that
    // they come by scalar loads (the 9 values of a node pair contiguous)
    const __attribute__((address_space(4))) T* K =
        (const __attribute__((address_space(4))) T*)(uintptr_t)(a.rows + (10 * t + la) * 10 * 16);
    T sr[NPL][3];
    static_for<0, NPL>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      sr[u][0] = sr[u][1] = sr[u][2] = 0;
    });
    static_for<0, 10>([&](auto bc) {
      constexpr int b = decltype(bc)::value;
      // point of local node b seen from the lane's node, and its staged entry
      constexpr int dx = PI + p2_hoff(t, b, 0) - hi, dy = PJ + p2_hoff(t, b, 1) - hj, dz = PK + p2_hoff(t, b, 2) - hk;
      constexpr int off = (4 * (dx & 1) + 2 * (dy & 1) + (dz & 1)) * SUB +
                          ((1 + floor2(dx)) * HY + (1 + floor2(dy))) * HZ + (1 + floor2(dz));
      typename Vec4<T>::type xb[NPL];
      static_for<0, NPL>([&](auto uc) {
        constexpr int u = decltype(uc)::value;
        xb[u] = xs[lbase + off + 4 * u];
      });
      const auto* Kb = K + 16 * b;
      static_for<0, 3>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        static_for<0, NPL>([&](auto uc) {
          constexpr int u = decltype(uc)::value;
          sr[u][0] = fma(Kb[c], xb[u][c], sr[u][0]);
          sr[u][1] = fma(Kb[3 + c], xb[u][c], sr[u][1]);
          sr[u][2] = fma(Kb[6 + c], xb[u][c], sr[u][2]);
        });
      });
    });
    // (a cell outside the box adds nothing; the branch also keeps each incidence's loads together)
    static_for<0, NPL>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      if (valid[u]) {
        acc[u][0] = fma(e[u], sr[u][0], acc[u][0]);
        acc[u][1] = fma(e[u], sr[u][1], acc[u][1]);
        acc[u][2] = fma(e[u], sr[u][2], acc[u][2]);
      }
    });
    // one incidence's loads in flight at a time (registers for two workgroups per compute unit)
    __builtin_amdgcn_sched_barrier(0);
  });
}

__host__ __device__ constexpr int kind_parity_rt(int w) {
  return w < 4 ? 0 : w == 4 ? 4 : w == 5 ? 2 : w == 6 ? 1 : w == 7 ? 6 : w == 8 ? 5 : w == 9 ? 3 : 7;
}

template <typename T, bool CG>
__global__ void __launch_bounds__(PB_NT) __attribute__((amdgpu_waves_per_eu(6))) ebe_box_k(EbeBox<T> a) {
  using T4 = typename Vec4<T>::type;
  __shared__ T4 xs[PB_NS];
  __shared__ T es[PB_NE];
  __shared__ T vpart[3][NPL][3][64];
  __shared__ double sm_out[PB_NT / 64];
  if (a.st.template begin_spmv<CG>()) return;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // lanes 16 g .. 16 g + 15 share lj = g: their staged entries 36 li + lk (+ constants) are distinct
  // modulo 16, so a 16-lane pass of a 16-byte LDS read touches every bank once
  const int lj = lane >> 4, li = (lane >> 2) & 3, lk = lane & 3;
  const int NJ = 2 * a.ny + 1, NK = 2 * a.nz + 1;
  const int G = gridDim.x, b = blockIdx.x;
  // XCD-aware: each XCD takes a contiguous range of bricks (neighbours share their halo in its L2)
  const int logical = (G % N_XCD == 0) ? (b % N_XCD) * (G / N_XCD) + b / N_XCD : b;
  double part = 0.0;
  for (int64_t br = logical; br < a.n_bricks; br += G) {
    const int bk = (int)(br % a.bz), bj = (int)((br / a.bz) % a.by), bi = (int)(br / ((int64_t)a.bz * a.by));
    const int gi = bi * PBX, gj = bj * PBY, gk = bk * PBZ;
    // stage x (constrained dofs as 0) of the points 2(g - 1) .. 2(g - 1) + 2P + 2 per axis, and E of
    // the voxels g - 1 .. g + P - 1 (0 outside the box), in rounds whose loads are issued first
    constexpr int RS = 2;  // points per thread and round
    for (int q0 = 0; q0 < PB_NX; q0 += RS * PB_NT) {
      int nd[RS];
      uint8_t m[RS];
#pragma unroll
      for (int r = 0; r < RS; ++r) {
        const int q = q0 + tid + r * PB_NT;
        nd[r] = -1;
        m[r] = 0;
        if (q < PB_NX) {
          const int I = 2 * gi - 2 + q / (SY * SZ), J = 2 * gj - 2 + (q / SZ) % SY, K = 2 * gk - 2 + q % SZ;
          if (I >= 0 && I <= 2 * a.nx && J >= 0 && J < NJ && K >= 0 && K < NK) {
            const int64_t L = ((int64_t)I * NJ + J) * NK + K;
            nd[r] = a.perm[L];
            m[r] = a.mlat[L];
          }
        }
      }
      T xv[RS][3];
#pragma unroll
      for (int r = 0; r < RS; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) xv[r][c] = nd[r] >= 0 ? a.x[3 * (int64_t)nd[r] + c] : T(0);
#pragma unroll
      for (int r = 0; r < RS; ++r) {
        const int q = q0 + tid + r * PB_NT;
        if (q < PB_NX) {
          T4 v;
          v[0] = (m[r] & 1) ? T(0) : xv[r][0];
          v[1] = (m[r] & 2) ? T(0) : xv[r][1];
          v[2] = (m[r] & 4) ? T(0) : xv[r][2];
          v[3] = 0;
          xs[box_slot(q / (SY * SZ), (q / SZ) % SY, q % SZ)] = v;
        }
      }
    }
    for (int q = tid; q < PB_NE; q += PB_NT) {
      const int t = q % 6, v = q / 6;
      const int ci = gi - 1 + v / (VY * VZ), cj = gj - 1 + (v / VZ) % VY, ck = gk - 1 + v % VZ;
      es[q] = ci >= 0 && ci < a.nx && cj >= 0 && cj < a.ny && ck >= 0 && ck < a.nz
                  ? a.E[(((int64_t)ci * a.ny + cj) * a.nz + ck) * 6 + t]
                  : T(0);
    }
    __syncthreads();
    T acc[NPL][3];
#pragma unroll
    for (int u = 0; u < NPL; ++u) acc[u][0] = acc[u][1] = acc[u][2] = 0;
    switch (w) {
      case 0: box_kind<0>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      case 1: box_kind<1>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      case 2: box_kind<2>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      case 3: box_kind<3>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      case 4: box_kind<4>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      case 5: box_kind<5>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      case 6: box_kind<6>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      case 7: box_kind<7>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      case 8: box_kind<8>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      case 9: box_kind<9>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
      default: box_kind<10>(a, xs, es, li, lj, lk, gi, gj, gk, acc); break;
    }
    // the vertex kinds' four partial sums, added in a fixed order by wavefront 0
    if (w >= 1 && w <= 3)
#pragma unroll
      for (int u = 0; u < NPL; ++u)
#pragma unroll
        for (int c = 0; c < 3; ++c) vpart[w - 1][u][c][lane] = acc[u][c];
    __syncthreads();
    if (w == 0)
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int u = 0; u < NPL; ++u)
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[u][c] += vpart[p][u][c][lane];
    if (w == 0 || w >= 4) {
      const int P = kind_parity_rt(w);
      const int I = 2 * (gi + li) + (P >> 2), J = 2 * (gj + lj) + ((P >> 1) & 1);
#pragma unroll
      for (int u = 0; u < NPL; ++u) {
        const int K = 2 * (gk + lk + 4 * u) + (P & 1);
        if (I <= 2 * a.nx && J < NJ && K < NK) {
          const int64_t L = ((int64_t)I * NJ + J) * NK + K;
          const int64_t nd = a.perm[L];
          const uint8_t m = a.mlat[L];
          // the node's own x was staged (as 0 on its constrained dofs, which read it from memory)
          const typename Vec4<T>::type xo = xs[P * SUB + ((li + 1) * HY + (lj + 1)) * HZ + lk + 4 * u + 1];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const T xd = ((m >> c) & 1) ? a.x[3 * nd + c] : xo[c];
            const T v = ((m >> c) & 1) ? xd : acc[u][c];
            a.y[3 * nd + c] = v;
            part += (double)xd * (double)v;
          }
        }
      }
    }
    // the next brick restages xs, es and vpart after every wavefront has read them
    __syncthreads();
  }
  if (a.part) store_out_1<PB_NT>(part, a.part, sm_out);
}

// the half-lattice map of a P2 mesh that is the structured Kuhn box: perm[L] = the node at point L,
// written from every (cell, local node); then checked: every (cell, local node) finds itself in
// perm and in the inverse map, and every point is covered
__device__ __forceinline__ int64_t box_point(int nx, int ny, int nz, int64_t e, int la) {
  const int t = (int)(e % 6);
  const int64_t v = e / 6;
  const int k = (int)(v % nz), j = (int)((v / nz) % ny), i = (int)(v / ((int64_t)nz * ny));
  int h[3];
  const int E[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
  for (int d = 0; d < 3; ++d)
    h[d] = la < 4 ? 2 * corner_bit(kuhn_corner(t, la), d)
                  : corner_bit(kuhn_corner(t, E[la - 4][0]), d) + corner_bit(kuhn_corner(t, E[la - 4][1]), d);
  const int64_t NJ = 2 * ny + 1, NK = 2 * nz + 1;
  return ((int64_t)(2 * i + h[0]) * NJ + (2 * j + h[1])) * NK + (2 * k + h[2]);
}

__global__ void box_map_fill_k(int nx, int ny, int nz, const int32_t* __restrict__ cells, int64_t n_cells,
                               int64_t n_nodes, int32_t* __restrict__ perm, int32_t* __restrict__ inv,
                               int32_t* __restrict__ bad) {
  int f = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_cells * 10; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = q / 10;
    const int la = (int)(q % 10);
    const int64_t L = box_point(nx, ny, nz, e, la);
    const int nd = cells[q];
    if (nd < 0 || nd >= n_nodes) {
      f = 1;
      continue;
    }
    perm[L] = nd;
    inv[nd] = (int32_t)L;
  }
  if (f) atomicOr(bad, 1);
}

__global__ void box_map_check_k(int nx, int ny, int nz, const int32_t* __restrict__ cells, int64_t n_cells,
                                int64_t n_points, const int32_t* __restrict__ perm, const int32_t* __restrict__ inv,
                                const uint8_t* __restrict__ dmask, uint8_t* __restrict__ mlat, int32_t* __restrict__ bad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int f = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_cells * 10; q += stride) {
    const int64_t L = box_point(nx, ny, nz, q / 10, (int)(q % 10));
    const int nd = cells[q];
    if (nd < 0 || nd >= n_points || perm[L] != nd || inv[nd] != (int32_t)L) f = 1;
  }
  for (int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; L < n_points; L += stride) {
    const int nd = perm[L];
    if (nd < 0) {
      f = 1;
      mlat[L] = 0;
    } else {
      mlat[L] = (uint8_t)((dmask[3 * (int64_t)nd] ? 1 : 0) | (dmask[3 * (int64_t)nd + 1] ? 2 : 0) |
                          (dmask[3 * (int64_t)nd + 2] ? 4 : 0));
    }
  }
  if (f) atomicOr(bad, 1);
}

// the coefficient table of the box SpMV from the class table (FP32, MFMA order) and the class of
// each Kuhn type
__global__ void box_rows_k(const float* __restrict__ table, jno_i6 cls, float* __restrict__ rows) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < BOX_ROWS; q += gridDim.x * blockDim.x) {
    const int t = q / 1600, la = (q / 160) % 10, b = (q / 16) % 10, j = q % 16;
    rows[q] = j < 9 ? table[(int64_t)cls.v[t] * 1024 + ebe_table_index(3 * la + j / 3, 3 * b + j % 3)] : 0.0f;
  }
}

__global__ void box_fill_k(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

int64_t ebe_box_bricks(const jno_ebe* e) {
  if (!e || e->box_n[0] <= 0) return 0;
  return cdiv((int64_t)e->box_n[0] + 1, (int64_t)PBX) * cdiv((int64_t)e->box_n[1] + 1, (int64_t)PBY) *
         cdiv((int64_t)e->box_n[2] + 1, (int64_t)PBZ);
}

template <typename T>
int ebe_box_launch(const jno_op* op, const T* x, T* y, double* part, PcgStateView st, int grid, bool cg,
                   hipStream_t s) {
  const jno_ebe* e = op->ebe;
  const T* E = std::is_same<T, float>::value ? (const T*)e->box_e32 : (const T*)e->box_e64;
  JNO_REQUIRE(e->box_rows32 && E && e->box_perm && e->box_mask, JNO_EINVAL, "ebe: the structured box data is incomplete");
  EbeBox<T> a{};
  a.nx = e->box_n[0];
  a.ny = e->box_n[1];
  a.nz = e->box_n[2];
  a.bx = (int)cdiv((int64_t)a.nx + 1, (int64_t)PBX);
  a.by = (int)cdiv((int64_t)a.ny + 1, (int64_t)PBY);
  a.bz = (int)cdiv((int64_t)a.nz + 1, (int64_t)PBZ);
  a.n_bricks = (int64_t)a.bx * a.by * a.bz;
  a.perm = e->box_perm;
  a.mlat = e->box_mask;
  a.E = E;
  a.rows = e->box_rows32;
  a.x = x;
  a.y = y;
  a.part = part;
  a.st = st;
  const int g = grid > 0 ? grid : (int)std::min<int64_t>(cdiv(a.n_bricks, (int64_t)N_XCD) * N_XCD, 1 << 20);
  if (cg)
    ebe_box_k<T, true><<<g, PB_NT, 0, s>>>(a);
  else
    ebe_box_k<T, false><<<g, PB_NT, 0, s>>>(a);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
// FP32 only: the FP64 coefficients (two scalar registers each) do not fit the scalar register file
// in this form; FP64 products (the residuals of mixed solves) take the general path
template int ebe_box_launch<float>(const jno_op*, const float*, float*, double*, PcgStateView, int, bool, hipStream_t);

}  // namespace jno

using namespace jno;

extern "C" int jno_ebe_box_map(int nx, int ny, int nz, const int32_t* d_cells10, int64_t n_cells, int64_t n_nodes,
                               const uint8_t* d_dmask, int32_t* d_perm, uint8_t* d_mask_lat, int32_t* d_ws,
                               int32_t* h_ok, void* stream) {
  JNO_REQUIRE(d_cells10 && d_dmask && d_perm && d_mask_lat && d_ws && h_ok, JNO_EINVAL, "jno_ebe_box_map: null pointer");
  JNO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, JNO_EINVAL, "jno_ebe_box_map: bad grid %d x %d x %d", nx, ny, nz);
  *h_ok = 0;
  const int64_t pts = (int64_t)(2 * nx + 1) * (2 * ny + 1) * (2 * nz + 1);
  if (n_cells != 6LL * nx * ny * nz || n_nodes != pts || pts >= (int64_t)1 << 31) return JNO_OK;
  hipStream_t s = as_stream(stream);
  int32_t* inv = d_ws;       // i32[n_nodes]
  int32_t* bad = d_ws + n_nodes;  // one flag
  box_fill_k<<<grid_for(pts, 256), 256, 0, s>>>(d_perm, pts, -1);
  JNO_LAUNCH_CHECK();
  box_fill_k<<<grid_for(n_nodes + 1, 256), 256, 0, s>>>(inv, n_nodes + 1, -1);
  JNO_LAUNCH_CHECK();
  box_fill_k<<<1, 64, 0, s>>>(bad, 1, 0);
  JNO_LAUNCH_CHECK();
  box_map_fill_k<<<grid_for(n_cells * 10, 256), 256, 0, s>>>(nx, ny, nz, d_cells10, n_cells, n_nodes, d_perm, inv,
                                                               bad);
  JNO_LAUNCH_CHECK();
  box_map_check_k<<<grid_for(std::max<int64_t>(n_cells * 10, pts), 256), 256, 0, s>>>(nx, ny, nz, d_cells10, n_cells,
                                                                                       pts, d_perm, inv, d_dmask,
                                                                                       d_mask_lat, bad);
  JNO_LAUNCH_CHECK();
  int32_t hb = 1;
  JNO_HIP(hipMemcpyAsync(&hb, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  *h_ok = hb == 0 ? 1 : 0;
  return JNO_OK;
}

extern "C" int jno_ebe_box_rows(const float* d_table32, int n_classes, const int32_t* h_class, float* d_rows,
                                void* stream) {
  JNO_REQUIRE(d_table32 && h_class && d_rows, JNO_EINVAL, "jno_ebe_box_rows: null pointer");
  jno_i6 c;
  for (int t = 0; t < 6; ++t) {
    JNO_REQUIRE(h_class[t] >= 0 && h_class[t] < n_classes, JNO_EINVAL, "jno_ebe_box_rows: bad class %d of Kuhn type %d",
                h_class[t], t);
    c.v[t] = h_class[t];
  }
  box_rows_k<<<grid_for(BOX_ROWS, 256), 256, 0, as_stream(stream)>>>(d_table32, c, d_rows);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
