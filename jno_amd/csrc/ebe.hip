// Element-by-element (matrix-free) operator of the P2 vector elasticity form and its SpMV, the
// analogue of the reference's ElementOperator (jno/utils/solver/element_operator.py:89-167: mv as a
// gather of each cell's local dofs, the dense block product, and a scatter-add of the local results).
//
// Storage: a class table of element blocks K_c (E = 1), one scale E_e per cell, the connectivity
// and the node incidence.  Cells whose vertex differences are bitwise equal have bitwise equal
// element matrices, so one block serves a whole class (a structured Kuhn box has 6); per cell the
// operator keeps 4 + 4 bytes and the 40-byte connectivity instead of 900 values.
//
// y = A x in two passes, no atomics:
//   ebe_elem_k   one wavefront per tile of 16 slots of one class: gathers the slots' 30 local dofs
//                of x (constrained dofs read as 0: P x), scales them by E_e and multiplies by K_c on
//                MFMA (16x16x4, FP32 or FP64): Y^T (16 cells x 32 rows) = U^T (16 x 32) K_c^T, eight
//                K steps by two row tiles; each (cell, local node) row triple goes straight to its
//                entry of the node-major incidence (the product buffer, 3 values per entry);
//   ebe_gather_k one thread per node sums its contiguous entries in ascending cell order (a fixed
//                order: results are bit-reproducible) and applies the elimination: y_d = x_d on
//                constrained dofs.  In the PCG it is the SpMV kernel KS (partial p.q).
// (A buffer of 32 values per slot read back through the incidence fetched each product line about
// five times from memory, once per touching node's workgroup: 117 MB per launch at C5 against 25 MB
// now.)
#include <type_traits>

#include "common.h"
#include "pcg_common.h"
#include "sparse.h"

namespace jno {

typedef float flt4 __attribute__((ext_vector_type(4)));
typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int EBE_PITCH = 32;  // product-buffer values per slot (30 used) and table row pitch
constexpr int EBE_TILE = 16;   // slots per tile (the MFMA M dimension)
constexpr int EBE_NT = 256;    // threads of the element pass: 4 tiles per workgroup

// MFMA 16x16x4 with A[m][k] in lane m + 16k and B[k][n] in lane n + 16k (both precisions).  The
// output layouts differ: FP32 lane l holds D[4(l/16) + t][l%16] in element t, FP64 holds
// D[4t + l/16][l%16] (measured with an in-bounds lane probe).
template <typename T>
struct Mma;
template <>
struct Mma<float> {
  using acc = flt4;
  __device__ static acc run(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row_of(int g, int t) { return 4 * g + t; }
};
template <>
struct Mma<double> {
  using acc = dbl4;
  __device__ static acc run(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row_of(int g, int t) { return 4 * t + g; }
};

template <typename T>
struct EbeElem {
  int64_t n_tiles;
  const int32_t* cells;
  const int32_t* order;
  const int32_t* tile_class;
  const uint32_t* slot_mask;
  const int32_t* qpos;  // [slot*10 + a]: incidence position of the slot's local node a (-1: padding)
  const T* table;
  const T* scale;
  const T* x;
  T* yloc;              // products in incidence order: 3 values per incidence entry
  const int32_t* stop;  // nullable: the PCG state's stop flag (launches after a stop exit at once)
};

template <typename T>
__device__ __forceinline__ void ebe_tile(const EbeElem<T>& a, int64_t tile, int lane, int j, int g, const T (*bv)[8],
                                         T (*sy)[33]) {
  const int64_t slot = tile * EBE_TILE + j;
  // A[m][k] = U^T[slot j][local dof 4ks + g] = E_e x_dof (0 for constrained, padding and dof >= 30);
  // the slot's nodes, mask and scale are stored in slot order, so the x gathers wait on one round
  // trip only
  T av[8];
  const int32_t* __restrict__ c = a.cells + slot * 10;
  const uint32_t m = a.slot_mask[slot];
  const T s = a.scale[slot];
  // the 8 node indices first, then the 8 gathers, so that each batch is one round trip
  int node[8];
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
    const int d = 4 * ks + g;
    node[ks] = c[(d < 30 ? d : 0) / 3];
  }
  T xv[8];
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
    const int d = 4 * ks + g;
    xv[ks] = a.x[3 * (int64_t)node[ks] + (d < 30 ? d : 0) % 3];
  }
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
    const int d = 4 * ks + g;
    const bool use = d < 30 && !((m >> (d < 30 ? d : 0)) & 1u);
    av[ks] = use ? s * xv[ks] : T(0);
  }
  typename Mma<T>::acc acc[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    acc[nt] = typename Mma<T>::acc{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) acc[nt] = Mma<T>::run(av[ks], bv[nt][ks], acc[nt]);
  }
  // D[m = slot][n = row] through the wavefront's LDS tile, then each (slot, local node a) row triple
  // 3a..3a+2 as one 12-byte store into its incidence entry (the positions of the tile's 160 pairs
  // are contiguous)
#pragma unroll
  for (int nt = 0; nt < 2; ++nt)
#pragma unroll
    for (int t = 0; t < 4; ++t) sy[Mma<T>::row_of(g, t)][16 * nt + j] = acc[nt][t];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
  const int32_t* qp = a.qpos + tile * EBE_TILE * 10;
  for (int pr = lane; pr < EBE_TILE * 10; pr += 64) {
    const int q = qp[pr];
    const int sl = pr / 10, loc = pr - 10 * sl;
    if (q >= 0) {
      T* y = a.yloc + 3 * (int64_t)q;
      y[0] = sy[sl][3 * loc];
      y[1] = sy[sl][3 * loc + 1];
      y[2] = sy[sl][3 * loc + 2];
    }
  }
  // the next tile of this wavefront rewrites the LDS tile after every lane has read it
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

constexpr int EBE_TPW = 2;  // consecutive tiles per wavefront (one operand load per class they share)

template <typename T>
__global__ void __launch_bounds__(EBE_NT) ebe_elem_k(EbeElem<T> a) {
  // XCD-aware: each XCD (workgroups are dealt to the 8 XCDs round-robin) takes one contiguous range
  // of tiles; the slots are ordered in space, so most incidence entries of a node are written from
  // one XCD and their lines reach memory once
  const int G = gridDim.x, b = blockIdx.x;
  const int logical = (G % N_XCD == 0) ? (b % N_XCD) * (G / N_XCD) + b / N_XCD : b;
  const int64_t wave = ((int64_t)logical * EBE_NT + threadIdx.x) / 64;
  if (wave * EBE_TPW >= a.n_tiles) return;  // wave-uniform
  if (a.stop && *a.stop) return;
  const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
  // B[k][n] = K_c^T[4ks + g][16nt + j] = K_c[16nt + j][4ks + g] (the table is zero past 30, and
  // stored in this operand order: lane + 64 (8 nt + ks)); reloaded only when the class changes
  __shared__ T sy_all[EBE_NT / 64][EBE_TILE][33];
  T(*sy)[33] = sy_all[threadIdx.x / 64];
  T bv[2][8];
  int cur = -1;
  for (int64_t tile = wave * EBE_TPW; tile < min(a.n_tiles, (wave + 1) * EBE_TPW); ++tile) {
    const int cls = a.tile_class[tile];
    if (cls != cur) {
      const T* K = a.table + (int64_t)cls * (EBE_PITCH * EBE_PITCH);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) bv[nt][ks] = K[(8 * nt + ks) * 64 + lane];
      cur = cls;
    }
    ebe_tile<T>(a, tile, lane, j, g, bv, sy);
  }
}


template <typename T>
struct EbeGather {
  int64_t n_nodes;
  const int32_t* inc_ptr;
  const uint8_t* dmask;
  const T* yloc;
  const T* x;
  T* y;
  double* part;  // nullable: per-workgroup partial x.y
  PcgStateView st;
};

// U incidence entries from q (those past qe read entry q and are not added), all loads issued
// before the first use
template <int U, typename T>
__device__ __forceinline__ void gather_batch(const EbeGather<T>& a, int q, int qe, T& y0, T& y1, T& y2) {
  T w[U][3];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const T* yl = a.yloc + 3 * (int64_t)(q + u < qe ? q + u : q);
    w[u][0] = yl[0];
    w[u][1] = yl[1];
    w[u][2] = yl[2];
  }
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (q + u < qe) {
      y0 += w[u][0];
      y1 += w[u][1];
      y2 += w[u][2];
    }
}

template <int NT, typename T, bool CG>
__global__ void __launch_bounds__(NT) ebe_gather_k(EbeGather<T> a) {
  __shared__ double sm_out[NT / 64];
  if (a.st.template begin_spmv<CG>()) return;
  // grid-stride over the nodes: consecutive nodes go to consecutive workgroups, so the heavy ones (a
  // vertex of a Kuhn lattice has 24 incidences, an edge node 4 to 6, and the vertices are numbered
  // first) spread over every XCD instead of filling one XCD's slab
  const int64_t n1 = a.n_nodes, step = (int64_t)gridDim.x * NT;
  int64_t nd = (int64_t)blockIdx.x * NT + threadIdx.x;
  double part = 0.0;
  for (; nd < n1; nd += step) {
    T y0 = 0, y1 = 0, y2 = 0;
    int q = a.inc_ptr[nd];
    const int qe = a.inc_ptr[nd + 1];
    // the node's entries are contiguous in the product buffer (incidence order, ascending cell):
    // long lists (a vertex of a Kuhn lattice has 24) in batches of 24, the rest (edge nodes: 4 to 6)
    // in batches of 8, each batch one round trip; the vertices are numbered first, so whole
    // wavefronts take the same path
    if (qe - q > 8)
      for (; q < qe; q += 24) gather_batch<24>(a, q, qe, y0, y1, y2);
    else
      for (; q < qe; q += 8) gather_batch<8>(a, q, qe, y0, y1, y2);
    const T yv[3] = {y0, y1, y2};
    T xd[3];
    uint8_t dm[3];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      xd[cc] = a.x[3 * nd + cc];
      dm[cc] = a.dmask[3 * nd + cc];
    }
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const T v = dm[cc] ? xd[cc] : yv[cc];
      a.y[3 * nd + cc] = v;
      part += (double)xd[cc] * (double)v;
    }
  }
  if (a.part) store_out_1<NT>(part, a.part, sm_out);
}

// exact diagonal: sum over a node's incidences of E_e K_c[3a+c][3a+c]; 1 on constrained dofs;
// with invert, the Jacobi inverse with its guard
__global__ void ebe_diag_k(int64_t n_nodes, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                           const int32_t* __restrict__ tile_class, const double* __restrict__ table,
                           const double* __restrict__ scale, const uint8_t* __restrict__ dmask, int invert,
                           double* __restrict__ out) {
  for (int64_t nd = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; nd < n_nodes;
       nd += (int64_t)gridDim.x * blockDim.x) {
    double dg[3] = {0.0, 0.0, 0.0};
    for (int q = inc_ptr[nd]; q < inc_ptr[nd + 1]; ++q) {
      const int v = inc[q];
      const int slot = v / 10, loc = v - 10 * slot;
      const double* K = table + (int64_t)tile_class[slot / EBE_TILE] * (EBE_PITCH * EBE_PITCH);
      const double s = scale[slot];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) dg[cc] += s * K[ebe_table_index(3 * loc + cc, 3 * loc + cc)];
    }
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const int64_t d = 3 * nd + cc;
      const double v = dmask[d] ? 1.0 : dg[cc];
      out[d] = !invert ? v : fabs(v) > 1e-30 ? 1.0 / v : 1.0;
    }
  }
}

__global__ void ebe_slots_k(int64_t n_slots, const int32_t* __restrict__ order, const int32_t* __restrict__ cells,
                            const uint8_t* __restrict__ dmask, int32_t* __restrict__ slot_cells,
                            uint32_t* __restrict__ mask) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n_slots; s += (int64_t)gridDim.x * blockDim.x) {
    const int e = order[s];
    uint32_t m = e >= 0 ? 0u : 0xffffffffu;
    for (int a = 0; a < 10; ++a) {
      const int nd = e >= 0 ? cells[(int64_t)e * 10 + a] : 0;
      slot_cells[s * 10 + a] = nd;
      if (e >= 0)
        for (int c = 0; c < 3; ++c)
          if (dmask[3 * (int64_t)nd + c]) m |= 1u << (3 * a + c);
    }
    mask[s] = m;
  }
}

__global__ void ebe_scales_k(int64_t n_slots, const int32_t* __restrict__ order, const double* __restrict__ E,
                             double* __restrict__ s64, float* __restrict__ s32) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n_slots; s += (int64_t)gridDim.x * blockDim.x) {
    const int e = order[s];
    const double v = e >= 0 ? E[e] : 0.0;
    s64[s] = v;
    if (s32) s32[s] = (float)v;
  }
}

__global__ void fill_i32_ebe_k(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void ebe_incidence_k(int64_t n, const int32_t* __restrict__ inc, const int32_t* __restrict__ slot_of,
                                int32_t* __restrict__ out) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int v = inc[q];
    const int e = v / 10;
    out[q] = slot_of[e] * 10 + (v - 10 * e);
  }
}

// position of every (slot, local node) entry in the node-major incidence: the inverse of inc
__global__ void ebe_qpos_k(int64_t n, const int32_t* __restrict__ inc, int32_t* __restrict__ qpos) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    qpos[inc[q]] = (int32_t)q;
}

__global__ void ebe_rhs_k(int64_t n, const uint8_t* __restrict__ dmask, double* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (dmask[i]) b[i] = 0.0;
}

// CSR values of the P2 elasticity operator from the class table (before the elimination): one
// wavefront per node sums E_e K_c over the node's incidences, in ascending cell order, into its three
// rows in LDS (each (cell, node pair)'s place from the plan's pair table), then writes each row once.
// The scatter of elast_p2_k reads and writes every value once per contributing cell instead.
constexpr int AC_WAVES = 4;
constexpr int AC_MAXR = 384;  // LDS capacity of a row, in values (rows of up to 128 nodes)
// With dmask (the constrained dofs, zero Dirichlet values) the symmetric elimination is fused as
// jno_dirichlet_eliminate applies it: contributions to constrained columns are left out (stored
// zeros), and a constrained row is written as 0 with 1 on its diagonal.
__global__ void __launch_bounds__(64 * AC_WAVES) assemble_p2_classes_k(
    int64_t n_nodes, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
    const int32_t* __restrict__ cells10, const int32_t* __restrict__ cell_class, const double* __restrict__ table,
    const double* __restrict__ E, const uint8_t* __restrict__ rank, const uint8_t* __restrict__ dmask,
    const int32_t* __restrict__ rowptr, double* __restrict__ vals) {
  __shared__ double acc[AC_WAVES][3][AC_MAXR];
  __shared__ int si[AC_WAVES][64];     // a batch of the node's incidences (cell * 10 + local node)
  __shared__ int sc[AC_WAVES][64];     // their cells' classes
  __shared__ double se[AC_WAVES][64];  // and moduli
  __shared__ uint32_t sm[AC_WAVES][64];  // and constrained-dof bits (bit 3b + j: local node b, comp j)
  const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
  double(*ra)[AC_MAXR] = acc[wid];
  for (int64_t nd = (int64_t)blockIdx.x * AC_WAVES + wid; nd < n_nodes; nd += (int64_t)gridDim.x * AC_WAVES) {
    const int r0 = rowptr[3 * nd], len = rowptr[3 * nd + 1] - r0;  // the node's three rows are equally long
    for (int i = lane; i < 3 * len; i += 64) ra[i / len][i % len] = 0.0;
    int self = 0;  // the node's own place in its rows
    const int q0 = inc_ptr[nd], q1 = inc_ptr[nd + 1];
    for (int qb = q0; qb < q1; qb += 64) {
      // the batch's incidence, class and modulus loads in flight together, one per lane
      if (qb + lane < q1) {
        const int v = inc[qb + lane];
        const int e = v / 10;
        si[wid][lane] = v;
        sc[wid][lane] = cell_class[e];
        se[wid][lane] = E[e];
        uint32_t bits = 0;
        if (dmask)
          for (int b = 0; b < 10; ++b) {
            const int64_t d = 3 * (int64_t)cells10[(int64_t)e * 10 + b];
            bits |= (dmask[d] ? 1u : 0u) << (3 * b) | (dmask[d + 1] ? 2u : 0u) << (3 * b) |
                    (dmask[d + 2] ? 4u : 0u) << (3 * b);
          }
        sm[wid][lane] = bits;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int qe = min(q1, qb + 64);
      // two incidences per step: both cells' loads in flight together, their sums added in order
      for (int q = qb; q < qe; q += 2) {
        const bool two = q + 1 < qe;
        int e[2], a[2];
        double s2[2];
        const double* __restrict__ K[2];
        uint32_t cbits[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int qq = two ? q + h : q;
          const int v = si[wid][qq - qb];
          e[h] = v / 10;
          a[h] = v - 10 * e[h];
          s2[h] = se[wid][qq - qb];
          K[h] = table + (int64_t)sc[wid][qq - qb] * 1024;
          cbits[h] = sm[wid][qq - qb];
        }
        self = rank[(int64_t)e[0] * 100 + 11 * a[0]];
        // the 90 values of rows 3a..3a+2: distinct places within one cell, so no two lanes collide
        for (int l = lane; l < 90; l += 64) {
          const int b = l / 9, i = (l / 3) % 3, j = l % 3;
          double val[2];
          int pos[2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            pos[h] = 3 * rank[(int64_t)e[h] * 100 + 10 * a[h] + b] + j;
            val[h] = s2[h] * K[h][ebe_table_index(3 * a[h] + i, 3 * b + j)];
          }
          if (!((cbits[0] >> (3 * b + j)) & 1u)) ra[i][pos[0]] += val[0];
          if (two && !((cbits[1] >> (3 * b + j)) & 1u)) ra[i][pos[1]] += val[1];
        }
      }
      // the next batch rewrites the staged incidences after every lane has read them
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    for (int i = lane; i < 3 * len; i += 64) {
      const int c = i / len, pos = i % len;
      double v = ra[c][pos];
      if (dmask && dmask[3 * nd + c]) v = pos == 3 * self + c ? 1.0 : 0.0;
      vals[rowptr[3 * nd + c] + pos] = v;
    }
  }
}

static int check_ebe(const jno_op* op) {
  JNO_REQUIRE(op && op->format == JNO_FMT_EBE && op->ebe, JNO_EINVAL, "ebe: not an element-by-element operator");
  const jno_ebe* e = op->ebe;
  JNO_REQUIRE(op->n == 3 * e->n_nodes, JNO_EINVAL, "ebe: the operator has 3 rows per node (%lld rows, %lld nodes)",
              (long long)op->n, (long long)e->n_nodes);
  JNO_REQUIRE(e->cells && e->order && e->tile_class && e->slot_mask && e->inc_ptr && e->inc && e->qpos && e->dmask,
              JNO_EINVAL, "ebe: null index array");
  JNO_REQUIRE(e->n_classes >= 1, JNO_EINVAL, "ebe: no geometry class");
  return JNO_OK;
}

template <typename T>
static int ebe_launch(const jno_op* op, const T* x, T* y, T* yloc, double* part, PcgStateView st, int grid,
                      int64_t rpb, int nt, bool cg, hipStream_t s) {
  int rc = check_ebe(op);
  if (rc) return rc;
  const jno_ebe* e = op->ebe;
  if constexpr (std::is_same<T, float>::value)
    if (ebe_box_bricks(e) > 0) return ebe_box_launch<T>(op, x, y, part, st, rpb > 0 ? grid : 0, cg, s);
  JNO_REQUIRE(yloc, JNO_EINVAL, "ebe: no product buffer (scratch)");
  const T* table = std::is_same<T, float>::value ? (const T*)e->table32 : (const T*)e->table64;
  const T* scale = std::is_same<T, float>::value ? (const T*)e->scale32 : (const T*)e->scale64;
  JNO_REQUIRE(table && scale, JNO_EINVAL, "ebe: the %s table or scale is missing",
              std::is_same<T, float>::value ? "FP32" : "FP64");
  if (e->n_tiles > 0) {
    EbeElem<T> ea{e->n_tiles, e->cells, e->order, e->tile_class, e->slot_mask, e->qpos, table, scale, x, yloc,
                  st.stop_flag()};
    // whole XCD slabs of workgroups (the surplus ones find no tile and exit)
    const int64_t blocks = cdiv(cdiv(cdiv(e->n_tiles, (int64_t)EBE_TPW), (int64_t)(EBE_NT / 64)), (int64_t)N_XCD) * N_XCD;
    ebe_elem_k<T><<<(unsigned)blocks, EBE_NT, 0, s>>>(ea);
    JNO_LAUNCH_CHECK();
  }
  EbeGather<T> ga{e->n_nodes, e->inc_ptr, e->dmask, yloc, x, y, part, st};
  const int g = rpb > 0 ? grid : grid_for(e->n_nodes, 256, 8192);
#define JNO_EBE_G(NT_)                                                   \
  if (cg)                                                                \
    ebe_gather_k<NT_, T, true><<<g, NT_, 0, s>>>(ga);                    \
  else                                                                   \
    ebe_gather_k<NT_, T, false><<<g, NT_, 0, s>>>(ga);
  switch (rpb > 0 ? nt : 256) {
    case 512: JNO_EBE_G(512) break;
    case 1024: JNO_EBE_G(1024) break;
    default: JNO_EBE_G(256) break;
  }
#undef JNO_EBE_G
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

template <typename T>
int ebe_spmv(const jno_op* op, const T* x, T* y, void* yloc, hipStream_t s) {
  return ebe_launch<T>(op, x, y, static_cast<T*>(yloc ? yloc : op->ebe ? op->ebe->scratch : nullptr), nullptr,
                       PcgStateView{}, 0, 0, 256, false, s);
}
template int ebe_spmv<float>(const jno_op*, const float*, float*, void*, hipStream_t);
template int ebe_spmv<double>(const jno_op*, const double*, double*, void*, hipStream_t);

template <typename T>
int ebe_spmv_pcg(const jno_op* op, const T* p, T* q, T* yloc, double* part, PcgStateView st, int grid, int64_t rpb,
                 int nt, bool cg, hipStream_t s) {
  return ebe_launch<T>(op, p, q, yloc, part, st, grid, rpb, nt, cg, s);
}
template int ebe_spmv_pcg<float>(const jno_op*, const float*, float*, float*, double*, PcgStateView, int, int64_t,
                                 int, bool, hipStream_t);
template int ebe_spmv_pcg<double>(const jno_op*, const double*, double*, double*, double*, PcgStateView, int,
                                  int64_t, int, bool, hipStream_t);

size_t ebe_scratch_bytes(const jno_op* op) {
  return op && op->format == JNO_FMT_EBE && op->ebe ? (size_t)op->ebe->n_cells * 30 * 8 : 0;
}

}  // namespace jno

using namespace jno;

extern "C" int jno_ebe_slots(int64_t n_slots, const int32_t* d_order, const int32_t* d_cells10, const uint8_t* d_dmask,
                             int32_t* d_slot_cells, uint32_t* d_slot_mask, void* stream) {
  JNO_REQUIRE(d_order && d_cells10 && d_dmask && d_slot_cells && d_slot_mask, JNO_EINVAL, "jno_ebe_slots: null pointer");
  if (n_slots == 0) return JNO_OK;
  ebe_slots_k<<<grid_for(n_slots, 256), 256, 0, as_stream(stream)>>>(n_slots, d_order, d_cells10, d_dmask,
                                                                      d_slot_cells, d_slot_mask);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_ebe_scales(int64_t n_slots, const int32_t* d_order, const double* d_E, double* d_s64, float* d_s32,
                              void* stream) {
  JNO_REQUIRE(d_order && d_E && d_s64, JNO_EINVAL, "jno_ebe_scales: null pointer");
  if (n_slots == 0) return JNO_OK;
  ebe_scales_k<<<grid_for(n_slots, 256), 256, 0, as_stream(stream)>>>(n_slots, d_order, d_E, d_s64, d_s32);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_ebe_incidence(int64_t n_entries, const int32_t* d_inc, const int32_t* d_slot_of, int64_t n_slots,
                                 int32_t* d_inc_out, int32_t* d_qpos, void* stream) {
  JNO_REQUIRE(d_inc && d_slot_of && d_inc_out && d_qpos, JNO_EINVAL, "jno_ebe_incidence: null pointer");
  hipStream_t s = as_stream(stream);
  if (n_slots > 0) {
    fill_i32_ebe_k<<<grid_for(10 * n_slots, 256), 256, 0, s>>>(d_qpos, 10 * n_slots, -1);
    JNO_LAUNCH_CHECK();
  }
  if (n_entries == 0) return JNO_OK;
  ebe_incidence_k<<<grid_for(n_entries, 256), 256, 0, s>>>(n_entries, d_inc, d_slot_of, d_inc_out);
  JNO_LAUNCH_CHECK();
  ebe_qpos_k<<<grid_for(n_entries, 256), 256, 0, s>>>(n_entries, d_inc_out, d_qpos);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_ebe_diagonal(const jno_op* h_op, int invert, double* d_out, void* stream) {
  int rc = check_ebe(h_op);
  if (rc) return rc;
  const jno_ebe* e = h_op->ebe;
  JNO_REQUIRE(d_out && e->table64 && e->scale64, JNO_EINVAL, "jno_ebe_diagonal: null pointer");
  if (e->n_nodes == 0) return JNO_OK;
  ebe_diag_k<<<grid_for(e->n_nodes, 256), 256, 0, as_stream(stream)>>>(e->n_nodes, e->inc_ptr, e->inc, e->tile_class,
                                                                        e->table64, e->scale64, e->dmask, invert,
                                                                        d_out);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_ebe_dirichlet_rhs(int64_t n_rows, const uint8_t* d_dmask, double* d_b, void* stream) {
  JNO_REQUIRE(d_dmask && d_b, JNO_EINVAL, "jno_ebe_dirichlet_rhs: null pointer");
  if (n_rows == 0) return JNO_OK;
  ebe_rhs_k<<<grid_for(n_rows, 256), 256, 0, as_stream(stream)>>>(n_rows, d_dmask, d_b);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_assemble_p2_classes(int64_t n_nodes, const int32_t* d_inc_ptr, const int32_t* d_inc,
                                       const int32_t* d_cells10, const int32_t* d_cell_class, const double* d_table64,
                                       const double* d_E, const uint8_t* d_rank, const uint8_t* d_dmask,
                                       const int32_t* d_rowptr, int max_row, double* d_vals, void* stream) {
  JNO_REQUIRE(d_inc_ptr && d_inc && d_cells10 && d_cell_class && d_table64 && d_E && d_rank && d_rowptr && d_vals,
              JNO_EINVAL, "jno_assemble_p2_classes: null pointer");
  JNO_REQUIRE(max_row >= 0 && max_row <= AC_MAXR, JNO_EINVAL,
              "jno_assemble_p2_classes: rows of %d values exceed the %d the kernel stages", max_row, AC_MAXR);
  if (n_nodes == 0) return JNO_OK;
  const int64_t g = std::min<int64_t>(cdiv(n_nodes, (int64_t)AC_WAVES), 1 << 16);
  assemble_p2_classes_k<<<(unsigned)g, 64 * AC_WAVES, 0, as_stream(stream)>>>(
      n_nodes, d_inc_ptr, d_inc, d_cells10, d_cell_class, d_table64, d_E, d_rank, d_dmask, d_rowptr, d_vals);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
