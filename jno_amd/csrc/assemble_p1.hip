// P1 scalar assembly (stiffness k grad u.grad v with load f, or mass u v) straight into the CSR
// values of the plan, with the symmetric Dirichlet elimination, the load vector and the Jacobi
// inverse diagonal fused into the same pass; and the P1 adjoint contraction dL/dk.
//
// Row-pull, atomic-free: one thread owns one row, walks the row's incident (element, local)
// entries in ascending element order (the plan sorted them), recomputes that element's geometry
// and adds the element-matrix row into register accumulators matched to the row's CSR columns.
// Every value is written exactly once; the summation order per entry is fixed, so results are
// bit-reproducible run to run.  (Half a wavefront per row, one lane per incident element with the
// contributions added in the same order through LDS, measured twice as slow.)
#include "common.h"
#include "kuhn.h"
#include "sparse.h"

namespace jno {

// Optional symmetric-DIA output of the assembly (offsets known from an earlier analysis of the
// same plan): the stored diagonals are written from the row's accumulators in the same pass, and
// *bad is set when a nonzero entry lies off the stored diagonals (the layout does not apply).  The
// lower diagonals need no check: the assembly is bitwise symmetric by construction (each entry
// pair is summed over the same elements in the same order, and the Dirichlet elimination is
// symmetric).
struct DiaOut {
  int ndiag;
  DiaOffsets off;
  double* d64;
  float* d32;
  int* bad;
};

struct TetGeom {
  double g[4][3];  // physical gradients of the barycentric basis
  double vol;      // |detJ| / 6
};

// J = [v1-v0, v2-v0, v3-v0] (columns).  Rows of inv(J) are (e2 x e3, e3 x e1, e1 x e2) / det,
// i.e. grad(lambda_1..3); grad(lambda_0) = -(sum).  (_cell_fields, fem_native.py:1461-1497)
__device__ __forceinline__ TetGeom tet_geom(const double* __restrict__ pts, int4 c) {
  const double* p0 = pts + 3 * (int64_t)c.x;
  const double* p1 = pts + 3 * (int64_t)c.y;
  const double* p2 = pts + 3 * (int64_t)c.z;
  const double* p3 = pts + 3 * (int64_t)c.w;
  double e1[3], e2[3], e3[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e1[i] = p1[i] - p0[i];
    e2[i] = p2[i] - p0[i];
    e3[i] = p3[i] - p0[i];
  }
  // cross products rounded term by term: a zero component of an axis-aligned edge stays an exact
  // zero, so couplings that vanish in exact arithmetic (the non-axis edges of a Kuhn lattice)
  // come out exactly 0.0 at any spacing
  double c23[3], c31[3], c12[3];
  cross_rn(e2, e3, c23);
  cross_rn(e3, e1, c31);
  cross_rn(e1, e2, c12);
  const double det = __dadd_rn(__dadd_rn(__dmul_rn(e1[0], c23[0]), __dmul_rn(e1[1], c23[1])), __dmul_rn(e1[2], c23[2]));
  const double inv = 1.0 / det;
  TetGeom t;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    t.g[1][i] = c23[i] * inv;
    t.g[2][i] = c31[i] * inv;
    t.g[3][i] = c12[i] * inv;
    t.g[0][i] = -(t.g[1][i] + t.g[2][i] + t.g[3][i]);
  }
  t.vol = fabs(det) / 6.0;
  return t;
}

// row a of the gradients and vertex a of the element, by selects (a runtime index into a register
// array would go through scratch memory)
__device__ __forceinline__ void pick_row(const TetGeom& t, int4 c, int a, double ga[3], int& na) {
#pragma unroll
  for (int i = 0; i < 3; ++i) ga[i] = a == 0 ? t.g[0][i] : a == 1 ? t.g[1][i] : a == 2 ? t.g[2][i] : t.g[3][i];
  na = a == 0 ? c.x : a == 1 ? c.y : a == 2 ? c.z : c.w;
}

template <int MAXR>
__global__ void __launch_bounds__(256) assemble_p1_k(
    int kind, const double* __restrict__ pts, const int4* __restrict__ cells, int64_t n_nodes,
    const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ col, const double* __restrict__ kcoef, double f_const,
    const double* __restrict__ f_nodal, const uint8_t* __restrict__ dmask, const double* __restrict__ g,
    double* __restrict__ vals, double* __restrict__ b, double* __restrict__ dinv, DiaOut dia) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_nodes;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int rb = rowptr[r];
    const int len = rowptr[r + 1] - rb;
    int cols[MAXR];
    double acc[MAXR];
#pragma unroll
    for (int s = 0; s < MAXR; ++s) {
      cols[s] = s < len ? col[rb + s] : -1;
      acc[s] = 0.0;
    }
    double bl = 0.0;
    const int ib = inc_ptr[r], ie = inc_ptr[r + 1];
    // one incident element's contributions to the row, in incidence order
    auto accum = [&](int ent, const int4 c, const TetGeom& t) {
      const int e = ent >> 2, a = ent & 3;
      const int nb[4] = {c.x, c.y, c.z, c.w};
      double ga[3];
      int na;
      pick_row(t, c, a, ga, na);
      if (kind == 0) {
        const double coef = kcoef[e] * t.vol;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
          // rounded term by term (no FMA contraction) so that K[r][c] and K[c][r] are bitwise equal
          const double dot = __dadd_rn(__dadd_rn(__dmul_rn(ga[0], t.g[bb][0]), __dmul_rn(ga[1], t.g[bb][1])),
                                       __dmul_rn(ga[2], t.g[bb][2]));
          const double v = __dmul_rn(coef, dot);
#pragma unroll
          for (int s = 0; s < MAXR; ++s) acc[s] += (cols[s] == nb[bb]) ? v : 0.0;
        }
        if (f_nodal) {
          const double fs = f_nodal[c.x] + f_nodal[c.y] + f_nodal[c.z] + f_nodal[c.w];
          bl += t.vol * (f_nodal[na] + fs) / 20.0;
        } else {
          bl += f_const * t.vol / 4.0;
        }
      } else {
        const double m = t.vol / 20.0;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
          const double v = (bb == a) ? 2.0 * m : m;
#pragma unroll
          for (int s = 0; s < MAXR; ++s) acc[s] += (cols[s] == nb[bb]) ? v : 0.0;
        }
      }
    };
    // two incident elements per step (rows of up to 16 entries; longer ones have no registers to
    // spare): both elements' index, vertex and coordinate loads are in flight together (the gather
    // chain incidence -> cell -> points is latency-bound); the sums keep the incidence order
    int q = ib;
    for (; MAXR <= 16 && q + 1 < ie; q += 2) {
      const int ent0 = inc[q], ent1 = inc[q + 1];
      const int4 c0 = cells[ent0 >> 2], c1 = cells[ent1 >> 2];
      const TetGeom t0 = tet_geom(pts, c0);
      const TetGeom t1 = tet_geom(pts, c1);
      accum(ent0, c0, t0);
      accum(ent1, c1, t1);
    }
    for (; q < ie; ++q) {
      const int ent = inc[q];
      const int4 c = cells[ent >> 2];
      accum(ent, c, tet_geom(pts, c));
    }
    // symmetric Dirichlet elimination (_apply_dirichlet_symmetric, fem_1d.py:667-686)
    if (dmask) {
      if (dmask[r]) {
#pragma unroll
        for (int s = 0; s < MAXR; ++s) acc[s] = (cols[s] == (int)r) ? 1.0 : 0.0;
        bl = g ? g[r] : 0.0;
      } else {
        double lift = 0.0;
#pragma unroll
        for (int s = 0; s < MAXR; ++s) {
          if (s < len && dmask[cols[s]]) {
            if (g) lift += acc[s] * g[cols[s]];
            acc[s] = 0.0;
          }
        }
        bl -= lift;
      }
    }
    double d = 0.0;
#pragma unroll
    for (int s = 0; s < MAXR; ++s) {
      if (s < len) vals[rb + s] = acc[s];
      if (cols[s] == (int)r) d += acc[s];
    }
    if (b) b[r] = bl;
    if (dinv) dinv[r] = fabs(d) > 1e-30 ? 1.0 / d : 1.0;
    if (dia.d64) {
      const int64_t ld = dia_ld(n_nodes);
      bool off_set = false;
#pragma unroll
      for (int s = 0; s < MAXR; ++s) {
        if (s < len && acc[s] != 0.0) {
          const int o = cols[s] - (int)r;
          const int ao = o < 0 ? -o : o;
          bool in_set = false;
          for (int dd = 0; dd < dia.ndiag; ++dd) in_set = in_set || dia.off.o[dd] == ao;
          off_set = off_set || !in_set;
        }
      }
      if (off_set) atomicOr(dia.bad, 1);
      for (int dd = 0; dd < dia.ndiag; ++dd) {
        double v = 0.0;
#pragma unroll
        for (int s = 0; s < MAXR; ++s)
          if (cols[s] - (int)r == dia.off.o[dd]) v += acc[s];
        dia.d64[(int64_t)dd * ld + r] = v;
        if (dia.d32) dia.d32[(int64_t)dd * ld + r] = (float)v;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Structured Kuhn box (the mesh of Shape.box(..).structured(n)): the row of node (i, j, k) sums the
// same element rows as assemble_p1_k, in the same ascending element order, without gathering the
// incidence, cells or points.  Every cell of type t (0..5) of a box has the vertex differences of
// the type-t cell of voxel 0 -- bitwise on dyadic spacings; on others they differ in the last bits
// from voxel to voxel, and voxel 0's stand for all -- so the six cells' |T| and gradient dot
// products (formed by tet_geom and the same term-by-term dots) are computed once per mesh and the
// kernel only reads k.  The 15 couplings of a node (offsets d in {-1,0,1}^3 whose nonzero
// components share a sign) are accumulated into registers by compile-time stencil slot, ordered
// like the CSR row (ascending node id = lexicographic d); a boundary row skips its missing slots.
// ---------------------------------------------------------------------------------------------
struct BoxGeom {
  double vol[6];
  double dot[6][4][4];
};

struct BoxArgs {
  int nx, ny, nz;
  BoxGeom geo;
  const int32_t* rowptr;
  const double* k;
  double f_const;
  const double* f_nodal;
  const uint8_t* dmask;
  const double* g;
  double* vals;
  double* b;
  double* dinv;
  int ndiag;
  int dia_slot[JNO_DIA_MAX];  // stencil slot of each stored diagonal (-1: none, the diagonal is 0)
  double* d64;
  float* d32;
};

constexpr int BOX_BLOCK = 256;
constexpr int BOX_GEO = 6 * 17;  // |T| of the six types, then their 4x4 dot products

// is the voxel of corner C (the row's node is that voxel's corner C) inside the box?
template <int C>
__device__ __forceinline__ bool box_inside(const BoxArgs& a, int i, int j, int kk) {
  const int vi = i - ((C >> 2) & 1), vj = j - ((C >> 1) & 1), vk = kk - (C & 1);
  return vi >= 0 && vi < a.nx && vj >= 0 && vj < a.ny && vk >= 0 && vk < a.nz;
}

// k of the types of corner C's voxel that hold the row's node (all six for corners 0 and 7, two for
// the others), as 16-byte loads where the types are adjacent: a voxel's six k values are contiguous.
// Loaded unconditionally (voxel 0 for a voxel outside the box) so that all of a row's loads are in
// flight together.
template <int C>
__device__ __forceinline__ void box_k(const BoxArgs& a, int i, int j, int kk, bool in, double (&kc)[8][6]) {
  const int vi = i - ((C >> 2) & 1), vj = j - ((C >> 1) & 1), vk = kk - (C & 1);
  const double* kp = a.k + (in ? 6 * (((int64_t)vi * a.ny + vj) * a.nz + vk) : 0);
  auto pair = [&](int t) {
    const double2 v = *reinterpret_cast<const double2*>(kp + t);
    kc[C][t] = v.x;
    kc[C][t + 1] = v.y;
  };
  if constexpr (C == 0 || C == 7) {
    pair(0);
    pair(2);
    pair(4);
  } else if constexpr (C == 4) {
    pair(0);
  } else if constexpr (C == 2) {
    pair(2);
  } else if constexpr (C == 1) {
    pair(4);
  } else {  // corners 6 (types 0, 2), 5 (types 1, 4), 3 (types 3, 5)
    constexpr int t0 = C == 6 ? 0 : C == 5 ? 1 : 3, t1 = C == 6 ? 2 : C == 5 ? 4 : 5;
    kc[C][t0] = kp[t0];
    kc[C][t1] = kp[t1];
  }
}

// the incidence Q of the row's node: its element's contributions to the 15 stencil slots and to the
// load.  An incidence outside the box (a boundary row) is masked: it reads a clamped k and adds
// +0.0, which leaves every sum bitwise as the unmasked kernel's (sums start at +0.0, so no -0.0
// arises either way).
template <int KIND, bool NODAL, int Q>
__device__ __forceinline__ void box_accumulate(const BoxArgs& a, const double* __restrict__ sg, const double (&kc)[8][6],
                                               const bool (&inc)[8], int64_t r, int64_t S1, int64_t S2,
                                               double (&acc)[15], double& bl) {
  if constexpr (Q < 24) {
    constexpr BoxInc L = box_inc(Q);
    const bool in = inc[L.c];
    const double vol = sg[L.t];
    const double* dot = sg + 6 + (L.t * 4 + L.a) * 4;
    if constexpr (KIND == 0) {
      const double coef = in ? __dmul_rn(kc[L.c][L.t], vol) : 0.0;
      static_for<0, 4>([&](auto BB) { acc[IncTerm<Q, BB>::slot] += __dmul_rn(coef, dot[BB]); });
      if constexpr (NODAL) {
        double fs = 0.0;
        static_for<0, 4>([&](auto BB) {
          using T = IncTerm<Q, BB>;
          fs += a.f_nodal[in ? r + T::di * S1 + T::dj * S2 + T::dk : r];
        });
        bl += in ? vol * (a.f_nodal[r] + fs) / 20.0 : 0.0;
      } else {
        bl += in ? a.f_const * vol / 4.0 : 0.0;
      }
    } else {
      const double m = in ? vol / 20.0 : 0.0;
      static_for<0, 4>([&](auto BB) { acc[IncTerm<Q, BB>::slot] += (BB == L.a) ? 2.0 * m : m; });
    }
    // the sums so far are complete here, and the geometry reads of the next incidences stay below
    // (an empty asm that takes the sums as operands and clobbers memory): with all 96 reads hoisted
    // to the top and the arithmetic sunk below them the kernel would hold ~300 registers
    if constexpr (Q % 2 == 1)
      __asm__ __volatile__(""
                           : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]),
                             "+v"(acc[6]), "+v"(acc[7]), "+v"(acc[8]), "+v"(acc[9]), "+v"(acc[10]), "+v"(acc[11]),
                             "+v"(acc[12]), "+v"(acc[13]), "+v"(acc[14]), "+v"(bl)
                           :
                           : "memory");
    box_accumulate<KIND, NODAL, Q + 1>(a, sg, kc, inc, r, S1, S2, acc, bl);
  }
}

// One row per thread, BOX_BLOCK consecutive rows per block.  The rows' CSR values are staged in
// LDS and written as one coalesced span [rowptr[r0], rowptr[r0 + BOX_BLOCK]).
template <int KIND, bool NODAL, bool DMASK>
__global__ void __launch_bounds__(BOX_BLOCK) assemble_box_k(BoxArgs a) {
  __shared__ double sg[BOX_GEO];
  __shared__ double sv[15 * BOX_BLOCK];
  const int nz1 = a.nz + 1, ny1 = a.ny + 1;
  const int64_t S2 = nz1, S1 = (int64_t)ny1 * nz1;
  const int64_t N = S1 * (a.nx + 1);
  const int tid = threadIdx.x;
  if (tid == 0) {  // compile-time indices: the table is read from the kernel arguments, not copied
#pragma unroll
    for (int t = 0; t < 6; ++t) sg[t] = a.geo.vol[t];
#pragma unroll
    for (int t = 0; t < 6; ++t)
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) sg[6 + (t * 4 + p) * 4 + q] = a.geo.dot[t][p][q];
  }
  const int64_t r0 = (int64_t)blockIdx.x * BOX_BLOCK;
  const int64_t r = r0 + tid;
  const int64_t base = a.rowptr[r0];
  const int64_t span = a.rowptr[r0 + BOX_BLOCK < N ? r0 + BOX_BLOCK : N] - base;
  __syncthreads();
  if (r < N) {
    const int kk = (int)(r % nz1);
    const int64_t w = r / nz1;
    const int j = (int)(w % ny1), i = (int)(w / ny1);
    const int64_t pos0 = a.rowptr[r] - base;  // issued first: it is needed only at the end
    double acc[15];
#pragma unroll
    for (int s = 0; s < 15; ++s) acc[s] = 0.0;
    double bl = 0.0;
    bool inc[8];
    double kc[8][6];
    static_for<0, 8>([&](auto C) { inc[C] = box_inside<C>(a, i, j, kk); });
    if constexpr (KIND == 0) static_for<0, 8>([&](auto C) { box_k<C>(a, i, j, kk, inc[C], kc); });
    box_accumulate<KIND, NODAL, 0>(a, sg, kc, inc, r, S1, S2, acc, bl);
    // the slots that exist (the neighbour is inside the box)
    bool ex[15];
    static_for<0, 15>([&](auto S) {
      using O = SlotOffset<S>;
      ex[S] = i + O::di >= 0 && i + O::di <= a.nx && j + O::dj >= 0 && j + O::dj <= a.ny && kk + O::dk >= 0 &&
              kk + O::dk <= a.nz;
    });
    // symmetric Dirichlet elimination (_apply_dirichlet_symmetric, fem_1d.py:667-686).  The mask (and
    // g) of every neighbour is loaded unconditionally from a clamped index, so that the 15 loads are
    // in flight together instead of one round trip per neighbour.
    if constexpr (DMASK) {
      bool dm[15];
      static_for<0, 15>([&](auto S) {
        using O = SlotOffset<S>;
        dm[S] = a.dmask[ex[S] ? r + O::di * S1 + O::dj * S2 + O::dk : r] != 0;
      });
      double gv[15];
      if (a.g) {
        static_for<0, 15>([&](auto S) {
          using O = SlotOffset<S>;
          gv[S] = a.g[ex[S] ? r + O::di * S1 + O::dj * S2 + O::dk : r];
        });
      }
      if (dm[7]) {
        static_for<0, 15>([&](auto S) { acc[S] = S == 7 ? 1.0 : 0.0; });
        bl = a.g ? gv[7] : 0.0;
      } else {
        double lift = 0.0;
        static_for<0, 15>([&](auto S) {
          if (ex[S] && dm[S]) {
            if (a.g) lift += acc[S] * gv[S];
            acc[S] = 0.0;
          }
        });
        bl -= lift;
      }
    }
    int64_t pos = pos0;
#pragma unroll
    for (int s = 0; s < 15; ++s)
      if (ex[s]) sv[pos++] = acc[s];
    if (a.b) a.b[r] = bl;
    if (a.dinv) a.dinv[r] = fabs(acc[7]) > 1e-30 ? 1.0 / acc[7] : 1.0;
    if (a.d64) {
      const int64_t ld = dia_ld(N);
      for (int dd = 0; dd < a.ndiag; ++dd) {
        const int want = a.dia_slot[dd];
        double v = 0.0;
#pragma unroll
        for (int s = 7; s < 15; ++s) v = (want == s) ? acc[s] : v;
        a.d64[(int64_t)dd * ld + r] = v;
        if (a.d32) a.d32[(int64_t)dd * ld + r] = (float)v;
      }
    }
  }
  __syncthreads();
  for (int64_t q = tid; q < span; q += BOX_BLOCK) a.vals[base + q] = sv[q];
}

// the six cell types of voxel 0: |T| and the term-by-term gradient dot products, as assemble_p1_k
// forms them
__global__ void box_geom_k(int ny, int nz, const double* __restrict__ pts, double* __restrict__ out) {
  const int t = threadIdx.x;
  if (t >= 6) return;
  const int64_t S2 = nz + 1, S1 = (int64_t)(ny + 1) * (nz + 1);
  int c[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int cc = kuhn_corner(t, a);
    c[a] = (int)(((cc >> 2) & 1) * S1 + ((cc >> 1) & 1) * S2 + (cc & 1));
  }
  const TetGeom g = tet_geom(pts, make_int4(c[0], c[1], c[2], c[3]));
  out[t] = g.vol;
  for (int a = 0; a < 4; ++a)
    for (int bb = 0; bb < 4; ++bb)
      out[6 + (t * 4 + a) * 4 + bb] =
          __dadd_rn(__dadd_rn(__dmul_rn(g.g[a][0], g.g[bb][0]), __dmul_rn(g.g[a][1], g.g[bb][1])),
                    __dmul_rn(g.g[a][2], g.g[bb][2]));
}

// is (points, cells) the structured box jno_mesh_kuhn_box makes (bitwise)?  *bad set when not
__global__ void box_check_k(int nx, int ny, int nz, double x0, double y0, double z0, double x1, double y1, double z1,
                            const double* __restrict__ pts, const int4* __restrict__ cells, int* __restrict__ bad) {
  const int64_t S2 = nz + 1, S1 = (int64_t)(ny + 1) * (nz + 1);
  const int64_t N = S1 * (nx + 1), T = (int64_t)6 * nx * ny * nz;
  bool ok = true;
  for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < N || id < T;
       id += (int64_t)gridDim.x * blockDim.x) {
    if (id < N) {
      const int k = (int)(id % S2);
      const int64_t w = id / S2;
      const int j = (int)(w % (ny + 1)), i = (int)(w / (ny + 1));
      ok = ok && pts[3 * id] == linspace_at(x0, x1, nx, i) && pts[3 * id + 1] == linspace_at(y0, y1, ny, j) &&
           pts[3 * id + 2] == linspace_at(z0, z1, nz, k);
    }
    if (id < T) {
      const int typ = (int)(id % 6);
      const int64_t v = id / 6;
      const int k = (int)(v % nz);
      const int64_t w = v / nz;
      const int j = (int)(w % ny), i = (int)(w / ny);
      const int4 c = cells[id];
      const int cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int q = kuhn_corner(typ, a);
        ok = ok && cc[a] == (int)((i + ((q >> 2) & 1)) * S1 + (j + ((q >> 1) & 1)) * S2 + (k + (q & 1)));
      }
    }
  }
  if (!ok) atomicOr(bad, 1);
}

__global__ void __launch_bounds__(256) adjoint_grad_p1_k(const double* __restrict__ pts, const int4* __restrict__ cells,
                                                         int64_t T, const double* __restrict__ u,
                                                         const double* __restrict__ lam,
                                                         const uint8_t* __restrict__ dmask,
                                                         double* __restrict__ grad) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < T; e += (int64_t)gridDim.x * blockDim.x) {
    const int4 c = cells[e];
    const TetGeom t = tet_geom(pts, c);
    const int nb[4] = {c.x, c.y, c.z, c.w};
    double gl[3] = {0, 0, 0}, gu[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const double la = (dmask && dmask[nb[a]]) ? 0.0 : lam[nb[a]];
      const double ua = u[nb[a]];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        gl[i] += t.g[a][i] * la;
        gu[i] += t.g[a][i] * ua;
      }
    }
    grad[e] = -t.vol * (gl[0] * gu[0] + gl[1] * gu[1] + gl[2] * gu[2]);
  }
}

template <int MAXR>
static void launch_assemble(hipStream_t s, int kind, const double* pts, const int4* cells, int64_t N,
                            const int32_t* inc_ptr, const int32_t* inc, const int32_t* rowptr, const int32_t* col,
                            const double* k, double fc, const double* fn, const uint8_t* dm, const double* g,
                            double* vals, double* b, double* dinv, const DiaOut& dia) {
  assemble_p1_k<MAXR><<<grid_for(N, 256, 8192), 256, 0, s>>>(kind, pts, cells, N, inc_ptr, inc, rowptr, col, k, fc,
                                                             fn, dm, g, vals, b, dinv, dia);
}

static int assemble_any(int kind, const double* d_points, const int32_t* d_cells, int64_t n_nodes,
                        const int32_t* d_inc_ptr, const int32_t* d_inc, const int32_t* d_rowptr,
                        const int32_t* d_col, int max_row, const double* d_k, double f_const,
                        const double* d_f_nodal, const uint8_t* d_dmask, const double* d_g, double* d_vals,
                        double* d_b, double* d_dinv, const DiaOut& dia, hipStream_t s) {
  const int4* cells = reinterpret_cast<const int4*>(d_cells);
  if (max_row <= 16)
    launch_assemble<16>(s, kind, d_points, cells, n_nodes, d_inc_ptr, d_inc, d_rowptr, d_col, d_k, f_const, d_f_nodal,
                        d_dmask, d_g, d_vals, d_b, d_dinv, dia);
  else if (max_row <= 32)
    launch_assemble<32>(s, kind, d_points, cells, n_nodes, d_inc_ptr, d_inc, d_rowptr, d_col, d_k, f_const, d_f_nodal,
                        d_dmask, d_g, d_vals, d_b, d_dinv, dia);
  else
    launch_assemble<64>(s, kind, d_points, cells, n_nodes, d_inc_ptr, d_inc, d_rowptr, d_col, d_k, f_const, d_f_nodal,
                        d_dmask, d_g, d_vals, d_b, d_dinv, dia);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

}  // namespace jno

using namespace jno;

extern "C" int jno_assemble_p1(int kind, const double* d_points, const int32_t* d_cells, int64_t n_cells,
                               int64_t n_nodes, const int32_t* d_inc_ptr, const int32_t* d_inc,
                               const int32_t* d_rowptr, const int32_t* d_col, int max_row, const double* d_k,
                               double f_const, const double* d_f_nodal, const uint8_t* d_dmask, const double* d_g,
                               double* d_vals, double* d_b, double* d_dinv, void* stream) {
  JNO_REQUIRE(kind == 0 || kind == 1, JNO_EINVAL, "jno_assemble_p1: kind must be 0 (stiffness) or 1 (mass)");
  JNO_REQUIRE(kind == 1 || d_k, JNO_EINVAL, "jno_assemble_p1: stiffness needs the per-element coefficient k");
  JNO_REQUIRE(max_row >= 0 && max_row <= 64, JNO_ECAPACITY,
              "jno_assemble_p1: a row has %d entries; the P1 row-pull kernel holds at most 64", max_row);
  JNO_REQUIRE(d_points && d_cells && d_rowptr && d_col && d_vals && d_inc_ptr && d_inc, JNO_EINVAL,
              "jno_assemble_p1: null pointer");
  (void)n_cells;
  if (n_nodes == 0) return JNO_OK;
  DiaOut none{};
  return assemble_any(kind, d_points, d_cells, n_nodes, d_inc_ptr, d_inc, d_rowptr, d_col, max_row, d_k, f_const,
                      d_f_nodal, d_dmask, d_g, d_vals, d_b, d_dinv, none, as_stream(stream));
}

extern "C" int jno_assemble_p1_dia(int kind, const double* d_points, const int32_t* d_cells, int64_t n_cells,
                                   int64_t n_nodes, const int32_t* d_inc_ptr, const int32_t* d_inc,
                                   const int32_t* d_rowptr, const int32_t* d_col, int max_row, const double* d_k,
                                   double f_const, const double* d_f_nodal, const uint8_t* d_dmask,
                                   const double* d_g, double* d_vals, double* d_b, double* d_dinv, int ndiag,
                                   const int32_t* h_offsets, double* d_dia64, float* d_dia32, int* h_ok, void* d_ws,
                                   size_t ws_bytes, void* stream) {
  JNO_REQUIRE(kind == 0 || kind == 1, JNO_EINVAL, "jno_assemble_p1_dia: kind must be 0 (stiffness) or 1 (mass)");
  JNO_REQUIRE(kind == 1 || d_k, JNO_EINVAL, "jno_assemble_p1_dia: stiffness needs the per-element coefficient k");
  JNO_REQUIRE(max_row >= 0 && max_row <= 64, JNO_ECAPACITY,
              "jno_assemble_p1_dia: a row has %d entries; the P1 row-pull kernel holds at most 64", max_row);
  JNO_REQUIRE(d_points && d_cells && d_rowptr && d_col && d_vals && d_inc_ptr && d_inc && d_dia64 && h_ok && d_ws,
              JNO_EINVAL, "jno_assemble_p1_dia: null pointer");
  JNO_REQUIRE(ndiag >= 1 && ndiag <= JNO_DIA_MAX && h_offsets && h_offsets[0] == 0, JNO_EINVAL,
              "jno_assemble_p1_dia: bad offsets (1..%d, the main diagonal first)", JNO_DIA_MAX);
  JNO_REQUIRE(ws_bytes >= 256, JNO_EWORKSPACE, "jno_assemble_p1_dia: workspace too small (256 bytes)");
  (void)n_cells;
  *h_ok = 0;
  if (n_nodes == 0) {
    *h_ok = 1;
    return JNO_OK;
  }
  hipStream_t s = as_stream(stream);
  DiaOut dia{};
  dia.ndiag = ndiag;
  for (int d = 0; d < JNO_DIA_MAX; ++d) dia.off.o[d] = d < ndiag ? h_offsets[d] : -1;
  dia.d64 = d_dia64;
  dia.d32 = d_dia32;
  dia.bad = static_cast<int*>(d_ws);
  JNO_HIP(hipMemsetAsync(dia.bad, 0, sizeof(int), s));
  int rc = assemble_any(kind, d_points, d_cells, n_nodes, d_inc_ptr, d_inc, d_rowptr, d_col, max_row, d_k, f_const,
                        d_f_nodal, d_dmask, d_g, d_vals, d_b, d_dinv, dia, s);
  if (rc) return rc;
  int h_bad = 1;
  JNO_HIP(hipMemcpyAsync(&h_bad, dia.bad, sizeof(int), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  *h_ok = h_bad ? 0 : 1;
  return JNO_OK;
}

extern "C" int jno_adjoint_grad_k_p1(const double* d_points, const int32_t* d_cells, int64_t n_cells,
                                     const double* d_u, const double* d_lam, const uint8_t* d_dmask, double* d_grad,
                                     void* stream) {
  JNO_REQUIRE(d_points && d_cells && d_u && d_lam && d_grad, JNO_EINVAL, "jno_adjoint_grad_k_p1: null pointer");
  if (n_cells == 0) return JNO_OK;
  adjoint_grad_p1_k<<<grid_for(n_cells, 256, 8192), 256, 0, as_stream(stream)>>>(
      d_points, reinterpret_cast<const int4*>(d_cells), n_cells, d_u, d_lam, d_dmask, d_grad);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_box_geometry(int nx, int ny, int nz, const double* d_points, double* h_geo, void* d_ws,
                                size_t ws_bytes, void* stream) {
  JNO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && d_points && h_geo && d_ws, JNO_EINVAL, "jno_box_geometry: bad argument");
  JNO_REQUIRE(ws_bytes >= 6 * 17 * sizeof(double), JNO_EWORKSPACE, "jno_box_geometry: workspace too small (816 bytes)");
  hipStream_t s = as_stream(stream);
  box_geom_k<<<1, 64, 0, s>>>(ny, nz, d_points, static_cast<double*>(d_ws));
  JNO_LAUNCH_CHECK();
  JNO_HIP(hipMemcpyAsync(h_geo, d_ws, 6 * 17 * sizeof(double), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  return JNO_OK;
}

extern "C" int jno_mesh_is_kuhn_box(int nx, int ny, int nz, const double* h_lo, const double* h_hi,
                                    const double* d_points, const int32_t* d_cells, int* h_ok, void* d_ws,
                                    size_t ws_bytes, void* stream) {
  JNO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && h_lo && h_hi && d_points && d_cells && h_ok && d_ws, JNO_EINVAL,
              "jno_mesh_is_kuhn_box: bad argument");
  JNO_REQUIRE(ws_bytes >= 256, JNO_EWORKSPACE, "jno_mesh_is_kuhn_box: workspace too small (256 bytes)");
  hipStream_t s = as_stream(stream);
  int* bad = static_cast<int*>(d_ws);
  JNO_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  const int64_t N = (int64_t)(nx + 1) * (ny + 1) * (nz + 1), T = (int64_t)6 * nx * ny * nz;
  box_check_k<<<grid_for(N > T ? N : T, 256, 8192), 256, 0, s>>>(nx, ny, nz, h_lo[0], h_lo[1], h_lo[2], h_hi[0],
                                                                 h_hi[1], h_hi[2], d_points,
                                                                 reinterpret_cast<const int4*>(d_cells), bad);
  JNO_LAUNCH_CHECK();
  int h_bad = 1;
  JNO_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  JNO_HIP(hipStreamSynchronize(s));
  *h_ok = h_bad ? 0 : 1;
  return JNO_OK;
}

extern "C" int jno_assemble_p1_box(int kind, int nx, int ny, int nz, const double* h_geo, const int32_t* d_rowptr,
                                   const double* d_k, double f_const, const double* d_f_nodal, const uint8_t* d_dmask,
                                   const double* d_g, double* d_vals, double* d_b, double* d_dinv, int ndiag,
                                   const int32_t* h_offsets, double* d_dia64, float* d_dia32, void* stream) {
  JNO_REQUIRE(kind == 0 || kind == 1, JNO_EINVAL, "jno_assemble_p1_box: kind must be 0 (stiffness) or 1 (mass)");
  JNO_REQUIRE(kind == 1 || d_k, JNO_EINVAL, "jno_assemble_p1_box: stiffness needs the per-element coefficient k");
  JNO_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, JNO_EINVAL, "jno_assemble_p1_box: the box needs at least 2 cells per axis");
  JNO_REQUIRE(h_geo && d_rowptr && d_vals, JNO_EINVAL, "jno_assemble_p1_box: null pointer");
  JNO_REQUIRE(!d_dia64 || (ndiag >= 1 && ndiag <= JNO_DIA_MAX && h_offsets && h_offsets[0] == 0), JNO_EINVAL,
              "jno_assemble_p1_box: bad DIA offsets (1..%d, the main diagonal first)", JNO_DIA_MAX);
  BoxArgs a{};
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  for (int t = 0; t < 6; ++t) {
    a.geo.vol[t] = h_geo[t];
    for (int p = 0; p < 4; ++p)
      for (int q = 0; q < 4; ++q) a.geo.dot[t][p][q] = h_geo[6 + (t * 4 + p) * 4 + q];
  }
  a.rowptr = d_rowptr;
  a.k = d_k;
  a.f_const = f_const;
  a.f_nodal = d_f_nodal;
  a.dmask = d_dmask;
  a.g = d_g;
  a.vals = d_vals;
  a.b = d_b;
  a.dinv = d_dinv;
  if (d_dia64) {
    // the stencil slot (upper half: 7..14) of each stored diagonal's offset
    const int64_t S2 = nz + 1, S1 = (int64_t)(ny + 1) * (nz + 1);
    a.ndiag = ndiag;
    for (int d = 0; d < JNO_DIA_MAX; ++d) {
      a.dia_slot[d] = -1;
      for (int sl = 7; d < ndiag && sl < 15; ++sl) {
        const int pk = slot_di(sl);
        if (((pk + 13) / 9 - 1) * S1 + (((pk + 13) / 3) % 3 - 1) * S2 + ((pk + 13) % 3 - 1) == h_offsets[d])
          a.dia_slot[d] = sl;
      }
    }
    a.d64 = d_dia64;
    a.d32 = d_dia32;
  }
  const int64_t N = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  const int64_t blocks = (N + BOX_BLOCK - 1) / BOX_BLOCK;
  hipStream_t s = as_stream(stream);
  const bool dm = d_dmask != nullptr;
  if (kind == 1)
    assemble_box_k<1, false, false><<<blocks, BOX_BLOCK, 0, s>>>(a);
  else if (d_f_nodal)
    dm ? assemble_box_k<0, true, true><<<blocks, BOX_BLOCK, 0, s>>>(a)
       : assemble_box_k<0, true, false><<<blocks, BOX_BLOCK, 0, s>>>(a);
  else
    dm ? assemble_box_k<0, false, true><<<blocks, BOX_BLOCK, 0, s>>>(a)
       : assemble_box_k<0, false, false><<<blocks, BOX_BLOCK, 0, s>>>(a);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
