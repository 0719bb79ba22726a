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
