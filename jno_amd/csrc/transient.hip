// Theta-method step of the linear semidiscrete block  M u' + A u = c + f(t)  on the plan's CSR
// pattern (M and A share it):
//   (M + theta dt A) u_next = (M - (1 - theta) dt A) u + dt c + dt (theta f(t+dt) + (1-theta) f(t))
// (SemidiscreteTimeBlock.step, jno/utils/solver/backend_blocks.py:329-351).  The Dirichlet edit of
// the transient block is the reference's _apply_dirichlet_transient (jno/utils/solver/fem_1d.py:
// 939-949): A and c carry the symmetric elimination already; M loses the constrained rows AND
// columns (no diagonal), so a constrained row of the step reads theta dt u_next[d] =
// -(1-theta) dt u[d] + dt g[d].
//
// The step operator K = Mz + theta dt A and the explicit operator E = Mz - (1-theta) dt A are formed
// once per (dt, theta) in one pass over the values (the constant-operator path of
// compose_transient_step_solvers, jno/utils/solver/solver_api.py:856-947), together with K's Jacobi
// inverse diagonal.  Each step then costs one fused pass for its right-hand side (on E's solver
// layout: symmetric DIA on structured meshes, else CSR) plus the PCG solve of K, which reuses the
// steady path's operator layouts and kernels unchanged.
#include "common.h"
#include "sparse.h"

namespace jno {

// one thread per row: Mz masked by the constrained rows/columns, K and E rounded as written
// (m + s*a, no contraction: K and E are bitwise symmetric when M and A are)
__global__ void __launch_bounds__(256) theta_ops_k(int64_t n, const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, const double* __restrict__ mv,
                                                   const double* __restrict__ av, const uint8_t* __restrict__ dmask,
                                                   double sk, double se, double* __restrict__ kv,
                                                   double* __restrict__ ev, double* __restrict__ kdinv) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const bool keep_r = !(dmask && dmask[r]);
    double d = 0.0;
    for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      const int c = col[j];
      const bool keep = keep_r && !(dmask && dmask[c]);
      const double m = keep ? mv[j] : 0.0;
      const double a = av[j];
      const double k = __dadd_rn(m, __dmul_rn(sk, a));
      kv[j] = k;
      if (ev) ev[j] = __dadd_rn(m, __dmul_rn(se, a));
      if (c == r) d += k;
    }
    // 1/diag with the |d| > 1e-30 guard of the step's Jacobi (backend_blocks.py:350-351)
    kdinv[r] = fabs(d) > 1e-30 ? 1.0 / d : 1.0;
  }
}

// The same operators on the diagonal layout, for a mass and an operator that both come with a
// symmetric DIA layout (the structured box: the mass's nd diagonals hold the operator's nda):
// per row and stored diagonal d of the mass, K_d = Mz_d + sk A_d and E_d = Mz_d + se A_d, rounded as
// written, so the values equal theta_ops_k's bitwise; amap[d] is the operator's diagonal at the
// mass's offset d (-1: none, a zero there).  The FP32 copy of K and 1/diag K come in the same pass.
// No CSR is read: the time-dependent step re-forms these every step.
__global__ void __launch_bounds__(256) theta_ops_dia_k(int64_t n, int nd, DiaOffsets off, DiaOffsets amap,
                                                       const double* __restrict__ mdia,
                                                       const double* __restrict__ adia,
                                                       const uint8_t* __restrict__ dmask, double sk, double se,
                                                       double* __restrict__ kdia, float* __restrict__ kdia32,
                                                       double* __restrict__ edia, double* __restrict__ kdinv) {
  const int64_t ld = dia_ld(n);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const bool ci = dmask && dmask[i];
    double kd = 0.0;
    for (int d = 0; d < nd; ++d) {
      const int64_t j = i + off.o[d];
      // the mass's entry (i, j) is kept when neither dof is constrained (past the last row the
      // stored upper value is 0 anyway)
      const bool keep = !ci && !(j < n && dmask && dmask[j]);
      const double m = keep ? mdia[d * ld + i] : 0.0;
      const int da = amap.o[d];
      const double a = da >= 0 ? adia[da * ld + i] : 0.0;
      const double k = __dadd_rn(m, __dmul_rn(sk, a));
      kdia[d * ld + i] = k;
      if (kdia32) kdia32[d * ld + i] = (float)k;
      if (edia) edia[d * ld + i] = __dadd_rn(m, __dmul_rn(se, a));
      if (d == 0) kd = k;
    }
    kdinv[i] = fabs(kd) > 1e-30 ? 1.0 / kd : 1.0;
  }
}

// y = E x [+ dt c] [+ dt (theta f1 + (1-theta) f0)], L lanes per row
template <int L>
__global__ void __launch_bounds__(256) theta_rhs_k(int64_t n, const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, const double* __restrict__ ev,
                                                   const double* __restrict__ x, const double* __restrict__ c,
                                                   const double* __restrict__ f0, const double* __restrict__ f1,
                                                   double dt, double theta, double* __restrict__ y) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int sub = threadIdx.x % L;
  for (int64_t r = gid / L; r < n; r += (int64_t)gridDim.x * blockDim.x / L) {
    const int b = rowptr[r], e = rowptr[r + 1];
    double s = 0.0;
    for (int j = b + sub; j < e; j += L) s += ev[j] * x[col[j]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
    if (sub == 0) {
      if (c) s += dt * c[r];
      if (f1) s += dt * (theta * f1[r] + (1.0 - theta) * f0[r]);
      y[r] = s;
    }
  }
}

// the same on the symmetric DIA layout of E (main and upper diagonals, the lower ones read through
// symmetry): branch-free, out-of-range neighbours read from a clamped row and multiplied by 0
template <int ND>
__global__ void __launch_bounds__(256) theta_rhs_dia_k(int64_t n, int nd, DiaOffsets off, const double* __restrict__ dia,
                                                       const double* __restrict__ x, const double* __restrict__ c,
                                                       const double* __restrict__ f0, const double* __restrict__ f1,
                                                       double dt, double theta, double* __restrict__ y) {
  const int64_t ld = dia_ld(n);
  const int ndr = ND > 0 ? ND : nd;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = dia[i] * x[i];
#pragma unroll
    for (int d = 1; d < (ND > 0 ? ND : JNO_DIA_MAX); ++d) {
      if (ND == 0 && d >= ndr) break;
      const int64_t o = off.o[d];
      const double* v = dia + d * ld;
      const bool up = i + o < n, lo = i - o >= 0;
      const double vu = v[i];  // zero where i + o >= n
      const double xu = x[up ? i + o : i];
      const double vl = v[lo ? i - o : i];
      const double xl = x[lo ? i - o : i];
      s += vu * xu;
      s += (lo ? vl : 0.0) * xl;
    }
    if (c) s += dt * c[i];
    if (f1) s += dt * (theta * f1[i] + (1.0 - theta) * f0[i]);
    y[i] = s;
  }
}

// w = theta u1 + (1-theta) u0 on free rows, g on constrained rows (0 without g)
__global__ void theta_mix_k(int64_t n, const double* __restrict__ u1, const double* __restrict__ u0,
                            const uint8_t* __restrict__ dmask, const double* __restrict__ g, double theta,
                            double* __restrict__ w) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = theta * u1[i] + (1.0 - theta) * u0[i];
    w[i] = (dmask && dmask[i]) ? (g ? g[i] : 0.0) : v;
  }
}

}  // namespace jno

using namespace jno;

extern "C" int jno_theta_operators(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                                   const double* d_mass, const double* d_stiff, const uint8_t* d_dmask, double dt,
                                   double theta, double* d_step, double* d_expl, double* d_step_dinv, void* stream) {
  JNO_REQUIRE(d_rowptr && d_col && d_mass && d_stiff && d_step && d_step_dinv, JNO_EINVAL,
              "jno_theta_operators: null pointer");
  JNO_REQUIRE(dt > 0.0 && theta > 0.0 && theta <= 1.0, JNO_EINVAL,
              "jno_theta_operators: need dt > 0 and 0 < theta <= 1 (got dt=%g, theta=%g)", dt, theta);
  if (n_rows == 0) return JNO_OK;
  theta_ops_k<<<grid_for(n_rows, 256, 8192), 256, 0, as_stream(stream)>>>(
      n_rows, d_rowptr, d_col, d_mass, d_stiff, d_dmask, theta * dt, -(1.0 - theta) * dt, d_step, d_expl, d_step_dinv);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_theta_operators_dia(int64_t n_rows, int ndiag, const int32_t* h_offsets, const double* d_mass_dia,
                                       int ndiag_stiff, const int32_t* h_offsets_stiff, const double* d_stiff_dia,
                                       const uint8_t* d_dmask, double dt, double theta, double* d_step_dia,
                                       float* d_step_dia32, double* d_expl_dia, double* d_step_dinv, void* stream) {
  JNO_REQUIRE(h_offsets && d_mass_dia && h_offsets_stiff && d_stiff_dia && d_step_dia && d_step_dinv, JNO_EINVAL,
              "jno_theta_operators_dia: null pointer");
  JNO_REQUIRE(dt > 0.0 && theta > 0.0 && theta <= 1.0, JNO_EINVAL,
              "jno_theta_operators_dia: need dt > 0 and 0 < theta <= 1 (got dt=%g, theta=%g)", dt, theta);
  JNO_REQUIRE(ndiag >= 1 && ndiag <= JNO_DIA_MAX && ndiag_stiff >= 1 && ndiag_stiff <= ndiag, JNO_EINVAL,
              "jno_theta_operators_dia: bad diagonal counts (%d, %d)", ndiag, ndiag_stiff);
  JNO_REQUIRE(h_offsets[0] == 0 && h_offsets_stiff[0] == 0, JNO_EINVAL,
              "jno_theta_operators_dia: the main diagonal is stored first");
  DiaOffsets off, amap;
  for (int d = 0; d < JNO_DIA_MAX; ++d) {
    off.o[d] = d < ndiag ? h_offsets[d] : 0;
    amap.o[d] = -1;
  }
  for (int e = 0; e < ndiag_stiff; ++e) {
    int hit = -1;
    for (int d = 0; d < ndiag; ++d)
      if (h_offsets[d] == h_offsets_stiff[e]) hit = d;
    JNO_REQUIRE(hit >= 0, JNO_EINVAL,
                "jno_theta_operators_dia: the operator's offset %d is not among the mass's diagonals",
                h_offsets_stiff[e]);
    amap.o[hit] = e;
  }
  if (n_rows == 0) return JNO_OK;
  theta_ops_dia_k<<<grid_for(n_rows, 256, 16384), 256, 0, as_stream(stream)>>>(
      n_rows, ndiag, off, amap, d_mass_dia, d_stiff_dia, d_dmask, theta * dt, -(1.0 - theta) * dt, d_step_dia,
      d_step_dia32, d_expl_dia, d_step_dinv);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_theta_rhs(const jno_op* h_expl, const double* d_u, const double* d_c, const double* d_f_now,
                             const double* d_f_next, double dt, double theta, double* d_rhs, void* stream) {
  JNO_REQUIRE(h_expl && d_u && d_rhs, JNO_EINVAL, "jno_theta_rhs: null pointer");
  JNO_REQUIRE(h_expl->vals64, JNO_EINVAL, "jno_theta_rhs: the explicit operator has no FP64 values");
  JNO_REQUIRE((d_f_now == nullptr) == (d_f_next == nullptr), JNO_EINVAL,
              "jno_theta_rhs: the forcing needs both f(t) and f(t+dt), or neither");
  JNO_REQUIRE(d_u != d_rhs, JNO_EINVAL, "jno_theta_rhs: rhs must not alias u");
  const int64_t n = h_expl->n;
  if (n == 0) return JNO_OK;
  hipStream_t s = as_stream(stream);
  const double* ev = h_expl->vals64;
  if (h_expl->format == JNO_FMT_DIA) {
    JNO_REQUIRE(h_expl->ndiag >= 1 && h_expl->ndiag <= JNO_DIA_MAX, JNO_EINVAL, "jno_theta_rhs: bad diagonal count");
    DiaOffsets off;
    for (int d = 0; d < JNO_DIA_MAX; ++d) off.o[d] = h_expl->offsets[d];
    const int g = grid_for(n, 256, 16384);
    switch (h_expl->ndiag) {
      case 4: theta_rhs_dia_k<4><<<g, 256, 0, s>>>(n, 4, off, ev, d_u, d_c, d_f_now, d_f_next, dt, theta, d_rhs); break;
      case 8: theta_rhs_dia_k<8><<<g, 256, 0, s>>>(n, 8, off, ev, d_u, d_c, d_f_now, d_f_next, dt, theta, d_rhs); break;
      default:
        theta_rhs_dia_k<0><<<g, 256, 0, s>>>(n, h_expl->ndiag, off, ev, d_u, d_c, d_f_now, d_f_next, dt, theta, d_rhs);
    }
  } else {
    JNO_REQUIRE(h_expl->format == JNO_FMT_CSR, JNO_EINVAL, "jno_theta_rhs: the explicit operator must be DIA or CSR");
    const int32_t* rp = h_expl->rowptr;
    const int32_t* cl = h_expl->col;
#define JNO_RHS_CSR(L_)                                                                                           \
  case L_:                                                                                                        \
    theta_rhs_k<L_><<<grid_for(n * L_, 256, 16384), 256, 0, s>>>(n, rp, cl, ev, d_u, d_c, d_f_now, d_f_next, dt, \
                                                                 theta, d_rhs);                                   \
    break;
    switch (h_expl->lanes) {
      JNO_RHS_CSR(1)
      JNO_RHS_CSR(2)
      JNO_RHS_CSR(4)
      JNO_RHS_CSR(8)
      JNO_RHS_CSR(16)
      JNO_RHS_CSR(32)
      JNO_RHS_CSR(64)
      default: JNO_REQUIRE(false, JNO_EINVAL, "jno_theta_rhs: lanes per row must be a power of two in 1..64");
    }
#undef JNO_RHS_CSR
  }
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_theta_mix(int64_t n, const double* d_u_next, const double* d_u_now, const uint8_t* d_dmask,
                             const double* d_g, double theta, double* d_w, void* stream) {
  JNO_REQUIRE(d_u_next && d_u_now && d_w, JNO_EINVAL, "jno_theta_mix: null pointer");
  if (n == 0) return JNO_OK;
  theta_mix_k<<<grid_for(n, 256, 8192), 256, 0, as_stream(stream)>>>(n, d_u_next, d_u_now, d_dmask, d_g, theta, d_w);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
