/*
 * jno_fem.h -- C-ABI of the MI355X (gfx950) finite-element hot path of jNO:
 * Lagrange P1/P2 tetrahedral assembly into CSR, symmetric Dirichlet elimination and
 * Jacobi-preconditioned CG (plain FP64, FP32-inner with FP64 refinement, batched), plus the
 * adjoint dL/dk contraction.
 *
 * Conventions
 *   - Every pointer argument named d_* is DEVICE memory owned by the caller.  Host pointers are
 *     named h_*.  Integer index arrays are int32 (the reference's BCOO index dtype).
 *   - `stream` is a hipStream_t passed as void*; NULL means the legacy default stream.
 *   - Entry points return 0 on success or a negative JNO_E* status; jno_last_error() gives the
 *     message of the calling thread's last failure.
 *   - Nothing here allocates device memory on the hot path: scratch comes in through `d_ws`
 *     whose size the matching *_workspace_size() query returns.
 *   - Reference citations are relative to the FhG-IISB/jNO 0.3.1 checkout.
 */
#ifndef JNO_FEM_H
#define JNO_FEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  JNO_OK = 0,
  JNO_EINVAL = -1,       /* bad argument (ValueError in Python)                                 */
  JNO_EHIP = -2,         /* HIP runtime error                                                   */
  JNO_ENOTCONV = -3,     /* iteration cap reached before the tolerance                          */
  JNO_ENONFINITE = -4,   /* NaN/Inf in the iteration                                            */
  JNO_EWORKSPACE = -5,   /* workspace too small                                                 */
  JNO_ECAPACITY = -6     /* a row or node exceeds a compiled capacity (e.g. > 64 P1 neighbours)  */
};

/* Message of the calling thread's last failure; returns its full length. */
int jno_last_error(char* buf, size_t len);
/* Library version string (static storage). */
const char* jno_version(void);

/* ---------------------------------------------------------------------------------------------
 * A1 mesh.  Replaces Geometries.equi_distant_box (jno/domain/geometries.py:204-301) as reached by
 * Shape.box(..).structured(n) (jno/geometry/shape.py:415-458): (n+1)^3 lattice, node id
 * (i*(ny+1)+j)*(nz+1)+k, six Kuhn tets per voxel in voxel-major order.
 * d_points: f64[(nx+1)(ny+1)(nz+1)*3]; d_cells: i32[6*nx*ny*nz*4].
 * ------------------------------------------------------------------------------------------- */
int jno_mesh_kuhn_box(int nx, int ny, int nz, const double* h_lo, const double* h_hi,
                      double* d_points, int32_t* d_cells, void* stream);

/* ---------------------------------------------------------------------------------------------
 * A7 sparsity pattern ("plan").  Replaces the static pattern hoist + compress_plan
 * (jno/utils/solver/fem_native.py:2264-2298, jno/utils/solver/fem_utils.py:4009-4060): the CSR
 * index structure = the (row, col)-sorted unique element pairs, dof = node*vec + comp.
 * Two phases: count fills d_rowptr (ndof+1) and the node->(element,local) incidence that assembly
 * reuses, and returns nnz in *h_nnz; fill writes d_col (nnz).
 * d_inc_ptr: i32[n_nodes+1]; d_inc: i32[T*nloc] (entries e*nloc+a, ascending per node).
 * ------------------------------------------------------------------------------------------- */
size_t jno_pattern_workspace_size(int64_t n_cells, int nloc, int64_t n_nodes);
int jno_pattern_count(const int32_t* d_cells, int64_t n_cells, int nloc, int vec, int64_t n_nodes,
                      int32_t* d_inc_ptr, int32_t* d_inc, int32_t* d_rowptr, int64_t* h_nnz,
                      int32_t* h_max_row, void* d_ws, size_t ws_bytes, void* stream);
int jno_pattern_fill(const int32_t* d_cells, int64_t n_cells, int nloc, int vec, int64_t n_nodes,
                     const int32_t* d_inc_ptr, const int32_t* d_inc, const int32_t* d_rowptr,
                     int32_t* d_col, void* stream);

/* The same plan for a structured Kuhn box (the mesh jno_mesh_kuhn_box makes, nx, ny, nz >= 2), in
 * closed form: the rowptr, col, inc_ptr and inc jno_pattern_count/fill produce for that mesh, in
 * one pass with no sort (no reference counterpart as a separate entry: compress_plan specialized
 * to the lattice).  jno_pattern_kuhn_box_nnz gives nnz and the longest row (15) for the
 * allocation; d_col: i32[nnz], d_inc: i32[24 nx ny nz].  The caller guarantees the mesh is the box
 * (jno_mesh_is_kuhn_box). */
int jno_pattern_kuhn_box_nnz(int nx, int ny, int nz, int64_t* h_nnz, int32_t* h_max_row);
int jno_pattern_kuhn_box(int nx, int ny, int nz, int32_t* d_rowptr, int32_t* d_col, int32_t* d_inc_ptr,
                         int32_t* d_inc, void* stream);

/* Opaque plan handles keyed by a content hash of the mesh's cells: the pattern of
 * jno_pattern_count/fill and the node incidence, built once per cell content and shared.  The
 * reference memoizes compress_plan on a digest of its index array
 * (jno/utils/solver/fem_utils.py:4027-4047); here the key is the device, the two content sums of
 * jno_hash_i32 and the sizes, and a hit is confirmed by comparing the cells with the plan's own copy
 * (a sum collision gets a plan of its own).  jno_plan_create returns the cached plan with its
 * reference count raised, or builds one (it allocates the plan's device memory: a set-up call, not
 * the hot path; it synchronizes the stream); jno_plan_get fills a view of its device arrays;
 * jno_plan_destroy drops one reference and frees the plan with the last (after the device is idle);
 * an unknown or already freed handle is JNO_EINVAL.  Thread-safe: a mutex guards the table. */
typedef struct jno_plan jno_plan;
typedef struct jno_plan_view {
  int64_t n_rows;          /* n_nodes * vec                                   */
  int64_t nnz;
  int64_t n_cells;
  int64_t n_nodes;
  int32_t nloc, vec;
  int32_t max_row;         /* longest row                                     */
  int32_t refs;            /* current reference count                         */
  const int32_t* rowptr;   /* i32[n_rows+1]                                   */
  const int32_t* col;      /* i32[nnz]                                        */
  const int32_t* inc_ptr;  /* i32[n_nodes+1]                                  */
  const int32_t* inc;      /* i32[n_cells*nloc]                               */
  uint64_t key[2];         /* the content sums                                */
} jno_plan_view;
int jno_plan_create(const int32_t* d_cells, int64_t n_cells, int nloc, int vec, int64_t n_nodes,
                    jno_plan** h_plan, void* stream);
int jno_plan_get(const jno_plan* plan, jno_plan_view* h_view);
int jno_plan_destroy(jno_plan* plan);
int jno_plan_live_count(int64_t* h_count);

/* Content key of an index array (the plan cache keys patterns on the cells' content, as
 * compress_plan memoizes on a digest of its index array, jno/utils/solver/fem_utils.py:4027-4060):
 * h_hash[2] = two independent 64-bit sums of a position-dependent mix of every entry; and an
 * equality check of two arrays (*h_equal = 1 when equal) that guards a cache hit.  d_ws: 256 bytes.
 * Both synchronize the stream. */
int jno_hash_i32(const int32_t* d_data, int64_t n, uint64_t* h_hash, void* d_ws, size_t ws_bytes, void* stream);
int jno_equal_i32(const int32_t* d_a, const int32_t* d_b, int64_t n, int* h_equal, void* d_ws, size_t ws_bytes,
                  void* stream);

/* ---------------------------------------------------------------------------------------------
 * A10 boundary.  Replaces _boundary_facets (jno/_fem.py:2695-2773) + _boundary_node_ids
 * (jno/utils/solver/fem_native.py:2563-2636): d_mask[n] = 1 for every node of a facet used by
 * exactly one cell (P2: the facet's edge nodes too).  jno_mask_plane keeps the masked nodes whose
 * coordinate `axis` equals `value` exactly (a named box face: left/right/... ).
 * ------------------------------------------------------------------------------------------- */
int jno_boundary_mask(const int32_t* d_cells, int64_t n_cells, int nloc, int64_t n_nodes,
                      const int32_t* d_inc_ptr, const int32_t* d_inc, uint8_t* d_mask, void* stream);
int jno_mask_plane(const double* d_points, int64_t n_nodes, int axis, double value,
                   const uint8_t* d_mask_in, uint8_t* d_mask_out, void* stream);
/* jno_boundary_mask's result for a structured Kuhn box in closed form: the nodes on the box
 * surface (the caller guarantees the mesh is the box). */
int jno_mask_kuhn_box(int nx, int ny, int nz, uint8_t* d_mask, void* stream);

/* ---------------------------------------------------------------------------------------------
 * A4-A6, A8, A9, A11, A14, A19  P1 scalar assembly.  Replaces FemLinearSystem.evaluate
 * (jno/trace/__init__.py:3612-3615 -> _assemble_at, jno/utils/solver/fem_native.py:3311-3318) for
 * the forms  kind=0: k grad u.grad v - f v  (stiffness + load)   kind=1: u v  (mass, k ignored).
 * Row-pull, atomic-free: every row sums its incident element rows in ascending element order.
 * d_k: f64[T] per-element coefficient (P0 parameter, jno/architectures/models.py:97-107).
 * f: if d_f_nodal is NULL the load is the constant f_const, else the P1 nodal field d_f_nodal[N].
 * d_dmask (nullable): u8[N] Dirichlet mask; d_g (nullable = zeros): f64[N] Dirichlet values.  With a
 * mask the symmetric elimination of _apply_dirichlet_symmetric (jno/utils/solver/fem_1d.py:667-686)
 * is fused: b -= A e(g), constrained rows/cols zeroed (pattern kept), unit diagonal, b[d] = g.
 * d_b (nullable) f64[N]; d_dinv (nullable) f64[N] Jacobi 1/diag with the |d| > 1e-30 guard
 * (jno/precond.py:105-109).  max_row = the plan's longest row (from jno_pattern_count).
 * ------------------------------------------------------------------------------------------- */
int jno_assemble_p1(int kind, const double* d_points, const int32_t* d_cells, int64_t n_cells,
                    int64_t n_nodes, const int32_t* d_inc_ptr, const int32_t* d_inc,
                    const int32_t* d_rowptr, const int32_t* d_col, int max_row, const double* d_k,
                    double f_const, const double* d_f_nodal, const uint8_t* d_dmask,
                    const double* d_g, double* d_vals, double* d_b, double* d_dinv, void* stream);

/* jno_assemble_p1 that also writes the symmetric DIA layout of the result (no reference
 * counterpart: it fuses jno_op_fill_dia into the assembly for an offset set found earlier on the
 * same plan).  d_dia64 (and d_dia32 when not null) get ndiag diagonals of ld = n rounded up to 64
 * rows, val[d][i] = A[i, i + h_offsets[d]].  *h_ok = 1 when every nonzero entry lies on a stored
 * diagonal (the layout reproduces A; the assembly is bitwise symmetric by construction), else 0
 * (then the DIA buffers must not be used; the CSR outputs are valid either way).  d_ws: 256
 * bytes of device scratch.  Synchronizes the stream to read the flag. */
int jno_assemble_p1_dia(int kind, const double* d_points, const int32_t* d_cells, int64_t n_cells,
                        int64_t n_nodes, const int32_t* d_inc_ptr, const int32_t* d_inc,
                        const int32_t* d_rowptr, const int32_t* d_col, int max_row, const double* d_k,
                        double f_const, const double* d_f_nodal, const uint8_t* d_dmask,
                        const double* d_g, double* d_vals, double* d_b, double* d_dinv, int ndiag,
                        const int32_t* h_offsets, double* d_dia64, float* d_dia32, int* h_ok, void* d_ws,
                        size_t ws_bytes, void* stream);

/* The same assembly on a structured Kuhn box (the mesh jno_mesh_kuhn_box makes with nx, ny, nz >= 2,
 * on the plan jno_pattern_count/fill built for it; no reference counterpart as a separate entry:
 * it is jno_assemble_p1 specialized to that mesh).  Each row sums the same element rows in the
 * same ascending element order without gathering the incidence, cells or points: the six cell
 * types' |T| and term-by-term gradient dot products come from h_geo (jno_box_geometry, once per
 * mesh), and only k is read per element.  On dyadic spacings the values equal jno_assemble_p1's
 * bitwise; on others every cell of a type takes voxel 0's geometry (a last-bit difference).
 * d_dia64 (nullable) / d_dia32: the DIA layout at h_offsets, which must hold every offset with a
 * nonzero coupling (the caller decides that from h_geo: no flag, no stream synchronization).
 * jno_box_geometry: h_geo[6*17] = per type |T| then the 4x4 dot products (d_ws 816 bytes).
 * jno_mesh_is_kuhn_box: *h_ok = 1 when the points and cells are bitwise those of the box
 * (d_ws 256 bytes).  Both synchronize the stream. */
int jno_box_geometry(int nx, int ny, int nz, const double* d_points, double* h_geo, void* d_ws, size_t ws_bytes,
                     void* stream);
int jno_mesh_is_kuhn_box(int nx, int ny, int nz, const double* h_lo, const double* h_hi, const double* d_points,
                         const int32_t* d_cells, int* h_ok, void* d_ws, size_t ws_bytes, void* stream);
int jno_assemble_p1_box(int kind, int nx, int ny, int nz, const double* h_geo, const int32_t* d_rowptr,
                        const double* d_k, double f_const, const double* d_f_nodal, const uint8_t* d_dmask,
                        const double* d_g, double* d_vals, double* d_b, double* d_dinv, int ndiag,
                        const int32_t* h_offsets, double* d_dia64, float* d_dia32, void* stream);

/* A11 on any assembled CSR (P2 elasticity uses it): bcoo_zero_rows_cols + bcoo_set_unit_diag
 * (jno/utils/solver/fem_utils.py:3332-3359) with the lift b -= A e(g) and b[d] = g.  In place. */
int jno_dirichlet_eliminate(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                            double* d_vals, double* d_b, const uint8_t* d_dmask, const double* d_g,
                            void* stream);

/* A14 matrix_diagonal + _Jacobi.materialize (jno/utils/solver/linear.py:88-94,
 * jno/precond.py:105-109): dinv = 1/d where |d| > 1e-30 else 1. */
int jno_jacobi(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
               double* d_dinv, void* stream);

/* A15 BCOO matvec LinearOperator.mv (jno/utils/solver/solver_api.py:137-147) on CSR: y = A x. */
int jno_spmv_f64(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
                 const double* d_x, double* d_y, void* stream);

/* A12 compress_eager (jno/utils/solver/fem_utils.py:4063-4091): drop exact-zero entries.
 * Count phase writes d_rowptr_out and *h_nnz_out; compact phase writes d_col_out/d_vals_out. */
int jno_drop_zeros_count(int64_t n_rows, const int32_t* d_rowptr, const double* d_vals,
                         int32_t* d_rowptr_out, int64_t* h_nnz_out, void* d_ws, size_t ws_bytes,
                         void* stream);
int jno_drop_zeros_fill(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                        const double* d_vals, const int32_t* d_rowptr_out, int32_t* d_col_out,
                        double* d_vals_out, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Solver operators.  A "solver operator" is an opaque device-side re-layout of an assembled CSR
 * for the PCG kernels (it is to the CSR what XLA's chosen layout is to a BCOO).  Formats:
 *   JNO_FMT_CSR  : CSR with `lanes` lanes per row (1..64), FP64 and FP32 value copies.
 *   JNO_FMT_DIA  : symmetric diagonal storage (main + upper diagonals; lower read by symmetry),
 *                  chosen when the operator has at most 2*JNO_DIA_MAX-1 distinct diagonals and
 *                  is structurally symmetric.  Index-free, fully coalesced.
 *   JNO_FMT_BSR3 : 3x3-block CSR for node-interleaved vector fields (vec = 3, e.g. P2
 *                  elasticity): rowptr/col index block rows and block columns (nodes); the 9
 *                  entries of block row R are stored entry-major, entry e = 3i+j of its block b at
 *                  9*rowptr[R] + e*len_R + b, so a lane per block reads each entry coalesced;
 *                  `lanes` = lanes per block row, nnz = 9 * blocks.  Stored zeros are kept.
 * jno_op_analyze inspects a CSR and reports whether DIA applies; jno_op_fill_dia lays it out;
 * jno_op_bsr3_count / jno_op_fill_bsr3 do the same for BSR3.
 * ------------------------------------------------------------------------------------------- */
enum { JNO_FMT_CSR = 0, JNO_FMT_DIA = 1, JNO_FMT_BSR3 = 2, JNO_FMT_EBE = 3 };
#define JNO_DIA_MAX 16

/* Launch tuning of the PCG iteration kernels, carried per call (no reference counterpart: these
 * are launch-granularity knobs of this engine, chosen by the autotuner).  A jno_op whose `tuning`
 * is NULL runs the library defaults (jno_pcg_set_tuning / _set_chunk / _set_pacing set those).
 * iter[0]: the FP64 iterations of jno_pcg_f64; iter[1]: the FP32 inner iterations of
 * jno_pcg_mixed.  The fields have the meaning of jno_pcg_set_tuning's arguments. */
typedef struct jno_launch_shape {
  int32_t grid_cap;
  int32_t k1_vector;
  int32_t block;
  int32_t split;
} jno_launch_shape;

typedef struct jno_tuning {
  jno_launch_shape iter[2];
  int32_t chunk;   /* iterations enqueued per host poll, 1..1024 */
  int32_t pacing;  /* 1: paced chunks for the one-kernel FP32 inner loops, 0: fixed chunks */
} jno_tuning;

/* Element-by-element (matrix-free) operator of a P2 vector (3 dofs per node) element form: the
 * analogue of the reference's ElementOperator (jno/utils/solver/element_operator.py:1-178), which
 * keeps one dense (n_local, n_local) block per cell plus the connectivity instead of a global sparse
 * matrix.  Here a block is E_e * K_c with K_c the block of the cell's geometry class (cells whose
 * vertex differences are bitwise equal share one: the 6 Kuhn tet shapes of a structured box), so
 * the operator stores a class table, one scale per cell and the connectivity.  y = A x is the
 * symmetrically eliminated operator: y = P K P x + (I - P) x with P the projector onto the free
 * dofs.  Element slots are grouped by class in tiles of 16 (one MFMA tile; padding slots hold -1).
 * All pointers are device memory. */
typedef struct jno_ebe {
  int64_t n_cells;              /* T                                                          */
  int64_t n_nodes;              /* N; the operator has 3 N rows (dof = 3 node + comp)         */
  int64_t n_tiles;              /* tiles of 16 element slots                                  */
  int32_t n_classes;            /* geometry classes                                           */
  int32_t pad;
  const int32_t* cells;         /* i32[10*16 n_tiles]: each slot's cell nodes (basix P2 order), in
                                   slot order (node 0 for padding)                             */
  const int32_t* order;         /* i32[16 n_tiles]: cell of each slot, -1 for padding          */
  const int32_t* tile_class;    /* i32[n_tiles]: class of each tile                            */
  const uint32_t* slot_mask;    /* u32[16 n_tiles]: bit d = local dof d of the slot is constrained
                                   (all bits for padding)                                      */
  const double* table64;        /* f64[n_classes*1024]: K_c[r][k] (zero past 30) at
                                   ((r/16)*8 + k/4)*64 + (k%4)*16 + r%16 (MFMA operand order)  */
  const float* table32;         /* f32 copy (FP32 inner iterations)                            */
  const double* scale64;        /* f64[16 n_tiles]: each slot's scale E_e (0 for padding)      */
  const float* scale32;         /* f32[16 n_tiles]                                             */
  const int32_t* inc_ptr;       /* i32[N+1]: node -> its incidence entries                     */
  const int32_t* inc;           /* i32[10 T]: slot*10 + local node, ascending cell per node    */
  const int32_t* qpos;          /* i32[10*16 n_tiles]: incidence position of (slot, local node),
                                   -1 for padding (the inverse of inc)                         */
  const uint8_t* dmask;         /* u8[3N]: constrained dofs                                    */
  void* scratch;                /* f64[30 T]: product buffer of jno_op_spmv_* (nullable)       */
  /* Structured Kuhn box (jno_ebe_box_map found the cells to be the P2 box of this grid, and every
   * cell of Kuhn type t to have the class box_class[t]): the SpMV runs on the half lattice, with
   * no product buffer.  box_n[0] = 0: the general path. */
  int32_t box_n[3];             /* voxels per axis                                             */
  int32_t box_class[6];         /* class (table block) of each Kuhn tet type                  */
  int32_t box_pad;
  const int32_t* box_perm;      /* i32[(2nx+1)(2ny+1)(2nz+1)]: node at each half-lattice point */
  const uint8_t* box_mask;      /* u8[same]: its constrained-dof bits (bit c: dof 3 node + c)  */
  const double* box_e64;        /* f64[T]: E per cell, in cell order                           */
  const float* box_e32;         /* f32[T]                                                      */
  const float* box_rows32;      /* f32[6*10*10*16]: K_{box_class[t]}[3a + r][3b + c] at
                                   ((10 t + a) 10 + b) 16 + 3 r + c (9 of 16 used)             */
} jno_ebe;

typedef struct jno_op {
  int32_t format;      /* JNO_FMT_*                                                     */
  int32_t lanes;       /* CSR: lanes per row                                            */
  int64_t n;           /* rows                                                          */
  int64_t nnz;         /* CSR nnz                                                       */
  const int32_t* rowptr;
  const int32_t* col;
  const double* vals64;   /* CSR FP64 values (or DIA FP64 values [ndiag][n])             */
  const float* vals32;    /* CSR FP32 values (or DIA FP32 values [ndiag][n]); nullable   */
  int32_t ndiag;       /* DIA: stored diagonals (main first, then ascending upper)       */
  int32_t offsets[JNO_DIA_MAX]; /* DIA: stored offsets (0 first)                        */
  const jno_tuning* tuning;     /* nullable: launch tuning of the solves on this operator */
  const jno_ebe* ebe;           /* JNO_FMT_EBE: the element data (vals64/vals32 = its tables) */
} jno_op;

/* Building an EBE operator (geometry classes are found on the host from the cells' Jacobians;
 * the rest runs here).
 * jno_p2_cell_jacobians: d_jac f64[9 T], the vertex differences x_{v+1} - x_0 of each cell (v-major),
 *   exactly as the element kernels form them: cells with bitwise-equal rows share an element matrix.
 * jno_ebe_class_blocks: K_c (E = 1) of the classes' representative cells d_rep[C] on FP64 MFMA, the
 *   same arithmetic as jno_assemble_p2_elasticity's element matrices, into d_table64[C*1024] (and
 *   the FP32 copy d_table32, nullable) in the table order of jno_ebe, zero past 30.
 * jno_ebe_slots: each slot's cell nodes in slot order (d_slot_cells i32[10 n_slots]) and its
 *   constrained-dof bits (d_dmask u8[3N]) into d_slot_mask.
 * jno_ebe_scales: each slot's scale from the per-cell d_E (f64[T]) into d_s64 and d_s32 (slot order,
 *   0 for padding), once per operator value.
 * jno_ebe_incidence: d_inc_out[q] = d_slot_of[e]*10 + a for each plan incidence entry e*10 + a,
 *   and its inverse d_qpos[slot*10 + a] = q (i32[10 n_slots], -1 for padding slots).
 * jno_ebe_diagonal: the exact diagonal of the operator (the reference's ElementOperator.diag,
 *   element_operator.py:150-157), or with invert = 1 its Jacobi inverse with the |d| > 1e-30 guard
 *   (jno/precond.py:105-109).
 * jno_ebe_dirichlet_rhs: b[d] = 0 on the constrained dofs (the elimination's b[d] = g with g = 0). */
int jno_p2_cell_jacobians(const double* d_points2, const int32_t* d_cells10, int64_t n_cells, double* d_jac,
                          void* stream);
int jno_ebe_class_blocks(const double* d_points2, const int32_t* d_cells10, int n_classes, const int32_t* d_rep,
                         double nu, double* d_table64, float* d_table32, void* stream);
int jno_ebe_slots(int64_t n_slots, const int32_t* d_order, const int32_t* d_cells10, const uint8_t* d_dmask,
                  int32_t* d_slot_cells, uint32_t* d_slot_mask, void* stream);
int jno_ebe_scales(int64_t n_slots, const int32_t* d_order, const double* d_E, double* d_s64, float* d_s32,
                   void* stream);
int jno_ebe_incidence(int64_t n_entries, const int32_t* d_inc, const int32_t* d_slot_of, int64_t n_slots,
                      int32_t* d_inc_out, int32_t* d_qpos, void* stream);
int jno_ebe_diagonal(const jno_op* h_op, int invert, double* d_out, void* stream);
int jno_ebe_dirichlet_rhs(int64_t n_rows, const uint8_t* d_dmask, double* d_b, void* stream);
/* The half-lattice map of a P2 mesh that is the structured Kuhn box of nx x ny x nz voxels (the P2
 * promotion of Shape.box(..).structured(n), jno/domain/geometries.py:204-301 with
 * jno/utils/solver/fem_utils.py:1979-2044): d_perm[L] = the node at half-lattice point
 * L = (I (2ny+1) + J)(2nz+1) + K and d_mask_lat[L] its constrained-dof bits from d_dmask (u8[3N]).
 * *h_ok = 1 when the cells are exactly that box (every cell e = 6 voxel + t holds the nodes of
 * Kuhn type t's local points, each node at one point, every point covered), else 0 (the general
 * element-by-element path applies).  d_ws: i32[N + 1].  Synchronizes the stream (one-shot set-up). */
int jno_ebe_box_map(int nx, int ny, int nz, const int32_t* d_cells10, int64_t n_cells, int64_t n_nodes,
                    const uint8_t* d_dmask, int32_t* d_perm, uint8_t* d_mask_lat, int32_t* d_ws, int32_t* h_ok,
                    void* stream);
/* The coefficient table of the structured-box SpMV: d_rows f32[6*10*10*16] from the FP32 class table of
 * jno_ebe (d_table32, n_classes blocks) and the class of each Kuhn type (h_class, 6 host ints). */
int jno_ebe_box_rows(const float* d_table32, int n_classes, const int32_t* h_class, float* d_rows, void* stream);

/* Distinct column-row offsets (up to JNO_DIA_MAX upper ones) and symmetry of a CSR.
 * *h_ndiag = number of distinct offsets >= 0 (0 is always counted), or -1 when there are more than
 * JNO_DIA_MAX or the matrix is not symmetric (bitwise, when d_vals is given; structurally when it
 * is NULL); h_offsets (JNO_DIA_MAX entries) receives them ascending from 0. */
size_t jno_op_analyze_workspace_size(void);
int jno_op_analyze(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
                   int32_t* h_ndiag, int32_t* h_offsets, void* d_ws, size_t ws_bytes, void* stream);
/* DIA fill from CSR: d_dia64 f64[ndiag*ld] (nullable), d_dia32 f32[ndiag*ld] (nullable), where
 * diagonal d starts at d*ld and ld = n rounded up to a multiple of 64 (aligned vector loads). */
int jno_op_fill_dia(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                    const double* d_vals, int ndiag, const int32_t* h_offsets, double* d_dia64,
                    float* d_dia32, void* stream);
/* jno_op_fill_dia plus a check that the layout reproduces the CSR: *h_ok = 1 when every stored
 * entry off the given diagonals is exactly 0 and every lower entry equals its stored upper mirror
 * bitwise (so the offsets of an earlier operator on the same pattern can be reused without a new
 * jno_op_analyze), else 0.  d_dia64 is required; d_ws: 256 bytes. */
int jno_op_fill_dia_checked(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                            const double* d_vals, int ndiag, const int32_t* h_offsets, double* d_dia64,
                            float* d_dia32, int* h_ok, void* d_ws, size_t ws_bytes, void* stream);
/* BSR3 applicability of a CSR: *h_nnzb = number of 3x3 blocks when every block row's three rows
 * hold the columns 3c+0, 3c+1, 3c+2 of the same ascending block columns c (the vec = 3 pattern of
 * jno_pattern_count/fill), else -1.  jno_op_fill_bsr3 writes the block rowptr (n_rows/3 + 1), the
 * block columns (nnzb) and the values (9*nnzb; FP64 and FP32 copies, each nullable). */
size_t jno_op_bsr3_workspace_size(void);
int jno_op_bsr3_count(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, int64_t* h_nnzb,
                      void* d_ws, size_t ws_bytes, void* stream);
int jno_op_fill_bsr3(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
                     int32_t* d_browptr, int32_t* d_bcol, double* d_bvals64, float* d_bvals32,
                     void* stream);
/* FP32 copy of CSR values. */
int jno_cast_f32(int64_t n, const double* d_in, float* d_out, void* stream);
/* y = A x with a solver operator, FP64 values (an EBE operator needs its ebe->scratch). */
int jno_op_spmv_f64(const jno_op* h_op, const double* d_x, double* d_y, void* stream);
int jno_op_spmv_f32(const jno_op* h_op, const float* d_x, float* d_y, void* stream);

/* ---------------------------------------------------------------------------------------------
 * A16 + A17  Jacobi-PCG.  Replaces jno.solve.cg -> jax.scipy.sparse.linalg.cg (jno/solve.py:191-206)
 * with M = jno.precond.jacobi(): x0 (or 0), r = b - A x0, z = M r, p = z; loop while r.r >
 * max(tol^2 b.b, atol^2) and k < maxiter.  The final relative residual ||b - A x||/||b|| (true,
 * FP64) is returned in *h_relres so the caller can apply the 1e-4 residual guard of
 * _maybe_residual_check (jno/utils/solver/solver_api.py:224-240).
 * d_x: in = x0 (if use_x0) / out = solution.  Returns JNO_ENOTCONV at maxiter (x still written).
 * ------------------------------------------------------------------------------------------- */
typedef struct jno_pcg_stats {
  int64_t iterations;      /* inner iterations (all precisions)          */
  int64_t refinements;     /* FP64 residual replacements (mixed only)    */
  double relres;           /* true FP64 ||b - A x|| / ||b||              */
  double relres_recursive; /* recursive residual at exit                 */
  double seconds;          /* wall time of the iteration on the stream   */
} jno_pcg_stats;

size_t jno_pcg_workspace_size(int64_t n, int mixed);
/* The workspace for a given operator: with mixed = 1 and a DIA operator it includes room for the
 * Jacobi-scaled FP32 copy of the operator that the scaled inner iteration (tuning mode 3) streams;
 * a workspace of jno_pcg_workspace_size(n, 1) bytes runs the unscaled inner iteration instead. */
size_t jno_pcg_workspace_size_op(const jno_op* h_op, int mixed);
/* Default iterations enqueued per host poll of the device-side convergence state (16), for
 * operators without their own jno_tuning.  No reference counterpart: a launch-granularity knob of
 * this engine. */
int jno_pcg_set_chunk(int chunk);
/* Paced chunks for the one-kernel FP32 inner loops (default on): the first chunks of each inner
 * loop are small and later ones follow the observed convergence rate, so fewer launches are
 * enqueued past a stop.  0 = fixed chunks of jno_pcg_set_chunk.  A library default for operators
 * without their own jno_tuning; no reference counterpart (a launch-scheduling knob). */
int jno_pcg_set_pacing(int enable);
/* Default launch shape of the iteration kernels, per iteration precision, for operators without
 * their own jno_tuning (no reference counterpart): precision = 0 for the FP64 iterations of jno_pcg_f64, 1 for
 * the FP32 inner iterations of jno_pcg_mixed; grid_cap = maximum workgroups (multiple of 8,
 * <= 4096); k1_vector = 1 for 16-byte row tiles in the DIA SpMV kernel, 0 for rows per thread,
 * 2 for four rows per thread, 3 for p staged in LDS (mode 4 only; elsewhere it is 0);
 * block = threads per workgroup (256/512/1024); split = 1 for three kernels per iteration (p
 * materialized, then a plain SpMV), 0 for two (p recomputed inside the SpMV kernel), 2 for the
 * single-reduction iteration, 3, 4 and 5 for the split, the two-kernel and the one-kernel
 * iteration on the Jacobi-scaled FP32 operator (FP32 inner iterations on DIA operators with room
 * for it in the workspace; elsewhere they are modes 1, 0 and 0).  With split = 5, k1_vector 0
 * selects the LDS-staged one-kernel iteration and 1 or 2 the register-resident one (4-row quads;
 * the far neighbours loaded before / after the head reduction), which takes 4-diagonal operators
 * with 2 or 3 diagonals within 256 rows at 256 or 512 threads; any other operator given that shape
 * runs block 512, grid cap 512, k1_vector 0.  The FP32 default is split 5, k1_vector 2, block 256,
 * grid cap 768; the FP64 default split 1, k1_vector 0, block 1024, grid cap 512.
 * jno_pcg_get_tuning reads the current shape. */
int jno_pcg_set_tuning(int precision, int grid_cap, int k1_vector, int block, int split);
int jno_pcg_get_tuning(int precision, int* grid_cap, int* k1_vector, int* block, int* split);
/* The library defaults as one jno_tuning (what a jno_op with tuning = NULL runs), and a check of a
 * tuning struct (JNO_EINVAL with the reason when a field is out of range).  The defaults are
 * copied under a lock at the start of each solve, so concurrent solves never see a torn shape. */
int jno_pcg_default_tuning(jno_tuning* h_out);
int jno_pcg_check_tuning(const jno_tuning* h_tuning);
/* Live kernel timing of the PCG iteration kernels (no reference counterpart: measurement hook),
 * per calling thread: it times the solves that thread runs.  When enabled, HIP events are recorded on the solve's stream around the kernels of every 8th
 * iteration, and the durations of the sampled iterations that did work are summed into
 * h_ms[0] (K0: p update, split iterations only), h_ms[1] (the SpMV kernel: K1 = p recompute +
 * SpMV + x update + p.q, or KS = SpMV + p.q) and h_ms[2] (K2: r update, r.z, r.r); *h_count is
 * the number of sampled iterations.  One-kernel iterations (mode 5) are timed as windows of 4
 * consecutive launches between two events, summed into h_ms[1] with 4 counted per window.
 * Enabling resets the sums. */
int jno_pcg_profile(int enable);
int jno_pcg_profile_read(double* h_ms, int64_t* h_count);
int jno_pcg_f64(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x, int use_x0,
                double tol, double atol, int64_t maxiter, jno_pcg_stats* h_stats, void* d_ws,
                size_t ws_bytes, void* stream);
/* FP32 inner PCG iterations with FP64 residual replacement ("reliable updates"): x accumulates in
 * FP64, the true residual b - A x is recomputed in FP64 with the FP64 operator whenever the FP32
 * recursive residual has dropped by `replace_factor` since the last replacement, and the stopping
 * test is the true FP64 residual <= max(tol ||b||, atol).  Needs h_op->vals32. */
int jno_pcg_mixed(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x,
                  int use_x0, double tol, double atol, int64_t maxiter, double replace_factor,
                  jno_pcg_stats* h_stats, void* d_ws, size_t ws_bytes, void* stream);

/* Allocation-free jno_pcg_mixed.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * Same arguments, kernels and results as jno_pcg_mixed, without the workspace: the solve runs in a
 * static device scratch of JNO_PCG_SCRATCH_BYTES per GPU (64 MiB: mixed solves on DIA operators of
 * up to about 1.19M rows, e.g. a 96^3 Kuhn lattice, and on CSR/BSR3 operators whose mixed workspace
 * jno_pcg_workspace_size_op(op, 1) fits), held from the first call on a GPU until the process ends,
 * and reads the device state back synchronously (no pinned memory, no events).  A larger operator
 * is rejected with JNO_ECAPACITY before any work is queued.  Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (threads or streams) race, so serialize them. */
#define JNO_PCG_SCRATCH_BYTES ((size_t)64 << 20)
int jno_pcg_mixed_noalloc(const jno_op* h_op, const double* d_dinv, const double* d_b, double* d_x,
                          int use_x0, double tol, double atol, int64_t maxiter, double replace_factor,
                          jno_pcg_stats* h_stats, void* stream);

/* ---------------------------------------------------------------------------------------------
 * A20  batched solves.  Replaces jax.vmap over the cg slot (LinearSolver traits vmap="native",
 * jno/utils/solver/solver_api.py:275-280; tests/test_fem_solver_slots.py:123-132), generalised to
 * systems of different sizes and operators: every item runs its own Jacobi-PCG recurrence (FP32
 * inner iterations, FP64 residual replacement, stop on the true FP64 residual) and its own device
 * state machine, so items do not wait for each other.  Items may share one operator (many right
 * hand sides).  steps_per_graph > 0 captures that many (iteration or replacement) steps in a
 * hipGraph and replays it; 0 launches the kernels directly.  h_stats receives one entry per item.
 * ------------------------------------------------------------------------------------------- */
typedef struct jno_batch_item {
  jno_op op;           /* needs vals32 and vals64                          */
  const double* dinv;  /* Jacobi inverse diagonal (FP64)                   */
  const double* b;     /* right-hand side                                  */
  double* x;           /* solution (x0 = 0)                                */
} jno_batch_item;

size_t jno_batch_workspace_size(int n_items, const jno_batch_item* h_items);
int jno_batch_pcg_mixed(int n_items, const jno_batch_item* h_items, double tol, double atol,
                        int64_t maxiter, double replace_factor, int steps_per_graph,
                        jno_pcg_stats* h_stats, void* d_ws, size_t ws_bytes, void* stream);

/* Launch shape of a batched solve, per call (no reference counterpart: the reference's vmap has no
 * launch shape).  one_kernel = 1 runs one kernel per step (the one-kernel iteration of the single
 * solver, carried to every item) when every item is a Jacobi-scaled 4-diagonal DIA operator whose
 * second offset is at most 128 (3-D Kuhn lattices up to 127 cells along the first axis); otherwise,
 * and with one_kernel = 0, the three-kernel step runs.  rows_per_entry (one-kernel step: a
 * multiple of 2048, at most 65536) and grid (a multiple of 8, at most 65536) shape its table of
 * (item, row range) entries; 0 takes the default (4096 rows, 4096 workgroups).  ran_one_kernel is
 * an output: 1 when the call ran the one-kernel step. */
typedef struct jno_batch_tuning {
  int32_t one_kernel;
  int32_t rows_per_entry;
  int32_t grid;
  int32_t ran_one_kernel;
} jno_batch_tuning;
/* jno_batch_pcg_mixed with a launch shape (h_tuning NULL: the defaults, as jno_batch_pcg_mixed).
 * Same workspace size. */
int jno_batch_pcg_mixed_tuned(int n_items, const jno_batch_item* h_items, double tol, double atol,
                              int64_t maxiter, double replace_factor, int steps_per_graph,
                              jno_batch_tuning* h_tuning, jno_pcg_stats* h_stats, void* d_ws,
                              size_t ws_bytes, void* stream);

/* ---------------------------------------------------------------------------------------------
 * A18  adjoint contraction.  Replaces the VJP chain of custom_linear_solve(symmetric=True) ->
 * BCOO data -> Dirichlet mask -> segment_sum -> jacfwd (jno/utils/solver/fem_native.py:2326-2345,
 * jno/utils/solver/fem_utils.py:3332-3336):  d_grad[e] = -|T_e| (G_e P lam_e).(G_e u_e) where P zeroes
 * the Dirichlet dofs (d_dmask, nullable).  Caller solves A lam = dL/du with jno_pcg_* first.
 * ------------------------------------------------------------------------------------------- */
int jno_adjoint_grad_k_p1(const double* d_points, const int32_t* d_cells, int64_t n_cells,
                          const double* d_u, const double* d_lam, const uint8_t* d_dmask,
                          double* d_grad, void* stream);

/* ---------------------------------------------------------------------------------------------
 * A2 P2 promotion.  Replaces _promote_to_degree (jno/utils/solver/fem_utils.py:1979-2044) with the
 * basix P2 edge order (jno/utils/solver/fem_topology.py:44): vertices keep ids 0..N-1, edge nodes
 * get N.. in cell-major / local-dof-minor first-appearance order.  Needs the P1 plan's incidence.
 * d_cells10: i32[T*10]; d_points2: f64 capacity (N + 6T)*3 (only the first *h_n_nodes2 rows are
 * written).
 * ------------------------------------------------------------------------------------------- */
size_t jno_promote_p2_workspace_size(int64_t n_cells);
int jno_promote_p2(const double* d_points, const int32_t* d_cells, int64_t n_cells, int64_t n_nodes,
                   const int32_t* d_inc_ptr, const int32_t* d_inc, int32_t* d_cells10,
                   double* d_points2, int64_t* h_n_nodes2, void* d_ws, size_t ws_bytes, void* stream);

/* Element coloring for the atomic-free scatter (no reference counterpart: the reference scatters
 * with XLA's segment_sum, jno/utils/solver/fem_native.py:2326).  Cells sharing a vertex get
 * different colors; d_order (T) lists cells grouped by color, h_color_ptr (max_colors+1) the
 * group bounds.  The incidence is the plan's (entries e*nloc+a). */
size_t jno_color_workspace_size(int64_t n_cells);
int jno_color_cells(const int32_t* d_cells, int64_t n_cells, int nloc, const int32_t* d_inc_ptr,
                    const int32_t* d_inc, int32_t* d_order, int32_t* h_color_ptr, int32_t max_colors,
                    int32_t* h_ncolors, void* d_ws, size_t ws_bytes, void* stream);

/* A4-A6, A8 for P2 vector (3 dof/node) linear elasticity: the element Jacobian of
 * E_e (lam tr eps(u) tr eps(v) + 2 mu eps(u):eps(v)) (weak form of
 * tests/test_topology_optimisation.py:59-66; jacfwd element kernel fem_native.py:2331-2345)
 * contracted on MFMA (FP64 16x16x4) and scattered color by color into the CSR values d_vals (nnz)
 * of the vec=3 plan.  lam, mu are the Lame parameters of E = 1 and Poisson ratio nu.  d_rank
 * (nullable): the pair table of jno_p2_pair_rank for this plan, which replaces the scatter's binary
 * searches (the values are the same either way). */
int jno_assemble_p2_elasticity(const double* d_points2, const int32_t* d_cells10, int64_t n_cells,
                               const int32_t* d_rowptr, const int32_t* d_col, int64_t nnz,
                               const double* d_E, double nu, const int32_t* d_order,
                               const int32_t* h_color_ptr, int ncolors, const uint8_t* d_rank,
                               double* d_vals, void* stream);
/* The place of each (cell, local node a, local node b) pair in node a's row of the vec = 3 plan
 * (d_rank u8[100 n_cells] at 100 e + 10 a + b: row 3 n_a + i holds node n_b's columns at
 * rowptr[3 n_a + i] + 3 place + j), found once per plan (the scatter in compress_plan's inverse map,
 * jno/utils/solver/fem_utils.py:4009-4060, kept per cell).  *h_ok = 0 when a row holds more than 255
 * nodes or a pair is missing (then the assembly searches).  d_ws: 256 bytes.  Synchronizes. */
int jno_p2_pair_rank(const int32_t* d_cells10, int64_t n_cells, const int32_t* d_rowptr, const int32_t* d_col,
                     uint8_t* d_rank, int32_t* h_ok, void* d_ws, size_t ws_bytes, void* stream);
/* A6, A8 (and A11 with d_dmask) for P2 elasticity when the cells repeat their shape: the CSR values
 * from the geometry-class table of jno_ebe_class_blocks (d_table64, blocks of E = 1 in jno_ebe's
 * table order) with d_cell_class[e] the class of cell e, d_E per cell and the pair table of
 * jno_p2_pair_rank: each node's three rows summed over its incidences (plan d_inc_ptr / d_inc,
 * entries e*10 + a; d_cells10 the cells) in ascending cell order in LDS and written once -- the
 * segment_sum of the triplets (jno/utils/solver/fem_native.py:2318-2329) as a row pull.  d_dmask
 * (u8[3N], nullable): the symmetric elimination with zero values fused, exactly as
 * jno_dirichlet_eliminate applies it to the matrix (_apply_dirichlet_symmetric, fem_1d.py:667-686);
 * the caller sets b[d] = 0.  max_row (values per row) at most 384. */
int jno_assemble_p2_classes(int64_t n_nodes, const int32_t* d_inc_ptr, const int32_t* d_inc,
                            const int32_t* d_cells10, const int32_t* d_cell_class, const double* d_table64,
                            const double* d_E, const uint8_t* d_rank, const uint8_t* d_dmask,
                            const int32_t* d_rowptr, int max_row, double* d_vals, void* stream);

/* A9 surface load: constant traction h_t[3] on the boundary facets lying on x[axis] == value,
 * int_F t phi = |F|/3 t on the facet's edge nodes (_surf_elem_res, fem_native.py:1882-1960).
 * Adds into d_b (3 dofs per node, node-major). */
int jno_traction_p2(const double* d_points2, const int32_t* d_cells10, int64_t n_cells,
                    const int32_t* d_inc_ptr, const int32_t* d_inc, int axis, double value,
                    const double* h_t, double* d_b, void* stream);

/* A18 for the compliance C = b.u (self-adjoint, lam = u): dC/dE_e = -u_e^T (K_e/E_e) u_e. */
int jno_compliance_grad_p2(const double* d_points2, const int32_t* d_cells10, int64_t n_cells,
                           double nu, const double* d_u, double* d_grad, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Theta-method transient step of the linear block M u' + A u = c + f(t) on one CSR pattern.
 * Replaces SemidiscreteTimeBlock.step's linear branch (jno/utils/solver/backend_blocks.py:329-351)
 * with the static step operator of compose_transient_step_solvers / _add_step_operator
 * (jno/utils/solver/solver_api.py:856-947) and the Dirichlet edit of _apply_dirichlet_transient
 * (jno/utils/solver/fem_1d.py:939-949).
 *
 * jno_theta_operators: d_mass, d_stiff (nnz, same pattern; d_stiff already symmetrically
 * eliminated) -> d_step = Mz + theta dt A, d_expl = Mz - (1-theta) dt A (nullable), d_step_dinv =
 * 1/diag(d_step) with the |d| > 1e-30 guard, where Mz is d_mass with the rows and columns of the
 * d_dmask dofs zeroed (d_dmask nullable: no constrained dofs).  dt > 0, 0 < theta <= 1.
 * ------------------------------------------------------------------------------------------- */
int jno_theta_operators(int64_t n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                        const double* d_mass, const double* d_stiff, const uint8_t* d_dmask, double dt,
                        double theta, double* d_step, double* d_expl, double* d_step_dinv, void* stream);
/* The same operators on the symmetric DIA layout (the time-dependent step on a structured box, where
 * the operator's assembly writes its DIA layout and the mass's is formed once): d_mass_dia holds
 * ndiag diagonals at h_offsets (main first), d_stiff_dia ndiag_stiff diagonals at h_offsets_stiff,
 * each one of the mass's offsets; both with leading dimension n_rows rounded up to 64.  Writes K
 * (d_step_dia, FP64, and d_step_dia32, FP32, nullable), E (d_expl_dia, nullable) on the mass's
 * diagonals and 1/diag K, with the values jno_theta_operators gives bitwise.  No reference
 * counterpart as a separate entry: the per-step operator of compose_transient_step_solvers
 * (jno/utils/solver/solver_api.py:923-935) on the diagonal layout. */
int jno_theta_operators_dia(int64_t n_rows, int ndiag, const int32_t* h_offsets, const double* d_mass_dia,
                            int ndiag_stiff, const int32_t* h_offsets_stiff, const double* d_stiff_dia,
                            const uint8_t* d_dmask, double dt, double theta, double* d_step_dia,
                            float* d_step_dia32, double* d_expl_dia, double* d_step_dinv, void* stream);

/* Right-hand side of one step: d_rhs = E u + dt c + dt (theta f(t+dt) + (1-theta) f(t))
 * (backend_blocks.py:342-344), E given as an operator layout (jno_op, FP64 values: symmetric DIA
 * or CSR) built from the d_expl values.  d_c, and the pair d_f_now/d_f_next, are nullable (terms
 * dropped).  d_rhs must not alias d_u.  The same call with d_c and the forcing null applies E
 * alone (the transposed step of the adjoint march; E is symmetric). */
int jno_theta_rhs(const jno_op* h_expl, const double* d_u, const double* d_c, const double* d_f_now,
                  const double* d_f_next, double dt, double theta, double* d_rhs, void* stream);

/* Adjoint of one step w.r.t. the element coefficient: d_w = theta u_next + (1-theta) u_now on free
 * dofs and g on the d_dmask dofs (0 when d_g is null); dt * jno_adjoint_grad_k_p1(w, mu) is then the
 * step's contribution to dL/dk for the step adjoint mu = K^-1 lam_next (no reference counterpart
 * as a kernel: the reference differentiates the march with jax.lax.scan's reverse mode,
 * backend_blocks.py:475-530). */
int jno_theta_mix(int64_t n, const double* d_u_next, const double* d_u_now, const uint8_t* d_dmask,
                  const double* d_g, double theta, double* d_w, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* JNO_FEM_H */
