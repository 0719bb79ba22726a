// A1 Kuhn lattice mesher and A10 boundary-node masks.
#include "common.h"

namespace jno {

// Six Kuhn tets per voxel; corner c = (di,dj,dk) = (c>>2&1, c>>1&1, c&1).
// Same table as Geometries.equi_distant_box (jno/domain/geometries.py:245).
__constant__ int8_t c_kuhn[6][4] = {{0, 4, 6, 7}, {0, 4, 5, 7}, {0, 2, 6, 7},
                                    {0, 2, 3, 7}, {0, 1, 5, 7}, {0, 1, 3, 7}};

__global__ void kuhn_points(int nx, int ny, int nz, double x0, double y0, double z0, double x1,
                            double y1, double z1, double* __restrict__ pts) {
  const int64_t N = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < N;
       id += (int64_t)gridDim.x * blockDim.x) {
    int k = (int)(id % (nz + 1));
    int64_t t = id / (nz + 1);
    int j = (int)(t % (ny + 1));
    int i = (int)(t / (ny + 1));
    pts[3 * id + 0] = linspace_at(x0, x1, nx, i);
    pts[3 * id + 1] = linspace_at(y0, y1, ny, j);
    pts[3 * id + 2] = linspace_at(z0, z1, nz, k);
  }
}

__global__ void kuhn_cells(int nx, int ny, int nz, int4* __restrict__ cells) {
  const int64_t T = (int64_t)6 * nx * ny * nz;
  const int64_t ny1 = ny + 1, nz1 = nz + 1;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < T;
       t += (int64_t)gridDim.x * blockDim.x) {
    int typ = (int)(t % 6);
    int64_t v = t / 6;
    int k = (int)(v % nz);
    int64_t w = v / nz;
    int j = (int)(w % ny);
    int i = (int)(w / ny);
    int c[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      int cc = c_kuhn[typ][a];
      c[a] = (int)(((i + ((cc >> 2) & 1)) * ny1 + (j + ((cc >> 1) & 1))) * nz1 + (k + (cc & 1)));
    }
    cells[t] = make_int4(c[0], c[1], c[2], c[3]);
  }
}

// Local facets of a tet (vertex triples) and, for P2, the local dofs of the three edges on each
// facet (basix edge order (2,3),(1,3),(1,2),(0,3),(0,2),(0,1) -> dofs 4..9).
__constant__ int8_t c_faces[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
__constant__ int8_t c_face_edges[4][3] = {{6, 8, 9}, {5, 7, 9}, {4, 7, 8}, {4, 5, 6}};

__global__ void fill_u8(uint8_t* __restrict__ p, int64_t n, uint8_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// One thread per (cell, facet): the facet is on the boundary iff no other cell incident to its
// first vertex holds all three of its vertices.
__global__ void boundary_facets_mark(const int32_t* __restrict__ cells, int64_t T, int nloc,
                                     const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc,
                                     uint8_t* __restrict__ mask) {
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < 4 * T;
       w += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = w >> 2;
    const int f = (int)(w & 3);
    const int32_t* ce = cells + e * nloc;
    const int a = ce[c_faces[f][0]], b = ce[c_faces[f][1]], c = ce[c_faces[f][2]];
    bool shared = false;
    for (int s = inc_ptr[a]; s < inc_ptr[a + 1] && !shared; ++s) {
      const int64_t e2 = inc[s] / nloc;
      if (e2 == e) continue;
      const int32_t* c2 = cells + e2 * nloc;
      bool hb = false, hc = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        hb |= (c2[q] == b);
        hc |= (c2[q] == c);
      }
      shared = hb && hc;
    }
    if (!shared) {
      mask[a] = 1;
      mask[b] = 1;
      mask[c] = 1;
      if (nloc == 10) {
#pragma unroll
        for (int q = 0; q < 3; ++q) mask[ce[c_face_edges[f][q]]] = 1;
      }
    }
  }
}

__global__ void mask_plane_k(const double* __restrict__ pts, int64_t n, int axis, double value,
                             const uint8_t* __restrict__ min, uint8_t* __restrict__ mout) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    mout[i] = (min[i] && pts[3 * i + axis] == value) ? 1 : 0;
}

}  // namespace jno

using namespace jno;

extern "C" int jno_mesh_kuhn_box(int nx, int ny, int nz, const double* h_lo, const double* h_hi,
                                 double* d_points, int32_t* d_cells, void* stream) {
  JNO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, JNO_EINVAL, "jno_mesh_kuhn_box: cell counts must be >= 1");
  const int64_t N = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  const int64_t T = (int64_t)6 * nx * ny * nz;
  JNO_REQUIRE(N < INT32_MAX && T * 4 < INT32_MAX, JNO_EINVAL, "jno_mesh_kuhn_box: mesh exceeds int32 ids");
  JNO_REQUIRE(d_points && d_cells && h_lo && h_hi, JNO_EINVAL, "jno_mesh_kuhn_box: null pointer");
  hipStream_t s = as_stream(stream);
  kuhn_points<<<grid_for(N, 256), 256, 0, s>>>(nx, ny, nz, h_lo[0], h_lo[1], h_lo[2], h_hi[0], h_hi[1],
                                                h_hi[2], d_points);
  JNO_LAUNCH_CHECK();
  kuhn_cells<<<grid_for(T, 256), 256, 0, s>>>(nx, ny, nz, reinterpret_cast<int4*>(d_cells));
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_boundary_mask(const int32_t* d_cells, int64_t n_cells, int nloc, int64_t n_nodes,
                                 const int32_t* d_inc_ptr, const int32_t* d_inc, uint8_t* d_mask,
                                 void* stream) {
  JNO_REQUIRE(nloc == 4 || nloc == 10, JNO_EINVAL, "jno_boundary_mask: nloc must be 4 (P1) or 10 (P2)");
  hipStream_t s = as_stream(stream);
  fill_u8<<<grid_for(n_nodes, 256), 256, 0, s>>>(d_mask, n_nodes, 0);
  JNO_LAUNCH_CHECK();
  if (n_cells == 0) return JNO_OK;
  boundary_facets_mark<<<grid_for(4 * n_cells, 256), 256, 0, s>>>(d_cells, n_cells, nloc, d_inc_ptr, d_inc,
                                                                    d_mask);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}

extern "C" int jno_mask_plane(const double* d_points, int64_t n_nodes, int axis, double value,
                              const uint8_t* d_mask_in, uint8_t* d_mask_out, void* stream) {
  JNO_REQUIRE(axis >= 0 && axis < 3, JNO_EINVAL, "jno_mask_plane: axis must be 0, 1 or 2");
  mask_plane_k<<<grid_for(n_nodes, 256), 256, 0, as_stream(stream)>>>(d_points, n_nodes, axis, value, d_mask_in,
                                                                      d_mask_out);
  JNO_LAUNCH_CHECK();
  return JNO_OK;
}
