// The structured Kuhn lattice (the mesh of Shape.box(..).structured(n)): node (i, j, k) has id
// (i (ny+1) + j)(nz+1) + k, voxel (i, j, k) id (i ny + j) nz + k, and the voxel's six tets (ids
// 6 voxel + t) join corners 0 and 7 along the six monotone paths, corner c = (c>>2&1, c>>1&1, c&1)
// (Geometries.equi_distant_box, jno/domain/geometries.py:244-258).  Compile-time tables of the
// lattice shared by the structured assembly and the structured plan.
#pragma once
#include <type_traits>

namespace jno {

__host__ __device__ constexpr int kuhn_corner(int t, int a) {
  constexpr int K[6][4] = {{0, 4, 6, 7}, {0, 4, 5, 7}, {0, 2, 6, 7}, {0, 2, 3, 7}, {0, 1, 5, 7}, {0, 1, 3, 7}};
  return K[t][a];
}

// stencil slot of the offset (di, dj, dk): its rank among the 15 in lexicographic order
__host__ __device__ constexpr int stencil_slot(int di, int dj, int dk) {
  int r = 0;
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
      for (int c = -1; c <= 1; ++c) {
        const bool pos = a >= 0 && b >= 0 && c >= 0, neg = a <= 0 && b <= 0 && c <= 0;
        if (!pos && !neg) continue;
        if (a == di && b == dj && c == dk) return r;
        ++r;
      }
  return -1;
}
__host__ __device__ constexpr int slot_di(int s) {
  int r = 0;
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
      for (int c = -1; c <= 1; ++c) {
        const bool pos = a >= 0 && b >= 0 && c >= 0, neg = a <= 0 && b <= 0 && c <= 0;
        if (!pos && !neg) continue;
        if (r++ == s) return a * 9 + b * 3 + c;  // packed (a, b, c)
      }
  return 0;
}

// q-th incident (voxel corner c, type t, local vertex a) of a node, in ascending element order:
// voxel (i - di, j - dj, k - dk) of corner c = (di, dj, dk) has the larger id the smaller c is, so
// corners descend from 7, and types ascend within a voxel
struct BoxInc {
  int c, t, a;
};
__host__ __device__ constexpr BoxInc box_inc(int q) {
  int n = 0;
  for (int c = 7; c >= 0; --c)
    for (int t = 0; t < 6; ++t)
      for (int a = 0; a < 4; ++a)
        if (kuhn_corner(t, a) == c) {
          if (n == q) return BoxInc{c, t, a};
          ++n;
        }
  return BoxInc{-1, -1, -1};
}

// f(std::integral_constant<int, B>), ..., f(std::integral_constant<int, E - 1>): every index a
// compile-time constant inside f
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// (di, dj, dk) of stencil slot S
template <int S>
struct SlotOffset {
  static constexpr int pk = slot_di(S);
  static constexpr int di = (pk + 13) / 9 - 1, dj = ((pk + 13) / 3) % 3 - 1, dk = (pk + 13) % 3 - 1;
};

template <int Q, int BB>
struct IncTerm {  // vertex BB of incidence Q's element, seen from the row's node
  static constexpr BoxInc L = box_inc(Q);
  static constexpr int cb = kuhn_corner(L.t, BB);
  static constexpr int di = ((cb >> 2) & 1) - ((L.c >> 2) & 1), dj = ((cb >> 1) & 1) - ((L.c >> 1) & 1),
                       dk = (cb & 1) - (L.c & 1);
  static constexpr int slot = stencil_slot(di, dj, dk);
};


}  // namespace jno
