// shape.h -- Gauss-point shape tables of the element kernels, one copy per
// translation unit that includes it (each device module holds its own
// constant memory); upload_shape_tables fills this unit's copy.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace dnlg {

// Element shape tables at the Gauss points: the closed forms tabulated in
// ElQua4NAx.C:63-110 and ElHex8N3D.C:93-231, generated on the host
// (api.hip, make_tables) and uploaded once per device.
static __constant__ double c_hexN[8][8];
static __constant__ double c_hexD[8][8][3];
static __constant__ double c_quaN[4][4];
static __constant__ double c_quaD[4][4][2];

template <int NN>
__device__ __forceinline__ double tabN(int p, int a) {
  if constexpr (NN == 8) return c_hexN[p][a];
  else return c_quaN[p][a];
}
template <int NN>
__device__ __forceinline__ double tabD(int p, int a, int j) {
  if constexpr (NN == 8) return c_hexD[p][a][j];
  else return c_quaD[p][a][j];
}

// Shape functions, their local derivatives and the weight of integration
// point p: the Gauss tables of the hex and quads (weights 1), or with TET
// the 4-node tetrahedron's single point (ElTet4N3D.C:62-71: the centroid,
// weight 1/6, N = 1/4, derivatives -1 at node 0 and the unit vectors at
// nodes 1-3).
template <int NN, bool TET>
__device__ __forceinline__ double shapeN(int p, int a) {
  if constexpr (TET) return 0.25;
  else return tabN<NN>(p, a);
}
template <int NN, bool TET>
__device__ __forceinline__ double shapeD(int p, int a, int k) {
  if constexpr (TET) return a == 0 ? -1.0 : (a - 1 == k ? 1.0 : 0.0);
  else return tabD<NN>(p, a, k);
}
template <bool TET>
__device__ __forceinline__ double ip_weight() {
  if constexpr (TET) return 1.0 / 6;
  else return 1.0;
}

// Local coordinate sign (+-1) of node a along direction k in the reference's
// numbering (ElHex8N3D.C:30-92, ElQua4NAx.C:30-62): hex x + for nodes 1, 2,
// 5, 6, y + for 2, 3, 6, 7, z + for 4-7; quad x + for 1, 2, y + for 2, 3.
template <int NN>
__host__ __device__ __forceinline__ double node_sign(int a, int k) {
  if (k == 0) return (a == 1 || a == 2 || a == 5 || a == 6) ? 1.0 : -1.0;
  if (k == 1) return (a == 2 || a == 3 || a == 6 || a == 7) ? 1.0 : -1.0;
  return a >= 4 ? 1.0 : -1.0;
}

static inline void upload_shape_tables(const double *hexN, const double *hexD, const double *quaN,
                                       const double *quaD) {
  hipMemcpyToSymbol(HIP_SYMBOL(c_hexN), hexN, sizeof(double) * 64);
  hipMemcpyToSymbol(HIP_SYMBOL(c_hexD), hexD, sizeof(double) * 192);
  hipMemcpyToSymbol(HIP_SYMBOL(c_quaN), quaN, sizeof(double) * 16);
  hipMemcpyToSymbol(HIP_SYMBOL(c_quaD), quaD, sizeof(double) * 32);
}

}  // namespace dnlg
