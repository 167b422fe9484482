// Device geometry helpers shared by the assembly and post-processing kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sucem {

// Affine map of a tetrahedron: J has columns x1-x0, x2-x0, x3-x0; Ji = J^-1.
__device__ inline void cell_jacobian(const double* xyz, const int32_t* cell, double J[3][3],
                                     double& det, double Ji[3][3]) {
  const double* x0 = xyz + 3 * (int64_t)cell[0];
  for (int col = 0; col < 3; ++col) {
    const double* xk = xyz + 3 * (int64_t)cell[col + 1];
    for (int r = 0; r < 3; ++r) J[r][col] = xk[r] - x0[r];
  }
  double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  double id = 1.0 / det;
  Ji[0][0] = c00 * id;
  Ji[1][0] = c01 * id;
  Ji[2][0] = c02 * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
}

}  // namespace sucem
