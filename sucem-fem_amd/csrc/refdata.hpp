// Reference-element data for first- and second-order Nedelec (first kind) tetrahedra.
//
// The nodal basis is the FIAT convention the reference's DOLFIN function space
// uses (ProblemConfigurations/EMProblem.py:100-105, element "Nedelec 1st kind
// H(curl)"): on the reference tet (0,0,0),(1,0,0),(0,1,0),(0,0,1) the dual is
// tangential point evaluation with unnormalised tangents -- per local edge
// (a,b) in UFC order (2,3),(1,3),(1,2),(0,3),(0,2),(0,1) at s = 1/2 (order 1) or
// s = 1/3, 2/3 (order 2) along v_a -> v_b, then (order 2) per local face
// (p,q,r) at the centroid with tangents v_q - v_p and v_r - v_p.
//
// Element tensors are split into six symmetric geometric components
// k = (0,0),(1,1),(2,2),(0,1),(0,2),(1,2) so that for a cell with Jacobian J
//   T_e = eps_r |det J|      sum_k (J^-1 J^-T)_k  M_k
//   S_e = (1/mu_r)/|det J|   sum_k (J^T J)_k      C_k
// and for its local face f
//   S0_e = (area_phys/area_ref_f) sum_k (J^-1 (I - n n^T) J^-T)_k  F_{f,k}
// with off-diagonal tensors holding R^{ab} + R^{ba}.
#pragma once
#include <vector>

namespace sucem {

constexpr int kMono = 10;   // 1, x, y, z, x^2, xy, xz, y^2, yz, z^2

struct RefData {
  int order = 0;
  int nloc = 0;        // 6 or 20
  int nfd = 0;         // DOFs with a tangential trace on one face: 3 or 8
  // basis[i][d][m] and curl[i][d][m]: reference basis as polynomials
  std::vector<double> basis, curl;
  // mass[k][i][j], stiff[k][i][j]  (6 * nloc * nloc)
  std::vector<double> mass, stiff;
  // face[f][k][a][b] over the face DOF list (4 * 6 * nfd * nfd)
  std::vector<double> face;
  // face_dofs[f][a]: local DOF index of face-DOF a of local face f (4 * nfd)
  std::vector<int> face_dofs;
  double ref_face_area[4];
  // Discrete gradient: grad[i][j] = nodal dual i applied to grad(phi_j) for the
  // Lagrange P_order basis phi_j (vertices, then (order 2) edge midpoints in UFC
  // edge order); the same on every cell (nloc x nlag).
  int nlag = 0;
  std::vector<double> grad;
  // Nodal duals: dual_lam[i][v] = barycentric coordinate v of dual point i, dual_t[i][k]
  // = its reference tangent (the physical tangent is J dual_t[i], J = [x1-x0 x2-x0 x3-x0]).
  std::vector<double> dual_lam, dual_t;
};

const RefData& ref_data(int order);

}  // namespace sucem
