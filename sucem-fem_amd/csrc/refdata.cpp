// Construction of the reference-element tensors (host, once per order).
#include "refdata.hpp"

#include <array>
#include <cmath>
#include <mutex>
#include <stdexcept>

namespace sucem {
namespace {

using Poly = std::array<double, kMono>;          // scalar polynomial, degree <= 2
using VPoly = std::array<Poly, 3>;               // vector polynomial

const int kExp[kMono][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                            {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};
const int kEdges[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
const int kFaces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
const double kVert[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
const int kSym[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};

int mono_index(int a, int b, int c) {
  for (int m = 0; m < kMono; ++m)
    if (kExp[m][0] == a && kExp[m][1] == b && kExp[m][2] == c) return m;
  throw std::logic_error("monomial degree too high");
}

Poly times_coord(const Poly& p, int axis) {
  Poly o{};
  for (int m = 0; m < kMono; ++m) {
    if (p[m] == 0) continue;
    int e[3] = {kExp[m][0], kExp[m][1], kExp[m][2]};
    e[axis]++;
    o[mono_index(e[0], e[1], e[2])] += p[m];
  }
  return o;
}

Poly deriv(const Poly& p, int axis) {
  Poly o{};
  for (int m = 0; m < kMono; ++m) {
    if (p[m] == 0 || kExp[m][axis] == 0) continue;
    int e[3] = {kExp[m][0], kExp[m][1], kExp[m][2]};
    double f = e[axis];
    e[axis]--;
    o[mono_index(e[0], e[1], e[2])] += f * p[m];
  }
  return o;
}

Poly mul(const Poly& a, const Poly& b) {
  Poly o{};
  for (int i = 0; i < kMono; ++i) {
    if (a[i] == 0) continue;
    for (int j = 0; j < kMono; ++j) {
      if (b[j] == 0) continue;
      o[mono_index(kExp[i][0] + kExp[j][0], kExp[i][1] + kExp[j][1], kExp[i][2] + kExp[j][2])] +=
          a[i] * b[j];
    }
  }
  return o;
}

Poly sub(const Poly& a, const Poly& b) {
  Poly o;
  for (int m = 0; m < kMono; ++m) o[m] = a[m] - b[m];
  return o;
}

VPoly cross_x(const VPoly& v) {  // x cross v
  return {sub(times_coord(v[2], 1), times_coord(v[1], 2)),
          sub(times_coord(v[0], 2), times_coord(v[2], 0)),
          sub(times_coord(v[1], 0), times_coord(v[0], 1))};
}

VPoly curl_of(const VPoly& v) {
  return {sub(deriv(v[2], 1), deriv(v[1], 2)), sub(deriv(v[0], 2), deriv(v[2], 0)),
          sub(deriv(v[1], 0), deriv(v[0], 1))};
}

double eval(const Poly& p, const double x[3]) {
  double s = 0;
  for (int m = 0; m < kMono; ++m)
    s += p[m] * std::pow(x[0], kExp[m][0]) * std::pow(x[1], kExp[m][1]) * std::pow(x[2], kExp[m][2]);
  return s;
}

double factorial(int n) {
  double f = 1;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

// Exact integral of x^a y^b z^c over the reference tetrahedron.
double tet_monomial_integral(int a, int b, int c) {
  return factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3);
}

double tet_product_integral(const Poly& p, const Poly& q) {
  double s = 0;
  for (int i = 0; i < kMono; ++i) {
    if (p[i] == 0) continue;
    for (int j = 0; j < kMono; ++j) {
      if (q[j] == 0) continue;
      s += p[i] * q[j] *
           tet_monomial_integral(kExp[i][0] + kExp[j][0], kExp[i][1] + kExp[j][1],
                                 kExp[i][2] + kExp[j][2]);
    }
  }
  return s;
}

// Solve V C = I by Gauss-Jordan with partial pivoting (n <= 20).
std::vector<double> invert(std::vector<double> V, int n) {
  std::vector<double> C(n * n, 0.0);
  for (int i = 0; i < n; ++i) C[i * n + i] = 1.0;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(V[r * n + col]) > std::fabs(V[piv * n + col])) piv = r;
    if (std::fabs(V[piv * n + col]) < 1e-14) throw std::logic_error("singular dual matrix");
    for (int k = 0; k < n; ++k) {
      std::swap(V[col * n + k], V[piv * n + k]);
      std::swap(C[col * n + k], C[piv * n + k]);
    }
    double d = V[col * n + col];
    for (int k = 0; k < n; ++k) { V[col * n + k] /= d; C[col * n + k] /= d; }
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      double f = V[r * n + col];
      if (f == 0) continue;
      for (int k = 0; k < n; ++k) { V[r * n + k] -= f * V[col * n + k]; C[r * n + k] -= f * C[col * n + k]; }
    }
  }
  return C;
}

RefData build(int order) {
  RefData R;
  R.order = order;
  R.nloc = order == 1 ? 6 : 20;
  R.nfd = order == 1 ? 3 : 8;
  const int n = R.nloc;

  // Primal polynomial basis of the Nedelec space.
  std::vector<VPoly> prim;
  auto unit_vec = [](int comp, int mono) {
    VPoly v{};
    v[comp][mono] = 1.0;
    return v;
  };
  if (order == 1) {
    for (int a = 0; a < 3; ++a) prim.push_back(unit_vec(a, 0));
    for (int a = 0; a < 3; ++a) prim.push_back(cross_x(unit_vec(a, 0)));
  } else {
    for (int a = 0; a < 3; ++a)
      for (int m = 0; m < 4; ++m) prim.push_back(unit_vec(a, m));
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        if (a == 2 && b == 2) continue;  // x cross (x e_x + y e_y + z e_z) = 0
        prim.push_back(cross_x(unit_vec(a, 1 + b)));
      }
  }

  // Dual functionals: (point, tangent).
  struct Dual { double p[3], t[3]; };
  std::vector<Dual> duals;
  for (int e = 0; e < 6; ++e) {
    const double* va = kVert[kEdges[e][0]];
    const double* vb = kVert[kEdges[e][1]];
    std::vector<double> svals = order == 1 ? std::vector<double>{0.5}
                                           : std::vector<double>{1.0 / 3.0, 2.0 / 3.0};
    for (double s : svals) {
      Dual d;
      for (int k = 0; k < 3; ++k) { d.t[k] = vb[k] - va[k]; d.p[k] = va[k] + s * d.t[k]; }
      duals.push_back(d);
    }
  }
  if (order == 2) {
    for (int f = 0; f < 4; ++f) {
      const double* vp = kVert[kFaces[f][0]];
      const double* vq = kVert[kFaces[f][1]];
      const double* vr = kVert[kFaces[f][2]];
      for (int which = 0; which < 2; ++which) {
        Dual d;
        for (int k = 0; k < 3; ++k) {
          d.p[k] = (vp[k] + vq[k] + vr[k]) / 3.0;
          d.t[k] = (which == 0 ? vq[k] : vr[k]) - vp[k];
        }
        duals.push_back(d);
      }
    }
  }
  R.dual_lam.assign(n * 4, 0.0);
  R.dual_t.assign(n * 3, 0.0);
  for (int i = 0; i < n; ++i) {
    const double* p = duals[i].p;
    R.dual_lam[4 * i] = 1.0 - p[0] - p[1] - p[2];
    for (int k = 0; k < 3; ++k) {
      R.dual_lam[4 * i + 1 + k] = p[k];
      R.dual_t[3 * i + k] = duals[i].t[k];
    }
    for (int v = 0; v < 4; ++v)
      if (std::fabs(R.dual_lam[4 * i + v]) < 1e-14) R.dual_lam[4 * i + v] = 0.0;
  }
  std::vector<double> V(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += eval(prim[j][k], duals[i].p) * duals[i].t[k];
      V[i * n + j] = s;
    }
  std::vector<double> C = invert(V, n);   // nodal_j = sum_k prim_k C[k][j]
  std::vector<VPoly> basis(n), curls(n);
  for (int j = 0; j < n; ++j) {
    VPoly b{};
    for (int k = 0; k < n; ++k)
      for (int d = 0; d < 3; ++d)
        for (int m = 0; m < kMono; ++m) b[d][m] += prim[k][d][m] * C[k * n + j];
    for (int d = 0; d < 3; ++d)
      for (int m = 0; m < kMono; ++m)
        if (std::fabs(b[d][m]) < 1e-13) b[d][m] = 0.0;
    basis[j] = b;
    curls[j] = curl_of(b);
  }
  R.basis.assign(n * 3 * kMono, 0.0);
  R.curl.assign(n * 3 * kMono, 0.0);
  for (int j = 0; j < n; ++j)
    for (int d = 0; d < 3; ++d)
      for (int m = 0; m < kMono; ++m) {
        R.basis[(j * 3 + d) * kMono + m] = basis[j][d][m];
        R.curl[(j * 3 + d) * kMono + m] = curls[j][d][m];
      }

  // Volume tensors (exact monomial integration).
  R.mass.assign(6 * n * n, 0.0);
  R.stiff.assign(6 * n * n, 0.0);
  for (int k = 0; k < 6; ++k) {
    int a = kSym[k][0], b = kSym[k][1];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double m = tet_product_integral(basis[i][a], basis[j][b]);
        double c = tet_product_integral(curls[i][a], curls[j][b]);
        if (a != b) {
          m += tet_product_integral(basis[i][b], basis[j][a]);
          c += tet_product_integral(curls[i][b], curls[j][a]);
        }
        R.mass[(k * n + i) * n + j] = m;
        R.stiff[(k * n + i) * n + j] = c;
      }
  }

  // Face DOF lists and face tensors (collapsed Gauss-Legendre, 4x4, exact to degree 5).
  R.face_dofs.assign(4 * R.nfd, 0);
  for (int f = 0; f < 4; ++f) {
    int cnt = 0;
    for (int e = 0; e < 6; ++e) {
      bool in_a = false, in_b = false;
      for (int q = 0; q < 3; ++q) {
        in_a |= kFaces[f][q] == kEdges[e][0];
        in_b |= kFaces[f][q] == kEdges[e][1];
      }
      if (in_a && in_b) {
        if (order == 1) R.face_dofs[f * R.nfd + cnt++] = e;
        else { R.face_dofs[f * R.nfd + cnt++] = 2 * e; R.face_dofs[f * R.nfd + cnt++] = 2 * e + 1; }
      }
    }
    if (order == 2) {
      R.face_dofs[f * R.nfd + cnt++] = 12 + 2 * f;
      R.face_dofs[f * R.nfd + cnt++] = 13 + 2 * f;
    }
    if (cnt != R.nfd) throw std::logic_error("face dof count");
  }
  const double gx[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563,
                        0.8611363115940526};
  const double gw[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461,
                        0.3478548451374538};
  R.face.assign(4 * 6 * R.nfd * R.nfd, 0.0);
  for (int f = 0; f < 4; ++f) {
    const double* P = kVert[kFaces[f][0]];
    const double* Q = kVert[kFaces[f][1]];
    const double* Rv = kVert[kFaces[f][2]];
    double e1[3], e2[3], cr[3];
    for (int k = 0; k < 3; ++k) { e1[k] = Q[k] - P[k]; e2[k] = Rv[k] - P[k]; }
    cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
    cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
    cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double area2 = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    R.ref_face_area[f] = 0.5 * area2;
    for (int i1 = 0; i1 < 4; ++i1)
      for (int i2 = 0; i2 < 4; ++i2) {
        double s = 0.5 * (gx[i1] + 1), t = 0.5 * (gx[i2] + 1);
        double u = s * (1 - t), v = t;
        double w = 0.25 * gw[i1] * gw[i2] * (1 - t) * area2;
        double x[3];
        for (int k = 0; k < 3; ++k) x[k] = P[k] + u * e1[k] + v * e2[k];
        double val[20][3];
        for (int a = 0; a < R.nfd; ++a) {
          int i = R.face_dofs[f * R.nfd + a];
          for (int d = 0; d < 3; ++d) val[a][d] = eval(basis[i][d], x);
        }
        for (int k = 0; k < 6; ++k) {
          int ca = kSym[k][0], cb = kSym[k][1];
          for (int a = 0; a < R.nfd; ++a)
            for (int b = 0; b < R.nfd; ++b) {
              double prod = val[a][ca] * val[b][cb];
              if (ca != cb) prod += val[a][cb] * val[b][ca];
              R.face[((f * 6 + k) * R.nfd + a) * R.nfd + b] += w * prod;
            }
        }
      }
  }
  // Discrete gradient of the Lagrange P_order basis.
  {
    Poly lam[4] = {};
    lam[0][0] = 1; lam[0][1] = -1; lam[0][2] = -1; lam[0][3] = -1;
    lam[1][1] = 1; lam[2][2] = 1; lam[3][3] = 1;
    std::vector<Poly> lag;
    if (order == 1) {
      for (int v = 0; v < 4; ++v) lag.push_back(lam[v]);
    } else {
      for (int v = 0; v < 4; ++v) {
        Poly two = lam[v];
        for (auto& c : two) c *= 2;
        two[0] -= 1;
        lag.push_back(mul(lam[v], two));        // lambda (2 lambda - 1)
      }
      for (int e = 0; e < 6; ++e) {
        Poly p = mul(lam[kEdges[e][0]], lam[kEdges[e][1]]);
        for (auto& c : p) c *= 4;
        lag.push_back(p);                       // 4 lambda_a lambda_b
      }
    }
    R.nlag = (int)lag.size();
    R.grad.assign(n * R.nlag, 0.0);
    for (int j = 0; j < R.nlag; ++j) {
      Poly g[3] = {deriv(lag[j], 0), deriv(lag[j], 1), deriv(lag[j], 2)};
      for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += eval(g[k], duals[i].p) * duals[i].t[k];
        R.grad[i * R.nlag + j] = std::fabs(s) < 1e-13 ? 0.0 : s;
      }
    }
  }
  return R;
}

}  // namespace

const RefData& ref_data(int order) {
  static std::once_flag once[2];
  static RefData data[2];
  if (order != 1 && order != 2) throw std::invalid_argument("order must be 1 or 2");
  std::call_once(once[order - 1], [order] { data[order - 1] = build(order); });
  return data[order - 1];
}

}  // namespace sucem
