// refelem.cpp -- equispaced Lagrange reference operators on the unit
// tetrahedron / triangle, built on the host once per solver.
//
// Mathematics of src/reference.mojo:519-660 and src/reference_2d.mojo:
// 130-306: nodal Lagrange basis on equispaced nodes, exact mass and
// stiffness integrals, weak-form derivative D_d = M^-1 S_d with
// S_d[i][j] = int dphi_i/dr_d phi_j, lift Lift_f = M^-1 L_f with L_f the
// face mass of the face-local Lagrange basis per unit reference face
// measure, and node weights w_i = int phi_i / |ref element|.
//
// Construction used here: Cartesian monomials r^a s^b t^c of total degree
// <= P span the space; the Lagrange coefficients come from an LU solve
// with the Vandermonde matrix; all integrals are exact monomial integrals
// over the unit simplex; everything is carried in long double and rounded
// to double once.  Node orderings (vertices, VTK edge order, face-interior,
// volume-interior) and the face-local orderings are the reference's
// conventions and are kept identical.
#include <cmath>
#include <map>
#include <stdexcept>
#include <tuple>

#include "host.hpp"

namespace mxm {
namespace {

using ld = long double;

// Square matrix LU with partial pivoting; solves A X = B in place of B.
void lu_solve(int n, std::vector<ld> A, std::vector<ld> &B, int nrhs) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    ld best = std::fabs(A[k * n + k]);
    for (int r = k + 1; r < n; ++r)
      if (std::fabs(A[r * n + k]) > best) {
        best = std::fabs(A[r * n + k]);
        p = r;
      }
    if (best == 0) throw std::runtime_error("refelem: singular matrix");
    if (p != k) {
      for (int c = 0; c < n; ++c) std::swap(A[k * n + c], A[p * n + c]);
      for (int c = 0; c < nrhs; ++c) std::swap(B[k * nrhs + c], B[p * nrhs + c]);
    }
    for (int r = k + 1; r < n; ++r) {
      ld f = A[r * n + k] / A[k * n + k];
      if (f == 0) continue;
      for (int c = k; c < n; ++c) A[r * n + c] -= f * A[k * n + c];
      for (int c = 0; c < nrhs; ++c) B[r * nrhs + c] -= f * B[k * nrhs + c];
    }
  }
  for (int k = n - 1; k >= 0; --k)
    for (int c = 0; c < nrhs; ++c) {
      ld s = B[k * nrhs + c];
      for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * B[j * nrhs + c];
      B[k * nrhs + c] = s / A[k * n + k];
    }
}

ld fact(int n) {
  ld f = 1;
  for (int k = 2; k <= n; ++k) f *= k;
  return f;
}

// int over the unit simplex of prod x_d^e_d = prod e_d! / (sum e + dim)!
ld simplex_integral(const int *e, int dim) {
  ld num = 1;
  int s = 0;
  for (int d = 0; d < dim; ++d) {
    num *= fact(e[d]);
    s += e[d];
  }
  return num / fact(s + dim);
}

// Integer barycentric labels in the reference's canonical node order.
std::vector<std::vector<int>> simplex_labels(int dim, int P) {
  const int nv = dim + 1;
  std::vector<std::vector<int>> out;
  auto vert = [&](int v) {
    std::vector<int> a(nv, 0);
    a[v] = P;
    return a;
  };
  for (int v = 0; v < nv; ++v) out.push_back(vert(v));
  if (P >= 2) {
    // tet edges in VTK order; triangle edges (0,1), (1,2), (2,0)
    static const int te[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    int ne = dim == 3 ? 6 : 3;
    for (int e = 0; e < ne; ++e)
      for (int k = 1; k < P; ++k) {
        std::vector<int> a(nv, 0);
        a[te[e][0]] = P - k;
        a[te[e][1]] = k;
        out.push_back(a);
      }
  }
  // interior of each 2-face (lex-descending on the first two of its
  // ascending vertices); in 2D the only 2-face is the triangle itself
  auto face_interior = [&](const std::vector<int> &verts) {
    for (int a0 = P - 2; a0 > 0; --a0)
      for (int a1 = P - a0 - 1; a1 > 0; --a1) {
        std::vector<int> a(nv, 0);
        a[verts[0]] = a0;
        a[verts[1]] = a1;
        a[verts[2]] = P - a0 - a1;
        out.push_back(a);
      }
  };
  if (P >= 3) {
    if (dim == 3) {
      for (int f = 0; f < 4; ++f) {
        std::vector<int> vs;
        for (int v = 0; v < 4; ++v)
          if (v != f) vs.push_back(v);
        face_interior(vs);
      }
    } else {
      face_interior({0, 1, 2});
    }
  }
  if (dim == 3 && P >= 4)
    for (int a0 = P - 2; a0 > 0; --a0)
      for (int a1 = P - a0 - 1; a1 > 0; --a1)
        for (int a2 = P - a0 - a1 - 1; a2 > 0; --a2) out.push_back({a0, a1, a2, P - a0 - a1 - a2});
  return out;
}

// Exponent tuples of the Cartesian monomials of degree <= P.
std::vector<std::vector<int>> monomials(int dim, int P) {
  std::vector<std::vector<int>> out;
  if (dim == 1) {
    for (int a = 0; a <= P; ++a) out.push_back({a});
  } else if (dim == 2) {
    for (int a = 0; a <= P; ++a)
      for (int b = 0; a + b <= P; ++b) out.push_back({a, b});
  } else {
    for (int a = 0; a <= P; ++a)
      for (int b = 0; a + b <= P; ++b)
        for (int c = 0; a + b + c <= P; ++c) out.push_back({a, b, c});
  }
  return out;
}

// Lagrange basis on nodes X (Cartesian, n x dim) in the monomial basis:
// returns C with phi_i = sum_k C[i][k] m_k, i.e. C = V^-T.
std::vector<ld> lagrange_coefs(const std::vector<std::vector<ld>> &X,
                               const std::vector<std::vector<int>> &mons, int dim) {
  const int n = (int)X.size();
  std::vector<ld> Vt(n * n), C(n * n, 0);  // C holds the identity, then (V^T)^-1
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      ld v = 1;
      for (int d = 0; d < dim; ++d)
        for (int r = 0; r < mons[k][d]; ++r) v *= X[j][d];
      Vt[k * n + j] = v;  // V^T[k][j] = m_k(x_j)
    }
  // phi_i(x_j) = (C V^T)[i][j] = delta_ij  ->  C = (V^T)^-1
  for (int i = 0; i < n; ++i) C[i * n + i] = 1;
  lu_solve(n, Vt, C, n);
  return C;
}

// Mass matrix of the Lagrange basis with coefficients C over the unit
// simplex (dim 2 or 3) or [0, 1] (dim 1).
std::vector<ld> mass(const std::vector<ld> &C, const std::vector<std::vector<int>> &mons, int dim) {
  const int n = (int)mons.size();
  std::vector<ld> G(n * n), T(n * n, 0), M(n * n, 0);
  for (int k = 0; k < n; ++k)
    for (int l = 0; l < n; ++l) {
      int e[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) e[d] = mons[k][d] + mons[l][d];
      G[k * n + l] = dim == 1 ? (ld)1 / (e[0] + 1) : simplex_integral(e, dim);
    }
  for (int i = 0; i < n; ++i)
    for (int l = 0; l < n; ++l) {
      ld s = 0;
      for (int k = 0; k < n; ++k) s += C[i * n + k] * G[k * n + l];
      T[i * n + l] = s;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      ld s = 0;
      for (int l = 0; l < n; ++l) s += T[i * n + l] * C[j * n + l];
      M[i * n + j] = s;
    }
  return M;
}

}  // namespace

RefElem build_refelem(int dim, int P) {
  if ((dim != 2 && dim != 3) || P < 1 || P > 6) throw std::runtime_error("refelem: unsupported order");
  RefElem re;
  re.dim = dim;
  re.P = P;
  re.NP = dim == 3 ? tet_nodes(P) : tri_nodes(P);
  re.NFP = dim == 3 ? tri_nodes(P) : P + 1;
  re.NF = dim + 1;
  const int NP = re.NP, NFP = re.NFP, NF = re.NF;

  auto labels = simplex_labels(dim, P);
  if ((int)labels.size() != NP) throw std::runtime_error("refelem: node count");
  std::map<std::vector<int>, int> index;
  for (int i = 0; i < NP; ++i) {
    index[labels[i]] = i;
    for (int v = 0; v <= dim; ++v) re.bary.push_back(labels[i][v]);
  }

  // Cartesian node coordinates (r, s[, t]) = (l1, l2[, l3]).
  std::vector<std::vector<ld>> X(NP, std::vector<ld>(dim));
  for (int i = 0; i < NP; ++i)
    for (int d = 0; d < dim; ++d) X[i][d] = (ld)labels[i][d + 1] / P;
  auto mons = monomials(dim, P);
  auto C = lagrange_coefs(X, mons, dim);
  auto M = mass(C, mons, dim);

  // Stiffness S_d[i][j] = sum_{k,l} C[i][k] C[j][l] int (d m_k/dr_d) m_l
  const int nm = (int)mons.size();
  re.D.assign((size_t)dim * NP * NP, 0.0);
  for (int d = 0; d < dim; ++d) {
    std::vector<ld> Gd(nm * nm, 0), T(NP * nm, 0), S(NP * NP, 0);
    for (int k = 0; k < nm; ++k) {
      if (mons[k][d] == 0) continue;
      for (int l = 0; l < nm; ++l) {
        int e[3] = {0, 0, 0};
        for (int dd = 0; dd < dim; ++dd) e[dd] = mons[k][dd] + mons[l][dd];
        e[d] -= 1;
        Gd[k * nm + l] = mons[k][d] * simplex_integral(e, dim);
      }
    }
    for (int i = 0; i < NP; ++i)
      for (int l = 0; l < nm; ++l) {
        ld s = 0;
        for (int k = 0; k < nm; ++k) s += C[i * nm + k] * Gd[k * nm + l];
        T[i * nm + l] = s;
      }
    for (int i = 0; i < NP; ++i)
      for (int j = 0; j < NP; ++j) {
        ld s = 0;
        for (int l = 0; l < nm; ++l) s += T[i * nm + l] * C[j * nm + l];
        S[i * NP + j] = s;
      }
    lu_solve(NP, M, S, NP);  // D_d = M^-1 S_d
    for (int i = 0; i < NP * NP; ++i) re.D[(size_t)d * NP * NP + i] = (double)S[i];
  }

  // Nodal differentiation Dn[d][i][j] = d phi_j / d r_d (x_i): the
  // derivative of each monomial at each node, through the coefficients.
  re.Dn.assign((size_t)dim * NP * NP, 0.0);
  for (int d = 0; d < dim; ++d)
    for (int i = 0; i < NP; ++i)
      for (int j = 0; j < NP; ++j) {
        ld s = 0;
        for (int k = 0; k < nm; ++k) {
          if (mons[k][d] == 0) continue;
          ld v = mons[k][d];
          for (int dd = 0; dd < dim; ++dd)
            for (int r = 0; r < mons[k][dd] - (dd == d ? 1 : 0); ++r) v *= X[i][dd];
          s += C[j * nm + k] * v;
        }
        re.Dn[((size_t)d * NP + i) * NP + j] = (double)s;
      }

  // Node weights w_i = int phi_i / |ref|.
  {
    const ld vol = dim == 3 ? (ld)1 / 6 : (ld)1 / 2;
    re.w.resize(NP);
    for (int i = 0; i < NP; ++i) {
      ld s = 0;
      for (int k = 0; k < nm; ++k) {
        int e[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) e[d] = mons[k][d];
        s += C[i * nm + k] * simplex_integral(e, dim);
      }
      re.w[i] = (double)(s / vol);
    }
  }

  // Face-local node maps.  3D face f (opposite vertex f): its vertices in
  // ascending order (vA, vB, vC), edges vA->vB, vB->vC, vC->vA, then the
  // interior lex-descending.  2D edge e = (va, vb) in {(0,1), (1,2), (2,0)}
  // walked from va to vb including both ends.
  re.fmap.assign((size_t)NF * NFP, -1);
  auto lookup = [&](const std::vector<int> &a) {
    auto it = index.find(a);
    if (it == index.end()) throw std::runtime_error("refelem: face node lookup");
    return it->second;
  };
  if (dim == 3) {
    for (int f = 0; f < 4; ++f) {
      int vs[3], c = 0;
      for (int v = 0; v < 4; ++v)
        if (v != f) vs[c++] = v;
      int pos = 0;
      for (int k = 0; k < 3; ++k) {
        std::vector<int> a(4, 0);
        a[vs[k]] = P;
        re.fmap[f * NFP + pos++] = lookup(a);
      }
      for (int k = 0; k < 3; ++k) {
        int va = vs[k], vb = vs[(k + 1) % 3];
        for (int s = 1; s < P; ++s) {
          std::vector<int> a(4, 0);
          a[va] = P - s;
          a[vb] = s;
          re.fmap[f * NFP + pos++] = lookup(a);
        }
      }
      for (int a0 = P - 2; a0 > 0; --a0)
        for (int a1 = P - a0 - 1; a1 > 0; --a1) {
          std::vector<int> a(4, 0);
          a[vs[0]] = a0;
          a[vs[1]] = a1;
          a[vs[2]] = P - a0 - a1;
          re.fmap[f * NFP + pos++] = lookup(a);
        }
    }
  } else {
    static const int ed[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (int e = 0; e < 3; ++e)
      for (int k = 0; k <= P; ++k) {
        std::vector<int> a(3, 0);
        a[ed[e][0]] += P - k;
        a[ed[e][1]] += k;
        re.fmap[e * NFP + k] = lookup(a);
      }
  }

  // Face mass per unit reference-face measure (the reference scales the
  // unit-triangle mass by 2 = 1/area; the unit edge has length 1).
  std::vector<ld> Mf;
  if (dim == 3) {
    auto fl = simplex_labels(2, P);
    std::vector<std::vector<ld>> Xf(NFP, std::vector<ld>(2));
    for (int m = 0; m < NFP; ++m) {
      Xf[m][0] = (ld)fl[m][1] / P;
      Xf[m][1] = (ld)fl[m][2] / P;
    }
    auto fm = monomials(2, P);
    Mf = mass(lagrange_coefs(Xf, fm, 2), fm, 2);
    for (auto &v : Mf) v *= 2;
  } else {
    std::vector<std::vector<ld>> Xf(NFP, std::vector<ld>(1));
    for (int m = 0; m < NFP; ++m) Xf[m][0] = (ld)m / P;
    auto fm = monomials(1, P);
    Mf = mass(lagrange_coefs(Xf, fm, 1), fm, 1);
  }
  re.Lift.assign((size_t)NF * NP * NFP, 0.0);
  for (int f = 0; f < NF; ++f) {
    std::vector<ld> L(NP * NFP, 0);
    for (int l = 0; l < NFP; ++l)
      for (int m = 0; m < NFP; ++m) L[re.fmap[f * NFP + l] * NFP + m] = Mf[l * NFP + m];
    lu_solve(NP, M, L, NFP);
    for (int i = 0; i < NP * NFP; ++i) re.Lift[(size_t)f * NP * NFP + i] = (double)L[i];
  }
  return re;
}

}  // namespace mxm
