// mesh.cpp -- host builders of the periodic Kuhn-tetrahedron mesh and the
// /-split triangle mesh, emitting the reference's tables
// (src/local_mesh.mojo:1241-1485, src/local_mesh_2d.mojo:103-395) with the
// same integer connectivity.
//
// The face-node maps are derived geometrically: every Lagrange node of the
// grid lies on the integer lattice  X = P*cell + sum_v a_v * corner_v  (in
// units of h/P), so the node a neighbour sees at a canonical face slot is
// found by matching lattice points, with the neighbour's cell taken across
// a periodic boundary without wrapping (exact, also with one cell per
// period), and the slot permutation of each element follows from its own
// face-local node map.
// Face normals and areas come from barycentric gradients of the side-0
// element (outward normal of face f is -grad(lambda_f)/|grad(lambda_f)|).
//
// Kept identical to the reference because they define the numbering:
//   the Kuhn split table              src/local_mesh.mojo:64-114
//   (tet, local face) -> (face type, owner offset, side)  168-217
//   canonical corners of each face type                   220-292
//   canonical face-slot order (vertices, edges, interior) 399-451
//   the 2D split, face types and reversed side-1 order
//                                     src/local_mesh_2d.mojo:138-320
//   the BC overlay numbering          src/local_mesh.mojo:946-1187,
//                                     src/local_mesh_2d.mojo:331-395
#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

#include "host.hpp"

namespace mxm {
namespace {

constexpr int KUHN[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7},
                            {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
// (face type, di, dj, dk, side) per (tet, local face)
constexpr int TFI[6][4][5] = {
    {{6, 0, 0, 0, 0}, {1, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, {11, 0, 0, -1, 1}},
    {{8, 0, 0, 0, 0}, {2, 0, 0, 0, 0}, {1, 0, 0, 0, 1}, {10, 0, 0, -1, 1}},
    {{9, 0, 0, 0, 0}, {3, 0, 0, 0, 0}, {2, 0, 0, 0, 1}, {6, -1, 0, 0, 1}},
    {{10, 0, 0, 0, 0}, {4, 0, 0, 0, 0}, {3, 0, 0, 0, 1}, {7, -1, 0, 0, 1}},
    {{11, 0, 0, 0, 0}, {5, 0, 0, 0, 0}, {4, 0, 0, 0, 1}, {9, 0, -1, 0, 1}},
    {{7, 0, 0, 0, 0}, {0, 0, 0, 0, 1}, {5, 0, 0, 0, 1}, {8, 0, -1, 0, 1}}};
constexpr int FTC[12][3] = {{0, 1, 7}, {0, 3, 7}, {0, 2, 7}, {0, 6, 7}, {0, 4, 7}, {0, 5, 7},
                            {1, 3, 7}, {1, 5, 7}, {2, 3, 7}, {2, 6, 7}, {4, 6, 7}, {4, 5, 7}};

inline std::array<int, 3> corner(int c) { return {c & 1, (c >> 1) & 1, (c >> 2) & 1}; }
inline int wrap(int a, int n) { return ((a % n) + n) % n; }

// Barycentric weights of canonical face slot m over the 3 canonical
// corners: vertices, edges 0->1, 1->2, 2->0, then interior lex-descending.
std::array<int, 3> canon_slot(int m, int P) {
  if (m < 3) {
    std::array<int, 3> b{0, 0, 0};
    b[m] = P;
    return b;
  }
  const int el = P - 1;
  if (m < 3 + 3 * el) {
    int which = (m - 3) / el, k = (m - 3) % el + 1;
    std::array<int, 3> b{0, 0, 0};
    b[which] = P - k;
    b[(which + 1) % 3] = k;
    return b;
  }
  int idx = 3 + 3 * el;
  for (int b0 = P - 2; b0 > 0; --b0)
    for (int b1 = P - b0 - 1; b1 > 0; --b1, ++idx)
      if (idx == m) return {b0, b1, P - b0 - b1};
  throw std::runtime_error("mesh: canonical slot out of range");
}

}  // namespace

HostMesh build_mesh_3d(int Nx, int Ny, int Nz, double lx, double ly, double lz, const int *bc_in,
                       int P, const RefElem &re) {
  if (Nx < 1 || Ny < 1 || Nz < 1) throw std::runtime_error("mesh: empty grid");
  HostMesh m;
  m.dim = 3;
  m.P = P;
  m.NP = re.NP;
  m.NFP = re.NFP;
  m.NF = 4;
  m.n[0] = Nx, m.n[1] = Ny, m.n[2] = Nz;
  m.L[0] = lx, m.L[1] = ly, m.L[2] = lz;
  for (int d = 0; d < 6; ++d) m.bc[d] = bc_in ? bc_in[d] : 0;
  const int *bc = m.bc;
  const int NP = re.NP, NFP = re.NFP;
  const long long cubes = (long long)Nx * Ny * Nz;
  if (cubes * 12 + 2LL * (Ny * Nz + Nx * Nz + Nx * Ny) > 0x7fffffffLL / (2 * NFP))
    throw std::runtime_error("mesh: too large for 32-bit tables");
  const int K = (int)(cubes * 6);
  const int nfp = (int)(cubes * 12);
  const int ex_x = bc[0] ? 2 * Ny * Nz : 0, ex_y = bc[2] ? 2 * Nx * Nz : 0,
            ex_z = bc[4] ? 2 * Nx * Ny : 0;
  const int Nf = nfp + ex_x + ex_y + ex_z;
  m.K = K;
  m.Nf = Nf;
  const double h[3] = {lx / Nx, ly / Ny, lz / Nz};

  m.elem_node_xyz.resize((size_t)K * NP * 3);
  m.elem_invJ.resize((size_t)K * 9);
  m.elem_inv_vol.resize(K);
  m.elem_faces.resize((size_t)K * 4);
  m.elem_face_side.resize((size_t)K * 4);
  m.elem_canon_to_ref.resize((size_t)K * 4 * NFP);
  m.face_elem.assign((size_t)Nf * 2, -1);
  m.face_elem_node.assign((size_t)Nf * 2 * NFP, -1);
  m.face_normal.resize((size_t)Nf * 3);
  m.face_area.resize(Nf);
  m.face_bc_type.assign(Nf, 0);

  // Per tet type: the inverse Jacobian (rows = grad r, grad s, grad t),
  // 1/det J, and per local face the outward unit normal and area.
  double invJ[6][9], invdet[6], fn[6][4][3], fa[6][4];
  for (int t = 0; t < 6; ++t) {
    double J[3][3];
    auto c0 = corner(KUHN[t][0]);
    for (int col = 0; col < 3; ++col) {
      auto c = corner(KUHN[t][col + 1]);
      for (int d = 0; d < 3; ++d) J[d][col] = (c[d] - c0[d]) * h[d];
    }
    double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                 J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (!(det > 0)) throw std::runtime_error("mesh: non-positive tet Jacobian");
    // inverse by cofactors: invJ[r][c] = cof(J)[c][r] / det
    double cof[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        cof[r][c] = J[r1][c1] * J[r2][c2] - J[r1][c2] * J[r2][c1];
      }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) invJ[t][r * 3 + c] = cof[c][r] / det;
    invdet[t] = 1.0 / det;
    const double vol = det / 6.0;
    for (int f = 0; f < 4; ++f) {
      double g[3];  // grad lambda_f
      for (int d = 0; d < 3; ++d)
        g[d] = f == 0 ? -(invJ[t][0 * 3 + d] + invJ[t][1 * 3 + d] + invJ[t][2 * 3 + d])
                      : invJ[t][(f - 1) * 3 + d];
      double gl = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
      for (int d = 0; d < 3; ++d) fn[t][f][d] = -g[d] / gl;
      fa[t][f] = 3.0 * vol * gl;
    }
  }

  auto node_lattice = [&](int cube_i, int cube_j, int cube_k, int t, int n) {
    std::array<int, 3> X{P * cube_i, P * cube_j, P * cube_k};
    for (int v = 0; v < 4; ++v) {
      auto c = corner(KUHN[t][v]);
      for (int d = 0; d < 3; ++d) X[d] += re.bary[n * 4 + v] * c[d];
    }
    return X;
  };

  // Elements: nodes, geometry, face ids and sides.
  for (int e = 0; e < K; ++e) {
    const int t = e % 6, cube = e / 6;
    const int k = cube / (Nx * Ny), j = (cube / Nx) % Ny, i = cube % Nx;
    for (int n = 0; n < NP; ++n) {
      auto X = node_lattice(i, j, k, t, n);
      for (int d = 0; d < 3; ++d) m.elem_node_xyz[((size_t)e * NP + n) * 3 + d] = X[d] * h[d] / P;
    }
    for (int q = 0; q < 9; ++q) m.elem_invJ[(size_t)e * 9 + q] = invJ[t][q];
    m.elem_inv_vol[e] = invdet[t];
    for (int lf = 0; lf < 4; ++lf) {
      const int *inf = TFI[t][lf];
      int owner = wrap(i + inf[1], Nx) + Nx * (wrap(j + inf[2], Ny) + Ny * wrap(k + inf[3], Nz));
      int fid = owner * 12 + inf[0];
      m.elem_faces[e * 4 + lf] = fid;
      m.elem_face_side[e * 4 + lf] = inf[4];
      if (m.face_elem[fid * 2 + inf[4]] != -1) throw std::runtime_error("mesh: face side claimed twice");
      m.face_elem[fid * 2 + inf[4]] = e;
    }
  }

  // Periodic faces: canonical slots, node maps, slot permutations, normals.
  for (int fid = 0; fid < nfp; ++fid) {
    const int ft = fid % 12, owner = fid / 12;
    const int ok = owner / (Nx * Ny), oj = (owner / Nx) % Ny, oi = owner % Nx;
    for (int s = 0; s < 2; ++s) {
      const int e = m.face_elem[fid * 2 + s];
      if (e < 0) throw std::runtime_error("mesh: face with a missing side");
      const int t = e % 6, cube = e / 6;
      const int ek = cube / (Nx * Ny), ej = (cube / Nx) % Ny, ei = cube % Nx;
      int lf = -1;
      for (int q = 0; q < 4; ++q)
        if (m.elem_faces[e * 4 + q] == fid && m.elem_face_side[e * 4 + q] == s) lf = q;
      // the element's cube relative to the owner, unwrapped, so that the
      // lattice points compare exactly (also with one cube per period)
      const int *inf = TFI[t][lf];
      const int ui = oi - inf[1], uj = oj - inf[2], uk = ok - inf[3];
      if (wrap(ui, Nx) != ei || wrap(uj, Ny) != ej || wrap(uk, Nz) != ek)
        throw std::runtime_error("mesh: face owner offset mismatch");
      for (int slot = 0; slot < NFP; ++slot) {
        auto b = canon_slot(slot, P);
        std::array<int, 3> Xm{P * oi, P * oj, P * ok};
        for (int q = 0; q < 3; ++q) {
          auto c = corner(FTC[ft][q]);
          for (int d = 0; d < 3; ++d) Xm[d] += b[q] * c[d];
        }
        int found = -1, rslot = -1;
        for (int r = 0; r < NFP; ++r) {
          int n = re.fmap[lf * NFP + r];
          if (node_lattice(ui, uj, uk, t, n) == Xm) {
            found = n;
            rslot = r;
            break;
          }
        }
        if (found < 0) throw std::runtime_error("mesh: face node not found");
        m.face_elem_node[((size_t)fid * 2 + s) * NFP + slot] = found;
        m.elem_canon_to_ref[((size_t)e * 4 + lf) * NFP + slot] = rslot;
      }
      if (s == 0) {
        for (int d = 0; d < 3; ++d) m.face_normal[(size_t)fid * 3 + d] = fn[t][lf][d];
        m.face_area[fid] = fa[t][lf];
      }
    }
  }

  // BC overlay.  "+" faces keep their ids and point side 1 at side 0;
  // "-" faces get fresh ids (blocks -x, -y, -z in that order) as one-sided
  // mirrors whose side 0 is the boundary element.
  auto plus = [&](int fid, int v) {
    m.face_elem[fid * 2 + 1] = m.face_elem[fid * 2 + 0];
    m.face_bc_type[fid] = v;
  };
  auto minus = [&](int elem, int new_fid, int v) {
    const int old = m.elem_faces[elem * 4 + 3];
    m.face_elem[new_fid * 2 + 0] = elem;
    m.face_elem[new_fid * 2 + 1] = elem;
    for (int q = 0; q < NFP; ++q) {
      int pat = m.face_elem_node[((size_t)old * 2 + 1) * NFP + q];
      m.face_elem_node[((size_t)new_fid * 2 + 0) * NFP + q] = pat;
      m.face_elem_node[((size_t)new_fid * 2 + 1) * NFP + q] = pat;
    }
    for (int d = 0; d < 3; ++d) m.face_normal[(size_t)new_fid * 3 + d] = -m.face_normal[(size_t)old * 3 + d];
    m.face_area[new_fid] = m.face_area[old];
    m.face_bc_type[new_fid] = v;
    m.elem_faces[elem * 4 + 3] = new_fid;
    m.elem_face_side[elem * 4 + 3] = 0;
  };
  auto cube_id = [&](int i, int j, int k) { return i + Nx * (j + Ny * k); };
  if (bc[1])
    for (int jj = 0; jj < Ny; ++jj)
      for (int kk = 0; kk < Nz; ++kk)
        for (int fo = 0; fo < 2; ++fo) plus(cube_id(Nx - 1, jj, kk) * 12 + 6 + fo, bc[1]);
  if (bc[3])
    for (int ii = 0; ii < Nx; ++ii)
      for (int kk = 0; kk < Nz; ++kk)
        for (int fo = 0; fo < 2; ++fo) plus(cube_id(ii, Ny - 1, kk) * 12 + 8 + fo, bc[3]);
  if (bc[5])
    for (int ii = 0; ii < Nx; ++ii)
      for (int jj = 0; jj < Ny; ++jj)
        for (int fo = 0; fo < 2; ++fo) plus(cube_id(ii, jj, Nz - 1) * 12 + 10 + fo, bc[5]);
  int base = nfp;
  if (bc[0]) {
    for (int jj = 0; jj < Ny; ++jj)
      for (int kk = 0; kk < Nz; ++kk)
        for (int fo = 0; fo < 2; ++fo)
          minus(cube_id(0, jj, kk) * 6 + 2 + fo, base + 2 * (jj * Nz + kk) + fo, bc[0]);
    base += ex_x;
  }
  if (bc[2]) {
    for (int ii = 0; ii < Nx; ++ii)
      for (int kk = 0; kk < Nz; ++kk)
        for (int fo = 0; fo < 2; ++fo)
          minus(cube_id(ii, 0, kk) * 6 + 5 - fo, base + 2 * (ii * Nz + kk) + fo, bc[2]);
    base += ex_y;
  }
  if (bc[4])
    for (int ii = 0; ii < Nx; ++ii)
      for (int jj = 0; jj < Ny; ++jj)
        for (int fo = 0; fo < 2; ++fo)
          minus(cube_id(ii, jj, 0) * 6 + 1 - fo, base + 2 * (ii * Ny + jj) + fo, bc[4]);
  return m;
}

HostMesh build_mesh_2d(int Nx, int Ny, double lx, double ly, const int *bc_in, int P, const RefElem &re) {
  if (Nx < 1 || Ny < 1) throw std::runtime_error("mesh: empty grid");
  HostMesh m;
  m.dim = 2;
  m.P = P;
  m.NP = re.NP;
  m.NFP = re.NFP;
  m.NF = 3;
  m.n[0] = Nx, m.n[1] = Ny;
  m.L[0] = lx, m.L[1] = ly;
  for (int d = 0; d < 4; ++d) m.bc[d] = bc_in ? bc_in[d] : 0;
  const int *bc = m.bc;
  const int NP = re.NP, NE = re.NFP;
  const long long cells = (long long)Nx * Ny;
  if (cells * 3 + Nx + Ny > 0x7fffffffLL / (2 * NE)) throw std::runtime_error("mesh: too large");
  const int K = (int)(cells * 2), nfp = (int)(cells * 3);
  const int ex_x = bc[0] ? Ny : 0, ex_y = bc[2] ? Nx : 0;
  const int Nf = nfp + ex_x + ex_y;
  m.K = K;
  m.Nf = Nf;
  const double dx = lx / Nx, dy = ly / Ny;
  // Vertex offsets of T0 = (00, 10, 11) and T1 = (00, 11, 01).
  static const int TV[2][3][2] = {{{0, 0}, {1, 0}, {1, 1}}, {{0, 0}, {1, 1}, {0, 1}}};
  // Face types: (side-0 tri, edge), (side-1 tri, edge), side-1 cell offset.
  static const int FT[3][6] = {{0, 2, 1, 0, 0, 0}, {0, 1, 1, 2, 1, 0}, {1, 1, 0, 0, 0, 1}};
  static const int ED[3][2] = {{0, 1}, {1, 2}, {2, 0}};

  m.elem_node_xyz.resize((size_t)K * NP * 2);
  m.elem_invJ.resize((size_t)K * 4);
  m.elem_inv_vol.resize(K);
  m.elem_faces.resize((size_t)K * 3);
  m.elem_face_side.resize((size_t)K * 3);
  m.elem_canon_to_ref.resize((size_t)K * 3 * NE);
  m.face_elem.assign((size_t)Nf * 2, -1);
  m.face_elem_node.assign((size_t)Nf * 2 * NE, -1);
  m.face_normal.resize((size_t)Nf * 2);
  m.face_area.resize(Nf);
  m.face_bc_type.assign(Nf, 0);

  auto lattice = [&](int ci, int cj, int t, int n) {
    std::array<int, 2> X{P * ci, P * cj};
    for (int v = 0; v < 3; ++v)
      for (int d = 0; d < 2; ++d) X[d] += re.bary[n * 3 + v] * TV[t][v][d];
    return X;
  };
  double invJ[2][4], invdet[2];
  for (int t = 0; t < 2; ++t) {
    double J00 = (TV[t][1][0] - TV[t][0][0]) * dx, J01 = (TV[t][2][0] - TV[t][0][0]) * dx;
    double J10 = (TV[t][1][1] - TV[t][0][1]) * dy, J11 = (TV[t][2][1] - TV[t][0][1]) * dy;
    double det = J00 * J11 - J01 * J10;
    if (!(det > 0)) throw std::runtime_error("mesh: non-positive triangle area");
    invJ[t][0] = J11 / det;
    invJ[t][1] = -J01 / det;
    invJ[t][2] = -J10 / det;
    invJ[t][3] = J00 / det;
    invdet[t] = 1.0 / det;
  }
  for (int e = 0; e < K; ++e) {
    const int t = e % 2, cell = e / 2, i = cell % Nx, j = cell / Nx;
    for (int n = 0; n < NP; ++n) {
      auto X = lattice(i, j, t, n);
      m.elem_node_xyz[((size_t)e * NP + n) * 2 + 0] = X[0] * dx / P;
      m.elem_node_xyz[((size_t)e * NP + n) * 2 + 1] = X[1] * dy / P;
    }
    for (int q = 0; q < 4; ++q) m.elem_invJ[(size_t)e * 4 + q] = invJ[t][q];
    m.elem_inv_vol[e] = invdet[t];
  }
  for (int j = 0; j < Ny; ++j)
    for (int i = 0; i < Nx; ++i) {
      const int cell = j * Nx + i;
      for (int ft = 0; ft < 3; ++ft) {
        const int fid = cell * 3 + ft;
        const int t0 = FT[ft][0], l0 = FT[ft][1], t1 = FT[ft][2], l1 = FT[ft][3];
        const int ni = wrap(i + FT[ft][4], Nx), nj = wrap(j + FT[ft][5], Ny);
        const int e0 = cell * 2 + t0, e1 = (nj * Nx + ni) * 2 + t1;
        m.face_elem[fid * 2 + 0] = e0;
        m.face_elem[fid * 2 + 1] = e1;
        m.elem_faces[e0 * 3 + l0] = fid;
        m.elem_face_side[e0 * 3 + l0] = 0;
        m.elem_faces[e1 * 3 + l1] = fid;
        m.elem_face_side[e1 * 3 + l1] = 1;
        // outward normal of side 0 along its CCW edge va -> vb
        const int *va = TV[t0][ED[l0][0]], *vb = TV[t0][ED[l0][1]];
        double exv = (vb[0] - va[0]) * dx, eyv = (vb[1] - va[1]) * dy;
        double len = std::sqrt(exv * exv + eyv * eyv);
        m.face_normal[(size_t)fid * 2 + 0] = eyv / len;
        m.face_normal[(size_t)fid * 2 + 1] = -exv / len;
        m.face_area[fid] = len;
        for (int slot = 0; slot < NE; ++slot) {
          std::array<int, 2> Xm{P * i + P * va[0] + slot * (vb[0] - va[0]),
                                P * j + P * va[1] + slot * (vb[1] - va[1])};
          for (int s = 0; s < 2; ++s) {
            const int e = s ? e1 : e0, lf = s ? l1 : l0, t = s ? t1 : t0;
            // the neighbour's cell unwrapped, so that the lattice points
            // compare exactly (also with one cell per period)
            const int ci = s ? i + FT[ft][4] : i, cj = s ? j + FT[ft][5] : j;
            int found = -1, rslot = -1;
            for (int r = 0; r < NE; ++r) {
              int n = re.fmap[lf * NE + r];
              auto X = lattice(ci, cj, t, n);
              if (X[0] == Xm[0] && X[1] == Xm[1]) {
                found = n;
                rslot = r;
                break;
              }
            }
            if (found < 0) throw std::runtime_error("mesh: edge node not found");
            m.face_elem_node[((size_t)fid * 2 + s) * NE + slot] = found;
            m.elem_canon_to_ref[((size_t)e * 3 + lf) * NE + slot] = rslot;
          }
        }
      }
    }
  if (bc[1])
    for (int j = 0; j < Ny; ++j) {
      int fid = (j * Nx + Nx - 1) * 3 + 1;
      m.face_elem[fid * 2 + 1] = m.face_elem[fid * 2];
      m.face_bc_type[fid] = bc[1];
    }
  if (bc[3])
    for (int i = 0; i < Nx; ++i) {
      int fid = ((Ny - 1) * Nx + i) * 3 + 2;
      m.face_elem[fid * 2 + 1] = m.face_elem[fid * 2];
      m.face_bc_type[fid] = bc[3];
    }
  auto minus = [&](int elem, int lf, int fid, double nx, double ny, double len, int v) {
    m.face_elem[fid * 2 + 0] = elem;
    m.face_elem[fid * 2 + 1] = elem;
    m.face_normal[(size_t)fid * 2 + 0] = nx;
    m.face_normal[(size_t)fid * 2 + 1] = ny;
    m.face_area[fid] = len;
    m.face_bc_type[fid] = v;
    m.elem_faces[elem * 3 + lf] = fid;
    m.elem_face_side[elem * 3 + lf] = 0;
    for (int q = 0; q < NE; ++q) {
      int nn = re.fmap[lf * NE + q];
      m.face_elem_node[((size_t)fid * 2 + 0) * NE + q] = nn;
      m.face_elem_node[((size_t)fid * 2 + 1) * NE + q] = nn;
      m.elem_canon_to_ref[((size_t)elem * 3 + lf) * NE + q] = q;
    }
  };
  if (bc[0])
    for (int j = 0; j < Ny; ++j) minus((j * Nx) * 2 + 1, 2, nfp + j, -1.0, 0.0, dy, bc[0]);
  if (bc[2])
    for (int i = 0; i < Nx; ++i) minus(i * 2 + 0, 0, nfp + ex_x + i, 0.0, -1.0, dx, bc[2]);
  return m;
}

Connectivity build_connectivity(const HostMesh &m, const RefElem &re) {
  const int NF = m.NF, NFP = m.NFP;
  Connectivity c;
  c.conn.assign((size_t)m.K * NF * 4, 0);
  std::vector<std::vector<int>> pats;
  std::vector<int> pat(NFP), inv(NFP);
  for (int r = 0; r < NF * NFP; ++r)
    if (re.fmap[r] < 0 || re.fmap[r] >= m.NP) throw std::runtime_error("conn: reference face map out of range");
  for (int e = 0; e < m.K; ++e)
    for (int lf = 0; lf < NF; ++lf) {
      const int fid = m.elem_faces[(size_t)e * NF + lf], side = m.elem_face_side[(size_t)e * NF + lf];
      if (fid < 0 || fid >= m.Nf || (side != 0 && side != 1)) throw std::runtime_error("conn: bad face");
      const int bc = m.face_bc_type[fid];
      if (bc < 0 || bc > 7) throw std::runtime_error("conn: bad bc kind");
      std::fill(inv.begin(), inv.end(), -1);
      for (int q = 0; q < NFP; ++q) {
        int r = m.elem_canon_to_ref[((size_t)e * NF + lf) * NFP + q];
        if (r < 0 || r >= NFP || inv[r] != -1) throw std::runtime_error("conn: canon_to_ref not a permutation");
        inv[r] = q;
      }
      for (int r = 0; r < NFP; ++r) {
        const int q = inv[r];
        if (m.face_elem_node[((size_t)fid * 2 + side) * NFP + q] != re.fmap[lf * NFP + r])
          throw std::runtime_error("conn: face node of an element disagrees with its reference face map");
        pat[r] = m.face_elem_node[((size_t)fid * 2 + (1 - side)) * NFP + q];
        if (pat[r] < 0 || pat[r] >= m.NP) throw std::runtime_error("conn: partner node out of range");
      }
      const int partner = m.face_elem[(size_t)fid * 2 + (1 - side)];
      if (partner < 0 || partner >= m.K) throw std::runtime_error("conn: partner element out of range");
      int id = -1;
      for (size_t p = 0; p < pats.size(); ++p)
        if (pats[p] == pat) id = (int)p;
      if (id < 0) {
        id = (int)pats.size();
        pats.push_back(pat);
      }
      int32_t *o = &c.conn[((size_t)e * NF + lf) * 4];
      o[0] = partner;
      o[1] = fid;
      o[2] = side | (bc << 1) | (id << 4);
      o[3] = 0;
    }
  c.nperm = (int)pats.size();
  for (auto &p : pats) c.perm_tab.insert(c.perm_tab.end(), p.begin(), p.end());
  return c;
}

Tiling build_tiling(const HostMesh &m, int epb, int NFP) {
  if (epb < 1 || epb > 126) throw std::runtime_error("tiling: tile size out of range");
  const int dim = m.dim, epc = dim == 3 ? 6 : 2;  // elements per lattice cell
  const int nx = m.n[0], ny = m.n[1], nz = dim == 3 ? m.n[2] : 1;
  if ((long long)nx * ny * nz * epc != m.K) throw std::runtime_error("tiling: element count is not a lattice");
  // Faces evaluated for a bx x by x bz block on a periodic lattice: pairs
  // inside the block once, faces to other blocks once per side.
  const int inner = dim == 3 ? 6 : 1, per_iface = dim == 3 ? 2 : 1;
  Tiling t;
  t.epb = epb;
  double best = 1e30;
  int best_faces = 0, best_cells = 0;
  for (int bz = 1; bz <= nz; ++bz)
    for (int by = 1; by <= ny; ++by)
      for (int bx = 1; bx <= nx; ++bx) {
        const int cells = bx * by * bz;
        if (cells * epc > epb) continue;
        const int ifaces = (bx - 1) * by * bz + bx * (by - 1) * bz + (dim == 3 ? bx * by * (bz - 1) : 0);
        const int sfaces = 2 * (by * bz + bx * bz + (dim == 3 ? bx * by : 0));
        const int faces = cells * inner + ifaces * per_iface + sfaces * per_iface;
        // wave-executions of the face phase plus about one for the per-node
        // phases of the block, per element
        const int waves = (faces * NFP + 63) / 64;
        const double cost = (double)(waves + 1) / (cells * epc);
        if (cost < best - 1e-12 || (cost < best + 1e-12 && (faces * best_cells < best_faces * cells ||
                                                           (faces * best_cells == best_faces * cells && cells > best_cells)))) {
          best = cost, best_faces = faces, best_cells = cells;
          t.block[0] = bx, t.block[1] = by, t.block[2] = bz;
        }
      }
  t.dev_of.assign(m.K, -1);
  if (best_cells == 0) {
    // a cell has more elements than a tile holds: consecutive elements
    t.block[0] = t.block[1] = t.block[2] = 0;
    t.ntiles = (m.K + epb - 1) / epb;
    t.ref_of.assign((size_t)t.ntiles * epb, -1);
    t.count.assign(t.ntiles, 0);
    for (int e = 0; e < m.K; ++e) {
      t.ref_of[e] = e;
      t.dev_of[e] = e;
      ++t.count[e / epb];
    }
    return t;
  }
  const int bx = t.block[0], by = t.block[1], bz = t.block[2];
  const int tx = (nx + bx - 1) / bx, ty = (ny + by - 1) / by, tz = (nz + bz - 1) / bz;
  t.ntiles = tx * ty * tz;
  t.ref_of.assign((size_t)t.ntiles * epb, -1);
  t.count.assign(t.ntiles, 0);
  // tiles x-fastest; cells inside a tile z, y, x order; then the cell's elements
  for (int TZ = 0; TZ < tz; ++TZ)
    for (int TY = 0; TY < ty; ++TY)
      for (int TX = 0; TX < tx; ++TX) {
        const int T = (TZ * ty + TY) * tx + TX;
        int el = 0;
        for (int k = TZ * bz; k < std::min(nz, (TZ + 1) * bz); ++k)
          for (int j = TY * by; j < std::min(ny, (TY + 1) * by); ++j)
            for (int i = TX * bx; i < std::min(nx, (TX + 1) * bx); ++i)
              for (int q = 0; q < epc; ++q) {
                const int e = ((k * ny + j) * nx + i) * epc + q;
                t.ref_of[(size_t)T * epb + el] = e;
                t.dev_of[e] = T * epb + el;
                ++el;
              }
        t.count[T] = el;
      }
  for (int e = 0; e < m.K; ++e)
    if (t.dev_of[e] < 0) throw std::runtime_error("tiling: element not placed");
  return t;
}

FaceJobs build_face_jobs(const HostMesh &m, const RefElem &re, const Connectivity &c, const Tiling &t) {
  const int NF = m.NF, NFP = m.NFP, epb = t.epb, ND = m.dim;
  if (m.NP > 255 || NFP > 255) throw std::runtime_error("jobs: node index does not fit 8 bits");
  FaceJobs J;
  std::vector<std::vector<int>> pats;
  std::vector<int> pat(NFP);
  // per tile: job records (4 ints) and the partner node of each slot
  std::vector<std::vector<int32_t>> tj(t.ntiles);
  std::vector<std::vector<int32_t>> tn(t.ntiles);
  for (int T = 0; T < t.ntiles; ++T) {
    const int nE = t.count[T];
    for (int el = 0; el < nE; ++el)
      for (int lf = 0; lf < NF; ++lf) {
        const int e = t.ref_of[(size_t)T * epb + el];
        const int32_t *cn = &c.conn[((size_t)e * NF + lf) * 4];
        const int partner = cn[0], fid = cn[1], side = cn[2] & 1, bc = (cn[2] >> 1) & 7, pid = cn[2] >> 4;
        const int pdev = t.dev_of[partner];
        const int pl = pdev - T * epb;
        // A pair needs both elements to reference the face: after a "-"
        // boundary overlay on an axis whose "+" side stays periodic, the
        // old periodic face keeps its side-0 element only.
        int plf = -1;
        if (bc == 0 && partner != e && pl >= 0 && pl < nE)
          for (int q = 0; q < NF; ++q)
            if (m.elem_faces[(size_t)partner * NF + q] == fid &&
                m.elem_face_side[(size_t)partner * NF + q] == 1 - side)
              plf = q;
        const bool pair = plf >= 0;
        if (pair && side == 1) continue;  // evaluated by the side-0 element
        int packed = el | (lf << 7);
        for (int r = 0; r < NFP; ++r) pat[r] = 0;
        if (pair) {
          for (int r = 0; r < NFP; ++r) {
            const int pn = c.perm_tab[(size_t)pid * NFP + r];
            int rs = -1;
            for (int q = 0; q < NFP; ++q)
              if (re.fmap[plf * NFP + q] == pn) rs = q;
            if (rs < 0) throw std::runtime_error("jobs: partner slot not found");
            pat[r] = rs;
          }
          int id = -1;
          for (size_t p = 0; p < pats.size(); ++p)
            if (pats[p] == pat) id = (int)p;
          if (id < 0) {
            id = (int)pats.size();
            pats.push_back(pat);
          }
          if (id >= (1 << 13)) throw std::runtime_error("jobs: too many slot patterns");
          packed |= ((pl + 1) << 9) | (plf << 16) | (id << 18);
        }
        tj[T].insert(tj[T].end(), {packed, pdev, fid, cn[2]});
        for (int r = 0; r < NFP; ++r) tn[T].push_back(c.perm_tab[(size_t)pid * NFP + r] | (pat[r] << 8));
      }
    J.maxj = std::max(J.maxj, (int)(tj[T].size() / 4));
  }
  J.npslot = (int)pats.size();
  // Every (element, face, slot) of a tile must be written exactly once.
  std::vector<int> hits((size_t)epb * NF * NFP);
  for (int T = 0; T < t.ntiles; ++T) {
    const int nE = t.count[T];
    std::fill(hits.begin(), hits.end(), 0);
    for (size_t j = 0; j < tj[T].size() / 4; ++j) {
      const int job = tj[T][j * 4], el = job & 127, lf = (job >> 7) & 3, pin = (job >> 9) & 127;
      if (el >= nE || (pin && pin - 1 >= nE)) throw std::runtime_error("jobs: element outside its tile");
      for (int r = 0; r < NFP; ++r) {
        ++hits[((size_t)el * NF + lf) * NFP + r];
        if (pin) ++hits[((size_t)(pin - 1) * NF + ((job >> 16) & 3)) * NFP + pats[job >> 18][r]];
      }
    }
    for (int k = 0; k < nE * NF * NFP; ++k)
      if (hits[k] != 1) throw std::runtime_error("jobs: face slot not covered exactly once");
  }
  // fixed-stride device tables
  const size_t slots = (size_t)t.ntiles * J.maxj;
  J.jobs.assign(slots * 4, 0);
  J.geo.assign(slots * 4, 0.0);
  J.ipr.assign(slots * NFP, 0);
  for (int T = 0; T < t.ntiles; ++T) {
    const size_t nj = tj[T].size() / 4;
    for (size_t j = 0; j < (size_t)J.maxj; ++j) {
      const size_t s = (size_t)T * J.maxj + j;
      if (j >= nj) {
        J.jobs[s * 4 + 0] = -1;
        J.jobs[s * 4 + 1] = T * epb;
        continue;
      }
      for (int k = 0; k < 4; ++k) J.jobs[s * 4 + k] = tj[T][j * 4 + k];
      const int fid = tj[T][j * 4 + 2];
      for (int d = 0; d < ND; ++d) J.geo[s * 4 + d] = m.face_normal[(size_t)fid * ND + d];
      J.geo[s * 4 + 3] = m.face_area[fid];
      for (int r = 0; r < NFP; ++r) J.ipr[s * NFP + r] = (uint16_t)tn[T][j * NFP + r];
    }
  }
  for (auto &p : pats) J.pslot_tab.insert(J.pslot_tab.end(), p.begin(), p.end());
  if (J.pslot_tab.empty()) J.pslot_tab.assign(NFP, 0);
  return J;
}

}  // namespace mxm
