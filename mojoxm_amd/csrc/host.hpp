// host.hpp -- host-side (CPU) data built once per solver: reference
// operators and the mesh in the reference's table layout, plus the
// element-centric connectivity the stage kernels read.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mxm {

inline int tet_nodes(int P) { return (P + 1) * (P + 2) * (P + 3) / 6; }
inline int tri_nodes(int P) { return (P + 1) * (P + 2) / 2; }

// Reference-element operators for one order P (src/reference.mojo:519-660,
// src/reference_2d.mojo:130-261).  Row-major:
//   D[d][i][j], Lift[f][i][m], w[i], fmap[f][m], bary[i][dim+1]
struct RefElem {
  int dim = 0, P = 0, NP = 0, NFP = 0, NF = 0;
  std::vector<double> D, Lift, w;
  std::vector<double> Dn;  // [dim][NP][NP] nodal differentiation: d phi_j / d r_k at node i
  std::vector<int> fmap;
  std::vector<int> bary;  // integer barycentric label of each node (sum = P)
};

// Builds the operators; throws std::runtime_error on failure.
RefElem build_refelem(int dim, int P);

// Mesh tables in the reference's layout (see include/mxm_dg.h MXM_T_*).
struct HostMesh {
  int dim = 0, P = 0, NP = 0, NFP = 0, NF = 0;
  int n[3] = {1, 1, 1};
  double L[3] = {1, 1, 1};
  int bc[6] = {0, 0, 0, 0, 0, 0};
  int K = 0, Nf = 0;
  std::vector<double> elem_node_xyz, elem_invJ, elem_inv_vol;
  std::vector<int32_t> elem_faces, elem_face_side, elem_canon_to_ref;
  std::vector<int32_t> face_elem, face_elem_node;
  std::vector<double> face_normal, face_area;
  std::vector<int32_t> face_bc_type;
};

HostMesh build_mesh_3d(int nx, int ny, int nz, double lx, double ly, double lz, const int *bc,
                       int P, const RefElem &re);
HostMesh build_mesh_2d(int nx, int ny, double lx, double ly, const int *bc, int P,
                       const RefElem &re);

// Element-centric connectivity for the stage kernels.  For element e and
// local face lf, conn[e*NF+lf] = {partner element, face id, flags, 0} with
// flags = side | bc << 1 | perm << 4, and perm_tab[perm*NFP + r] is the
// partner's element-local node that meets this element's reference face
// slot r.  For side 1 the flux is evaluated with the partner as the left
// state, exactly as the reference evaluates every face from side 0.
struct Connectivity {
  std::vector<int32_t> conn;  // [K*NF*4]
  std::vector<int32_t> perm_tab;
  int nperm = 0;
};

// Validates that the tables are consistent (the node each element sees at
// a face slot is its own reference-face node) and compresses them.
Connectivity build_connectivity(const HostMesh &m, const RefElem &re);

// Element tiles of the stage kernel and the device element order.  A tile
// is a block of lattice cells (squares of 2 triangles, cubes of 6 tets),
// at most `epb` elements, whose shape minimises the face-flux work per
// element (faces with both sides in the tile are evaluated once).  Device
// element slot T*epb + el holds element ref_of[T*epb + el] (-1: padding),
// so every tile is contiguous in device memory; dev_of is the inverse.
// When one cell has more elements than a tile (3D, P = 5), tiles are runs
// of consecutive elements and block = {0, 0, 0}.
struct Tiling {
  int epb = 0, ntiles = 0;
  int block[3] = {1, 1, 1};  // cells per tile along x, y, z
  std::vector<int32_t> dev_of, ref_of, count;
  int slots() const { return ntiles * epb; }
};
Tiling build_tiling(const HostMesh &m, int epb, int NFP);

// Face-flux work lists of the stage kernel.  Within a tile, a face whose
// two sides both lie in the tile is evaluated once (by its side-0
// element), which also writes the negated flux into the partner's slots.
// Every tile owns `maxj` job slots (padding: packed = -1), so the records a
// thread needs are addressed from the tile index alone:
//   jobs[T*maxj + j]  el | lf << 7 | (partner_local + 1 or 0) << 9 |
//                     partner_lf << 16 | pslot << 18, the partner's device
//                     slot, the face id, the connectivity flags
//   geo[(T*maxj + j)*4 + 0..3]  face normal (x, y, z; 0 beyond ND), area
//   ipr[(T*maxj + j)*NFP + r]   partner node meeting slot r | (partner's
//                     reference slot for the paired write) << 8
// pslot_tab[pslot*NFP + r] is the partner's reference slot meeting slot r.
struct FaceJobs {
  int maxj = 0;
  std::vector<int32_t> jobs, pslot_tab;
  std::vector<double> geo;
  std::vector<uint16_t> ipr;
  int npslot = 0;
};
FaceJobs build_face_jobs(const HostMesh &m, const RefElem &re, const Connectivity &c, const Tiling &t);

}  // namespace mxm
