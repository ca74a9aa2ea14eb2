/*
 * mesh.c -- oracle restatement of the Kuhn-tetrahedron mesh
 * (src/local_mesh.mojo) and the /-split triangle mesh
 * (src/local_mesh_2d.mojo) in the reference's table layout
 * (TEST INFRASTRUCTURE, see oracle.h).
 *
 * The integer connectivity follows the reference's own algorithm: per
 * tet-type and per face-type tables built from the Lagrange exponent
 * bookkeeping (_compute_tables, src/local_mesh.mojo:498-769), stamped over
 * the grid (build_elements_kernel 777-845, build_faces_kernel 848-898),
 * then the BC overlay (946-1187).  Geometry uses the same formulas in
 * FP64 (the reference evaluates them in Float32).
 *
 * Pinned by: the invariants of test/local_mesh_2d_test.mojo:36-106 and
 * their 3D analogues (face_elem/elem_faces round trip, coincident face
 * node coordinates on both sides modulo periodic wrap, closed element
 * surfaces, outward normals), the literal P=2 face table of
 * src/reference.mojo:699-752, and an independently derived (lattice
 * geometric) builder in the shipped library that must agree bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* src/local_mesh.mojo:64-114 */
static const int KUHN[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7},
                               {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
static int cdx(int c) { return (c == 1 || c == 3 || c == 5 || c == 7) ? 1 : 0; }
static int cdy(int c) { return (c == 2 || c == 3 || c == 6 || c == 7) ? 1 : 0; }
static int cdz(int c) { return c >= 4 ? 1 : 0; }

/* tet_face_info, src/local_mesh.mojo:168-217: (face_type, di, dj, dk, side) */
static const int TFI[6][4][5] = {
    {{6, 0, 0, 0, 0}, {1, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, {11, 0, 0, -1, 1}},
    {{8, 0, 0, 0, 0}, {2, 0, 0, 0, 0}, {1, 0, 0, 0, 1}, {10, 0, 0, -1, 1}},
    {{9, 0, 0, 0, 0}, {3, 0, 0, 0, 0}, {2, 0, 0, 0, 1}, {6, -1, 0, 0, 1}},
    {{10, 0, 0, 0, 0}, {4, 0, 0, 0, 0}, {3, 0, 0, 0, 1}, {7, -1, 0, 0, 1}},
    {{11, 0, 0, 0, 0}, {5, 0, 0, 0, 0}, {4, 0, 0, 0, 1}, {9, 0, -1, 0, 1}},
    {{7, 0, 0, 0, 0}, {0, 0, 0, 0, 1}, {5, 0, 0, 0, 1}, {8, 0, -1, 0, 1}}};

/* face_type_corner, src/local_mesh.mojo:220-292 */
static const int FTC[12][3] = {{0, 1, 7}, {0, 3, 7}, {0, 2, 7}, {0, 6, 7},
                               {0, 4, 7}, {0, 5, 7}, {1, 3, 7}, {1, 5, 7},
                               {2, 3, 7}, {2, 6, 7}, {4, 6, 7}, {4, 5, 7}};

/* Lagrange tet exponents, src/local_mesh.mojo:332-388 (same ordering as
 * src/reference.mojo:125-200). */
static int lagrange_tet_exponents(int P, int *e) {
    static const int ea[6] = {0, 1, 2, 0, 1, 2};
    static const int eb[6] = {1, 2, 0, 3, 3, 3};
    int n = 0;
    for (int v = 0; v < 4; ++v) {
        for (int d = 0; d < 4; ++d) e[4 * n + d] = (d == v) ? P : 0;
        ++n;
    }
    if (P >= 2)
        for (int ed = 0; ed < 6; ++ed)
            for (int k = 1; k < P; ++k) {
                for (int d = 0; d < 4; ++d)
                    e[4 * n + d] = (ea[ed] == d) ? P - k : ((eb[ed] == d) ? k : 0);
                ++n;
            }
    if (P >= 3)
        for (int f = 0; f < 4; ++f) {
            int ni[3], c = 0;
            for (int i = 0; i < 4; ++i)
                if (i != f) ni[c++] = i;
            for (int a0v = P - 2; a0v > 0; --a0v)
                for (int a1v = P - a0v - 1; a1v > 0; --a1v) {
                    int a2v = P - a0v - a1v;
                    for (int d = 0; d < 4; ++d)
                        e[4 * n + d] =
                            (ni[0] == d) ? a0v
                                         : ((ni[1] == d) ? a1v : ((ni[2] == d) ? a2v : 0));
                    ++n;
                }
        }
    if (P >= 4)
        for (int a0v = P - 2; a0v > 0; --a0v)
            for (int a1v = P - a0v - 1; a1v > 0; --a1v)
                for (int a2v = P - a0v - a1v - 1; a2v > 0; --a2v) {
                    e[4 * n + 0] = a0v;
                    e[4 * n + 1] = a1v;
                    e[4 * n + 2] = a2v;
                    e[4 * n + 3] = P - a0v - a1v - a2v;
                    ++n;
                }
    return n;
}

static int find_tet_node(const int *ex, int N, const int *a) {
    for (int i = 0; i < N; ++i)
        if (ex[4 * i] == a[0] && ex[4 * i + 1] == a[1] && ex[4 * i + 2] == a[2] &&
            ex[4 * i + 3] == a[3])
            return i;
    return -1;
}

/* src/local_mesh.mojo:399-451 */
static int canon_face_barycentric(int m, int P, int *b) {
    if (m < 3) {
        b[0] = b[1] = b[2] = 0;
        b[m] = P;
        return 0;
    }
    int el = P - 1;
    if (m < 3 + 3 * el) {
        int which = (m - 3) / el, k = (m - 3) % el + 1;
        if (which == 0) {
            b[0] = P - k; b[1] = k; b[2] = 0;
        } else if (which == 1) {
            b[0] = 0; b[1] = P - k; b[2] = k;
        } else {
            b[0] = k; b[1] = 0; b[2] = P - k;
        }
        return 0;
    }
    int idx = 3 + 3 * el;
    for (int b0 = P - 2; b0 > 0; --b0)
        for (int b1 = P - b0 - 1; b1 > 0; --b1) {
            if (idx == m) {
                b[0] = b0; b[1] = b1; b[2] = P - b0 - b1;
                return 0;
            }
            ++idx;
        }
    return -1;
}

typedef struct {
    double tet_invJ[6 * 9], tet_inv_6V[6];
    double *tet_node_rel; /* [6*NP*3] */
    int tet_face_type[24], tet_face_off[24 * 3], tet_face_side[24];
    int *tet_canon_to_ref; /* [24*NFP] */
    int ft_side0_tet[12], ft_side1_tet[12], ft_side1_off[36];
    int *ft_elem_node; /* [12*2*NFP] */
    double ft_normal[36], ft_area[12];
} tables3d;

/* _compute_tables, src/local_mesh.mojo:498-769 */
static int compute_tables(double dx, double dy, double dz, int P, tables3d *T) {
    int NP = orc_num_tet_nodes(P), NFP = orc_num_tri_nodes(P);
    int32_t *ref_face = (int32_t *)malloc(sizeof(int32_t) * 4 * NFP);
    if (orc_refelem3d(P, NULL, NULL, NULL, ref_face, NULL)) return -1;
    int *ex = (int *)malloc(sizeof(int) * 4 * NP);
    lagrange_tet_exponents(P, ex);
    T->tet_node_rel = (double *)malloc(sizeof(double) * 6 * NP * 3);
    T->tet_canon_to_ref = (int *)malloc(sizeof(int) * 24 * NFP);
    T->ft_elem_node = (int *)malloc(sizeof(int) * 24 * NFP);
    int *cnode = (int *)malloc(sizeof(int) * 24 * NFP);

    for (int t = 0; t < 6; ++t) {
        double vx[4], vy[4], vz[4];
        for (int k = 0; k < 4; ++k) {
            int c = KUHN[t][k];
            vx[k] = cdx(c);
            vy[k] = cdy(c);
            vz[k] = cdz(c);
        }
        double J[9];
        J[0] = (vx[1] - vx[0]) * dx; J[1] = (vx[2] - vx[0]) * dx; J[2] = (vx[3] - vx[0]) * dx;
        J[3] = (vy[1] - vy[0]) * dy; J[4] = (vy[2] - vy[0]) * dy; J[5] = (vy[3] - vy[0]) * dy;
        J[6] = (vz[1] - vz[0]) * dz; J[7] = (vz[2] - vz[0]) * dz; J[8] = (vz[3] - vz[0]) * dz;
        double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                     J[2] * (J[3] * J[7] - J[4] * J[6]);
        if (det <= 0.0) return -2;
        double V = det / 6.0;
        T->tet_inv_6V[t] = 1.0 / (6.0 * V);
        double *iJ = &T->tet_invJ[t * 9];
        iJ[0] = (J[4] * J[8] - J[5] * J[7]) / det;
        iJ[1] = (J[2] * J[7] - J[1] * J[8]) / det;
        iJ[2] = (J[1] * J[5] - J[2] * J[4]) / det;
        iJ[3] = (J[5] * J[6] - J[3] * J[8]) / det;
        iJ[4] = (J[0] * J[8] - J[2] * J[6]) / det;
        iJ[5] = (J[2] * J[3] - J[0] * J[5]) / det;
        iJ[6] = (J[3] * J[7] - J[4] * J[6]) / det;
        iJ[7] = (J[1] * J[6] - J[0] * J[7]) / det;
        iJ[8] = (J[0] * J[4] - J[1] * J[3]) / det;
        double Pf = (double)P;
        for (int n = 0; n < NP; ++n) {
            double w0 = ex[4 * n] / Pf, w1 = ex[4 * n + 1] / Pf, w2 = ex[4 * n + 2] / Pf,
                   w3 = ex[4 * n + 3] / Pf;
            double *o = &T->tet_node_rel[(t * NP + n) * 3];
            o[0] = w0 * vx[0] + w1 * vx[1] + w2 * vx[2] + w3 * vx[3];
            o[1] = w0 * vy[0] + w1 * vy[1] + w2 * vy[2] + w3 * vy[3];
            o[2] = w0 * vz[0] + w1 * vz[1] + w2 * vz[2] + w3 * vz[3];
        }
    }

    for (int t = 0; t < 6; ++t)
        for (int f = 0; f < 4; ++f) {
            const int *info = TFI[t][f];
            int ft = info[0], di = info[1], dj = info[2], dk = info[3];
            T->tet_face_type[t * 4 + f] = ft;
            T->tet_face_off[(t * 4 + f) * 3 + 0] = di;
            T->tet_face_off[(t * 4 + f) * 3 + 1] = dj;
            T->tet_face_off[(t * 4 + f) * 3 + 2] = dk;
            T->tet_face_side[t * 4 + f] = info[4];
            int fv[3], c = 0;
            for (int k = 0; k < 4; ++k)
                if (k != f) fv[c++] = k;
            int owner_cube[3];
            for (int kk = 0; kk < 3; ++kk) {
                int cc = KUHN[t][fv[kk]];
                if (di == -1) cc ^= 1;
                if (dj == -1) cc ^= 2;
                if (dk == -1) cc ^= 4;
                owner_cube[kk] = cc;
            }
            int c2r_vertex[3];
            for (int k = 0; k < 3; ++k) {
                c2r_vertex[k] = -1;
                for (int ii = 0; ii < 3; ++ii)
                    if (owner_cube[ii] == FTC[ft][k]) {
                        c2r_vertex[k] = ii;
                        break;
                    }
                if (c2r_vertex[k] < 0) return -3;
            }
            for (int m = 0; m < NFP; ++m) {
                int b[3];
                if (canon_face_barycentric(m, P, b)) return -4;
                int wfv[3] = {0, 0, 0};
                for (int k = 0; k < 3; ++k) wfv[c2r_vertex[k]] = b[k];
                int a[4] = {0, 0, 0, 0};
                for (int ii = 0; ii < 3; ++ii) a[fv[ii]] = wfv[ii];
                int tn = find_tet_node(ex, NP, a);
                if (tn < 0) return -5;
                cnode[(t * 4 + f) * NFP + m] = tn;
                int ref_m = -1;
                for (int mm = 0; mm < NFP; ++mm)
                    if (ref_face[f * NFP + mm] == tn) {
                        ref_m = mm;
                        break;
                    }
                if (ref_m < 0) return -6;
                T->tet_canon_to_ref[(t * 4 + f) * NFP + m] = ref_m;
            }
        }

    int side0_lf[12], side1_lf[12];
    memset(T->ft_side1_off, 0, sizeof(T->ft_side1_off));
    for (int t = 0; t < 6; ++t)
        for (int f = 0; f < 4; ++f) {
            const int *info = TFI[t][f];
            int ft = info[0];
            if (info[4] == 0) {
                T->ft_side0_tet[ft] = t;
                side0_lf[ft] = f;
            } else {
                T->ft_side1_tet[ft] = t;
                side1_lf[ft] = f;
                T->ft_side1_off[ft * 3 + 0] = -info[1];
                T->ft_side1_off[ft * 3 + 1] = -info[2];
                T->ft_side1_off[ft * 3 + 2] = -info[3];
            }
        }

    for (int ft = 0; ft < 12; ++ft) {
        int c0 = FTC[ft][0], c1 = FTC[ft][1], c2 = FTC[ft][2];
        double p0x = cdx(c0) * dx, p0y = cdy(c0) * dy, p0z = cdz(c0) * dz;
        double p1x = cdx(c1) * dx, p1y = cdy(c1) * dy, p1z = cdz(c1) * dz;
        double p2x = cdx(c2) * dx, p2y = cdy(c2) * dy, p2z = cdz(c2) * dz;
        double e01x = p1x - p0x, e01y = p1y - p0y, e01z = p1z - p0z;
        double e02x = p2x - p0x, e02y = p2y - p0y, e02z = p2z - p0z;
        double nx = e01y * e02z - e01z * e02y;
        double ny = e01z * e02x - e01x * e02z;
        double nz = e01x * e02y - e01y * e02x;
        double nlen = sqrt(nx * nx + ny * ny + nz * nz);
        double area = 0.5 * nlen;
        nx = nx / nlen;
        ny = ny / nlen;
        nz = nz / nlen;
        int v4 = KUHN[T->ft_side0_tet[ft]][side0_lf[ft]];
        double v4x = cdx(v4) * dx, v4y = cdy(v4) * dy, v4z = cdz(v4) * dz;
        double dotn = (v4x - p0x) * nx + (v4y - p0y) * ny + (v4z - p0z) * nz;
        if (dotn > 0.0) {
            nx = -nx;
            ny = -ny;
            nz = -nz;
        }
        T->ft_normal[ft * 3 + 0] = nx;
        T->ft_normal[ft * 3 + 1] = ny;
        T->ft_normal[ft * 3 + 2] = nz;
        T->ft_area[ft] = area;
        int t0 = T->ft_side0_tet[ft], f0 = side0_lf[ft];
        int t1 = T->ft_side1_tet[ft], f1 = side1_lf[ft];
        for (int m = 0; m < NFP; ++m) {
            T->ft_elem_node[(ft * 2 + 0) * NFP + m] = cnode[(t0 * 4 + f0) * NFP + m];
            T->ft_elem_node[(ft * 2 + 1) * NFP + m] = cnode[(t1 * 4 + f1) * NFP + m];
        }
    }
    free(ref_face);
    free(ex);
    free(cnode);
    return 0;
}

static void free_tables(tables3d *T) {
    free(T->tet_node_rel);
    free(T->tet_canon_to_ref);
    free(T->ft_elem_node);
}

/* LocalMesh face-count logic, src/local_mesh.mojo:1241-1270 (bnd_off=0). */
int orc_mesh3d_counts(int nx, int ny, int nz, const int32_t *bc, int *ne, int *nf) {
    *ne = nx * ny * nz * 6;
    int extra = 0;
    if (bc && bc[0]) extra += 2 * ny * nz;
    if (bc && bc[2]) extra += 2 * nx * nz;
    if (bc && bc[4]) extra += 2 * nx * ny;
    *nf = nx * ny * nz * 12 + extra;
    return 0;
}

/* Minus-axis overlay, src/local_mesh.mojo:1050-1187. */
static void minus_overlay(orc_mesh_t *m, const tables3d *T, int cube, int tet, int ft,
                          int new_fid, int bcv) {
    int NFP = m->NFP, elem = cube * 6 + tet;
    m->face_elem[new_fid * 2 + 0] = elem;
    m->face_elem[new_fid * 2 + 1] = elem;
    for (int k = 0; k < NFP; ++k) {
        int pat = T->ft_elem_node[(ft * 2 + 1) * NFP + k];
        m->face_elem_node[(new_fid * 2 + 0) * NFP + k] = pat;
        m->face_elem_node[(new_fid * 2 + 1) * NFP + k] = pat;
    }
    for (int d = 0; d < 3; ++d) m->face_normal[new_fid * 3 + d] = -T->ft_normal[ft * 3 + d];
    m->face_area[new_fid] = T->ft_area[ft];
    m->face_bc_type[new_fid] = bcv;
    m->elem_faces[elem * 4 + 3] = new_fid;
    m->elem_face_side[elem * 4 + 3] = 0;
}

int orc_mesh3d_build(orc_mesh_t *m, const int32_t *bc) {
    static const int32_t zero[6] = {0, 0, 0, 0, 0, 0};
    if (!bc) bc = zero;
    int P = m->P, Nx = m->nx, Ny = m->ny, Nz = m->nz;
    int NP = orc_num_tet_nodes(P), NFP = orc_num_tri_nodes(P);
    m->ndim = 3;
    m->NP = NP;
    m->NFP = NFP;
    m->NF = 4;
    int ne, nf;
    orc_mesh3d_counts(Nx, Ny, Nz, bc, &ne, &nf);
    m->num_elements = ne;
    m->num_faces = nf;
    double dx = m->lx / Nx, dy = m->ly / Ny, dz = m->lz / Nz;
    tables3d T;
    int rc = compute_tables(dx, dy, dz, P, &T);
    if (rc) return rc;

    /* build_elements_kernel, src/local_mesh.mojo:777-845 */
    for (int elem = 0; elem < ne; ++elem) {
        int t = elem % 6, cell = elem / 6;
        int k = cell / (Nx * Ny), rem = cell - k * Nx * Ny, j = rem / Nx, i = rem - j * Nx;
        double ox = i * dx, oy = j * dy, oz = k * dz;
        for (int n = 0; n < NP; ++n) {
            const double *r = &T.tet_node_rel[(t * NP + n) * 3];
            m->elem_node_xyz[(elem * NP + n) * 3 + 0] = ox + r[0] * dx;
            m->elem_node_xyz[(elem * NP + n) * 3 + 1] = oy + r[1] * dy;
            m->elem_node_xyz[(elem * NP + n) * 3 + 2] = oz + r[2] * dz;
        }
        for (int d = 0; d < 9; ++d) m->elem_invJ[elem * 9 + d] = T.tet_invJ[t * 9 + d];
        m->elem_inv_vol[elem] = T.tet_inv_6V[t];
        for (int lf = 0; lf < 4; ++lf) {
            int ft = T.tet_face_type[t * 4 + lf];
            int oi = (i + T.tet_face_off[(t * 4 + lf) * 3 + 0] + Nx) % Nx;
            int oj = (j + T.tet_face_off[(t * 4 + lf) * 3 + 1] + Ny) % Ny;
            int ok = (k + T.tet_face_off[(t * 4 + lf) * 3 + 2] + Nz) % Nz;
            int owner = oi + Nx * (oj + Ny * ok);
            m->elem_faces[elem * 4 + lf] = owner * 12 + ft;
            m->elem_face_side[elem * 4 + lf] = T.tet_face_side[t * 4 + lf];
            for (int q = 0; q < NFP; ++q)
                m->elem_canon_to_ref[(elem * 4 + lf) * NFP + q] =
                    T.tet_canon_to_ref[(t * 4 + lf) * NFP + q];
        }
    }
    /* build_faces_kernel, src/local_mesh.mojo:848-898 */
    int nfp = Nx * Ny * Nz * 12;
    for (int face = 0; face < nfp; ++face) {
        int ft = face % 12, cell = face / 12;
        int k = cell / (Nx * Ny), rem = cell - k * Nx * Ny, j = rem / Nx, i = rem - j * Nx;
        m->face_elem[face * 2 + 0] = cell * 6 + T.ft_side0_tet[ft];
        int si = (i + T.ft_side1_off[ft * 3 + 0] + Nx) % Nx;
        int sj = (j + T.ft_side1_off[ft * 3 + 1] + Ny) % Ny;
        int sk = (k + T.ft_side1_off[ft * 3 + 2] + Nz) % Nz;
        m->face_elem[face * 2 + 1] = (si + Nx * (sj + Ny * sk)) * 6 + T.ft_side1_tet[ft];
        for (int d = 0; d < 3; ++d) m->face_normal[face * 3 + d] = T.ft_normal[ft * 3 + d];
        m->face_area[face] = T.ft_area[ft];
        m->face_bc_type[face] = 0;
        for (int q = 0; q < NFP; ++q) {
            m->face_elem_node[(face * 2 + 0) * NFP + q] = T.ft_elem_node[(ft * 2 + 0) * NFP + q];
            m->face_elem_node[(face * 2 + 1) * NFP + q] = T.ft_elem_node[(ft * 2 + 1) * NFP + q];
        }
    }
    for (int f = nfp; f < nf; ++f) m->face_bc_type[f] = 0;

    /* BC overlay, src/local_mesh.mojo:1375-1470 (bnd_off = 0).  "+" faces
     * first, then the "-" mirror faces appended at nf_periodic.. */
    if (bc[1])
        for (int tid = 0; tid < 2 * Ny * Nz; ++tid) {
            int fo = tid % 2, jk = tid / 2, kk = jk % Nz, jj = jk / Nz;
            int fid = ((Nx - 1) + Nx * (jj + Ny * kk)) * 12 + 6 + fo;
            m->face_elem[fid * 2 + 1] = m->face_elem[fid * 2 + 0];
            m->face_bc_type[fid] = bc[1];
        }
    if (bc[3])
        for (int tid = 0; tid < 2 * Nx * Nz; ++tid) {
            int fo = tid % 2, ik = tid / 2, kk = ik % Nz, ii = ik / Nz;
            int fid = (ii + Nx * ((Ny - 1) + Ny * kk)) * 12 + 8 + fo;
            m->face_elem[fid * 2 + 1] = m->face_elem[fid * 2 + 0];
            m->face_bc_type[fid] = bc[3];
        }
    if (bc[5])
        for (int tid = 0; tid < 2 * Nx * Ny; ++tid) {
            int fo = tid % 2, ij = tid / 2, jj = ij % Ny, ii = ij / Ny;
            int fid = (ii + Nx * (jj + Ny * (Nz - 1))) * 12 + 10 + fo;
            m->face_elem[fid * 2 + 1] = m->face_elem[fid * 2 + 0];
            m->face_bc_type[fid] = bc[5];
        }
    int extra_mx = bc[0] ? 2 * Ny * Nz : 0, extra_my = bc[2] ? 2 * Nx * Nz : 0;
    int mx_base = nfp, my_base = mx_base + extra_mx, mz_base = my_base + extra_my;
    if (bc[0])
        for (int tid = 0; tid < 2 * Ny * Nz; ++tid) {
            int fo = tid % 2, jk = tid / 2, kk = jk % Nz, jj = jk / Nz;
            minus_overlay(m, &T, 0 + Nx * (jj + Ny * kk), 2 + fo, 6 + fo, mx_base + tid, bc[0]);
        }
    if (bc[2])
        for (int tid = 0; tid < 2 * Nx * Nz; ++tid) {
            int fo = tid % 2, ik = tid / 2, kk = ik % Nz, ii = ik / Nz;
            minus_overlay(m, &T, ii + Nx * (0 + Ny * kk), 5 - fo, 8 + fo, my_base + tid, bc[2]);
        }
    if (bc[4])
        for (int tid = 0; tid < 2 * Nx * Ny; ++tid) {
            int fo = tid % 2, ij = tid / 2, jj = ij % Ny, ii = ij / Ny;
            minus_overlay(m, &T, ii + Nx * (jj + Ny * 0), 1 - fo, 10 + fo, mz_base + tid, bc[4]);
        }
    free_tables(&T);
    return 0;
}

/* ---------------------------------------------------------------- 2D -- */

int orc_mesh2d_counts(int nx, int ny, const int32_t *bc, int *ne, int *nf) {
    *ne = nx * ny * 2;
    *nf = nx * ny * 3 + ((bc && bc[0]) ? ny : 0) + ((bc && bc[2]) ? nx : 0);
    return 0;
}

/* LocalMesh2D.__init__, src/local_mesh_2d.mojo:103-395 */
int orc_mesh2d_build(orc_mesh_t *m, const int32_t *bc) {
    static const int32_t zero[4] = {0, 0, 0, 0};
    if (!bc) bc = zero;
    int P = m->P, Nx = m->nx, Ny = m->ny;
    int NP = orc_num_tri_nodes(P), NE = P + 1;
    m->ndim = 2;
    m->NP = NP;
    m->NFP = NE;
    m->NF = 3;
    int ne, nf;
    orc_mesh2d_counts(Nx, Ny, bc, &ne, &nf);
    m->num_elements = ne;
    m->num_faces = nf;
    double dx = m->lx / (double)Nx, dy = m->ly / (double)Ny;
    double *npos = (double *)malloc(sizeof(double) * NP * 2);
    int32_t *e2e = (int32_t *)malloc(sizeof(int32_t) * 3 * NE);
    if (orc_refelem2d(P, NULL, NULL, NULL, e2e, npos)) return -1;
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            double bx = (double)i * dx, by = (double)j * dy;
            for (int t = 0; t < 2; ++t) {
                int elem = (j * Nx + i) * 2 + t;
                double v0x = bx, v0y = by, v1x, v1y, v2x, v2y;
                if (t == 0) {
                    v1x = bx + dx; v1y = by; v2x = bx + dx; v2y = by + dy;
                } else {
                    v1x = bx + dx; v1y = by + dy; v2x = bx; v2y = by + dy;
                }
                for (int n = 0; n < NP; ++n) {
                    double r = npos[n * 2], s = npos[n * 2 + 1], l0 = 1.0 - r - s;
                    m->elem_node_xyz[(elem * NP + n) * 2 + 0] = l0 * v0x + r * v1x + s * v2x;
                    m->elem_node_xyz[(elem * NP + n) * 2 + 1] = l0 * v0y + r * v1y + s * v2y;
                }
                double J00 = v1x - v0x, J01 = v2x - v0x, J10 = v1y - v0y, J11 = v2y - v0y;
                double detJ = J00 * J11 - J01 * J10;
                if (detJ <= 0.0) return -2;
                double inv = 1.0 / detJ;
                m->elem_invJ[elem * 4 + 0] = J11 * inv;
                m->elem_invJ[elem * 4 + 1] = -J01 * inv;
                m->elem_invJ[elem * 4 + 2] = -J10 * inv;
                m->elem_invJ[elem * 4 + 3] = J00 * inv;
                m->elem_inv_vol[elem] = inv;
            }
        }
    for (int f = 0; f < nf; ++f) m->face_bc_type[f] = 0;
    static const int s0t[3] = {0, 0, 1}, s0e[3] = {2, 1, 1};
    static const int s1t[3] = {1, 1, 0}, s1e[3] = {0, 2, 0};
    static const int fdi[3] = {0, 1, 0}, fdj[3] = {0, 0, 1};
    double diag = sqrt(dx * dx + dy * dy);
    double fnx[3] = {-dy / diag, 1.0, 0.0}, fny[3] = {dx / diag, 0.0, 1.0};
    double flen[3] = {diag, dy, dx};
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            int cell = j * Nx + i;
            for (int ft = 0; ft < 3; ++ft) {
                int fid = cell * 3 + ft;
                int e0 = cell * 2 + s0t[ft];
                int ni = (i + fdi[ft] + Nx) % Nx, nj = (j + fdj[ft] + Ny) % Ny;
                int e1 = (nj * Nx + ni) * 2 + s1t[ft];
                m->face_elem[fid * 2 + 0] = e0;
                m->face_elem[fid * 2 + 1] = e1;
                m->face_normal[fid * 2 + 0] = fnx[ft];
                m->face_normal[fid * 2 + 1] = fny[ft];
                m->face_area[fid] = flen[ft];
                int l0 = s0e[ft], l1 = s1e[ft];
                m->elem_faces[e0 * 3 + l0] = fid;
                m->elem_face_side[e0 * 3 + l0] = 0;
                m->elem_faces[e1 * 3 + l1] = fid;
                m->elem_face_side[e1 * 3 + l1] = 1;
                for (int q = 0; q < NE; ++q) {
                    m->face_elem_node[(fid * 2 + 0) * NE + q] = e2e[l0 * NE + q];
                    m->face_elem_node[(fid * 2 + 1) * NE + q] = e2e[l1 * NE + (NE - 1 - q)];
                    m->elem_canon_to_ref[(e0 * 3 + l0) * NE + q] = q;
                    m->elem_canon_to_ref[(e1 * 3 + l1) * NE + q] = NE - 1 - q;
                }
            }
        }
    if (bc[1])
        for (int j = 0; j < Ny; ++j) {
            int fid = (j * Nx + (Nx - 1)) * 3 + 1;
            m->face_elem[fid * 2 + 1] = m->face_elem[fid * 2 + 0];
            m->face_bc_type[fid] = bc[1];
        }
    if (bc[3])
        for (int i = 0; i < Nx; ++i) {
            int fid = ((Ny - 1) * Nx + i) * 3 + 2;
            m->face_elem[fid * 2 + 1] = m->face_elem[fid * 2 + 0];
            m->face_bc_type[fid] = bc[3];
        }
    int nfp = Nx * Ny * 3, extra_mx = bc[0] ? Ny : 0;
    if (bc[0])
        for (int j = 0; j < Ny; ++j) {
            int elem = (j * Nx) * 2 + 1, fid = nfp + j, lf = 2;
            m->face_elem[fid * 2 + 0] = elem;
            m->face_elem[fid * 2 + 1] = elem;
            m->face_normal[fid * 2 + 0] = -1.0;
            m->face_normal[fid * 2 + 1] = 0.0;
            m->face_area[fid] = dy;
            m->face_bc_type[fid] = bc[0];
            m->elem_faces[elem * 3 + lf] = fid;
            m->elem_face_side[elem * 3 + lf] = 0;
            for (int q = 0; q < NE; ++q) {
                int nn = e2e[lf * NE + q];
                m->face_elem_node[(fid * 2 + 0) * NE + q] = nn;
                m->face_elem_node[(fid * 2 + 1) * NE + q] = nn;
                m->elem_canon_to_ref[(elem * 3 + lf) * NE + q] = q;
            }
        }
    if (bc[2])
        for (int i = 0; i < Nx; ++i) {
            int elem = i * 2 + 0, fid = nfp + extra_mx + i, lf = 0;
            m->face_elem[fid * 2 + 0] = elem;
            m->face_elem[fid * 2 + 1] = elem;
            m->face_normal[fid * 2 + 0] = 0.0;
            m->face_normal[fid * 2 + 1] = -1.0;
            m->face_area[fid] = dx;
            m->face_bc_type[fid] = bc[2];
            m->elem_faces[elem * 3 + lf] = fid;
            m->elem_face_side[elem * 3 + lf] = 0;
            for (int q = 0; q < NE; ++q) {
                int nn = e2e[lf * NE + q];
                m->face_elem_node[(fid * 2 + 0) * NE + q] = nn;
                m->face_elem_node[(fid * 2 + 1) * NE + q] = nn;
                m->elem_canon_to_ref[(elem * 3 + lf) * NE + q] = q;
            }
        }
    free(npos);
    free(e2e);
    return 0;
}
