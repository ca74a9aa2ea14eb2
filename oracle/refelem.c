/*
 * refelem.c -- oracle restatement of the equispaced Lagrange reference
 * elements (TEST INFRASTRUCTURE, see oracle.h).
 *
 * Follows src/reference.mojo (tetrahedron) and src/reference_2d.mojo
 * (triangle) step by step: homogeneous barycentric monomial basis,
 * Vandermonde inversion by Gauss-Jordan, exact Dirichlet integration,
 * M, S, D_ref = M^-1 S, Lift_ref = M^-1 L_raw, node_weights.  The
 * arithmetic is carried in long double and rounded once to double, so
 * the operators are the exact ones to FP64 rounding (the reference builds
 * them in FP64 and rounds to FP32).
 *
 * Pinned by: exact rational (Fraction) construction in
 * tests/golden/make_golden.py (all entries, P=1..5), the P=2 closed forms
 * of test/reference_element_test.mojo:133-172 and
 * test/reference_element_2d_test.mojo:111-137, the hand-coded P=2 face
 * table of src/reference.mojo:699-752, and the mat_inv known answers of
 * test/reference_element_test.mojo:175-250.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef long double real;

typedef struct {
    int a0, a1, a2, a3;
    real c;
} mono;

typedef struct {
    mono *m;
    int n, cap;
} poly;

static void poly_push(poly *p, int a0, int a1, int a2, int a3, real c) {
    if (p->n == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 16;
        p->m = (mono *)realloc(p->m, sizeof(mono) * p->cap);
    }
    p->m[p->n].a0 = a0;
    p->m[p->n].a1 = a1;
    p->m[p->n].a2 = a2;
    p->m[p->n].a3 = a3;
    p->m[p->n].c = c;
    p->n++;
}

static void poly_free(poly *p) {
    free(p->m);
    p->m = NULL;
    p->n = p->cap = 0;
}

int orc_num_tet_nodes(int P) { return (P + 1) * (P + 2) * (P + 3) / 6; }
int orc_num_tri_nodes(int P) { return (P + 1) * (P + 2) / 2; }

/* src/reference.mojo:71-78 */
static poly poly_mul(const poly *a, const poly *b) {
    poly out = {0};
    for (int i = 0; i < a->n; ++i)
        for (int j = 0; j < b->n; ++j)
            poly_push(&out, a->m[i].a0 + b->m[j].a0, a->m[i].a1 + b->m[j].a1,
                      a->m[i].a2 + b->m[j].a2, a->m[i].a3 + b->m[j].a3,
                      a->m[i].c * b->m[j].c);
    return out;
}

/* src/reference.mojo:81-85 */
static real factorial(int n) {
    real f = 1.0L;
    for (int k = 2; k <= n; ++k) f *= (real)k;
    return f;
}

/* src/reference.mojo:89-96: Dirichlet integral over the unit tet. */
static real integrate_tet(const poly *p) {
    real s = 0.0L;
    for (int i = 0; i < p->n; ++i) {
        const mono *m = &p->m[i];
        real num = factorial(m->a0) * factorial(m->a1) * factorial(m->a2) *
                   factorial(m->a3);
        s += m->c * num / factorial(m->a0 + m->a1 + m->a2 + m->a3 + 3);
    }
    return s;
}

/* src/reference.mojo:101-108: over the unit triangle. */
static real integrate_tri(const poly *p) {
    real s = 0.0L;
    for (int i = 0; i < p->n; ++i) {
        const mono *m = &p->m[i];
        real num = factorial(m->a0) * factorial(m->a1) * factorial(m->a2);
        s += m->c * num / factorial(m->a0 + m->a1 + m->a2 + 2);
    }
    return s;
}

/* Node exponents, src/reference.mojo:125-200 (tet) and 203-238 (tri).
 * Writes (a0,a1,a2,a3) quadruples into e[4*N]; returns N. */
static int tet_node_exponents(int P, int *e) {
    static const int ea[6] = {0, 1, 2, 0, 1, 2};
    static const int eb[6] = {1, 2, 0, 3, 3, 3};
    int n = 0;
#define PUSH4(x0, x1, x2, x3)                                                  \
    do {                                                                       \
        e[4 * n + 0] = (x0);                                                   \
        e[4 * n + 1] = (x1);                                                   \
        e[4 * n + 2] = (x2);                                                   \
        e[4 * n + 3] = (x3);                                                   \
        ++n;                                                                   \
    } while (0)
    PUSH4(P, 0, 0, 0);
    PUSH4(0, P, 0, 0);
    PUSH4(0, 0, P, 0);
    PUSH4(0, 0, 0, P);
    if (P < 2) return n;
    for (int ed = 0; ed < 6; ++ed) {
        for (int k = 1; k < P; ++k) {
            int x[4];
            for (int d = 0; d < 4; ++d)
                x[d] = (ea[ed] == d) ? P - k : ((eb[ed] == d) ? k : 0);
            PUSH4(x[0], x[1], x[2], x[3]);
        }
    }
    if (P < 3) return n;
    for (int f = 0; f < 4; ++f) {
        int ni[3], c = 0;
        for (int i = 0; i < 4; ++i)
            if (i != f) ni[c++] = i;
        for (int a0v = P - 2; a0v > 0; --a0v) {
            for (int a1v = P - a0v - 1; a1v > 0; --a1v) {
                int a2v = P - a0v - a1v;
                int x[4];
                for (int d = 0; d < 4; ++d)
                    x[d] = (ni[0] == d) ? a0v
                                        : ((ni[1] == d) ? a1v
                                                        : ((ni[2] == d) ? a2v : 0));
                PUSH4(x[0], x[1], x[2], x[3]);
            }
        }
    }
    if (P < 4) return n;
    for (int a0v = P - 2; a0v > 0; --a0v)
        for (int a1v = P - a0v - 1; a1v > 0; --a1v)
            for (int a2v = P - a0v - a1v - 1; a2v > 0; --a2v)
                PUSH4(a0v, a1v, a2v, P - a0v - a1v - a2v);
    return n;
}

static int tri_node_exponents(int P, int *e) {
    static const int ea[3] = {0, 1, 2};
    static const int eb[3] = {1, 2, 0};
    int n = 0;
    PUSH4(P, 0, 0, 0);
    PUSH4(0, P, 0, 0);
    PUSH4(0, 0, P, 0);
    if (P < 2) return n;
    for (int ed = 0; ed < 3; ++ed) {
        for (int k = 1; k < P; ++k) {
            int x[3];
            for (int d = 0; d < 3; ++d)
                x[d] = (ea[ed] == d) ? P - k : ((eb[ed] == d) ? k : 0);
            PUSH4(x[0], x[1], x[2], 0);
        }
    }
    if (P < 3) return n;
    for (int a0v = P - 2; a0v > 0; --a0v)
        for (int a1v = P - a0v - 1; a1v > 0; --a1v)
            PUSH4(a0v, a1v, P - a0v - a1v, 0);
    return n;
#undef PUSH4
}

/* Homogeneous degree-P monomials, src/reference.mojo:250-266. */
static int tet_monomials(int P, int *e) {
    int n = 0;
    for (int a0 = 0; a0 <= P; ++a0)
        for (int a1 = 0; a1 <= P - a0; ++a1)
            for (int a2 = 0; a2 <= P - a0 - a1; ++a2) {
                e[4 * n + 0] = a0;
                e[4 * n + 1] = a1;
                e[4 * n + 2] = a2;
                e[4 * n + 3] = P - a0 - a1 - a2;
                ++n;
            }
    return n;
}

static int tri_monomials(int P, int *e) {
    int n = 0;
    for (int a0 = 0; a0 <= P; ++a0)
        for (int a1 = 0; a1 <= P - a0; ++a1) {
            e[4 * n + 0] = a0;
            e[4 * n + 1] = a1;
            e[4 * n + 2] = P - a0 - a1;
            e[4 * n + 3] = 0;
            ++n;
        }
    return n;
}

/* Gauss-Jordan with partial pivoting, src/reference.mojo:289-331. */
int orc_mat_inv(const double *m, int n, double *inv) {
    int cols = 2 * n;
    double *a = (double *)malloc(sizeof(double) * n * cols);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) a[i * cols + j] = m[i * n + j];
        for (int j = 0; j < n; ++j) a[i * cols + n + j] = (i == j) ? 1.0 : 0.0;
    }
    for (int i = 0; i < n; ++i) {
        int pivot = i;
        double best = 0.0;
        for (int r = i; r < n; ++r) {
            double v = a[r * cols + i];
            if (v < 0.0) v = -v;
            if (v > best) {
                best = v;
                pivot = r;
            }
        }
        if (best == 0.0) {
            free(a);
            return -1;
        }
        if (pivot != i)
            for (int c = 0; c < cols; ++c) {
                double t = a[i * cols + c];
                a[i * cols + c] = a[pivot * cols + c];
                a[pivot * cols + c] = t;
            }
        double pv = a[i * cols + i];
        for (int c = 0; c < cols; ++c) a[i * cols + c] /= pv;
        for (int r = 0; r < n; ++r) {
            if (r == i) continue;
            double f = a[r * cols + i];
            if (f == 0.0) continue;
            for (int c = 0; c < cols; ++c) a[r * cols + c] -= f * a[i * cols + c];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) inv[i * n + j] = a[i * cols + n + j];
    free(a);
    return 0;
}

/* The same elimination carried out in extended precision. */
static int mat_inv_l(const real *m, int n, real *inv) {
    int cols = 2 * n;
    real *a = (real *)malloc(sizeof(real) * n * cols);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) a[i * cols + j] = m[i * n + j];
        for (int j = 0; j < n; ++j) a[i * cols + n + j] = (i == j) ? 1.0 : 0.0;
    }
    for (int i = 0; i < n; ++i) {
        int pivot = i;
        real best = 0.0;
        for (int r = i; r < n; ++r) {
            real v = a[r * cols + i];
            if (v < 0.0) v = -v;
            if (v > best) {
                best = v;
                pivot = r;
            }
        }
        if (best == 0.0) {
            free(a);
            return -1;
        }
        if (pivot != i)
            for (int c = 0; c < cols; ++c) {
                real t = a[i * cols + c];
                a[i * cols + c] = a[pivot * cols + c];
                a[pivot * cols + c] = t;
            }
        real pv = a[i * cols + i];
        for (int c = 0; c < cols; ++c) a[i * cols + c] /= pv;
        for (int r = 0; r < n; ++r) {
            if (r == i) continue;
            real f = a[r * cols + i];
            if (f == 0.0) continue;
            for (int c = 0; c < cols; ++c) a[r * cols + c] -= f * a[i * cols + c];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) inv[i * n + j] = a[i * cols + n + j];
    free(a);
    return 0;
}

/* Basis coefficients C[i][k] = Vinv[k][i], src/reference.mojo:346-387. */
static int basis_coefs(const int *nodes, const int *monos, int N, int P,
                       real *C) {
    real *V = (real *)calloc((size_t)N * N, sizeof(real));
    real *Vi = (real *)malloc(sizeof(real) * N * N);
    real Pf = (real)P;
    for (int j = 0; j < N; ++j) {
        real l[4];
        for (int d = 0; d < 4; ++d) l[d] = (real)nodes[4 * j + d] / Pf;
        for (int k = 0; k < N; ++k) {
            real v = 1.0;
            for (int d = 0; d < 4; ++d)
                for (int r = 0; r < monos[4 * k + d]; ++r) v *= l[d];
            V[j * N + k] = v;
        }
    }
    int rc = mat_inv_l(V, N, Vi);
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < N; ++k) C[i * N + k] = Vi[k * N + i];
    free(V);
    free(Vi);
    return rc;
}

/* src/reference.mojo:390-400: drops zero coefficients. */
static poly coefs_to_poly(const real *C, int i, const int *monos, int N) {
    poly out = {0};
    for (int k = 0; k < N; ++k) {
        real c = C[i * N + k];
        if (c == 0.0) continue;
        poly_push(&out, monos[4 * k], monos[4 * k + 1], monos[4 * k + 2],
                  monos[4 * k + 3], c);
    }
    return out;
}

/* d/dr_k with l0 = 1 - r - s - t, src/reference.mojo:411-427. */
static poly poly_drdx(const poly *p, int k) {
    real dl1 = (k == 0) ? 1.0 : 0.0;
    real dl2 = (k == 1) ? 1.0 : 0.0;
    real dl3 = (k == 2) ? 1.0 : 0.0;
    poly out = {0};
    for (int i = 0; i < p->n; ++i) {
        const mono *m = &p->m[i];
        if (m->a0 > 0)
            poly_push(&out, m->a0 - 1, m->a1, m->a2, m->a3,
                      m->c * (real)m->a0 * -1.0);
        if (m->a1 > 0 && dl1 != 0.0)
            poly_push(&out, m->a0, m->a1 - 1, m->a2, m->a3,
                      m->c * (real)m->a1 * dl1);
        if (m->a2 > 0 && dl2 != 0.0)
            poly_push(&out, m->a0, m->a1, m->a2 - 1, m->a3,
                      m->c * (real)m->a2 * dl2);
        if (m->a3 > 0 && dl3 != 0.0)
            poly_push(&out, m->a0, m->a1, m->a2, m->a3 - 1,
                      m->c * (real)m->a3 * dl3);
    }
    return out;
}

static int lookup_node(const int *nodes, int N, int a0, int a1, int a2,
                       int a3) {
    for (int i = 0; i < N; ++i)
        if (nodes[4 * i] == a0 && nodes[4 * i + 1] == a1 &&
            nodes[4 * i + 2] == a2 && nodes[4 * i + 3] == a3)
            return i;
    return -1;
}

/* src/reference.mojo:453-511 */
static void face_to_element_node(int P, const int *tet, int NPn, int32_t *out) {
    int NFP = orc_num_tri_nodes(P);
    for (int i = 0; i < 4 * NFP; ++i) out[i] = -1;
    for (int f = 0; f < 4; ++f) {
        int vA = (f == 0) ? 1 : 0;
        int vB = (f <= 1) ? 2 : 1;
        int vC = (f <= 2) ? 3 : 2;
        int vv[3] = {vA, vB, vC};
        for (int l = 0; l < 3; ++l) {
            int x[4] = {0, 0, 0, 0};
            x[vv[l]] = P;
            out[f * NFP + l] = lookup_node(tet, NPn, x[0], x[1], x[2], x[3]);
        }
        if (P < 2) continue;
        int pos = 3;
        int eA[3] = {vA, vB, vC}, eB[3] = {vB, vC, vA};
        for (int e = 0; e < 3; ++e)
            for (int k = 1; k < P; ++k) {
                int x[4];
                for (int d = 0; d < 4; ++d)
                    x[d] = (eA[e] == d) ? P - k : ((eB[e] == d) ? k : 0);
                out[f * NFP + pos++] =
                    lookup_node(tet, NPn, x[0], x[1], x[2], x[3]);
            }
        if (P < 3) continue;
        for (int a0v = P - 2; a0v > 0; --a0v)
            for (int a1v = P - a0v - 1; a1v > 0; --a1v) {
                int a2v = P - a0v - a1v;
                int x[4];
                for (int d = 0; d < 4; ++d)
                    x[d] = (vA == d) ? a0v
                                     : ((vB == d) ? a1v : ((vC == d) ? a2v : 0));
                out[f * NFP + pos++] =
                    lookup_node(tet, NPn, x[0], x[1], x[2], x[3]);
            }
    }
}

/* ReferenceElement[P].__init__, src/reference.mojo:562-660 */
int orc_refelem3d(int P, double *D, double *Lift, double *w, int32_t *f2e,
                  double *node_pos) {
    if (P < 1 || P > 6) return -1;
    int NP = orc_num_tet_nodes(P), NFP = orc_num_tri_nodes(P);
    int *tet = (int *)malloc(sizeof(int) * 4 * NP);
    int *tri = (int *)malloc(sizeof(int) * 4 * NFP);
    int *tm = (int *)malloc(sizeof(int) * 4 * NP);
    int *rm = (int *)malloc(sizeof(int) * 4 * NFP);
    if (tet_node_exponents(P, tet) != NP || tri_node_exponents(P, tri) != NFP)
        return -2;
    tet_monomials(P, tm);
    tri_monomials(P, rm);
    if (node_pos)
        for (int i = 0; i < NP; ++i) {
            node_pos[3 * i + 0] = (double)tet[4 * i + 1] / (double)P;
            node_pos[3 * i + 1] = (double)tet[4 * i + 2] / (double)P;
            node_pos[3 * i + 2] = (double)tet[4 * i + 3] / (double)P;
        }
    real *Ct = (real *)malloc(sizeof(real) * NP * NP);
    real *Cr = (real *)malloc(sizeof(real) * NFP * NFP);
    if (basis_coefs(tet, tm, NP, P, Ct) || basis_coefs(tri, rm, NFP, P, Cr))
        return -3;
    poly *phi = (poly *)calloc(NP, sizeof(poly));
    poly *psi = (poly *)calloc(NFP, sizeof(poly));
    for (int i = 0; i < NP; ++i) phi[i] = coefs_to_poly(Ct, i, tm, NP);
    for (int i = 0; i < NFP; ++i) psi[i] = coefs_to_poly(Cr, i, rm, NFP);

    real *M = (real *)malloc(sizeof(real) * NP * NP);
    real *Mi = (real *)malloc(sizeof(real) * NP * NP);
    for (int i = 0; i < NP; ++i)
        for (int j = 0; j < NP; ++j) {
            poly pr = poly_mul(&phi[i], &phi[j]);
            M[i * NP + j] = integrate_tet(&pr);
            poly_free(&pr);
        }
    if (mat_inv_l(M, NP, Mi)) return -4;

    real *S = (real *)malloc(sizeof(real) * 3 * NP * NP);
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < NP; ++i) {
            poly dpi = poly_drdx(&phi[i], k);
            for (int j = 0; j < NP; ++j) {
                poly pr = poly_mul(&dpi, &phi[j]);
                S[(k * NP + i) * NP + j] = integrate_tet(&pr);
                poly_free(&pr);
            }
            poly_free(&dpi);
        }
    if (D)
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < NP; ++i)
                for (int j = 0; j < NP; ++j) {
                    real s = 0.0;
                    for (int ip = 0; ip < NP; ++ip)
                        s += Mi[i * NP + ip] * S[(k * NP + ip) * NP + j];
                    D[(k * NP + i) * NP + j] = (double)s;
                }

    real *Mt = (real *)malloc(sizeof(real) * NFP * NFP);
    for (int i = 0; i < NFP; ++i)
        for (int j = 0; j < NFP; ++j) {
            poly pr = poly_mul(&psi[i], &psi[j]);
            Mt[i * NFP + j] = integrate_tri(&pr);
            poly_free(&pr);
        }
    int32_t *fmap = (int32_t *)malloc(sizeof(int32_t) * 4 * NFP);
    face_to_element_node(P, tet, NP, fmap);
    if (f2e) memcpy(f2e, fmap, sizeof(int32_t) * 4 * NFP);

    real *L = (real *)calloc((size_t)4 * NP * NFP, sizeof(real));
    for (int f = 0; f < 4; ++f)
        for (int l = 0; l < NFP; ++l) {
            int i = fmap[f * NFP + l];
            if (i < 0) return -5;
            for (int m = 0; m < NFP; ++m)
                L[(f * NP + i) * NFP + m] = 2.0 * Mt[l * NFP + m];
        }
    if (Lift)
        for (int f = 0; f < 4; ++f)
            for (int i = 0; i < NP; ++i)
                for (int m = 0; m < NFP; ++m) {
                    real s = 0.0;
                    for (int ip = 0; ip < NP; ++ip)
                        s += Mi[i * NP + ip] * L[(f * NP + ip) * NFP + m];
                    Lift[(f * NP + i) * NFP + m] = (double)s;
                }
    if (w) {
        real vref = 1.0L / 6.0L;
        for (int i = 0; i < NP; ++i) w[i] = (double)(integrate_tet(&phi[i]) / vref);
    }
    for (int i = 0; i < NP; ++i) poly_free(&phi[i]);
    for (int i = 0; i < NFP; ++i) poly_free(&psi[i]);
    free(phi);
    free(psi);
    free(tet);
    free(tri);
    free(tm);
    free(rm);
    free(Ct);
    free(Cr);
    free(M);
    free(Mi);
    free(S);
    free(Mt);
    free(fmap);
    free(L);
    return 0;
}

/* 1D edge mass on [0,1], src/reference_2d.mojo:278-306 */
static void edge_mass(int P, real *Me) {
    int NE = P + 1;
    real *V = (real *)malloc(sizeof(real) * NE * NE);
    real *C = (real *)malloc(sizeof(real) * NE * NE);
    for (int k = 0; k < NE; ++k) {
        real xi = (real)k / (real)P, pw = 1.0;
        for (int j = 0; j < NE; ++j) {
            V[k * NE + j] = pw;
            pw *= xi;
        }
    }
    mat_inv_l(V, NE, C);
    for (int k = 0; k < NE; ++k)
        for (int l = 0; l < NE; ++l) {
            real s = 0.0;
            for (int i = 0; i < NE; ++i)
                for (int j = 0; j < NE; ++j)
                    s += C[i * NE + k] * C[j * NE + l] / (real)(i + j + 1);
            Me[k * NE + l] = s;
        }
    free(V);
    free(C);
}

/* ReferenceElement2D[P].__init__, src/reference_2d.mojo:168-261;
 * edge map src/reference_2d.mojo:87-120. */
int orc_refelem2d(int P, double *D, double *Lift, double *w, int32_t *e2e,
                  double *node_pos) {
    if (P < 1 || P > 8) return -1;
    int NP = orc_num_tri_nodes(P), NE = P + 1;
    int *tri = (int *)malloc(sizeof(int) * 4 * NP);
    int *rm = (int *)malloc(sizeof(int) * 4 * NP);
    if (tri_node_exponents(P, tri) != NP) return -2;
    tri_monomials(P, rm);
    if (node_pos)
        for (int i = 0; i < NP; ++i) {
            node_pos[2 * i + 0] = (double)tri[4 * i + 1] / (double)P;
            node_pos[2 * i + 1] = (double)tri[4 * i + 2] / (double)P;
        }
    real *C = (real *)malloc(sizeof(real) * NP * NP);
    if (basis_coefs(tri, rm, NP, P, C)) return -3;
    poly *phi = (poly *)calloc(NP, sizeof(poly));
    for (int i = 0; i < NP; ++i) phi[i] = coefs_to_poly(C, i, rm, NP);
    real *M = (real *)malloc(sizeof(real) * NP * NP);
    real *Mi = (real *)malloc(sizeof(real) * NP * NP);
    for (int i = 0; i < NP; ++i)
        for (int j = 0; j < NP; ++j) {
            poly pr = poly_mul(&phi[i], &phi[j]);
            M[i * NP + j] = integrate_tri(&pr);
            poly_free(&pr);
        }
    if (mat_inv_l(M, NP, Mi)) return -4;
    if (w)
        for (int i = 0; i < NP; ++i) w[i] = (double)(2.0L * integrate_tri(&phi[i]));
    real *S = (real *)malloc(sizeof(real) * 2 * NP * NP);
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < NP; ++i) {
            poly dpi = poly_drdx(&phi[i], k);
            for (int j = 0; j < NP; ++j) {
                poly pr = poly_mul(&dpi, &phi[j]);
                S[(k * NP + i) * NP + j] = integrate_tri(&pr);
                poly_free(&pr);
            }
            poly_free(&dpi);
        }
    if (D)
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < NP; ++i)
                for (int j = 0; j < NP; ++j) {
                    real s = 0.0;
                    for (int ip = 0; ip < NP; ++ip)
                        s += Mi[i * NP + ip] * S[(k * NP + ip) * NP + j];
                    D[(k * NP + i) * NP + j] = (double)s;
                }
    real *Me = (real *)malloc(sizeof(real) * NE * NE);
    edge_mass(P, Me);
    int32_t *emap = (int32_t *)malloc(sizeof(int32_t) * 3 * NE);
    static const int va[3] = {0, 1, 2}, vb[3] = {1, 2, 0};
    for (int e = 0; e < 3; ++e)
        for (int k = 0; k < NE; ++k) {
            int x[3];
            for (int d = 0; d < 3; ++d)
                x[d] = (va[e] == d) ? P - k : ((vb[e] == d) ? k : 0);
            int i = lookup_node(tri, NP, x[0], x[1], x[2], 0);
            if (i < 0) return -5;
            emap[e * NE + k] = i;
        }
    if (e2e) memcpy(e2e, emap, sizeof(int32_t) * 3 * NE);
    real *L = (real *)calloc((size_t)3 * NP * NE, sizeof(real));
    for (int e = 0; e < 3; ++e)
        for (int l = 0; l < NE; ++l) {
            int i = emap[e * NE + l];
            for (int m = 0; m < NE; ++m) L[(e * NP + i) * NE + m] = Me[l * NE + m];
        }
    if (Lift)
        for (int e = 0; e < 3; ++e)
            for (int i = 0; i < NP; ++i)
                for (int m = 0; m < NE; ++m) {
                    real s = 0.0;
                    for (int ip = 0; ip < NP; ++ip)
                        s += Mi[i * NP + ip] * L[(e * NP + ip) * NE + m];
                    Lift[(e * NP + i) * NE + m] = (double)s;
                }
    for (int i = 0; i < NP; ++i) poly_free(&phi[i]);
    free(phi);
    free(tri);
    free(rm);
    free(C);
    free(M);
    free(Mi);
    free(S);
    free(Me);
    free(emap);
    free(L);
    return 0;
}

/* Nodal differentiation: Dn[k][i][j] = d phi_j / d r_k at node i, from the
 * same barycentric Lagrange basis (src/reference.mojo:346-427,
 * src/reference_2d.mojo:87-128) by evaluating the derivative polynomials
 * (poly_drdx) at the nodes.  The reference has no such matrix: its volume
 * operator is the weak D_ref = M^-1 S; the diagnostics need the strong
 * nodal derivative for the vorticity.  ndim = 2 or 3; Dn[ndim*NP*NP]. */
static real poly_eval(const poly *p, const real *l) {
    real s = 0.0;
    for (int t = 0; t < p->n; ++t) {
        const mono *m = &p->m[t];
        real v = m->c;
        for (int r = 0; r < m->a0; ++r) v *= l[0];
        for (int r = 0; r < m->a1; ++r) v *= l[1];
        for (int r = 0; r < m->a2; ++r) v *= l[2];
        for (int r = 0; r < m->a3; ++r) v *= l[3];
        s += v;
    }
    return s;
}

int orc_nodal_deriv(int ndim, int P, double *Dn) {
    if ((ndim != 2 && ndim != 3) || P < 1 || P > 6) return -1;
    int NP = ndim == 3 ? orc_num_tet_nodes(P) : orc_num_tri_nodes(P);
    int *nodes = (int *)malloc(sizeof(int) * 4 * NP);
    int *monos = (int *)malloc(sizeof(int) * 4 * NP);
    if ((ndim == 3 ? tet_node_exponents(P, nodes) : tri_node_exponents(P, nodes)) != NP) return -2;
    if (ndim == 3)
        tet_monomials(P, monos);
    else
        tri_monomials(P, monos);
    real *C = (real *)malloc(sizeof(real) * NP * NP);
    if (basis_coefs(nodes, monos, NP, P, C)) return -3;
    for (int j = 0; j < NP; ++j) {
        poly phi = coefs_to_poly(C, j, monos, NP);
        for (int k = 0; k < ndim; ++k) {
            poly d = poly_drdx(&phi, k);
            for (int i = 0; i < NP; ++i) {
                real l[4];
                for (int t = 0; t < 4; ++t) l[t] = (real)nodes[4 * i + t] / (real)P;
                Dn[((size_t)k * NP + i) * NP + j] = (double)poly_eval(&d, l);
            }
            poly_free(&d);
        }
        poly_free(&phi);
    }
    free(nodes);
    free(monos);
    free(C);
    return 0;
}
