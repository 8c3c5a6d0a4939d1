/*
 * tracefem_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the P1
 * trace-FEM assembly over level-set-cut Kuhn tetrahedra.
 *
 * Restates, in plain C and in the reference's own operation order, the DROPS
 * routines cited at each function.  It deliberately keeps the reference's
 * data structures (explicit vertex lists, the edge-merge map of
 * MergeCutPolicyCL, per-row builder maps) so that the GPU library, which is
 * designed differently, is checked against an independent restatement.
 *
 * Compile with -ffp-contract=off for parity runs (the reference evaluates
 * every product and sum separately in the order used here).
 */
#include "tracefem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Tables                                                                    */
/* ------------------------------------------------------------------------ */

/* Vertices of the 6 edges of the reference tet, geom/topo.cpp:45-46. */
static const int VOE[6][2] = {{0, 1}, {0, 2}, {1, 2}, {0, 3}, {1, 3}, {2, 3}};

/* Corner offsets (d1,d2,d3) of the 6 Kuhn tets of a brick cell, in the vertex
 * order of geom/builder.cpp:155-175. */
static const int KUHN[6][4][3] = {
    {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}},
    {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {1, 1, 1}},
    {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}},
    {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}},
    {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}},
    {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}}};

static int LAT_B2[10][4];    /* lattice-2 vertices, barycentric x 2 */
static double LAT_B[10][4];  /* lattice-2 vertices, barycentric */
static int LAT_T[8][4];      /* lattice-2 sub-tets */
static int LAT_OF_P2[10];    /* lattice vertex of P2 dof (vertices 0-3, edges 0-5) */

static int RP_SIZE[81], RP_TRI[81][2][3], RP_BND[81];

static double Q2_N[4][3], Q2_W[4];     /* Quad2_2DDataCL, num/discretize.cpp:234-238 */
static double Q5_N[7][3], Q5_W[7];     /* Quad5_2DDataCL, num/discretize.cpp:263-296 */
static double V5_N[15][4];             /* Quad5DataCL nodes, num/discretize.cpp:181-203 */
static double V5_C1, V5_C2;            /* Quad5CL::quad weights, num/discretize.tpp:754-762 */

static int g_inited = 0;

static int lattice_vertex_index(int i, int j, int k)
{
    /* PrincipalLatticeCL::vertex_index, geom/principallattice.h:62-64 */
    return (i > 0 ? ((i + 2) * (i + 1) * i) / 6 : 0) + (j > 0 ? ((j + 1) * j) / 2 : 0) + k;
}

static int in_ref_tetra(const int v[3], int n) { return n >= v[0] && v[0] >= v[1] && v[1] >= v[2]; }

static void init_tables(void)
{
    if (g_inited)
        return;
    /* PrincipalLatticeCL ctor, geom/principallattice.cpp:33-44 (n = 2) */
    int idx = 0;
    for (int x = 0; x <= 2; ++x)
        for (int y = 0; y <= x; ++y)
            for (int z = 0; z <= y; ++z) {
                LAT_B2[idx][0] = 2 - x;
                LAT_B2[idx][1] = x - y;
                LAT_B2[idx][2] = y - z;
                LAT_B2[idx][3] = z;
                for (int c = 0; c < 4; ++c)
                    LAT_B[idx][c] = (double)LAT_B2[idx][c] / 2.0;
                ++idx;
            }
    /* create_tetras(0, 2), geom/principallattice.cpp:46-80 */
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int nt = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k <= j; ++k) {
                int t[4][3];
                t[0][0] = i; t[0][1] = j; t[0][2] = k;
                for (int l = 0; l < 6; ++l) {
                    const int *p = perm[l];
                    memcpy(t[1], t[0], sizeof t[0]);
                    ++t[1][p[0]];
                    if (!in_ref_tetra(t[1], 2)) continue;
                    memcpy(t[2], t[1], sizeof t[0]);
                    ++t[2][p[1]];
                    if (!in_ref_tetra(t[2], 2)) continue;
                    memcpy(t[3], t[2], sizeof t[0]);
                    ++t[3][p[2]];
                    if (!in_ref_tetra(t[3], 2)) continue;
                    for (int q = 0; q < 4; ++q)
                        LAT_T[nt][q] = lattice_vertex_index(t[q][0], t[q][1], t[q][2]);
                    ++nt;
                }
            }
    /* P2 dof -> lattice vertex (p2_dof_on_lattice_2, geom/principallattice.cpp:103-106),
     * derived from the barycentric coordinates */
    for (int l = 0; l < 10; ++l) {
        int nz = 0, a = -1, b = -1;
        for (int c = 0; c < 4; ++c)
            if (LAT_B2[l][c]) {
                if (a < 0) a = c; else b = c;
                ++nz;
            }
        if (nz == 1)
            LAT_OF_P2[a] = l;
        else
            for (int e = 0; e < 6; ++e)
                if (VOE[e][0] == a && VOE[e][1] == b)
                    LAT_OF_P2[4 + e] = l;
    }
    /* RefPatchBuilderCL<3>::StaticInit/assign, geom/reftetracut.cpp:40-60,
     * with SignTraitsCL<3>::assign/compute_cuts, geom/signtraits.h:112-139 */
    for (int i = 0; i < 81; ++i) {
        RP_SIZE[i] = 0;
        RP_BND[i] = 0;
    }
    int ls[4];
    for (ls[0] = -1; ls[0] < 2; ++ls[0])
        for (ls[1] = -1; ls[1] < 2; ++ls[1])
            for (ls[2] = -1; ls[2] < 2; ++ls[2])
                for (ls[3] = -1; ls[3] < 2; ++ls[3]) {
                    if (!ls[0] && !ls[1] && !ls[2] && !ls[3])
                        continue;
                    const int pid = ls[0] + 3 * ls[1] + 9 * ls[2] + 27 * ls[3] + 40;
                    int nrv = 0, nr = 0, rep[4] = {0, 0, 0, 0};
                    const int sum = ls[0] + ls[1] + ls[2] + ls[3];
                    if (abs(sum) != 4) {
                        for (int v = 0; v < 4; ++v)
                            if (ls[v] == 0) rep[nrv++] = v;
                        nr = nrv;
                        for (int e = 0; e < 6; ++e)
                            if (ls[VOE[e][0]] * ls[VOE[e][1]] == -1) rep[nr++] = e + 4;
                    }
                    const int ntri = nr >= 3 ? nr - 2 : 0;
                    RP_SIZE[pid] = ntri;
                    for (int s = 0; s < ntri; ++s)
                        for (int j = 0; j < 3; ++j)
                            RP_TRI[pid][s][j] = rep[s + j];
                    RP_BND[pid] = (nrv == 3);
                }
    /* Quad2_2DDataCL */
    memset(Q2_N, 0, sizeof Q2_N);
    Q2_N[0][0] = 1.; Q2_N[1][1] = 1.; Q2_N[2][2] = 1.;
    Q2_N[3][0] = Q2_N[3][1] = Q2_N[3][2] = 1. / 3.;
    Q2_W[0] = Q2_W[1] = Q2_W[2] = 1. / 24.;
    Q2_W[3] = 0.375;
    /* Quad5_2DDataCL */
    const double s15 = sqrt(15.0);
    Q5_N[0][0] = Q5_N[0][1] = Q5_N[0][2] = 1. / 3.;
    {
        const double A1 = (6.0 - s15) / 21.0, B1 = (9.0 + 2.0 * s15) / 21.0;
        const double A2 = (6.0 + s15) / 21.0, B2 = (9.0 - 2.0 * s15) / 21.0;
        const double n1[3][3] = {{A1, A1, B1}, {A1, B1, A1}, {B1, A1, A1}};
        const double n2[3][3] = {{A2, A2, B2}, {A2, B2, A2}, {B2, A2, A2}};
        memcpy(Q5_N[1], n1, sizeof n1);
        memcpy(Q5_N[4], n2, sizeof n2);
    }
    Q5_W[0] = 9. / 80.;
    Q5_W[1] = Q5_W[2] = Q5_W[3] = (155. - s15) / 2400.;
    Q5_W[4] = Q5_W[5] = Q5_W[6] = (155. + s15) / 2400.;
    /* Quad5DataCL nodes */
    for (int c = 0; c < 4; ++c) V5_N[0][c] = 0.25;
    {
        const double A1 = (7.0 - s15) / 34.0, B1 = (13.0 + 3.0 * s15) / 34.0;
        const double A2 = (7.0 + s15) / 34.0, B2 = (13.0 - 3.0 * s15) / 34.0;
        const double A3 = (10.0 - 2.0 * s15) / 40.0, B3 = (10.0 + 2.0 * s15) / 40.0;
        for (int q = 0; q < 4; ++q)
            for (int c = 0; c < 4; ++c) {
                V5_N[1 + q][c] = (c == 3 - q) ? B1 : A1;
                V5_N[5 + q][c] = (c == 3 - q) ? B2 : A2;
            }
        const double n3[6][4] = {{A3, A3, B3, B3}, {A3, B3, A3, B3}, {A3, B3, B3, A3},
                                 {B3, A3, A3, B3}, {B3, A3, B3, A3}, {B3, B3, A3, A3}};
        memcpy(V5_N[9], n3, sizeof n3);
    }
    V5_C1 = (2665.0 + 14.0 * s15) / 226800.0;
    V5_C2 = (2665.0 - 14.0 * s15) / 226800.0;
    g_inited = 1;
}

void orc_lattice_tables(int32_t bary2[10][4], int32_t tets[8][4])
{
    init_tables();
    for (int i = 0; i < 10; ++i)
        for (int c = 0; c < 4; ++c) bary2[i][c] = LAT_B2[i][c];
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 4; ++c) tets[i][c] = LAT_T[i][c];
}

void orc_quad_tables(double q5n[7][3], double q5w[7], double v5n[15][4], double v5c[2])
{
    init_tables();
    memcpy(q5n, Q5_N, sizeof Q5_N);
    memcpy(q5w, Q5_W, sizeof Q5_W);
    memcpy(v5n, V5_N, sizeof V5_N);
    v5c[0] = V5_C1;
    v5c[1] = V5_C2;
}

void orc_refpatch_table(int32_t size[81], int32_t tri[81][2][3], int32_t bnd[81])
{
    init_tables();
    for (int i = 0; i < 81; ++i) {
        size[i] = RP_SIZE[i];
        bnd[i] = RP_BND[i];
        for (int s = 0; s < 2; ++s)
            for (int j = 0; j < 3; ++j) tri[i][s][j] = s < RP_SIZE[i] ? RP_TRI[i][s][j] : -1;
    }
}

/* ------------------------------------------------------------------------ */
/* Mesh                                                                      */
/* ------------------------------------------------------------------------ */

void orc_vertex_coord(const orc_mesh_t *m, int64_t i1, int64_t i2, int64_t i3, double x[3])
{
    /* geom/builder.cpp:108-120: off_k = 1/n_k * e_k; x = orig + i1 off1 + i2 off2 + i3 off3 */
    for (int c = 0; c < 3; ++c) {
        const double o1 = 1.0 / (double)m->n[0] * m->e[0][c];
        const double o2 = 1.0 / (double)m->n[1] * m->e[1][c];
        const double o3 = 1.0 / (double)m->n[2] * m->e[2][c];
        x[c] = ((m->orig[c] + (double)i1 * o1) + (double)i2 * o2) + (double)i3 * o3;
    }
}

void orc_halfgrid_coord(const orc_mesh_t *m, int64_t j1, int64_t j2, int64_t j3, double x[3])
{
    /* P2 node = vertex or barycentre of a Kuhn edge (BaryCenter, misc/container.h:395-400) */
    double a[3], b[3];
    orc_vertex_coord(m, j1 / 2, j2 / 2, j3 / 2, a);
    if (!(j1 & 1) && !(j2 & 1) && !(j3 & 1)) {
        memcpy(x, a, sizeof a);
        return;
    }
    orc_vertex_coord(m, (j1 + 1) / 2, (j2 + 1) / 2, (j3 + 1) / 2, b);
    for (int c = 0; c < 3; ++c) x[c] = .5 * (a[c] + b[c]);
}

typedef struct {
    int64_t i[3];   /* cell index */
    int k;          /* Kuhn tet type */
    double X[4][3]; /* vertex coordinates */
    int64_t vid[4]; /* global vertex ids */
    double ls[10];  /* level set on lattice-2 vertices */
} tet_t;

/* level-set values only (enough to classify) */
static void load_tet_ls(const orc_mesh_t *m, const double *phi, int64_t t, tet_t *T)
{
    const int64_t n1 = m->n[0], n2 = m->n[1];
    const int64_t cell = t / 6;
    T->k = (int)(t % 6);
    T->i[0] = cell % n1;
    T->i[1] = (cell / n1) % n2;
    T->i[2] = cell / (n1 * n2);
    const int64_t h1 = 2 * n1 + 1, h2 = 2 * n2 + 1;
    for (int l = 0; l < 10; ++l) {
        int64_t j[3];
        for (int c = 0; c < 3; ++c) {
            j[c] = 2 * T->i[c];
            for (int q = 0; q < 4; ++q) j[c] += LAT_B2[l][q] * KUHN[T->k][q][c];
        }
        /* evaluate_on_vertexes(LocalP2CL, lattice-2) = P2 nodal values, num/lattice-eval.h:41-62 */
        T->ls[l] = phi[(j[2] * h2 + j[1]) * h1 + j[0]];
    }
}

/* level-set values, vertex coordinates and vertex ids */
static void load_tet(const orc_mesh_t *m, const double *phi, int64_t t, tet_t *T)
{
    const int64_t n1 = m->n[0], n2 = m->n[1];
    load_tet_ls(m, phi, t, T);
    for (int q = 0; q < 4; ++q) {
        const int *d = KUHN[T->k][q];
        const int64_t v1 = T->i[0] + d[0], v2 = T->i[1] + d[1], v3 = T->i[2] + d[2];
        orc_vertex_coord(m, v1, v2, v3, T->X[q]);
        T->vid[q] = (v3 * (n2 + 1) + v2) * (n1 + 1) + v1;
    }
}

static int sgn(double d) { return d > 0. ? 1 : (d < 0. ? -1 : 0); } /* misc/utils.h:810-813 */

/* equalSigns, geom/subtriangulation.tpp:32-36 */
static int equal_signs(const double *f, int n)
{
    int sum = 0;
    for (int i = 0; i < n; ++i) sum += sgn(f[i]);
    return abs(sum) == n;
}

static int has_zero(const double *f, int n)
{
    for (int i = 0; i < n; ++i)
        if (f[i] == 0.) return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Interface patch (SPatchCL<3>::make_patch<MergeCutPolicyCL>)               */
/* ------------------------------------------------------------------------ */

#define MAXV 40
#define MAXF 16

typedef struct {
    int nf, nv;
    int facet[MAXF][3];
    int bnd[MAXF];
    double vb[MAXV][4];
    double vw[MAXV][3];
    double absdet[MAXF];
    double normal[MAXF][3];
    int deficient[MAXF];
} patch_t;

/* geom/subtriangulation.tpp:194-231 (make_patch), :149-177 (make_sub_facet),
 * geom/subtriangulation.h:514-527 (MergeCutPolicyCL), misc/container.h:403-410 (ConvexComb);
 * nsub lattice tets sub[] over lattice vertices with barycentric coordinates B[] and values ls[] */
static int make_patch_on(const double *ls, int nsub, const int (*sub)[4], const double (*B)[4], patch_t *p)
{
    int ekey[MAXV][2];
    double cutb[MAXV][4];
    int ncut = 0;
    int copied[10];
    double zerob[10][4];
    int nzero = 0;
    int zf[3 * MAXF], zj[3 * MAXF], nzu = 0;
    for (int i = 0; i < 10; ++i) copied[i] = -1;
    p->nf = 0;
    for (int lt = 0; lt < nsub; ++lt) {
        const int *v = sub[lt];
        double lls[4];
        int s[4];
        for (int q = 0; q < 4; ++q) {
            lls[q] = ls[v[q]];
            s[q] = sgn(lls[q]);
        }
        if (!s[0] && !s[1] && !s[2] && !s[3])
            return ORC_E_ZERO_PATTERN; /* RefPatchCL::instance, geom/reftetracut.h:148-155 */
        const int pid = s[0] + 3 * s[1] + 9 * s[2] + 27 * s[3] + 40;
        for (int tr = 0; tr < RP_SIZE[pid]; ++tr) {
            int out[3];
            for (int j = 0; j < 3; ++j) {
                const int rep = RP_TRI[pid][tr][j];
                if (rep < 4) {
                    const int lv = v[rep];
                    if (copied[lv] < 0) {
                        copied[lv] = nzero;
                        memcpy(zerob[nzero], B[lv], sizeof zerob[0]);
                        ++nzero;
                    }
                    out[j] = copied[lv];
                    zf[nzu] = p->nf;
                    zj[nzu] = j;
                    ++nzu;
                } else {
                    const int e = rep - 4, v0 = VOE[e][0], v1 = VOE[e][1];
                    const int a = v[v0], b = v[v1];
                    const int ka = a < b ? a : b, kb = a < b ? b : a;
                    int found = -1;
                    for (int c = 0; c < ncut; ++c)
                        if (ekey[c][0] == ka && ekey[c][1] == kb) { found = c; break; }
                    if (found < 0) {
                        const double alpha = lls[v0] / (lls[v0] - lls[v1]);
                        for (int c = 0; c < 4; ++c)
                            cutb[ncut][c] = (1.0 - alpha) * B[a][c] + alpha * B[b][c];
                        ekey[ncut][0] = ka;
                        ekey[ncut][1] = kb;
                        found = ncut++;
                    }
                    out[j] = found;
                }
            }
            memcpy(p->facet[p->nf], out, sizeof out);
            p->bnd[p->nf] = RP_BND[pid];
            ++p->nf;
        }
    }
    for (int u = 0; u < nzu; ++u) p->facet[zf[u]][zj[u]] += ncut;
    p->nv = ncut + nzero;
    memcpy(p->vb, cutb, sizeof(double) * 4 * ncut);
    memcpy(p->vb + ncut, zerob, sizeof(double) * 4 * nzero);
    return ORC_OK;
}

/* the patch of the P2 level set on lattice 2 (its eight sub-tets) */
static int make_patch(const double ls[10], patch_t *p)
{
    return make_patch_on(ls, 8, (const int (*)[4])LAT_T, (const double (*)[4])LAT_B, p);
}

/* the patch on lattice 1: the tet itself with the level set at its four vertices (lattice-1 vertex order
 * is the tet's vertex order, geom/principallattice.cpp:33-44) */
static int make_patch_l1(const double ls4[4], patch_t *p)
{
    static const int sub1[1][4] = {{0, 1, 2, 3}};
    static const double B1[4][4] = {{1., 0., 0., 0.}, {0., 1., 0., 0.}, {0., 0., 1., 0.}, {0., 0., 0., 1.}};
    return make_patch_on(ls4, 1, sub1, B1, p);
}

/* Bary2WorldCoordCL: SMatrixCL<3,4> * b, geom/simplex.h:867-881, misc/container.h:966-973 */
static void bary_to_world(const double X[4][3], const double b[4], double w[3])
{
    for (int r = 0; r < 3; ++r) {
        double s = 0.0;
        for (int i = 0; i < 4; ++i) s += X[i][r] * b[i];
        w[r] = s;
    }
}

/* SPatchCL::compute_normals, geom/subtriangulation.cpp:140-170 with
 * QRDecompCL<3,2>::prepare_solve / apply_Q / Determinant_R, misc/container.h:1210-1290 */
static void patch_geometry(const double X[4][3], patch_t *p)
{
    for (int v = 0; v < p->nv; ++v) bary_to_world(X, p->vb[v], p->vw[v]);
    for (int f = 0; f < p->nf; ++f) {
        double a[3][2], d[2], beta[2];
        const double *w0 = p->vw[p->facet[f][0]];
        for (int i = 1; i < 3; ++i) {
            const double *wi = p->vw[p->facet[f][i]];
            for (int r = 0; r < 3; ++r) a[r][i - 1] = wi[r] - w0[r];
        }
        int deficient = 0;
        for (int j = 0; j < 2 && !deficient; ++j) {
            double sigma = 0.;
            for (int i = j; i < 3; ++i) sigma += a[i][j] * a[i][j];
            if (sigma == 0.) { deficient = 1; break; }
            d[j] = (a[j][j] < 0 ? 1. : -1.) * sqrt(sigma);
            beta[j] = 1. / (d[j] * a[j][j] - sigma);
            a[j][j] -= d[j];
            for (int k = j + 1; k < 2; ++k) {
                double sp = 0.;
                for (int i = j; i < 3; ++i) sp += a[i][j] * a[i][k];
                sp *= beta[j];
                for (int i = j; i < 3; ++i) a[i][k] += a[i][j] * sp;
            }
        }
        p->deficient[f] = deficient;
        if (deficient) {
            p->absdet[f] = 0.;
            p->normal[f][0] = p->normal[f][1] = p->normal[f][2] = 0.;
            continue;
        }
        double det = 1.0;
        det *= d[0];
        det *= d[1];
        p->absdet[f] = fabs(det);
        double t[3] = {0., 0., 1.};
        for (int j = 1; j >= 0; --j) {
            double sp = 0.;
            for (int i = j; i < 3; ++i) sp += a[i][j] * t[i];
            sp *= beta[j];
            for (int i = j; i < 3; ++i) t[i] += a[i][j] * sp;
        }
        memcpy(p->normal[f], t, sizeof t);
    }
}

/* make_CompositeQuadDomainCodim1<QuadData, Codim1Absdet, 3>, num/quadrature.tpp:265-312:
 * node = sum_j vertex(facet[j]) * Node[i][j]; weight = (0.5 if boundary facet) * absdet * W[i] */
static int composite_nodes(const patch_t *p, int nn, const double N[][3], const double *W,
                           double (*node)[4], double *weight)
{
    int k = 0;
    for (int f = 0; f < p->nf; ++f) {
        const double *v0 = p->vb[p->facet[f][0]], *v1 = p->vb[p->facet[f][1]], *v2 = p->vb[p->facet[f][2]];
        double a = p->bnd[f] ? 0.5 : 1.;
        a *= p->absdet[f];
        for (int i = 0; i < nn; ++i, ++k) {
            for (int c = 0; c < 4; ++c) {
                double s = v0[c] * N[i][0];
                s += v1[c] * N[i][1];
                s += v2[c] * N[i][2];
                node[k][c] = s;
            }
            weight[k] = a * W[i];
        }
    }
    return k;
}

/* P1DiscCL::GetGradients(Point3DCL[4], double&, const TetraCL&), num/discretize.h:1724-1746 */
static double p1_gradients(const double X[4][3], double H[4][3])
{
    double M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[j][i] = X[i + 1][j] - X[0][j];
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
                     - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
                     + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    H[1][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    H[1][1] = (M[2][1] * M[0][2] - M[0][1] * M[2][2]) / det;
    H[1][2] = (M[0][1] * M[1][2] - M[1][1] * M[0][2]) / det;
    H[2][0] = (M[2][0] * M[1][2] - M[1][0] * M[2][2]) / det;
    H[2][1] = (M[0][0] * M[2][2] - M[2][0] * M[0][2]) / det;
    H[2][2] = (M[1][0] * M[0][2] - M[0][0] * M[1][2]) / det;
    H[3][0] = (M[1][0] * M[2][1] - M[2][0] * M[1][1]) / det;
    H[3][1] = (M[2][0] * M[0][1] - M[0][0] * M[2][1]) / det;
    H[3][2] = (M[0][0] * M[1][1] - M[1][0] * M[0][1]) / det;
    for (int c = 0; c < 3; ++c) H[0][c] = ((-H[1][c]) - H[2][c]) - H[3][c];
    return det;
}

/* TetraCL::GetVolume, geom/simplex.cpp:522-531 */
static double tet_volume(const double X[4][3])
{
    double v[3][3];
    for (int i = 1; i < 4; ++i)
        for (int c = 0; c < 3; ++c) v[i - 1][c] = X[i][c] - X[0][c];
    return fabs(v[0][0] * (v[1][1] * v[2][2] - v[1][2] * v[2][1])
              - v[0][1] * (v[1][0] * v[2][2] - v[1][2] * v[2][0])
              + v[0][2] * (v[1][0] * v[2][1] - v[1][1] * v[2][0])) / 6.0;
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* FE_P2CL::H0..H9 at barycentric p, num/fe.h:287-296, summed as in
 * LinearCombinationCL::do_it, num/fe.h:44-53 */
static double p2_eval(const double c[10], const double p[4])
{
    const double h[10] = {p[0] * (2. * p[0] - 1.), p[1] * (2. * p[1] - 1.), p[2] * (2. * p[2] - 1.),
                          p[3] * (2. * p[3] - 1.), 4. * p[0] * p[1],         4. * p[0] * p[2],
                          4. * p[1] * p[2],         4. * p[0] * p[3],         4. * p[1] * p[3],
                          4. * p[2] * p[3]};
    double s = h[0] * c[0];
    for (int d = 1; d < 10; ++d) s += h[d] * c[d];
    return s;
}

/* Normalised gradient of the P2 level set at barycentric point b.
 * grad phi_h = sum_k G_k grad(lambda_k) with G_k = v_k (4 b_k - 1) + 4 sum_{e=(k,l)} v_e b_l
 * (the normal used by LocalAssembler::build3DNormal, surfactant/ifacetransp.h:1073-1076). */
static void p2_normal(const double v[10], const double b[4], const double G1[4][3], double n[3])
{
    double G[4];
    for (int k = 0; k < 4; ++k) G[k] = v[k] * (4.0 * b[k] - 1.0);
    for (int e = 0; e < 6; ++e) {
        const int a = VOE[e][0], c = VOE[e][1];
        G[a] += 4.0 * (v[4 + e] * b[c]);
        G[c] += 4.0 * (v[4 + e] * b[a]);
    }
    double g[3];
    for (int d = 0; d < 3; ++d) g[d] = ((G[0] * G1[0][d] + G[1] * G1[1][d]) + G[2] * G1[2][d]) + G[3] * G1[3][d];
    const double nrm = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    for (int d = 0; d < 3; ++d) n[d] = g[d] / nrm;
}

static double rhs_fun(int kind, const double *param, const double x[3])
{
    switch (kind) {
    case ORC_RHS_3XY_MINUS_X3: return 3.0 * x[0] * x[1] - x[0] * x[0] * x[0];
    case ORC_RHS_CONST: return param ? param[0] : 0.;
    default: return 0.;
    }
}

/* All per-tet local quantities of one cut tet with a non-empty patch. */
/* LocalNormalLaplaceBulkP1CL::setup, surfactant/ifacetransp.h:2005-2051 (the same formula for the
 * InterfaceCommonDataP1CL and the NarrowBandCommonDataP1CL overloads): D * Quad5(U_i U_j, |det|/6),
 * U_i = n . grad(phi_i) on the whole tet */
static void local_stab(const tet_t *T, const double coef[3], double stab[16])
{
    double G1[4][3];
    const double det = p1_gradients(T->X, G1);
    {
        double v[10];
        for (int d = 0; d < 10; ++d) v[d] = T->ls[LAT_OF_P2[d]];
        double U[4][15];
        for (int q = 0; q < 15; ++q) {
            double n[3];
            p2_normal(v, V5_N[q], G1, n);
            for (int i = 0; i < 4; ++i) U[i][q] = dot3(n, G1[i]);
        }
        const double ad = fabs(det) / 6;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j <= i; ++j) {
                double f[15];
                for (int q = 0; q < 15; ++q) f[q] = U[i][q] * U[j][q];
                const double val = ad * (((8. / 405. * f[0] + V5_C1 * (((f[1] + f[2]) + f[3]) + f[4]))
                                          + V5_C2 * (((f[5] + f[6]) + f[7]) + f[8]))
                                         + 5. / 567. * (((((f[9] + f[10]) + f[11]) + f[12]) + f[13]) + f[14]));
                stab[i * 4 + j] = stab[j * 4 + i] = coef[2] * val;
            }
    }
}

static void local_quantities(const tet_t *T, const patch_t *p, const double coef[3], int rhs_kind,
                             const double *rhs_param, orc_local_t *o)
{
    double node[MAXF * 7][4], w[MAXF * 7];
    o->nfacet = p->nf;
    o->nvert = p->nv;
    /* ---- LocalInterfaceMassP1CL::setup, surfactant/ifacetransp.h:893-903 (Quad2 composite) ---- */
    int nq = composite_nodes(p, 4, Q2_N, Q2_W, node, w);
    for (int i = 0; i < 4; ++i) {
        double s = 0.;
        for (int k = 0; k < nq; ++k) s += w[k] * (node[k][i] * node[k][i]);
        o->mass[i * 4 + i] = s;
        for (int j = 0; j < i; ++j) {
            double t = 0.;
            for (int k = 0; k < nq; ++k) t += w[k] * (node[k][j] * node[k][i]);
            o->mass[i * 4 + j] = o->mass[j * 4 + i] = coef[0] * t;
        }
    }
    /* interface numbering integral of phi_i^2, misc/problem.cpp:452-456, 487-492 */
    for (int i = 0; i < 4; ++i) {
        double c[10] = {0};
        c[i] = 1.;
        for (int e = 0; e < 6; ++e)
            if (VOE[e][0] == i || VOE[e][1] == i) c[4 + e] = 0.25;
        double s = 0.;
        for (int k = 0; k < nq; ++k) s += w[k] * p2_eval(c, node[k]);
        o->hat_sq[i] = s;
    }
    /* area: quad_2D(1) on the Quad5 composite rule, as tests/principallattice.cpp:318-323 */
    {
        double n5[MAXF * 7][4], w5[MAXF * 7];
        const int n5q = composite_nodes(p, 7, Q5_N, Q5_W, n5, w5);
        double s = 0.;
        for (int k = 0; k < n5q; ++k) s += w5[k] * 1.;
        o->area = s;
    }
    /* ---- LocalLaplaceBeltramiP1CL::setup, surfactant/ifacetransp.h:1671-1686 ---- */
    double G1[4][3];
    const double det = p1_gradients(T->X, G1);
    {
        double q[4][MAXF][3];
        for (int i = 0; i < 4; ++i)
            for (int f = 0; f < p->nf; ++f) {
                const double d = dot3(G1[i], p->normal[f]);
                for (int c = 0; c < 3; ++c) q[i][f][c] = G1[i][c] - d * p->normal[f][c];
            }
        const double fac = coef[1] * 0.5;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j <= i; ++j) {
                double s = 0.;
                for (int f = 0; f < p->nf; ++f) s += dot3(q[i][f], q[j][f]) * p->absdet[f];
                o->lb[i * 4 + j] = o->lb[j * 4 + i] = fac * s;
            }
    }
    local_stab(T, coef, o->stab);
    /* ---- LocalVectorP1CL::setup, surfactant/ifacetransp.h:785-797 (Quad5 composite) ---- */
    if (rhs_kind != ORC_RHS_NONE) {
        nq = composite_nodes(p, 7, Q5_N, Q5_W, node, w);
        double qf[MAXF * 7];
        for (int k = 0; k < nq; ++k) {
            double x[3];
            bary_to_world(T->X, node[k], x);
            qf[k] = rhs_fun(rhs_kind, rhs_param, x);
        }
        for (int i = 0; i < 4; ++i) {
            double s = 0.;
            for (int k = 0; k < nq; ++k) s += w[k] * (qf[k] * node[k][i]);
            o->rhs[i] = s;
        }
    } else {
        o->rhs[0] = o->rhs[1] = o->rhs[2] = o->rhs[3] = 0.;
    }
}

/* ------------------------------------------------------------------------ */
/* Public entry points                                                       */
/* ------------------------------------------------------------------------ */

static int check_mesh(const orc_mesh_t *m)
{
    return m && m->n[0] > 0 && m->n[1] > 0 && m->n[2] > 0;
}

/* Per-cell classification of the 6 Kuhn tets (equalSigns on the 10 lattice values,
 * geom/subtriangulation.tpp:32-36); bit k of mask[cell] = tet k is cut, bit 6 = a cut
 * tet has an exact-zero node.  Independent per cell (OpenMP in the baseline build). */
static void cell_masks(const orc_mesh_t *m, const double *phi, uint8_t *mask)
{
    const int64_t n1 = m->n[0], n2 = m->n[1], n3 = m->n[2];
    const int64_t h1 = 2 * n1 + 1, h2 = 2 * n2 + 1;
    int64_t off[6][10];
    for (int k = 0; k < 6; ++k)
        for (int l = 0; l < 10; ++l) {
            int64_t o[3] = {0, 0, 0};
            for (int c = 0; c < 3; ++c)
                for (int q = 0; q < 4; ++q) o[c] += LAT_B2[l][q] * KUHN[k][q][c];
            off[k][l] = (o[2] * h2 + o[1]) * h1 + o[0];
        }
    const int64_t ncell = n1 * n2 * n3;
#pragma omp parallel for schedule(static)
    for (int64_t cell = 0; cell < ncell; ++cell) {
        const int64_t i1 = cell % n1, i2 = (cell / n1) % n2, i3 = cell / (n1 * n2);
        const double *base = phi + ((2 * i3) * h2 + 2 * i2) * h1 + 2 * i1;
        uint8_t mk = 0;
        for (int k = 0; k < 6; ++k) {
            int sum = 0, zero = 0;
            for (int l = 0; l < 10; ++l) {
                const double v = base[off[k][l]];
                sum += sgn(v);
                zero |= v == 0.;
            }
            if (abs(sum) != 10) {
                mk |= (uint8_t)(1u << k);
                if (zero) mk |= 64u;
            }
        }
        mask[cell] = mk;
    }
}

int orc_classify(const orc_mesh_t *m, const double *phi, int zero_policy,
                 int64_t cap, int64_t *ncut, int64_t *cut_tets)
{
    init_tables();
    if (!check_mesh(m) || !phi || !ncut) return ORC_E_ARG;
    const int64_t ncell = (int64_t)m->n[0] * m->n[1] * m->n[2];
    uint8_t *mask = (uint8_t *)malloc(ncell > 0 ? ncell : 1);
    if (!mask) return ORC_E_NOMEM;
    cell_masks(m, phi, mask);
    int64_t c = 0;
    int err = ORC_OK;
    for (int64_t cell = 0; cell < ncell && !err; ++cell) {
        const uint8_t mk = mask[cell];
        if (!(mk & 63)) continue;
        if (zero_policy == ORC_ZERO_THROW && (mk & 64)) { err = ORC_E_ZERO_NODE; break; }
        for (int k = 0; k < 6; ++k) {
            if (!(mk & (1u << k))) continue;
            if (cut_tets) {
                if (c >= cap) { err = ORC_E_CAPACITY; break; }
                cut_tets[c] = cell * 6 + k;
            }
            ++c;
        }
    }
    free(mask);
    if (err) return err;
    *ncut = c;
    return ORC_OK;
}

int orc_tet_local(const orc_mesh_t *m, const double *phi, int64_t tet, const double coef[3],
                  int rhs_kind, const double *rhs_param, int zero_policy, orc_local_t *out)
{
    init_tables();
    if (!check_mesh(m) || !phi || !out) return ORC_E_ARG;
    memset(out, 0, sizeof *out);
    tet_t T;
    load_tet(m, phi, tet, &T);
    if (equal_signs(T.ls, 10)) return ORC_OK;
    if (zero_policy == ORC_ZERO_THROW && has_zero(T.ls, 10)) return ORC_E_ZERO_NODE;
    patch_t p;
    int err = make_patch(T.ls, &p);
    if (err) return err;
    if (p.nf == 0) return ORC_OK;
    patch_geometry(T.X, &p);
    local_quantities(&T, &p, coef, rhs_kind, rhs_param, out);
    return ORC_OK;
}

int orc_tet_patch(const orc_mesh_t *m, const double *phi, int64_t tet, int zero_policy, orc_patch_t *out)
{
    init_tables();
    if (!check_mesh(m) || !phi || !out) return ORC_E_ARG;
    memset(out, 0, sizeof *out);
    tet_t T;
    load_tet(m, phi, tet, &T);
    memcpy(out->ls, T.ls, sizeof T.ls);
    if (equal_signs(T.ls, 10)) return ORC_OK;
    if (zero_policy == ORC_ZERO_THROW && has_zero(T.ls, 10)) return ORC_E_ZERO_NODE;
    patch_t p;
    int err = make_patch(T.ls, &p);
    if (err) return err;
    patch_geometry(T.X, &p);
    out->nfacet = p.nf;
    for (int f = 0; f < p.nf; ++f) {
        out->bnd[f] = p.bnd[f];
        out->absdet[f] = p.absdet[f];
        for (int c = 0; c < 3; ++c) out->normal[f][c] = p.normal[f][c];
        for (int j = 0; j < 3; ++j) {
            for (int c = 0; c < 3; ++c) out->world[f][j][c] = p.vw[p.facet[f][j]][c];
            for (int c = 0; c < 4; ++c) out->bary[f][j][c] = p.vb[p.facet[f][j]][c];
        }
    }
    return ORC_OK;
}

/* IdxDescCL::CreateNumbOnInterface -> CreateNumbOnInterfaceVertex, misc/problem.cpp:443-503.
 * Tets whose 10 values share one strict sign are skipped (distance() != 0), so the serial
 * first-encounter loop runs over the cut tets in mesh order. */
int orc_number_p1if(const orc_mesh_t *m, const double *phi, double omit_bound, int zero_policy,
                    int32_t *dof_of_vertex, int64_t *ndof)
{
    init_tables();
    if (!check_mesh(m) || !phi || !dof_of_vertex || !ndof) return ORC_E_ARG;
    const int64_t nv = (int64_t)(m->n[0] + 1) * (m->n[1] + 1) * (m->n[2] + 1);
    for (int64_t v = 0; v < nv; ++v) dof_of_vertex[v] = -1;
    int64_t ncut = 0;
    int err = orc_classify(m, phi, zero_policy, 0, &ncut, NULL);
    if (err) return err;
    int64_t *cut = (int64_t *)malloc(sizeof(int64_t) * (ncut > 0 ? ncut : 1));
    if (!cut) return ORC_E_NOMEM;
    orc_classify(m, phi, zero_policy, ncut, &ncut, cut);
    int64_t counter = 0;
    for (int64_t c = 0; c < ncut && !err; ++c) {
        tet_t T;
        load_tet(m, phi, cut[c], &T);
        const double limit = omit_bound < 0. ? 0. : omit_bound * pow(tet_volume(T.X) * 6, 4. / 3.);
        patch_t p;
        err = make_patch(T.ls, &p);
        if (err) break;
        if (p.nf == 0) continue;
        patch_geometry(T.X, &p);
        double node[MAXF * 4][4], w[MAXF * 4];
        const int nq = composite_nodes(&p, 4, Q2_N, Q2_W, node, w);
        for (int i = 0; i < 4; ++i) {
            if (dof_of_vertex[T.vid[i]] >= 0) continue;
            double cc[10] = {0};
            cc[i] = 1.;
            for (int e = 0; e < 6; ++e)
                if (VOE[e][0] == i || VOE[e][1] == i) cc[4 + e] = 0.25;
            double s = 0.;
            for (int k = 0; k < nq; ++k) s += w[k] * p2_eval(cc, node[k]);
            if (s > limit) dof_of_vertex[T.vid[i]] = (int32_t)counter++;
        }
    }
    free(cut);
    if (err) return err;
    *ndof = counter;
    return ORC_OK;
}

/* ---- row builder with the semantics of SparseMatBuilderCL (num/spmat.h:551-677):
 *      every touched (i,j) is stored (explicit zeros kept), values are summed in
 *      insertion order starting from 0.0, rows are sorted by column on Build() ---- */
#define ROWCAP 32
typedef struct {
    int32_t n;
    int32_t col[ROWCAP];
    double v[3][ROWCAP];
} brow_t;

static int builder_add(brow_t *r, int32_t col, const double val[3])
{
    int k = 0;
    for (; k < r->n; ++k)
        if (r->col[k] == col) break;
    if (k == r->n) {
        if (r->n == ROWCAP) return ORC_E_CAPACITY;
        r->col[k] = col;
        r->v[0][k] = r->v[1][k] = r->v[2][k] = 0.0;
        ++r->n;
    }
    for (int o = 0; o < 3; ++o) r->v[o][k] += val[o];
    return ORC_OK;
}

/* Build(), num/spmat.h:621-677: row sizes, prefix sum, per-row sort by column */
static int build_rows(const brow_t *rows, int64_t ndof, int64_t cap_nnz, int64_t *nnz, int64_t *rowptr,
                      int32_t *colind, double *val_mass, double *val_lb, double *val_stab)
{
    rowptr[0] = 0;
    for (int64_t r = 0; r < ndof; ++r) rowptr[r + 1] = rowptr[r] + rows[r].n;
    *nnz = rowptr[ndof];
    if (*nnz > cap_nnz) return ORC_E_CAPACITY;
    if (!colind) return ORC_OK;
    for (int64_t r = 0; r < ndof; ++r) {
        const brow_t *R = &rows[r];
        int32_t ord[ROWCAP];
        for (int k = 0; k < R->n; ++k) ord[k] = k;
        for (int a = 1; a < R->n; ++a) {
            int32_t x = ord[a];
            int b = a - 1;
            while (b >= 0 && R->col[ord[b]] > R->col[x]) { ord[b + 1] = ord[b]; --b; }
            ord[b + 1] = x;
        }
        for (int k = 0; k < R->n; ++k) {
            const int64_t dst = rowptr[r] + k;
            colind[dst] = R->col[ord[k]];
            if (val_mass) val_mass[dst] = R->v[0][ord[k]];
            if (val_lb) val_lb[dst] = R->v[1][ord[k]];
            if (val_stab) val_stab[dst] = R->v[2][ord[k]];
        }
    }
    return ORC_OK;
}

typedef struct {
    int64_t t;
    int err;
    orc_local_t loc;
    int64_t vid[4];
    int nonempty;
} cutrec_t;

int orc_assemble_p1(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                    const double coef[3], int rhs_kind, const double *rhs_param, int zero_policy,
                    int64_t cap_nnz, int64_t *nnz, int64_t *rowptr, int32_t *colind,
                    double *val_mass, double *val_lb, double *val_stab, double *rhs)
{
    init_tables();
    if (!check_mesh(m) || !phi || !dof_of_vertex || !nnz || !rowptr || !coef) return ORC_E_ARG;
    int64_t ncut = 0;
    int err = orc_classify(m, phi, zero_policy, 0, &ncut, NULL);
    if (err) return err;
    int64_t *cut = (int64_t *)malloc(sizeof(int64_t) * (ncut > 0 ? ncut : 1));
    cutrec_t *rec = (cutrec_t *)malloc(sizeof(cutrec_t) * (ncut > 0 ? ncut : 1));
    brow_t *rows = (brow_t *)calloc(ndof > 0 ? ndof : 1, sizeof(brow_t));
    if (!cut || !rec || !rows) {
        free(cut); free(rec); free(rows);
        return ORC_E_NOMEM;
    }
    orc_classify(m, phi, zero_policy, ncut, &ncut, cut);
    /* local matrices: independent per tet (the reference runs them per colour class in parallel) */
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < ncut; ++c) {
        tet_t T;
        load_tet(m, phi, cut[c], &T);
        patch_t p;
        rec[c].t = cut[c];
        rec[c].nonempty = 0;
        rec[c].err = make_patch(T.ls, &p);
        for (int i = 0; i < 4; ++i) rec[c].vid[i] = T.vid[i];
        if (rec[c].err || p.nf == 0) continue;
        rec[c].nonempty = 1;
        patch_geometry(T.X, &p);
        local_quantities(&T, &p, coef, rhs_kind, rhs_param, &rec[c].loc);
    }
    if (rhs)
        for (int64_t i = 0; i < ndof; ++i) rhs[i] = 0.;
    /* scatter in mesh order: update_global_matrix, surfactant/ifacetransp.tpp:31-43,
     * InterfaceVectorAccuCL::visit, surfactant/ifacetransp.h:630-637 */
    for (int64_t c = 0; c < ncut && !err; ++c) {
        if (rec[c].err) { err = rec[c].err; break; }
        if (!rec[c].nonempty) continue; /* cdata.empty() */
        int32_t numb[4];
        for (int i = 0; i < 4; ++i) numb[i] = dof_of_vertex[rec[c].vid[i]];
        const orc_local_t *L = &rec[c].loc;
        for (int i = 0; i < 4 && !err; ++i) {
            if (numb[i] < 0) continue;
            for (int j = 0; j < 4; ++j) {
                if (numb[j] < 0) continue;
                const double v3[3] = {L->mass[i * 4 + j], L->lb[i * 4 + j], L->stab[i * 4 + j]};
                err = builder_add(&rows[numb[i]], numb[j], v3);
                if (err) break;
            }
        }
        if (rhs)
            for (int i = 0; i < 4; ++i)
                if (numb[i] >= 0) rhs[numb[i]] += L->rhs[i];
    }
    if (!err) err = build_rows(rows, ndof, cap_nnz, nnz, rowptr, colind, val_mass, val_lb, val_stab);
    free(cut);
    free(rec);
    free(rows);
    return err;
}

/* ------------------------------------------------------------------------ */
/* Colour classes and the colour-parallel accumulate() of the OpenMP build    */
/* ------------------------------------------------------------------------ */

/* ColorClassesCL::compute_color_classes, geom/multigrid.cpp:1238-1319: tets in mesh order, each
 * given the smallest colour not yet carried by one of its four vertices (a Kuhn tet shares a
 * vertex with at most 70 others: two 64-bit words per vertex).  Classes are sorted by tet (the
 * reference sorts by address, the creation order of a fresh brick).  class_start[c] .. [c+1]
 * index class_tets; returns the number of colours, at most 128. */
int orc_color_classes(const orc_mesh_t *m, int64_t *ncolors, int64_t *class_start, int64_t *class_tets)
{
    init_tables();
    if (!check_mesh(m) || !ncolors || !class_start || !class_tets) return ORC_E_ARG;
    const int64_t n1 = m->n[0], n2 = m->n[1], n3 = m->n[2];
    const int64_t nt = 6 * n1 * n2 * n3, nv = (n1 + 1) * (n2 + 1) * (n3 + 1);
    uint64_t *used = (uint64_t *)calloc((size_t)(2 * nv), sizeof(uint64_t));
    uint8_t *color = (uint8_t *)malloc((size_t)(nt > 0 ? nt : 1));
    if (!used || !color) {
        free(used);
        free(color);
        return ORC_E_NOMEM;
    }
    int64_t count[129] = {0};
    int err = ORC_OK;
    for (int64_t t = 0; t < nt && !err; ++t) {
        const int64_t cell = t / 6, k = t % 6;
        const int64_t i1 = cell % n1, i2 = (cell / n1) % n2, i3 = cell / (n1 * n2);
        int64_t vid[4];
        uint64_t u[2] = {0, 0};
        for (int q = 0; q < 4; ++q) {
            const int *d = KUHN[k][q];
            vid[q] = ((i3 + d[2]) * (n2 + 1) + i2 + d[1]) * (n1 + 1) + i1 + d[0];
            u[0] |= used[2 * vid[q]];
            u[1] |= used[2 * vid[q] + 1];
        }
        int c = 0;
        while (c < 128 && (u[c >> 6] >> (c & 63) & 1u)) ++c;
        if (c == 128) { err = ORC_E_CAPACITY; break; }
        color[t] = (uint8_t)c;
        ++count[c];
        for (int q = 0; q < 4; ++q) used[2 * vid[q] + (c >> 6)] |= (uint64_t)1 << (c & 63);
    }
    int64_t nc = 0;
    while (nc < 128 && count[nc] > 0) ++nc;
    if (!err) {
        class_start[0] = 0;
        for (int64_t c = 0; c < nc; ++c) class_start[c + 1] = class_start[c] + count[c];
        int64_t fill[128];
        for (int64_t c = 0; c < nc; ++c) fill[c] = class_start[c];
        for (int64_t t = 0; t < nt; ++t) class_tets[fill[color[t]]++] = t;  /* ascending within a class */
        *ncolors = nc;
    }
    free(used);
    free(color);
    return err;
}

/* accumulate() with colour classes (num/accumulator.h:161-188, 215-228): for each colour, the
 * threads visit its tets (schedule(dynamic)), each visit being InterfaceCommonDataP1CL::visit
 * (lattice values, equalSigns, patch) followed by the matrix and vector accumulators'
 * local setup and update_global_matrix into the shared builders.  Tets of one colour share no
 * vertex, so no builder row is written by two threads; entries are summed in colour order (the
 * reference's OpenMP rounding, not its serial mesh order). */
int orc_assemble_p1_colored(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                            const double coef[3], int rhs_kind, const double *rhs_param, int zero_policy,
                            int64_t ncolors, const int64_t *class_start, const int64_t *class_tets,
                            int64_t cap_nnz, int64_t *nnz, int64_t *rowptr, int32_t *colind,
                            double *val_mass, double *val_lb, double *val_stab, double *rhs)
{
    init_tables();
    if (!check_mesh(m) || !phi || !dof_of_vertex || !nnz || !rowptr || !coef || !class_start || !class_tets)
        return ORC_E_ARG;
    brow_t *rows = (brow_t *)calloc(ndof > 0 ? ndof : 1, sizeof(brow_t));
    if (!rows) return ORC_E_NOMEM;
    if (rhs)
        for (int64_t i = 0; i < ndof; ++i) rhs[i] = 0.;
    int err = ORC_OK;
    for (int64_t c = 0; c < ncolors; ++c) {
        const int64_t b0 = class_start[c], b1 = class_start[c + 1];
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t j = b0; j < b1; ++j) {
            int e = ORC_OK;
            tet_t T;
            load_tet_ls(m, phi, class_tets[j], &T);
            if (equal_signs(T.ls, 10)) continue;
            if (zero_policy == ORC_ZERO_THROW && has_zero(T.ls, 10)) e = ORC_E_ZERO_NODE;
            patch_t p;
            if (!e) {
                load_tet(m, phi, class_tets[j], &T);
                e = make_patch(T.ls, &p);
            }
            if (!e && p.nf > 0) {
                orc_local_t L;
                patch_geometry(T.X, &p);
                local_quantities(&T, &p, coef, rhs_kind, rhs_param, &L);
                int32_t numb[4];
                for (int i = 0; i < 4; ++i) numb[i] = dof_of_vertex[T.vid[i]];
                for (int i = 0; i < 4 && !e; ++i) {
                    if (numb[i] < 0) continue;
                    for (int jj = 0; jj < 4 && !e; ++jj) {
                        if (numb[jj] < 0) continue;
                        const double v3[3] = {L.mass[i * 4 + jj], L.lb[i * 4 + jj], L.stab[i * 4 + jj]};
                        e = builder_add(&rows[numb[i]], numb[jj], v3);
                    }
                }
                if (rhs)
                    for (int i = 0; i < 4; ++i)
                        if (numb[i] >= 0) rhs[numb[i]] += L.rhs[i];
            }
            if (e) {
#pragma omp critical(orc_color_err)
                if (!err) err = e;
            }
        }
        if (err) break;
    }
    if (!err) err = build_rows(rows, ndof, cap_nnz, nnz, rowptr, colind, val_mass, val_lb, val_stab);
    free(rows);
    return err;
}

/* ------------------------------------------------------------------------ */
/* Narrow band (SurfactantNarrowBandStblP1CL::Update, surfactant/ifacetransp.cpp:1110-1172)         */
/* ------------------------------------------------------------------------ */

/* distance() of the 10 lattice values, geom/subtriangulation.tpp:38-46: min |f| when all share one
 * strict sign, else 0 (a cut tet; with an exact zero the reference throws) */
static int tet_distance(const double ls[10], double *d)
{
    if (equal_signs(ls, 10)) {
        double m = fabs(ls[0]);
        for (int l = 1; l < 10; ++l)
            if (fabs(ls[l]) < m) m = fabs(ls[l]);
        *d = m;
        return ORC_OK;
    }
    *d = 0.;
    return has_zero(ls, 10) ? ORC_E_ZERO_NODE : ORC_OK;
}

/* Tets with distance() <= width in mesh order (NarrowBandCommonDataP1CL::visit,
 * surfactant/ifacetransp.h:688-700); is_cut[c] = 1 for the cut ones */
int orc_classify_band(const orc_mesh_t *m, const double *phi, double width, int zero_policy, int64_t cap,
                      int64_t *nband, int64_t *band_tets, uint8_t *is_cut)
{
    init_tables();
    if (!check_mesh(m) || !phi || !nband || width < 0.) return ORC_E_ARG;
    const int64_t nt = 6LL * m->n[0] * m->n[1] * m->n[2];
    int64_t c = 0;
    for (int64_t t = 0; t < nt; ++t) {
        tet_t T;
        load_tet_ls(m, phi, t, &T);
        double d;
        const int err = tet_distance(T.ls, &d);
        if (err && zero_policy == ORC_ZERO_THROW) return err;
        if (d > width) continue;
        if (band_tets) {
            if (c >= cap) return ORC_E_CAPACITY;
            band_tets[c] = t;
            if (is_cut) is_cut[c] = equal_signs(T.ls, 10) ? 0 : 1;
        }
        ++c;
    }
    *nband = c;
    return ORC_OK;
}

/* IdxDescCL::CreateNumbering(..., width) -> CreateNumbNearInterfaceVertex, misc/problem.cpp:509-555:
 * every vertex of every band tet, first encounter in mesh order (no omit_bound test) */
int orc_number_p1band(const orc_mesh_t *m, const double *phi, double width, int zero_policy,
                      int32_t *dof_of_vertex, int64_t *ndof)
{
    init_tables();
    if (!check_mesh(m) || !phi || !dof_of_vertex || !ndof || width < 0.) return ORC_E_ARG;
    const int64_t nv = (int64_t)(m->n[0] + 1) * (m->n[1] + 1) * (m->n[2] + 1);
    for (int64_t v = 0; v < nv; ++v) dof_of_vertex[v] = -1;
    const int64_t nt = 6LL * m->n[0] * m->n[1] * m->n[2];
    int64_t counter = 0;
    for (int64_t t = 0; t < nt; ++t) {
        tet_t T;
        load_tet(m, phi, t, &T);
        double d;
        const int err = tet_distance(T.ls, &d);
        if (err && zero_policy == ORC_ZERO_THROW) return err;
        if (d > width) continue;
        for (int i = 0; i < 4; ++i)
            if (dof_of_vertex[T.vid[i]] < 0) dof_of_vertex[T.vid[i]] = (int32_t)counter++;
    }
    *ndof = counter;
    return ORC_OK;
}

/* The accumulate() of SurfactantNarrowBandStblP1CL::Update on the narrow-band numbering: mass,
 * Laplace-Beltrami and load through InterfaceCommonDataP1CL (cut tets with a non-empty patch; their
 * own builders, so their own pattern), the normal stabilisation through NarrowBandMatrixAccuP1CL
 * over every band tet (its own pattern, surfactant/ifacetransp.h:746-755).  Each builder sums its
 * contributions in mesh order. */
int orc_assemble_p1band(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                        double width, const double coef[3], int rhs_kind, const double *rhs_param, int zero_policy,
                        int64_t cap_nnz, int64_t *nnz_s, int64_t *rowptr_s, int32_t *colind_s, double *val_stab,
                        int64_t *nnz_m, int64_t *rowptr_m, int32_t *colind_m, double *val_mass, double *val_lb,
                        double *rhs)
{
    init_tables();
    if (!check_mesh(m) || !phi || !dof_of_vertex || !nnz_s || !rowptr_s || !nnz_m || !rowptr_m || !coef)
        return ORC_E_ARG;
    int64_t nb = 0;
    int err = orc_classify_band(m, phi, width, zero_policy, 0, &nb, NULL, NULL);
    if (err) return err;
    int64_t *band = (int64_t *)malloc(sizeof(int64_t) * (nb > 0 ? nb : 1));
    uint8_t *cutf = (uint8_t *)malloc(nb > 0 ? nb : 1);
    brow_t *rs = (brow_t *)calloc(ndof > 0 ? ndof : 1, sizeof(brow_t));
    brow_t *rm = (brow_t *)calloc(ndof > 0 ? ndof : 1, sizeof(brow_t));
    if (!band || !cutf || !rs || !rm) {
        free(band); free(cutf); free(rs); free(rm);
        return ORC_E_NOMEM;
    }
    orc_classify_band(m, phi, width, zero_policy, nb, &nb, band, cutf);
    if (rhs)
        for (int64_t i = 0; i < ndof; ++i) rhs[i] = 0.;
    for (int64_t c = 0; c < nb && !err; ++c) {
        tet_t T;
        load_tet(m, phi, band[c], &T);
        int32_t numb[4];
        for (int i = 0; i < 4; ++i) numb[i] = dof_of_vertex[T.vid[i]];
        orc_local_t L;
        memset(&L, 0, sizeof L);
        int nonempty = 0;
        if (cutf[c]) {
            patch_t p;
            err = make_patch(T.ls, &p);
            if (err) break;
            if (p.nf > 0) {
                patch_geometry(T.X, &p);
                local_quantities(&T, &p, coef, rhs_kind, rhs_param, &L);
                nonempty = 1;
            }
        }
        double stab[16];
        local_stab(&T, coef, stab);
        for (int i = 0; i < 4 && !err; ++i) {
            if (numb[i] < 0) continue;
            for (int j = 0; j < 4 && !err; ++j) {
                if (numb[j] < 0) continue;
                if (nonempty) {
                    const double v3[3] = {L.mass[i * 4 + j], L.lb[i * 4 + j], 0.};
                    err = builder_add(&rm[numb[i]], numb[j], v3);
                }
                if (!err) {
                    const double s3[3] = {stab[i * 4 + j], 0., 0.};
                    err = builder_add(&rs[numb[i]], numb[j], s3);
                }
            }
        }
        if (rhs && nonempty)
            for (int i = 0; i < 4; ++i)
                if (numb[i] >= 0) rhs[numb[i]] += L.rhs[i];
    }
    if (!err) err = build_rows(rs, ndof, cap_nnz, nnz_s, rowptr_s, colind_s, val_stab, NULL, NULL);
    if (!err) err = build_rows(rm, ndof, cap_nnz, nnz_m, rowptr_m, colind_m, val_mass, val_lb, NULL);
    free(band);
    free(cutf);
    free(rs);
    free(rm);
    return err;
}

/* tests/principallattice.cpp:296-330: patch + make_CompositeQuad5Domain2D + quad_2D(1) over every tet */
int orc_surface_area(const orc_mesh_t *m, const double *phi, int zero_policy, double *area)
{
    init_tables();
    if (!check_mesh(m) || !phi || !area) return ORC_E_ARG;
    const int64_t nt = 6LL * m->n[0] * m->n[1] * m->n[2];
    double surf = 0.;
    for (int64_t t = 0; t < nt; ++t) {
        tet_t T;
        load_tet_ls(m, phi, t, &T);
        if (equal_signs(T.ls, 10)) continue;
        if (zero_policy == ORC_ZERO_THROW && has_zero(T.ls, 10)) return ORC_E_ZERO_NODE;
        load_tet(m, phi, t, &T);
        patch_t p;
        int err = make_patch(T.ls, &p);
        if (err) return err;
        if (p.nf == 0) continue;
        patch_geometry(T.X, &p);
        double node[MAXF * 7][4], w[MAXF * 7];
        const int nq = composite_nodes(&p, 7, Q5_N, Q5_W, node, w);
        double s = 0.;
        for (int k = 0; k < nq; ++k) s += w[k] * 1.;
        surf += s;
    }
    *area = surf;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* P2 interface space on the planar patch                                    */
/* ------------------------------------------------------------------------ */

/* FE_P2CL::H0..H9 at barycentric p (num/fe.h:287-296): vertices l(2l-1), edges 4 l_a l_b */
static void p2_basis(const double p[4], double h[10])
{
    for (int k = 0; k < 4; ++k) h[k] = p[k] * (2. * p[k] - 1.);
    for (int e = 0; e < 6; ++e) h[4 + e] = 4. * p[VOE[e][0]] * p[VOE[e][1]];
}

/* gradients of the P2 basis at p from the barycentric derivatives:
 * grad h_k = (4 l_k - 1) grad l_k, grad h_e = 4 l_b grad l_a + 4 l_a grad l_b */
static void p2_gradients(const double p[4], const double G1[4][3], double g[10][3])
{
    for (int k = 0; k < 4; ++k) {
        const double f = 4.0 * p[k] - 1.0;
        for (int d = 0; d < 3; ++d) g[k][d] = f * G1[k][d];
    }
    for (int e = 0; e < 6; ++e) {
        const int a = VOE[e][0], b = VOE[e][1];
        for (int d = 0; d < 3; ++d) g[4 + e][d] = 4.0 * p[b] * G1[a][d] + 4.0 * p[a] * G1[b][d];
    }
}

static int64_t p2_node_index(const orc_mesh_t *m, const tet_t *T, int d)
{
    const int64_t h1 = 2LL * m->n[0] + 1, h2 = 2LL * m->n[1] + 1;
    const int l = LAT_OF_P2[d];
    int64_t j[3];
    for (int c = 0; c < 3; ++c) {
        j[c] = 2 * T->i[c];
        for (int q = 0; q < 4; ++q) j[c] += LAT_B2[l][q] * KUHN[T->k][q][c];
    }
    return (j[2] * h2 + j[1]) * h1 + j[0];
}

/* CreateNumbOnInterfaceP2 (misc/problem.cpp:571-636): first-encounter in tet order over the 10
 * P2 nodes (vertices, then edges (0,1),(0,2),(1,2),(0,3),(1,3),(2,3)); a node gets the next number
 * iff the Quad5 composite integral of its squared P2 basis function exceeds the limit. */
int orc_number_p2if(const orc_mesh_t *m, const double *phi, double omit_bound, int zero_policy,
                    int32_t *dof_of_node, int64_t *ndof)
{
    init_tables();
    if (!check_mesh(m) || !phi || !dof_of_node || !ndof) return ORC_E_ARG;
    const int64_t nn = (2LL * m->n[0] + 1) * (2LL * m->n[1] + 1) * (2LL * m->n[2] + 1);
    for (int64_t v = 0; v < nn; ++v) dof_of_node[v] = -1;
    int64_t ncut = 0;
    int err = orc_classify(m, phi, zero_policy, 0, &ncut, NULL);
    if (err) return err;
    int64_t *cut = (int64_t *)malloc(sizeof(int64_t) * (ncut > 0 ? ncut : 1));
    if (!cut) return ORC_E_NOMEM;
    orc_classify(m, phi, zero_policy, ncut, &ncut, cut);
    int64_t counter = 0;
    for (int64_t c = 0; c < ncut && !err; ++c) {
        tet_t T;
        load_tet(m, phi, cut[c], &T);
        const double limit = omit_bound < 0. ? 0. : omit_bound * pow(tet_volume(T.X) * 6, 4. / 3.);
        patch_t p;
        err = make_patch(T.ls, &p);
        if (err) break;
        if (p.nf == 0) continue;
        patch_geometry(T.X, &p);
        double node[MAXF * 7][4], w[MAXF * 7];
        const int nq = composite_nodes(&p, 7, Q5_N, Q5_W, node, w);
        for (int i = 0; i < 10; ++i) {
            const int64_t g = p2_node_index(m, &T, i);
            if (dof_of_node[g] >= 0) continue;
            double s = 0.;
            for (int k = 0; k < nq; ++k) {
                double h[10];
                p2_basis(node[k], h);
                s += w[k] * (h[i] * h[i]);
            }
            if (s > limit) dof_of_node[g] = (int32_t)counter++;
        }
    }
    free(cut);
    if (err) return err;
    *ndof = counter;
    return ORC_OK;
}

/* Local P2 quantities of one cut tet with a non-empty patch (Quad5 composite rule, normal = the
 * normalised P2 level-set gradient at each point, as LocalAssembler::buildNormal,
 * surfactant/ifacetransp.h:1062-1068):
 *   mass  M_ij = sum_q w_q h_j h_i                         (LocalAssembler::M_P2P2, :1579-1588)
 *   lb    A_ij = D sum_q w_q (P grad h_j).(P grad h_i)      (the P2 form of A_P1P1 with
 *                                                           buildSurfGradP2, :1049-1053)
 *   stab  C_ij = D |det| sum_q5 w (n.grad h_i)(n.grad h_j) (the volume integral, as C_n_P1P1,
 *                                                           :1545-1555, 15-point Quad5 rule)
 *   rhs   b_i = sum_q w_q f(x_q) h_i;  hat_sq_i = sum_q w_q h_i h_i (the numbering integrals) */
static void local_quantities_p2(const tet_t *T, const patch_t *p, const double coef[3], int rhs_kind,
                                const double *rhs_param, orc_local_p2_t *o)
{
    double node[MAXF * 7][4], w[MAXF * 7];
    const int nq = composite_nodes(p, 7, Q5_N, Q5_W, node, w);
    double G1[4][3];
    const double det = p1_gradients(T->X, G1);
    double v[10];
    for (int d = 0; d < 10; ++d) v[d] = T->ls[LAT_OF_P2[d]];
    memset(o, 0, sizeof *o);
    o->nfacet = p->nf;
    for (int k = 0; k < nq; ++k) {
        double h[10], g[10][3], sg[10][3], nrm[3];
        p2_basis(node[k], h);
        p2_gradients(node[k], G1, g);
        p2_normal(v, node[k], G1, nrm);
        for (int i = 0; i < 10; ++i) {
            const double d = dot3(g[i], nrm);
            for (int c = 0; c < 3; ++c) sg[i][c] = g[i][c] - d * nrm[c];
        }
        double f = 0.;
        if (rhs_kind != ORC_RHS_NONE) {
            double x[3];
            bary_to_world(T->X, node[k], x);
            f = rhs_fun(rhs_kind, rhs_param, x);
        }
        for (int i = 0; i < 10; ++i) {
            o->hat_sq[i] += w[k] * (h[i] * h[i]);
            if (rhs_kind != ORC_RHS_NONE) o->rhs[i] += w[k] * (f * h[i]);
            for (int j = 0; j <= i; ++j) {
                o->mass[i * 10 + j] += w[k] * (h[j] * h[i]);
                o->lb[i * 10 + j] += w[k] * dot3(sg[j], sg[i]);
            }
        }
    }
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j <= i; ++j) {
            o->lb[i * 10 + j] = coef[1] * o->lb[i * 10 + j];
            o->mass[i * 10 + j] = coef[0] * o->mass[i * 10 + j];
            o->mass[j * 10 + i] = o->mass[i * 10 + j];
            o->lb[j * 10 + i] = o->lb[i * 10 + j];
        }
    /* volume stabilisation on the 15-point Quad5 rule (weights of Quad5CL::quad) */
    double U[10][15];
    for (int q = 0; q < 15; ++q) {
        double g[10][3], nrm[3];
        p2_gradients(V5_N[q], G1, g);
        p2_normal(v, V5_N[q], G1, nrm);
        for (int i = 0; i < 10; ++i) U[i][q] = dot3(nrm, g[i]);
    }
    const double ad = fabs(det);
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j <= i; ++j) {
            double f[15];
            for (int q = 0; q < 15; ++q) f[q] = U[i][q] * U[j][q];
            const double val = ad * (((8. / 405. * f[0] + V5_C1 * (((f[1] + f[2]) + f[3]) + f[4]))
                                      + V5_C2 * (((f[5] + f[6]) + f[7]) + f[8]))
                                     + 5. / 567. * (((((f[9] + f[10]) + f[11]) + f[12]) + f[13]) + f[14]));
            o->stab[i * 10 + j] = o->stab[j * 10 + i] = coef[2] * val;
        }
}

int orc_tet_local_p2(const orc_mesh_t *m, const double *phi, int64_t tet, const double coef[3], int rhs_kind,
                     const double *rhs_param, int zero_policy, orc_local_p2_t *out)
{
    init_tables();
    if (!check_mesh(m) || !phi || !out || !coef) return ORC_E_ARG;
    memset(out, 0, sizeof *out);
    tet_t T;
    load_tet(m, phi, tet, &T);
    if (equal_signs(T.ls, 10)) return ORC_OK;
    if (zero_policy == ORC_ZERO_THROW && has_zero(T.ls, 10)) return ORC_E_ZERO_NODE;
    patch_t p;
    int err = make_patch(T.ls, &p);
    if (err) return err;
    if (p.nf == 0) return ORC_OK;
    patch_geometry(T.X, &p);
    local_quantities_p2(&T, &p, coef, rhs_kind, rhs_param, out);
    return ORC_OK;
}

#define ROWCAP2 160
typedef struct {
    int32_t n;
    int32_t col[ROWCAP2];
    double v[3][ROWCAP2];
} brow2_t;

/* accumulate() with the P2 forms and SparseMatBuilderCL semantics (num/spmat.h:551-677) */
int orc_assemble_p2(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_node, int64_t ndof,
                    const double coef[3], int rhs_kind, const double *rhs_param, int zero_policy,
                    int64_t cap_nnz, int64_t *nnz, int64_t *rowptr, int32_t *colind,
                    double *val_mass, double *val_lb, double *val_stab, double *rhs)
{
    init_tables();
    if (!check_mesh(m) || !phi || !dof_of_node || !nnz || !rowptr || !coef) return ORC_E_ARG;
    int64_t ncut = 0;
    int err = orc_classify(m, phi, zero_policy, 0, &ncut, NULL);
    if (err) return err;
    int64_t *cut = (int64_t *)malloc(sizeof(int64_t) * (ncut > 0 ? ncut : 1));
    orc_local_p2_t *loc = (orc_local_p2_t *)malloc(sizeof(orc_local_p2_t) * (ncut > 0 ? ncut : 1));
    int32_t (*numb)[10] = (int32_t (*)[10])malloc(sizeof(int32_t) * 10 * (ncut > 0 ? ncut : 1));
    int *st = (int *)calloc(ncut > 0 ? ncut : 1, sizeof(int));
    brow2_t *rows = (brow2_t *)calloc(ndof > 0 ? ndof : 1, sizeof(brow2_t));
    if (!cut || !loc || !numb || !st || !rows) {
        free(cut); free(loc); free(numb); free(st); free(rows);
        return ORC_E_NOMEM;
    }
    orc_classify(m, phi, zero_policy, ncut, &ncut, cut);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < ncut; ++c) {
        tet_t T;
        load_tet(m, phi, cut[c], &T);
        patch_t p;
        int e = make_patch(T.ls, &p);
        if (e) { st[c] = e; continue; }
        if (p.nf == 0) { st[c] = -1; continue; }
        patch_geometry(T.X, &p);
        local_quantities_p2(&T, &p, coef, rhs_kind, rhs_param, &loc[c]);
        for (int i = 0; i < 10; ++i) numb[c][i] = dof_of_node[p2_node_index(m, &T, i)];
    }
    if (rhs)
        for (int64_t i = 0; i < ndof; ++i) rhs[i] = 0.;
    for (int64_t c = 0; c < ncut && !err; ++c) {
        if (st[c] > 0) { err = st[c]; break; }
        if (st[c] < 0) continue;
        for (int i = 0; i < 10 && !err; ++i) {
            const int32_t r = numb[c][i];
            if (r < 0) continue;
            brow2_t *R = &rows[r];
            for (int j = 0; j < 10; ++j) {
                const int32_t cj = numb[c][j];
                if (cj < 0) continue;
                int k = 0;
                for (; k < R->n; ++k)
                    if (R->col[k] == cj) break;
                if (k == R->n) {
                    if (R->n == ROWCAP2) { err = ORC_E_CAPACITY; break; }
                    R->col[k] = cj;
                    R->v[0][k] = R->v[1][k] = R->v[2][k] = 0.0;
                    ++R->n;
                }
                R->v[0][k] += loc[c].mass[i * 10 + j];
                R->v[1][k] += loc[c].lb[i * 10 + j];
                R->v[2][k] += loc[c].stab[i * 10 + j];
            }
            if (rhs) rhs[r] += loc[c].rhs[i];
        }
    }
    if (!err) {
        rowptr[0] = 0;
        for (int64_t r = 0; r < ndof; ++r) rowptr[r + 1] = rowptr[r] + rows[r].n;
        *nnz = rowptr[ndof];
        if (*nnz > cap_nnz) err = ORC_E_CAPACITY;
    }
    if (!err && colind) {
        for (int64_t r = 0; r < ndof; ++r) {
            brow2_t *R = &rows[r];
            int32_t ord[ROWCAP2];
            for (int k = 0; k < R->n; ++k) ord[k] = k;
            for (int a = 1; a < R->n; ++a) {
                int32_t x = ord[a];
                int b = a - 1;
                while (b >= 0 && R->col[ord[b]] > R->col[x]) { ord[b + 1] = ord[b]; --b; }
                ord[b + 1] = x;
            }
            for (int k = 0; k < R->n; ++k) {
                const int64_t dst = rowptr[r] + k;
                colind[dst] = R->col[ord[k]];
                if (val_mass) val_mass[dst] = R->v[0][ord[k]];
                if (val_lb) val_lb[dst] = R->v[1][ord[k]];
                if (val_stab) val_stab[dst] = R->v[2][ord[k]];
            }
        }
    }
    free(cut); free(loc); free(numb); free(st); free(rows);
    return err;
}

/* ------------------------------------------------------------------------ */
/* QuaQua-mapped P2 geometry: InterfaceCommonDataP2CL + QuaQuaMapperCL with  */
/* LocalMassP2CL, LocalLaplaceBeltramiP2CL, LocalVectorP2CL                   */
/* ------------------------------------------------------------------------ */

/* Householder QR of a small row-major R x C matrix: QRDecompCL<R,C>::prepare_solve,
 * apply_reflection, apply_Q and Solve, misc/container.h:1210-1300 (std::pow(x, 2) taken as x * x,
 * the correctly rounded square). */
typedef struct {
    int R, C;
    double a[16], d[4], beta[4];
} qr_t;
#define QA(q, i, j) ((q)->a[(i) * (q)->C + (j)])

static int qr_prepare(qr_t *q) /* 1: rank deficient (the reference throws) */
{
    for (int j = 0; j < q->C; ++j) {
        double sigma = 0.;
        for (int i = j; i < q->R; ++i) sigma += QA(q, i, j) * QA(q, i, j);
        if (sigma == 0.) return 1;
        q->d[j] = (QA(q, j, j) < 0 ? 1. : -1.) * sqrt(sigma);
        q->beta[j] = 1. / (q->d[j] * QA(q, j, j) - sigma);
        QA(q, j, j) -= q->d[j];
        for (int k = j + 1; k < q->C; ++k) {
            double sp = 0.;
            for (int i = j; i < q->R; ++i) sp += QA(q, i, j) * QA(q, i, k);
            sp *= q->beta[j];
            for (int i = j; i < q->R; ++i) QA(q, i, k) += QA(q, i, j) * sp;
        }
    }
    return 0;
}

static void qr_reflect(const qr_t *q, int j, double *b)
{
    double sp = 0.;
    for (int i = j; i < q->R; ++i) sp += QA(q, i, j) * b[i];
    sp *= q->beta[j];
    for (int i = j; i < q->R; ++i) b[i] += QA(q, i, j) * sp;
}

static void qr_apply_Q(const qr_t *q, double *b)
{
    for (int j = q->C - 1; j >= 0; --j) qr_reflect(q, j, b);
}

static void qr_solve(const qr_t *q, double *b)
{
    for (int j = 0; j < q->C; ++j) qr_reflect(q, j, b);
    for (int i = q->C - 1; i >= 0; --i) {
        for (int j = i + 1; j < q->C; ++j) b[i] -= QA(q, i, j) * b[j];
        b[i] /= q->d[i];
    }
}

/* GetTrafoTr, geom/multigrid.h:627-647 */
static void trafo_tr(const double X[4][3], double T[3][3], double *det)
{
    double M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[j][i] = X[i + 1][j] - X[0][j];
    *det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
         - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
         + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    const double d = *det;
    T[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / d;
    T[0][1] = (M[2][0] * M[1][2] - M[1][0] * M[2][2]) / d;
    T[0][2] = (M[1][0] * M[2][1] - M[2][0] * M[1][1]) / d;
    T[1][0] = (M[2][1] * M[0][2] - M[0][1] * M[2][2]) / d;
    T[1][1] = (M[0][0] * M[2][2] - M[2][0] * M[0][2]) / d;
    T[1][2] = (M[2][0] * M[0][1] - M[0][0] * M[2][1]) / d;
    T[2][0] = (M[0][1] * M[1][2] - M[1][1] * M[0][2]) / d;
    T[2][1] = (M[1][0] * M[0][2] - M[0][0] * M[1][2]) / d;
    T[2][2] = (M[0][0] * M[1][1] - M[1][0] * M[0][1]) / d;
}

/* FE_P2CL::DH0Ref .. DH9Ref, num/fe.tpp:122-190 */
static void p2_dhref(int i, double v1, double v2, double v3, double g[3])
{
    switch (i) {
    case 0: g[0] = g[1] = g[2] = 4. * (v1 + v2 + v3) - 3.; break;
    case 1: g[0] = 4. * v1 - 1.; g[1] = g[2] = 0.; break;
    case 2: g[0] = g[2] = 0.; g[1] = 4 * v2 - 1.; break;
    case 3: g[0] = 0.; g[1] = 0.; g[2] = 4. * v3 - 1.; break;
    case 4: g[0] = 4. * (1. - (2. * v1 + v2 + v3)); g[1] = g[2] = -4. * v1; break;
    case 5: g[0] = g[2] = -4. * v2; g[1] = 4. * (1. - (2. * v2 + v1 + v3)); break;
    case 6: g[0] = 4. * v2; g[1] = 4. * v1; g[2] = 0.; break;
    case 7: g[0] = g[1] = -4. * v3; g[2] = 4. * (1. - (2. * v3 + v1 + v2)); break;
    case 8: g[0] = 4. * v3; g[1] = 0.; g[2] = 4. * v1; break;
    default: g[0] = 0.; g[1] = 4. * v3; g[2] = 4. * v2; break;
    }
}

/* 3x3 matrix times vector, misc/container.h:975-984 */
static void mv3(const double A[3][3], const double v[3], double r[3])
{
    for (int i = 0; i < 3; ++i) r[i] = A[i][0] * v[0] + A[i][1] * v[1] + A[i][2] * v[2];
}

/* inner_prod of 3-vectors, misc/container.h:460-464; norm = sqrt of it */
static double ip3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double norm3(const double a[3]) { return sqrt(ip3(a, a)); }

/* one tet of the mapped geometry: the cached quantities of base_point_newton_cacheCL::set_tetra
 * (levelset/levelsetmapper.cpp:77-100) and the World2BaryCoordCL factorisation (geom/simplex.cpp:456-484) */
typedef struct {
    double X[4][3];
    double T[3][3], det, h;
    double G[10][4][3];  /* P2DiscCL::GetGradients (num/discretize.h:1430-1431): T * DH_iRef at the vertices */
    double v[10];        /* level set in P2 dof order */
    double gh[10][3];    /* recovered gradient at the P2 nodes */
    qr_t w2b;
} qqtet_t;

static void qq_gradients(const double X[4][3], double T[3][3], double *det, double G[10][4][3])
{
    static const double ref[4][3] = {{0., 0., 0.}, {1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
    trafo_tr(X, T, det);
    for (int i = 0; i < 10; ++i)
        for (int k = 0; k < 4; ++k) {
            double g[3];
            p2_dhref(i, ref[k][0], ref[k][1], ref[k][2], g);
            mv3((const double (*)[3])T, g, G[i][k]);
        }
}

/* LocalP1CL<Point3DCL> at barycentric b (LinearCombinationCL::do_it, num/fe.h:44-53) */
static void p1v_eval(const double G[4][3], const double b[4], double r[3])
{
    for (int c = 0; c < 3; ++c) r[c] = b[0] * G[0][c] + b[1] * G[1][c] + b[2] * G[2][c] + b[3] * G[3][c];
}

/* LocalP2CL<Point3DCL> at barycentric b */
static void p2v_eval(const double g[10][3], const double b[4], double r[3])
{
    double h[10];
    p2_basis(b, h);
    for (int c = 0; c < 3; ++c) {
        double s = h[0] * g[0][c];
        for (int d = 1; d < 10; ++d) s += h[d] * g[d][c];
        r[c] = s;
    }
}

/* LocalP2GradientCL::set_tetra, num/gradient_recovery.cpp:62-76: the gradient of the tet's P2 level
 * set at its vertices, and at the edge midpoints the mean of the two vertex values */
static void local_p2_gradient(const double G[10][4][3], const double v[10], double out[10][3])
{
    double p1[4][3] = {{0.}};
    for (int i = 0; i < 10; ++i)
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c) p1[k][c] += v[i] * G[i][k][c];
    for (int k = 0; k < 4; ++k) memcpy(out[k], p1[k], sizeof p1[k]);
    for (int e = 0; e < 6; ++e)
        for (int c = 0; c < 3; ++c) out[4 + e][c] = 0.5 * (p1[VOE[e][0]][c] + p1[VOE[e][1]][c]);
}

/* averaging_P2_gradient_recovery (num/gradient_recovery.cpp:82-89) through OswaldProjectionP2AccuCL::visit
 * (num/oswald_projection.h:102-125): every tet of the mesh in id order, each node's value the running mean
 * ((n - 1)/n) avg + (1/n) local over the tets that contain it.  grad: 3 doubles per half-grid node. */
int orc_p2_recovered_gradient(const orc_mesh_t *m, const double *phi, double *grad)
{
    init_tables();
    if (!check_mesh(m) || !phi || !grad) return ORC_E_ARG;
    const int64_t nn = (2LL * m->n[0] + 1) * (2LL * m->n[1] + 1) * (2LL * m->n[2] + 1);
    const int64_t ntet = 6LL * m->n[0] * m->n[1] * m->n[2];
    double *cnt = (double *)calloc((size_t)nn, sizeof(double));
    if (!cnt) return ORC_E_NOMEM;
    for (int64_t i = 0; i < 3 * nn; ++i) grad[i] = 0.;
    for (int64_t t = 0; t < ntet; ++t) {
        tet_t T;
        load_tet(m, phi, t, &T);
        double Tm[3][3], det, G[10][4][3], v[10], loc[10][3];
        qq_gradients((const double (*)[3])T.X, Tm, &det, G);
        for (int d = 0; d < 10; ++d) v[d] = T.ls[LAT_OF_P2[d]];
        local_p2_gradient((const double (*)[4][3])G, v, loc);
        for (int i = 0; i < 10; ++i) {
            const int64_t g = p2_node_index(m, &T, i);
            const double n = (cnt[g] += 1.);
            for (int c = 0; c < 3; ++c) grad[3 * g + c] = ((n - 1.) / n) * grad[3 * g + c] + (1. / n) * loc[i][c];
        }
    }
    free(cnt);
    return ORC_OK;
}

static int qq_tet(const orc_mesh_t *m, const double *phi, const double *grad, int64_t t, qqtet_t *Q, tet_t *T)
{
    load_tet(m, phi, t, T);
    memcpy(Q->X, T->X, sizeof Q->X);
    qq_gradients((const double (*)[3])T->X, Q->T, &Q->det, Q->G);
    Q->h = cbrt(fabs(Q->det));
    for (int d = 0; d < 10; ++d) {
        Q->v[d] = T->ls[LAT_OF_P2[d]];
        const int64_t g = p2_node_index(m, T, d);
        for (int c = 0; c < 3; ++c) Q->gh[d][c] = grad[3 * g + c];
    }
    Q->w2b.R = Q->w2b.C = 4;
    for (int v = 0; v < 4; ++v) {
        for (int i = 0; i < 3; ++i) QA(&Q->w2b, i, v) = T->X[v][i];
        QA(&Q->w2b, 3, v) = 1.;
    }
    return qr_prepare(&Q->w2b) ? ORC_E_ARG : ORC_OK;
}

/* the state of QuaQuaMapperFunctionCL (levelset/levelsetmapper.cpp:112-290): point (y, s), its barycentric
 * coordinates in the tet, the recovered gradient there and F = (x0 - y - s gh, -ls) */
typedef struct {
    double xcur[4], bxb[4], gh[3], F[4], x0[3];
    int have_dF;
    qr_t dFqr;
    double dF[4][4];
} qqfun_t;

static void qqfun_compute_F(const qqtet_t *Q, qqfun_t *f)
{
    for (int i = 0; i < 3; ++i) f->F[i] = f->x0[i] - f->xcur[i] - f->xcur[3] * f->gh[i];
    f->F[3] = -p2_eval(Q->v, f->bxb);
}

/* set_point: the new point's barycentric coordinates in the same tet.  locate_new_point
 * (levelsetmapper.cpp:463-490) searches a neighbourhood only while its tetra pointer is null, which it
 * never is here, so points that leave the tet keep it: the level set and the recovered gradient are
 * extrapolated from the starting tet. */
static void qqfun_set_point(const qqtet_t *Q, qqfun_t *f, const double x[4])
{
    if (f->xcur[0] == x[0] && f->xcur[1] == x[1] && f->xcur[2] == x[2] && f->xcur[3] == x[3]) return;
    double b[4] = {x[0], x[1], x[2], 1.};
    qr_solve(&Q->w2b, b);
    f->have_dF = 0;
    memcpy(f->xcur, x, sizeof f->xcur);
    memcpy(f->bxb, b, sizeof b);
    p2v_eval((const double (*)[3])Q->gh, f->bxb, f->gh);
    qqfun_compute_F(Q, f);
}

/* compute_dF and compute_dFinv, levelsetmapper.cpp:222-270 */
static int qqfun_dF(const qqtet_t *Q, qqfun_t *f)
{
    if (f->have_dF) return 0;
    double g_ls[3] = {0., 0., 0.}, dgh[3][3] = {{0.}};
    for (int i = 0; i < 10; ++i) {
        double t[3];
        p1v_eval((const double (*)[3])Q->G[i], f->bxb, t);
        for (int c = 0; c < 3; ++c) g_ls[c] += Q->v[i] * t[c];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) dgh[r][c] += Q->gh[i][r] * t[c];
    }
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) f->dF[i][j] = -f->xcur[3] * dgh[i][j];
        f->dF[i][i] -= 1.;
        f->dF[i][3] = -f->gh[i];
        f->dF[3][i] = -g_ls[i];
    }
    f->dF[3][3] = 0.;
    f->dFqr.R = f->dFqr.C = 4;
    memcpy(f->dFqr.a, f->dF, sizeof f->dF);
    if (qr_prepare(&f->dFqr)) return 1;
    f->have_dF = 1;
    return 0;
}

static void mv4(const double A[4][4], const double v[4], double r[4])
{
    for (int i = 0; i < 4; ++i) r[i] = A[i][0] * v[0] + A[i][1] * v[1] + A[i][2] * v[2] + A[i][3] * v[3];
}

static double dot4(const double a[4], const double b[4])
{
    double s = 0.0;
    for (int i = 0; i < 4; ++i) s += a[i] * b[i];
    return s;
}

/* the base point of barycentric point xb of the tet: QuaQuaMapperCL::base_point_newton
 * (levelsetmapper.cpp:551-566) with newton_solve (num/newton.h:64-115).  Returns the base point's
 * barycentric coordinates (in the tet), dh and the Newton iteration count; 1 on a singular Jacobian. */
static int qq_base_point(const qqtet_t *Q, const orc_quaqua_t *P, const double xb[4], double bxb[4], double *dh,
                         int32_t *iters)
{
    qqfun_t f;
    memset(&f, 0, sizeof f);
    memcpy(f.bxb, xb, sizeof f.bxb);
    bary_to_world(Q->X, xb, f.x0);
    f.xcur[0] = f.x0[0]; f.xcur[1] = f.x0[1]; f.xcur[2] = f.x0[2]; f.xcur[3] = 0.;
    p2v_eval((const double (*)[3])Q->gh, f.bxb, f.gh);
    f.F[0] = f.F[1] = f.F[2] = 0.;
    f.F[3] = -p2_eval(Q->v, f.bxb);
    double x[4];
    memcpy(x, f.xcur, sizeof x);
    const double min_step_length = 1e-7;
    int iter;
    for (iter = 0;; ++iter) {
        qqfun_set_point(Q, &f, x);
        const double normF = sqrt(dot4(f.F, f.F));
        if (iter >= P->maxiter || normF < P->tol) break;
        if (qqfun_dF(Q, &f)) return 1;
        double dx[4], Ddx[4];
        memcpy(dx, f.F, sizeof dx);
        qr_solve(&f.dFqr, dx);
        mv4((const double (*)[4])f.dF, dx, Ddx);
        const double armijo_slope = P->armijo_c * dot4(f.F, Ddx) / normF;
        const double dxn = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        double l = 0.5 * Q->h / dxn;
        if (!(l < 1.)) l = 1.;
        for (int j = 0; j < P->max_damping && l >= min_step_length; ++j, l *= 0.5) {
            double xt[4];
            for (int i = 0; i < 4; ++i) xt[i] = x[i] - l * dx[i];
            qqfun_set_point(Q, &f, xt);
            if (sqrt(dot4(f.F, f.F)) < normF + armijo_slope * l) break;
        }
        if (l < min_step_length) break;
        for (int i = 0; i < 4; ++i) x[i] -= l * dx[i];
    }
    memcpy(bxb, f.bxb, sizeof f.bxb);
    *dh = x[3] * norm3(f.gh);
    if (iters) *iters = iter;
    return 0;
}

/* QuaQuaMapperCL::jacobian, levelsetmapper.cpp:590-638: dph = (I + dh Q dn)^-1 Q at the base point */
static int qq_jacobian(const qqtet_t *Q, const double bxb[4], double dh, double dph[3][3])
{
    double gh[3], gp[10][3], n[3] = {0., 0., 0.}, dgh[3][3] = {{0.}};
    p2v_eval((const double (*)[3])Q->gh, bxb, gh);
    const double ghn = norm3(gh);
    double qn[3];
    for (int c = 0; c < 3; ++c) qn[c] = gh[c] / ghn;
    for (int i = 0; i < 10; ++i) {
        p1v_eval((const double (*)[3])Q->G[i], bxb, gp[i]);
        for (int c = 0; c < 3; ++c) n[c] += Q->v[i] * gp[i][c];
    }
    const double nn = norm3(n);
    for (int c = 0; c < 3; ++c) n[c] /= nn;
    for (int i = 0; i < 10; ++i)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) dgh[r][c] += Q->gh[i][r] * gp[i][c];
    double tq[3];
    for (int col = 0; col < 3; ++col) {
        double s = 0.0;
        for (int i = 0; i < 3; ++i) s += dgh[i][col] * qn[i];
        tq[col] = s;
    }
    double dn[3][3], Qm[3][3];
    const double ig = 1. / ghn;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) dn[r][c] = ig * (dgh[r][c] - qn[r] * tq[c]);
    const double qnn = ip3(qn, n);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Qm[r][c] = (r == c ? 1. : 0.) - (qn[r] / qnn) * n[c];
    qr_t A;
    A.R = A.C = 3;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int i = 0; i < 3; ++i) s += (dh * Qm[r][i]) * dn[i][c];
            QA(&A, r, c) = (r == c ? 1. : 0.) + s;
        }
    if (qr_prepare(&A)) return 1;
    for (int i = 0; i < 3; ++i) {
        double t[3] = {Qm[0][i], Qm[1][i], Qm[2][i]};
        qr_solve(&A, t);
        for (int r = 0; r < 3; ++r) dph[r][i] = t[r];
    }
    return 0;
}

/* W^-1 of gradient_trafo (ifacetransp.cpp:894-918): W = dph^T + nl (n/(nl.n) - dph nl), nl the normal of
 * the patch's first facet, n the unit level-set gradient at the base point (QuaQuaMapperCL::local_ls_grad) */
static int qq_winv(const qqtet_t *Q, const double bxb[4], const double dph[3][3], const double nl[3], double Wi[3][3])
{
    double n[3] = {0., 0., 0.};
    for (int i = 0; i < 10; ++i) {
        double t[3];
        p1v_eval((const double (*)[3])Q->G[i], bxb, t);
        for (int c = 0; c < 3; ++c) n[c] += Q->v[i] * t[c];
    }
    const double nn = norm3(n);
    for (int c = 0; c < 3; ++c) n[c] /= nn;
    double dnl[3], tv[3];
    mv3(dph, nl, dnl);
    const double nln = ip3(nl, n);
    for (int c = 0; c < 3; ++c) tv[c] = n[c] / nln - dnl[c];
    qr_t W;
    W.R = W.C = 3;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) QA(&W, r, c) = dph[c][r] + nl[r] * tv[c];
    if (qr_prepare(&W)) return 1;
    for (int j = 0; j < 3; ++j) {
        double t[3] = {0., 0., 0.};
        t[j] = 1.;
        qr_solve(&W, t);
        for (int r = 0; r < 3; ++r) Wi[r][j] = t[r];
    }
    return 0;
}

/* W^-1 at barycentric point xb of the tet for the plane normal nl (row-major) */
int orc_quaqua_winv(const orc_mesh_t *m, const double *phi, const double *grad, int64_t tet, const double xb[4],
                    const double nl[3], const orc_quaqua_t *P, double winv[9])
{
    init_tables();
    if (!check_mesh(m) || !phi || !grad || !xb || !nl || !P || !winv) return ORC_E_ARG;
    qqtet_t Q;
    tet_t T;
    if (qq_tet(m, phi, grad, tet, &Q, &T)) return ORC_E_ARG;
    double bxb[4], dh, J[3][3], Wi[3][3];
    if (qq_base_point(&Q, P, xb, bxb, &dh, NULL) || qq_jacobian(&Q, bxb, dh, J) ||
        qq_winv(&Q, bxb, (const double (*)[3])J, nl, Wi))
        return ORC_E_NOT_CONVERGED;
    memcpy(winv, Wi, sizeof Wi);
    return ORC_OK;
}

int orc_quaqua_point(const orc_mesh_t *m, const double *phi, const double *grad, int64_t tet, const double xb[4],
                     const orc_quaqua_t *P, double bxb[4], double y[3], double *dh, double dph[9], int32_t *iters)
{
    init_tables();
    if (!check_mesh(m) || !phi || !grad || !xb || !P || !bxb || !y || !dh || !dph) return ORC_E_ARG;
    qqtet_t Q;
    tet_t T;
    if (qq_tet(m, phi, grad, tet, &Q, &T)) return ORC_E_ARG;
    if (qq_base_point(&Q, P, xb, bxb, dh, iters)) return ORC_E_NOT_CONVERGED;
    bary_to_world(Q.X, bxb, y);
    double J[3][3];
    if (qq_jacobian(&Q, bxb, *dh, J)) return ORC_E_NOT_CONVERGED;
    memcpy(dph, J, sizeof J);
    return ORC_OK;
}

/* Local P2 matrices of one tet on the mapped surface: the lattice-1 patch and its composite Quad5 rule
 * (InterfaceCommonDataP2CL::visit, surfactant/ifacetransp.h:258-276), |det| of the map per point
 * (ProjectedQuadDomain2DCL::assign, ifacetransp.cpp:850-892), W^-1 (gradient_trafo, ifacetransp.cpp:894-918)
 * and LocalMassP2CL / LocalLaplaceBeltramiP2CL / LocalVectorP2CL (ifacetransp.h:2302-2403; the
 * Laplace-Beltrami form carries no coefficient there).  out->nfacet = 0: the patch is empty. */
static int local_p2qq(const qqtet_t *Q, const tet_t *T, const orc_quaqua_t *P, int rhs_kind, const double *rhs_param,
                      orc_local_p2_t *o)
{
    memset(o, 0, sizeof *o);
    double ls4[4];
    for (int k = 0; k < 4; ++k) ls4[k] = Q->v[k];
    patch_t p;
    int err = make_patch_l1(ls4, &p);
    if (err) return err;
    if (p.nf == 0) return ORC_OK;
    patch_geometry(Q->X, &p);
    for (int f = 0; f < p.nf; ++f)
        if (p.deficient[f]) return ORC_E_NOT_CONVERGED;
    o->nfacet = p.nf;
    double node[MAXF * 7][4], w[MAXF * 7];
    const int nq = composite_nodes(&p, 7, Q5_N, Q5_W, node, w);
    double H[MAXF * 7][10], Gq[MAXF * 7][10][3], ad[MAXF * 7], fq[MAXF * 7];
    double U[3][2] = {{0.}};
    for (int k = 0; k < nq; ++k) {
        const int f = k / 7;
        if (k % 7 == 0) {
            qr_t M;
            M.R = 3; M.C = 2;
            const double *w0 = p.vw[p.facet[f][0]], *w1 = p.vw[p.facet[f][1]], *w2 = p.vw[p.facet[f][2]];
            for (int r = 0; r < 3; ++r) {
                QA(&M, r, 0) = w1[r] - w0[r];
                QA(&M, r, 1) = w2[r] - w0[r];
            }
            if (qr_prepare(&M)) return ORC_E_NOT_CONVERGED;
            for (int j = 0; j < 2; ++j) {
                double t[3] = {0., 0., 0.};
                t[j] = 1.;
                qr_apply_Q(&M, t);
                for (int r = 0; r < 3; ++r) U[r][j] = t[r];
            }
        }
        double bxb[4], dh, dph[3][3];
        if (qq_base_point(Q, P, node[k], bxb, &dh, NULL)) return ORC_E_NOT_CONVERGED;
        if (qq_jacobian(Q, bxb, dh, dph)) return ORC_E_NOT_CONVERGED;
        /* |det|: sqrt(det Gram(dph U)), GramMatrix misc/container.h:935-950 */
        double DU[3][2];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 2; ++c) {
                double s = 0.0;
                for (int i = 0; i < 3; ++i) s += dph[r][i] * U[i][c];
                DU[r][c] = s;
            }
        double g00 = 0.0, g10 = 0.0, g11 = 0.0;
        for (int i = 0; i < 3; ++i) g00 += DU[i][0] * DU[i][0];
        for (int i = 0; i < 3; ++i) g10 += DU[i][1] * DU[i][0];
        for (int i = 0; i < 3; ++i) g11 += DU[i][1] * DU[i][1];
        ad[k] = sqrt(g00 * g11 - g10 * g10);
        double Wi[3][3];
        if (qq_winv(Q, bxb, (const double (*)[3])dph, p.normal[0], Wi)) return ORC_E_NOT_CONVERGED;
        /* the P2 gradients at the planar point, projected on the facet's plane, times W^-1 */
        const double *nf = p.normal[f];
        for (int i = 0; i < 10; ++i) {
            double g[3], t[3];
            p1v_eval((const double (*)[3])Q->G[i], node[k], g);
            const double d = ip3(nf, g);
            for (int c = 0; c < 3; ++c) t[c] = g[c] - d * nf[c];
            mv3((const double (*)[3])Wi, t, Gq[k][i]);
        }
        p2_basis(node[k], H[k]);
        fq[k] = 0.;
        if (rhs_kind != ORC_RHS_NONE) {
            double yw[3];
            bary_to_world(Q->X, bxb, yw);
            fq[k] = rhs_fun(rhs_kind, rhs_param, yw);
        }
    }
    for (int i = 0; i < 10; ++i) {
        for (int j = 0; j <= i; ++j) {
            double sm = 0.0, sl = 0.0;
            for (int k = 0; k < nq; ++k) {
                sm += w[k] * ((ad[k] * H[k][j]) * H[k][i]);
                sl += w[k] * (ad[k] * ip3(Gq[k][j], Gq[k][i]));
            }
            o->mass[i * 10 + j] = o->mass[j * 10 + i] = sm;
            o->lb[i * 10 + j] = o->lb[j * 10 + i] = sl;
        }
        double sr = 0.0;
        if (rhs_kind != ORC_RHS_NONE)
            for (int k = 0; k < nq; ++k) sr += w[k] * ((ad[k] * fq[k]) * H[k][i]);
        o->rhs[i] = sr;
    }
    return ORC_OK;
}

int orc_tet_local_p2qq(const orc_mesh_t *m, const double *phi, const double *grad, int64_t tet,
                       const orc_quaqua_t *P, int rhs_kind, const double *rhs_param, int zero_policy,
                       orc_local_p2_t *out)
{
    (void)zero_policy;
    init_tables();
    if (!check_mesh(m) || !phi || !grad || !P || !out) return ORC_E_ARG;
    qqtet_t Q;
    tet_t T;
    if (qq_tet(m, phi, grad, tet, &Q, &T)) return ORC_E_ARG;
    return local_p2qq(&Q, &T, P, rhs_kind, rhs_param, out);
}

/* InterfaceMatrixAccuCL<LocalMassP2CL / LocalLaplaceBeltramiP2CL> and InterfaceVectorAccuCL<LocalVectorP2CL>
 * over InterfaceCommonDataP2CL (surfactant/surfactant.cpp:1600-1640): the tets whose lattice-1 patch is
 * not empty, in tet order, their couplings added into the P2 numbering's rows (an entry exists where
 * such a tet couples two DOFs).  grad: the recovered gradient (orc_p2_recovered_gradient). */
int orc_assemble_p2qq(const orc_mesh_t *m, const double *phi, const double *grad, const int32_t *dof_of_node,
                      int64_t ndof, const orc_quaqua_t *P, int rhs_kind, const double *rhs_param, int zero_policy,
                      int64_t cap_nnz, int64_t *nnz, int64_t *rowptr, int32_t *colind, double *val_mass,
                      double *val_lb, double *rhs)
{
    init_tables();
    if (!check_mesh(m) || !phi || !grad || !dof_of_node || !P || !nnz || !rowptr) return ORC_E_ARG;
    int64_t ncut = 0;
    int err = orc_classify(m, phi, zero_policy, 0, &ncut, NULL);
    if (err) return err;
    int64_t *cut = (int64_t *)malloc(sizeof(int64_t) * (ncut > 0 ? ncut : 1));
    orc_local_p2_t *loc = (orc_local_p2_t *)malloc(sizeof(orc_local_p2_t) * (ncut > 0 ? ncut : 1));
    int32_t (*numb)[10] = (int32_t (*)[10])malloc(sizeof(int32_t) * 10 * (ncut > 0 ? ncut : 1));
    int *st = (int *)calloc(ncut > 0 ? ncut : 1, sizeof(int));
    brow2_t *rows = (brow2_t *)calloc(ndof > 0 ? ndof : 1, sizeof(brow2_t));
    if (!cut || !loc || !numb || !st || !rows) {
        free(cut); free(loc); free(numb); free(st); free(rows);
        return ORC_E_NOMEM;
    }
    orc_classify(m, phi, zero_policy, ncut, &ncut, cut);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t c = 0; c < ncut; ++c) {
        qqtet_t Q;
        tet_t T;
        if (qq_tet(m, phi, grad, cut[c], &Q, &T)) { st[c] = ORC_E_ARG; continue; }
        const int e = local_p2qq(&Q, &T, P, rhs_kind, rhs_param, &loc[c]);
        if (e) { st[c] = e; continue; }
        if (loc[c].nfacet == 0) { st[c] = -1; continue; }
        for (int i = 0; i < 10; ++i) numb[c][i] = dof_of_node[p2_node_index(m, &T, i)];
    }
    if (rhs)
        for (int64_t i = 0; i < ndof; ++i) rhs[i] = 0.;
    for (int64_t c = 0; c < ncut && !err; ++c) {
        if (st[c] > 0) { err = st[c]; break; }
        if (st[c] < 0) continue;
        for (int i = 0; i < 10 && !err; ++i) {
            const int32_t r = numb[c][i];
            if (r < 0) continue;
            brow2_t *R = &rows[r];
            for (int j = 0; j < 10; ++j) {
                const int32_t cj = numb[c][j];
                if (cj < 0) continue;
                int k = 0;
                for (; k < R->n; ++k)
                    if (R->col[k] == cj) break;
                if (k == R->n) {
                    if (R->n == ROWCAP2) { err = ORC_E_CAPACITY; break; }
                    R->col[k] = cj;
                    R->v[0][k] = R->v[1][k] = R->v[2][k] = 0.0;
                    ++R->n;
                }
                R->v[0][k] += loc[c].mass[i * 10 + j];
                R->v[1][k] += loc[c].lb[i * 10 + j];
            }
            if (rhs) rhs[r] += loc[c].rhs[i];
        }
    }
    if (!err) {
        rowptr[0] = 0;
        for (int64_t r = 0; r < ndof; ++r) rowptr[r + 1] = rowptr[r] + rows[r].n;
        *nnz = rowptr[ndof];
        if (*nnz > cap_nnz) err = ORC_E_CAPACITY;
    }
    if (!err && colind) {
        for (int64_t r = 0; r < ndof; ++r) {
            brow2_t *R = &rows[r];
            int32_t ord[ROWCAP2];
            for (int k = 0; k < R->n; ++k) ord[k] = k;
            for (int a = 1; a < R->n; ++a) {
                int32_t x = ord[a];
                int b = a - 1;
                while (b >= 0 && R->col[ord[b]] > R->col[x]) { ord[b + 1] = ord[b]; --b; }
                ord[b + 1] = x;
            }
            for (int k = 0; k < R->n; ++k) {
                const int64_t dst = rowptr[r] + k;
                colind[dst] = R->col[ord[k]];
                if (val_mass) val_mass[dst] = R->v[0][ord[k]];
                if (val_lb) val_lb[dst] = R->v[1][ord[k]];
            }
        }
    }
    free(cut); free(loc); free(numb); free(st); free(rows);
    return err;
}

/* ------------------------------------------------------------------------ */
/* Cahn-Hilliard step pieces (CahnHilliardcGP1CL, resting interface)          */
/* ------------------------------------------------------------------------ */

/* Mobility_function, surfactant/ifacetransp.cpp:3334-3339 */
static double mobility_fun(double x)
{
    const double v = (1. - x) * x;
    return v > 0. ? v : 0.;
}

/* Potential_prime_function, surfactant/ifacetransp.cpp:3360-3364 */
static double potential_prime(double x) { return (x - 1.) * x * (x - 0.5); }

/* LocalLaplaceMobilityP1CL::setup, surfactant/ifacetransp.h:1731-1766: Quad5 composite rule,
 * mobility = P1 interpolant of Mobility_function at the vertices (LocalP1CL, evaluated as
 * FE_P1CL::val, num/fe.h:150-153), projected P1 gradients with the normal at each quadrature
 * point; the normal is the normalised gradient of the P2 level set. */
static void local_mobility(const tet_t *T, const patch_t *p, const double cv[4], double out[16])
{
    double node[MAXF * 7][4], w[MAXF * 7];
    const int nq = composite_nodes(p, 7, Q5_N, Q5_W, node, w);
    double G1[4][3];
    p1_gradients(T->X, G1);
    double v[10];
    for (int d = 0; d < 10; ++d) v[d] = T->ls[LAT_OF_P2[d]];
    double mv[4];
    for (int i = 0; i < 4; ++i) mv[i] = mobility_fun(cv[i]);
    double qq[4][MAXF * 7][3], qm[MAXF * 7];
    for (int k = 0; k < nq; ++k) {
        double nrm[3];
        p2_normal(v, node[k], G1, nrm);
        for (int j = 0; j < 4; ++j) {
            const double d = dot3(G1[j], nrm);
            for (int c = 0; c < 3; ++c) qq[j][k][c] = G1[j][c] - d * nrm[c];
        }
        const double *b = node[k];
        qm[k] = b[0] * mv[0] + b[1] * mv[1] + b[2] * mv[2] + b[3] * mv[3];
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.;
            for (int k = 0; k < nq; ++k) s += w[k] * (qm[k] * dot3(qq[i][k], qq[j][k]));
            out[i * 4 + j] = s;
        }
}

/* The P2 counterpart: mobility = P2 interpolant sum_d Mob(c_d) h_d of the nodal mobilities
 * (LinearCombinationCL order, num/fe.h:44-53), projected P2 gradients (buildSurfGradP2,
 * surfactant/ifacetransp.h:1049-1053) with the P2 level-set normal, Quad5 composite rule. */
static void local_mobility_p2(const tet_t *T, const patch_t *p, const double cv[10], double out[100])
{
    double node[MAXF * 7][4], w[MAXF * 7];
    const int nq = composite_nodes(p, 7, Q5_N, Q5_W, node, w);
    double G1[4][3];
    p1_gradients(T->X, G1);
    double v[10], mv[10];
    for (int d = 0; d < 10; ++d) {
        v[d] = T->ls[LAT_OF_P2[d]];
        mv[d] = mobility_fun(cv[d]);
    }
    for (int e = 0; e < 100; ++e) out[e] = 0.;
    for (int k = 0; k < nq; ++k) {
        double h[10], g[10][3], sg[10][3], nrm[3];
        p2_basis(node[k], h);
        p2_gradients(node[k], G1, g);
        p2_normal(v, node[k], G1, nrm);
        for (int i = 0; i < 10; ++i) {
            const double d = dot3(g[i], nrm);
            for (int c = 0; c < 3; ++c) sg[i][c] = g[i][c] - d * nrm[c];
        }
        double qm = h[0] * mv[0];
        for (int d = 1; d < 10; ++d) qm += h[d] * mv[d];
        for (int i = 0; i < 10; ++i)
            for (int j = 0; j < 10; ++j) out[i * 10 + j] += w[k] * (qm * dot3(sg[i], sg[j]));
    }
}

int orc_tet_mobility_p2(const orc_mesh_t *m, const double *phi, int64_t tet, const double cv[10], int zero_policy,
                        double out[100])
{
    init_tables();
    if (!check_mesh(m) || !phi || !cv || !out) return ORC_E_ARG;
    memset(out, 0, 100 * sizeof(double));
    tet_t T;
    load_tet(m, phi, tet, &T);
    if (equal_signs(T.ls, 10)) return ORC_OK;
    if (zero_policy == ORC_ZERO_THROW && has_zero(T.ls, 10)) return ORC_E_ZERO_NODE;
    patch_t p;
    int err = make_patch(T.ls, &p);
    if (err) return err;
    if (p.nf == 0) return ORC_OK;
    patch_geometry(T.X, &p);
    local_mobility_p2(&T, &p, cv, out);
    return ORC_OK;
}

int orc_tet_mobility(const orc_mesh_t *m, const double *phi, int64_t tet, const double cv[4], int zero_policy,
                     double out[16])
{
    init_tables();
    if (!check_mesh(m) || !phi || !cv || !out) return ORC_E_ARG;
    memset(out, 0, 16 * sizeof(double));
    tet_t T;
    load_tet(m, phi, tet, &T);
    if (equal_signs(T.ls, 10)) return ORC_OK;
    if (zero_policy == ORC_ZERO_THROW && has_zero(T.ls, 10)) return ORC_E_ZERO_NODE;
    patch_t p;
    int err = make_patch(T.ls, &p);
    if (err) return err;
    if (p.nf == 0) return ORC_OK;
    patch_geometry(T.X, &p);
    local_mobility(&T, &p, cv, out);
    return ORC_OK;
}

/* CahnHilliardcGP1CL::Update (LaplaceM, surfactant/ifacetransp.cpp:1560-1561) and InitStep with
 * theta = 1 (:1610-1641): the mobility matrix from concentration c on the given pattern, and the
 * vectors timeder = LocalMatVecP1CL<LocalInterfaceMassP1CL>(c) and
 * well = LocalMatVecP1CL<LocalInterfaceMassP1CL>(f'(c)) (surfactant/ifacetransp.h:800-828), each
 * accumulated in mesh order. */
static int ch_assemble_sp(int space, const orc_mesh_t *m, const double *phi, const int32_t *dof_of, int64_t ndof,
                          const int64_t *rowptr, const int32_t *colind, const double *c, int zero_policy,
                          double *lm, double *timeder, double *well)
{
    init_tables();
    if (!check_mesh(m) || !phi || !dof_of || !rowptr || !colind || !c || !lm || !timeder || !well)
        return ORC_E_ARG;
    const int nl = space == 2 ? 10 : 4;
    int64_t ncut = 0;
    int err = orc_classify(m, phi, zero_policy, 0, &ncut, NULL);
    if (err) return err;
    int64_t *cut = (int64_t *)malloc(sizeof(int64_t) * (ncut > 0 ? ncut : 1));
    typedef struct { int err, nonempty; int32_t numb[10]; double mass[100], mob[100]; } chrec_t;
    chrec_t *rec = (chrec_t *)malloc(sizeof(chrec_t) * (ncut > 0 ? ncut : 1));
    if (!cut || !rec) { free(cut); free(rec); return ORC_E_NOMEM; }
    orc_classify(m, phi, zero_policy, ncut, &ncut, cut);
    const double one3[3] = {1., 1., 1.};
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < ncut; ++q) {
        tet_t T;
        load_tet(m, phi, cut[q], &T);
        patch_t p;
        rec[q].nonempty = 0;
        rec[q].err = make_patch(T.ls, &p);
        if (rec[q].err || p.nf == 0) continue;
        rec[q].nonempty = 1;
        patch_geometry(T.X, &p);
        double cv[10];
        for (int i = 0; i < nl; ++i) {
            rec[q].numb[i] = space == 2 ? dof_of[p2_node_index(m, &T, i)] : dof_of[T.vid[i]];
            cv[i] = rec[q].numb[i] >= 0 ? c[rec[q].numb[i]] : 0.;
        }
        if (space == 2) {
            orc_local_p2_t L;
            local_quantities_p2(&T, &p, one3, ORC_RHS_NONE, NULL, &L);
            memcpy(rec[q].mass, L.mass, sizeof L.mass);
            local_mobility_p2(&T, &p, cv, rec[q].mob);
        } else {
            orc_local_t L;
            local_quantities(&T, &p, one3, ORC_RHS_NONE, NULL, &L);
            memcpy(rec[q].mass, L.mass, sizeof L.mass);
            local_mobility(&T, &p, cv, rec[q].mob);
        }
    }
    for (int64_t e = 0; e < rowptr[ndof]; ++e) lm[e] = 0.;
    for (int64_t i = 0; i < ndof; ++i) timeder[i] = well[i] = 0.;
    for (int64_t q = 0; q < ncut && !err; ++q) {
        if (rec[q].err) { err = rec[q].err; break; }
        if (!rec[q].nonempty) continue;
        const int32_t *nb = rec[q].numb;
        double xc[10], xf[10];
        for (int j = 0; j < nl; ++j) {
            xc[j] = nb[j] >= 0 ? c[nb[j]] : 0.;
            xf[j] = potential_prime(xc[j]);
        }
        for (int i = 0; i < nl; ++i) {
            if (nb[i] < 0) continue;
            double vt = 0., vw = 0.;
            for (int j = 0; j < nl; ++j) {
                if (nb[j] < 0) continue;
                vt += rec[q].mass[i * nl + j] * xc[j];
                vw += rec[q].mass[i * nl + j] * xf[j];
                int64_t e = rowptr[nb[i]];
                while (e < rowptr[nb[i] + 1] && colind[e] != nb[j]) ++e;
                if (e == rowptr[nb[i] + 1]) { err = ORC_E_ARG; break; }
                lm[e] += rec[q].mob[i * nl + j];
            }
            timeder[nb[i]] += vt;
            well[nb[i]] += vw;
        }
    }
    free(cut);
    free(rec);
    return err;
}

int orc_ch_assemble(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                    const int64_t *rowptr, const int32_t *colind, const double *c, int zero_policy,
                    double *lm, double *timeder, double *well)
{
    return ch_assemble_sp(1, m, phi, dof_of_vertex, ndof, rowptr, colind, c, zero_policy, lm, timeder, well);
}

/* The same on the P2 interface space: dof_of_node over the half-grid nodes, P2 local mass and
 * P2 mobility matrix (local_mobility_p2), 10 local dofs per cut tet. */
int orc_ch_assemble_p2(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_node, int64_t ndof,
                       const int64_t *rowptr, const int32_t *colind, const double *c, int zero_policy,
                       double *lm, double *timeder, double *well)
{
    return ch_assemble_sp(2, m, phi, dof_of_node, ndof, rowptr, colind, c, zero_policy, lm, timeder, well);
}

/* The step system of CahnHilliardcGP1CL::DoStep (surfactant/ifacetransp.cpp:1823-1856), theta = 1,
 * no wind, no sources: block CSR (SparseMatBaseCL(A, B, C, D), num/spmat.h:719-751), right-hand
 * side [M c; -M f'(c) + S M c] and the block-Jacobi diagonal [diag Precond3; diag Precond4]
 * (:1567-1570), each LinComb evaluated in the reference's order. */
/* c_r == NULL: the implicit-Euler step from c.  Otherwise a BDF2 step: mobility, M f'(c) and
 * S M c from the extrapolation c (= c_e) and the time-derivative part of rhs3 = M c_r, with dt the
 * effective step 1/alpha; the M c_r products are LocalMatVec accumulations like M c. */
static int ch_system_sp(int space, const orc_mesh_t *m, const double *phi, const int32_t *dof_of, int64_t ndof,
                        const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                        const double *stab, const double *c, const double *c_r, double dt, double eps, double rho,
                        double sigma, double S, int64_t *brp, int32_t *bcol, double *bval, double *rhs,
                        double *pcdiag)
{
    const int64_t nnz = rowptr[ndof];
    double *lm = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    double *td = (double *)malloc(sizeof(double) * (ndof > 0 ? ndof : 1));
    double *wl = (double *)malloc(sizeof(double) * (ndof > 0 ? ndof : 1));
    double *tr = c_r ? (double *)malloc(sizeof(double) * (ndof > 0 ? ndof : 1)) : td;
    double *lm2 = c_r ? (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1)) : lm;
    double *wl2 = c_r ? (double *)malloc(sizeof(double) * (ndof > 0 ? ndof : 1)) : wl;
    if (!lm || !td || !wl || !tr || !lm2 || !wl2) {
        free(lm); free(td); free(wl);
        if (c_r) { free(tr); free(lm2); free(wl2); }
        return ORC_E_NOMEM;
    }
    int err = ch_assemble_sp(space, m, phi, dof_of, ndof, rowptr, colind, c, ORC_ZERO_THROW, lm, td, wl);
    if (!err && c_r) err = ch_assemble_sp(space, m, phi, dof_of, ndof, rowptr, colind, c_r, ORC_ZERO_THROW, lm2, tr, wl2);
    if (!err) {
        const double sdt = sigma * dt, rdt = rho * dt, dtr = dt * rho, e2 = eps * eps, re2 = rho * eps * eps,
                     e2r = eps * eps * rho, meps = -eps;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < ndof; ++i) {
            const int64_t e0 = rowptr[i], len = rowptr[i + 1] - e0;
            const int64_t top = 2 * e0, bot = 2 * nnz + 2 * e0;
            brp[i] = top;
            brp[ndof + i] = bot;
            for (int64_t k = 0; k < len; ++k) {
                const int64_t e = e0 + k;
                const int32_t cj = colind[e];
                bcol[top + k] = cj;
                bval[top + k] = sdt * lm[e] + rdt * stab[e];
                bcol[top + len + k] = (int32_t)(cj + ndof);
                bval[top + len + k] = 1.0 * (1.0 * mass[e] + 0.0 * 0.0) + dt * 0.0;
                bcol[bot + k] = cj;
                bval[bot + k] = 0.0 * lb[e] + meps * mass[e];
                bcol[bot + len + k] = (int32_t)(cj + ndof);
                bval[bot + len + k] = 1.0 * (e2 * lb[e] + S * mass[e]) + re2 * stab[e];
                if (cj == i) {
                    pcdiag[i] = 1.0 * (1.0 * mass[e] + sdt * lm[e]) + dtr * stab[e];
                    pcdiag[ndof + i] = 1.0 * (1.0 * mass[e] + e2 * lb[e]) + e2r * stab[e];
                }
            }
            rhs[i] = 1.0 * (tr[i] + dt * 0.0);
            rhs[ndof + i] = ((-1.0) * wl[i] + S * td[i]) + 0.0;
        }
        brp[2 * ndof] = 4 * nnz;
    }
    free(lm); free(td); free(wl);
    if (c_r) { free(tr); free(lm2); free(wl2); }
    return err;
}

int orc_ch_system(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                  const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                  const double *stab, const double *c, double dt, double eps, double rho, double sigma, double S,
                  int64_t *brp, int32_t *bcol, double *bval, double *rhs, double *pcdiag)
{
    return ch_system_sp(1, m, phi, dof_of_vertex, ndof, rowptr, colind, mass, lb, stab, c, NULL, dt, eps, rho, sigma,
                        S, brp, bcol, bval, rhs, pcdiag);
}

int orc_ch_system_p2(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_node, int64_t ndof,
                     const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                     const double *stab, const double *c, double dt, double eps, double rho, double sigma,
                     double S, int64_t *brp, int32_t *bcol, double *bval, double *rhs, double *pcdiag)
{
    return ch_system_sp(2, m, phi, dof_of_node, ndof, rowptr, colind, mass, lb, stab, c, NULL, dt, eps, rho, sigma,
                        S, brp, bcol, bval, rhs, pcdiag);
}

/* one step: system + right-preconditioned GMRES(restart) from x = [mu; c] (in/out) */
static int ch_step_sp(int space, const orc_mesh_t *m, const double *phi, const int32_t *dof_of, int64_t ndof,
                      const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                      const double *stab, double *x, double dt, double eps, double rho, double sigma, double S,
                      int32_t restart, int32_t *max_iter, double *tol, int32_t rel)
{
    const int64_t nnz = rowptr[ndof], n2 = 2 * ndof;
    int64_t *brp = (int64_t *)malloc(sizeof(int64_t) * (n2 + 1));
    int32_t *bcol = (int32_t *)malloc(sizeof(int32_t) * (4 * nnz + 1));
    double *bval = (double *)malloc(sizeof(double) * (4 * nnz + 1));
    double *rhs = (double *)malloc(sizeof(double) * (n2 + 1));
    double *pcd = (double *)malloc(sizeof(double) * (n2 + 1));
    int err = (!brp || !bcol || !bval || !rhs || !pcd) ? ORC_E_NOMEM : ORC_OK;
    if (!err)
        err = ch_system_sp(space, m, phi, dof_of, ndof, rowptr, colind, mass, lb, stab, x + ndof, NULL, dt, eps,
                           rho, sigma, S, brp, bcol, bval, rhs, pcd);
    if (!err) err = orc_gmres_diag(n2, brp, bcol, bval, pcd, rhs, x, restart, max_iter, tol, rel, 0);
    free(brp); free(bcol); free(bval); free(rhs); free(pcd);
    return err;
}

int orc_ch_step(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                const double *stab, double *x, double dt, double eps, double rho, double sigma, double S,
                int32_t restart, int32_t *max_iter, double *tol, int32_t rel)
{
    return ch_step_sp(1, m, phi, dof_of_vertex, ndof, rowptr, colind, mass, lb, stab, x, dt, eps, rho, sigma, S,
                      restart, max_iter, tol, rel);
}

int orc_ch_step_p2(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_node, int64_t ndof,
                   const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                   const double *stab, double *x, double dt, double eps, double rho, double sigma, double S,
                   int32_t restart, int32_t *max_iter, double *tol, int32_t rel)
{
    return ch_step_sp(2, m, phi, dof_of_node, ndof, rowptr, colind, mass, lb, stab, x, dt, eps, rho, sigma, S,
                      restart, max_iter, tol, rel);
}

/* ---- BDF1/BDF2 with the adaptive step size, surfphasesep/surfphasesep.cpp:112-123, 396-491 ---- */

/* r = dt / dt_prev.  The BDF2 row of the time derivative, alpha chi - F_chi with
 * alpha = (1 + 2r) / ((1 + r) dt) and F_chi = (1 + r)/dt M chi_n - r^2/((1 + r) dt) M chi_{n-1}
 * (:451-468), scaled by dt_eff = 1/alpha to the implicit-Euler layout of the step system:
 *   M c + dt_eff A mu = M c_r,  c_r = w3 c_n + w4 c_{n-1}.
 * The extrapolation chi_extrap = (1 + r) chi_n - r chi_{n-1} (:452) gives c_e = w1 c_n + w2 c_{n-1}.
 * w = {dt_eff, w1, w2, w3, w4}. */
void orc_bdf2_weights(double dt, double dt_prev, double w[5])
{
    const double r = dt / dt_prev;
    w[0] = (1. + r) * dt / (1. + 2. * r);
    w[1] = 1. + r;
    w[2] = -r;
    w[3] = (1. + r) * (1. + r) / (1. + 2. * r);
    w[4] = -(r * r) / (1. + 2. * r);
}

/* the step-size function F(e, dt) (:117-123): the initial step for e == 0, else
 * rho sqrt(tol / e) dt clamped to [min_step, max_step] */
double orc_bdf_step_size(const orc_bdf_params_t *p, double e, double dt)
{
    if (!e) return p->step_size;
    const double f = p->rho * sqrt(p->tol / e) * dt;
    if (f > p->max_step) return p->max_step;
    if (f < p->min_step) return p->min_step;
    return f;
}

int orc_ch_system_bdf(int space, const orc_mesh_t *m, const double *phi, const int32_t *dof_of, int64_t ndof,
                      const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                      const double *stab, const double *c_e, const double *c_r, double dt_eff, double eps,
                      double rho, double sigma, double S, int64_t *brp, int32_t *bcol, double *bval, double *rhs,
                      double *pcdiag)
{
    if ((space != 1 && space != 2) || !c_r) return ORC_E_ARG;
    return ch_system_sp(space, m, phi, dof_of, ndof, rowptr, colind, mass, lb, stab, c_e, c_r, dt_eff, eps, rho,
                        sigma, S, brp, bcol, bval, rhs, pcdiag);
}

/* One time step t_{i-1} -> t_i of the loop at :397-491.  x = [mu; c] holds the accepted state,
 * cprev the concentration one step back (read from the second step on), st the loop's carried
 * values.  Every attempt: dt = F(e, dt) on the first step or when adaptive; a BDF1 solve from the
 * solver's previous vector (the reference keeps one lhs across solves, :319); on the first step
 * the BDF1 solution is taken and the loop ends (:446-450); else a BDF2 solve started from the BDF1
 * solution and the error e = sqrt((d, M d) / (c2, M c2)), d = c2 - c1 (:477-479).  Retried while
 * adaptive and e > tol and min_step < dt < max_step (:483).  On acceptance the last solve's
 * [mu; c] becomes the state (:485-488).  The reference retries without bound; max_tries > 0 ends
 * the step with ORC_E_NOT_CONVERGED after that many attempts (state unchanged). */
int orc_ch_bdf_step(int space, const orc_mesh_t *m, const double *phi, const int32_t *dof_of, int64_t ndof,
                    const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                    const double *stab, double *x, double *cprev, orc_bdf_state_t *st, const orc_bdf_params_t *bp,
                    double eps, double rho, double sigma, double S, int32_t restart, int32_t max_iter, double tol,
                    int32_t rel, orc_bdf_stat_t *stat)
{
    if ((space != 1 && space != 2) || !x || !cprev || !st || !bp || ndof < 0) return ORC_E_ARG;
    const int64_t n = ndof, n2 = 2 * ndof, nnz = rowptr[ndof];
    int64_t *brp = (int64_t *)malloc(sizeof(int64_t) * (n2 + 1));
    int32_t *bcol = (int32_t *)malloc(sizeof(int32_t) * (4 * nnz + 1));
    double *bval = (double *)malloc(sizeof(double) * (4 * nnz + 1));
    double *rhs = (double *)malloc(sizeof(double) * (n2 + 1));
    double *pcd = (double *)malloc(sizeof(double) * (n2 + 1));
    double *xl = (double *)malloc(sizeof(double) * (n2 + 1));
    double *c1 = (double *)malloc(sizeof(double) * (n + 1));
    double *ce = (double *)malloc(sizeof(double) * (n + 1));
    double *cr = (double *)malloc(sizeof(double) * (n + 1));
    int err = (!brp || !bcol || !bval || !rhs || !pcd || !xl || !c1 || !ce || !cr) ? ORC_E_NOMEM : ORC_OK;
    const int64_t i = st->step + 1;
    const double t_prev = st->t, dt_prev = st->dt;
    double dt = st->dt, e = st->e, resid = 0.;
    int32_t tries = 0, it1 = 0, it2 = 0;
    if (!err) memcpy(xl, x, sizeof(double) * n2);
    while (!err) {
        if (bp->max_tries > 0 && tries == bp->max_tries) { err = ORC_E_NOT_CONVERGED; break; }
        ++tries;
        if (i == 1 || bp->adaptive) dt = orc_bdf_step_size(bp, e, dt);
        /* BDF1: the implicit-Euler system from chi_n with alpha = 1/dt */
        err = ch_system_sp(space, m, phi, dof_of, n, rowptr, colind, mass, lb, stab, x + n, NULL, dt, eps, rho,
                           sigma, S, brp, bcol, bval, rhs, pcd);
        int32_t it = max_iter;
        double tl = tol;
        if (!err) err = orc_gmres_diag(n2, brp, bcol, bval, pcd, rhs, xl, restart, &it, &tl, rel, 0);
        it1 += it;
        resid = tl;
        if (err) break;
        memcpy(c1, xl + n, sizeof(double) * n);
        if (i == 1) break;
        double w[5];
        orc_bdf2_weights(dt, dt_prev, w);
        for (int64_t k = 0; k < n; ++k) {
            ce[k] = w[1] * x[n + k] + w[2] * cprev[k];
            cr[k] = w[3] * x[n + k] + w[4] * cprev[k];
        }
        err = ch_system_sp(space, m, phi, dof_of, n, rowptr, colind, mass, lb, stab, ce, cr, w[0], eps, rho, sigma,
                           S, brp, bcol, bval, rhs, pcd);
        it = max_iter;
        tl = tol;
        if (!err) err = orc_gmres_diag(n2, brp, bcol, bval, pcd, rhs, xl, restart, &it, &tl, rel, 0);
        it2 += it;
        resid = tl;
        if (err) break;
        double dd = 0., cc = 0.;
        for (int64_t r = 0; r < n; ++r) {
            double md = 0., mc = 0.;
            for (int64_t q = rowptr[r]; q < rowptr[r + 1]; ++q) {
                const int32_t j = colind[q];
                md += mass[q] * (xl[n + j] - c1[j]);
                mc += mass[q] * xl[n + j];
            }
            dd += (xl[n + r] - c1[r]) * md;
            cc += xl[n + r] * mc;
        }
        e = sqrt(dd / cc);
        if (!(bp->adaptive && e > bp->tol && bp->min_step < dt && dt < bp->max_step)) break;
    }
    if (!err) {
        memcpy(cprev, x + n, sizeof(double) * n);
        memcpy(x, xl, sizeof(double) * n2);
        st->step = i;
        st->t = t_prev + dt;
        st->dt = dt;
        st->e = e;
    }
    if (stat) {
        stat->tries = tries;
        stat->iters_bdf1 = it1;
        stat->iters_bdf2 = it2;
        stat->resid = resid;
        stat->dt = dt;
        stat->e = e;
    }
    free(brp); free(bcol); free(bval); free(rhs); free(pcd); free(xl); free(c1); free(ce); free(cr);
    return err;
}

/* ------------------------------------------------------------------------ */
/* Sparse linear algebra                                                     */
/* ------------------------------------------------------------------------ */

/* y_Ax, num/spmat.h:1683-1711 */
void orc_spmv(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
              const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double sum = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) sum += val[k] * x[colind[k]];
        y[i] = sum;
    }
}

/* KahanInnerProd, num/spmat.h:318-335 (compiled without associative math) */
#if defined(__GNUC__) && !defined(__clang__)
__attribute__((optimize("no-associative-math")))
#endif
double orc_kahan_dot(int64_t n, const double *x, const double *y)
{
    volatile double sum = 0., c = 0.;
    for (int64_t i = 0; i < n; ++i) {
        const double yy = x[i] * y[i] - c;
        const double t = sum + yy;
        c = (t - sum) - yy;
        sum = t;
    }
    return sum;
}

/* SolveGSstep<PB_JAC0> with omega = 1, num/precond.h:155-170: diagonal found by a row scan */
static void jacobi_apply(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
                         const double *b, double *x)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        int64_t nz = rowptr[i];
        for (const int64_t end = rowptr[i + 1]; nz < end && colind[nz] != i; ++nz) ;
        x[i] = 1.0 * b[i] / val[nz];
    }
}

/* PCG, num/krylovsolver.h:107-196, with DummyExchangeCL dots (parallel/exchange.h:58-84) */
int orc_pcg_jacobi(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
                   const double *b, double *x, int32_t *max_iter, double *tol, int32_t rel)
{
    double *p = (double *)calloc(n, sizeof(double)), *z = (double *)malloc(n * sizeof(double)),
           *q = (double *)malloc(n * sizeof(double)), *r = (double *)malloc(n * sizeof(double));
    if (!p || !z || !q || !r) { free(p); free(z); free(q); free(r); return ORC_E_NOMEM; }
    orc_spmv(n, rowptr, colind, val, x, q);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
    double resid = sqrt(orc_kahan_dot(n, r, r));
    double normb = sqrt(orc_kahan_dot(n, b, b));
    if (normb == 0.0 || rel == 0) normb = 1.0;
    resid = resid / normb;
    int ret = ORC_E_NOT_CONVERGED;
    if (resid <= *tol) {
        *tol = resid;
        *max_iter = 0;
        ret = ORC_OK;
        goto done;
    }
    jacobi_apply(n, rowptr, colind, val, r, z);
    double rho = orc_kahan_dot(n, z, r), rho_1;
    for (int64_t i = 0; i < n; ++i) p[i] = z[i];
    for (int it = 1; it <= *max_iter; ++it) {
        orc_spmv(n, rowptr, colind, val, p, q);
        const double lambda = orc_kahan_dot(n, q, p);
        const double alpha = rho / lambda;
        for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i];
        for (int64_t i = 0; i < n; ++i) r[i] += (-alpha) * q[i];
        resid = sqrt(orc_kahan_dot(n, r, r)) / normb;
        if (resid <= *tol) {
            *tol = resid;
            *max_iter = it;
            ret = ORC_OK;
            goto done;
        }
        jacobi_apply(n, rowptr, colind, val, r, z);
        rho_1 = rho;
        rho = orc_kahan_dot(n, z, r);
        const double beta = rho / rho_1;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    *tol = resid;
done:
    free(p); free(z); free(q); free(r);
    return ret;
}

static double plain_dot(int64_t n, const double *x, const double *y)
{
    double s = 0.; /* std::inner_product, num/spmat.h:164-170 */
    for (int64_t i = 0; i < n; ++i) s = s + x[i] * y[i];
    return s;
}

/* GMRES_GeneratePlaneRotation / ApplyPlaneRotation, num/krylovsolver.h:258-280 */
static void gen_rot(double dx, double dy, double *cs, double *sn)
{
    if (dy == 0.0) { *cs = 1.0; *sn = 0.0; }
    else { const double r = sqrt(dx * dx + dy * dy); *cs = dx / r; *sn = dy / r; }
}
static void apply_rot(double *dx, double *dy, double cs, double sn)
{
    const double tmp = cs * *dx + sn * *dy;
    *dy = -sn * *dx + cs * *dy;
    *dx = tmp;
}

/* GMRES_Update, num/krylovsolver.h:283-296 (H stored column-major, ldh = m) */
static void gmres_update(int64_t n, double *x, int k, const double *H, int m, const double *s, double *const *v)
{
    double *y = (double *)malloc(sizeof(double) * m);
    for (int i = 0; i < m; ++i) y[i] = s[i];
    for (int i = k; i >= 0; --i) {
        y[i] /= H[i + (size_t)i * m];
        for (int j = i - 1; j >= 0; --j) y[j] -= H[j + (size_t)i * m] * y[i];
    }
    for (int i = 0; i <= k; ++i)
        for (int64_t l = 0; l < n; ++l) x[l] += y[i] * v[i][l];
    free(y);
}

/* serial GMRES, num/krylovsolver.h:320-463 (restart m performs m-1 Arnoldi steps per cycle) */
/* z_i = 1.0 * b_i / d_i: JACPcCL applied with a given diagonal (the diagonal of each block's
 * preconditioning matrix for the block-Jacobi form of DiagBlockPreCL, num/oseenprecond.h:1051-1073) */
static void diag_apply(int64_t n, const double *d, const double *b, double *x)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] = 1.0 * b[i] / d[i];
}

static int gmres_core(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
                      const double *pcdiag, const double *b, double *x, int32_t restart, int32_t *max_iter,
                      double *tol, int32_t rel, int32_t left_pc, int32_t use_pc);

int orc_gmres_jacobi(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
                     const double *b, double *x, int32_t restart, int32_t *max_iter, double *tol,
                     int32_t rel, int32_t left_pc, int32_t use_pc)
{
    return gmres_core(n, rowptr, colind, val, NULL, b, x, restart, max_iter, tol, rel, left_pc, use_pc);
}

int orc_gmres_diag(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
                   const double *pcdiag, const double *b, double *x, int32_t restart, int32_t *max_iter,
                   double *tol, int32_t rel, int32_t left_pc)
{
    if (!pcdiag) return ORC_E_ARG;
    return gmres_core(n, rowptr, colind, val, pcdiag, b, x, restart, max_iter, tol, rel, left_pc, 1);
}

static int gmres_core(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
                      const double *pcdiag, const double *b, double *x, int32_t restart, int32_t *max_iter,
                      double *tol, int32_t rel, int32_t left_pc, int32_t use_pc)
{
    int m = restart <= *max_iter ? restart : *max_iter;
    if (m < 2) m = 2;
    double *H = (double *)calloc((size_t)m * m, sizeof(double));
    double *s = (double *)calloc(m, sizeof(double)), *cs = (double *)calloc(m, sizeof(double)),
           *sn = (double *)calloc(m, sizeof(double));
    double *w = (double *)malloc(n * sizeof(double)), *r = (double *)malloc(n * sizeof(double));
    double *z = (double *)malloc(n * sizeof(double)), *t = (double *)malloc(n * sizeof(double));
    double **v = (double **)malloc(sizeof(double *) * m);
    for (int i = 0; i < m; ++i) v[i] = (double *)malloc(n * sizeof(double));
    double beta, normb, resid;
    int ret = ORC_E_NOT_CONVERGED;
#define PC_APPLY(dst, src) \
    do { if (use_pc && pcdiag) diag_apply(n, pcdiag, (src), (dst)); \
         else if (use_pc) jacobi_apply(n, rowptr, colind, val, (src), (dst)); \
         else memcpy((dst), (src), n * sizeof(double)); } while (0)
    if (!left_pc) {
        orc_spmv(n, rowptr, colind, val, x, t);
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - t[i];
        beta = sqrt(plain_dot(n, r, r));
        normb = sqrt(plain_dot(n, b, b));
    } else {
        orc_spmv(n, rowptr, colind, val, x, t);
        for (int64_t i = 0; i < n; ++i) t[i] = b[i] - t[i];
        PC_APPLY(r, t);
        beta = sqrt(plain_dot(n, r, r));
        PC_APPLY(w, b);
        normb = sqrt(plain_dot(n, w, w));
    }
    if (normb == 0.0 || rel == 0) normb = 1.0;
    resid = beta / normb;
    if (resid <= *tol) {
        *tol = resid;
        *max_iter = 0;
        ret = ORC_OK;
        goto out;
    }
    int j = 1;
    while (j <= *max_iter) {
        for (int64_t l = 0; l < n; ++l) v[0][l] = r[l] * (1.0 / beta);
        for (int i = 0; i < m; ++i) s[i] = 0.0;
        s[0] = beta;
        int i;
        for (i = 0; i < m - 1 && j <= *max_iter; ++i, ++j) {
            if (!left_pc) {
                PC_APPLY(t, v[i]);
                orc_spmv(n, rowptr, colind, val, t, w);
            } else {
                orc_spmv(n, rowptr, colind, val, v[i], t);
                PC_APPLY(w, t);
            }
            for (int k = 0; k <= i; ++k) {
                H[k + (size_t)i * m] = plain_dot(n, w, v[k]);
                const double hk = H[k + (size_t)i * m];
                for (int64_t l = 0; l < n; ++l) w[l] -= hk * v[k][l];
            }
            H[(i + 1) + (size_t)i * m] = sqrt(plain_dot(n, w, w));
            {
                const double inv = 1.0 / H[(i + 1) + (size_t)i * m];
                for (int64_t l = 0; l < n; ++l) v[i + 1][l] = w[l] * inv;
            }
            for (int k = 0; k < i; ++k)
                apply_rot(&H[k + (size_t)i * m], &H[k + 1 + (size_t)i * m], cs[k], sn[k]);
            gen_rot(H[i + (size_t)i * m], H[i + 1 + (size_t)i * m], &cs[i], &sn[i]);
            apply_rot(&H[i + (size_t)i * m], &H[i + 1 + (size_t)i * m], cs[i], sn[i]);
            apply_rot(&s[i], &s[i + 1], cs[i], sn[i]);
            resid = fabs(s[i + 1]) / normb;
            if (resid <= *tol) {
                if (!left_pc) {
                    for (int64_t l = 0; l < n; ++l) z[l] = 0.;
                    gmres_update(n, z, i, H, m, s, v);
                    PC_APPLY(t, z);
                    for (int64_t l = 0; l < n; ++l) x[l] += t[l];
                } else {
                    gmres_update(n, x, i, H, m, s, v);
                }
                *tol = resid;
                *max_iter = j;
                ret = ORC_OK;
                goto out;
            }
        }
        if (!left_pc) {
            for (int64_t l = 0; l < n; ++l) z[l] = 0.;
            gmres_update(n, z, i - 1, H, m, s, v);
            PC_APPLY(t, z);
            for (int64_t l = 0; l < n; ++l) x[l] += t[l];
            orc_spmv(n, rowptr, colind, val, x, t);
            for (int64_t l = 0; l < n; ++l) r[l] = b[l] - t[l];
        } else {
            gmres_update(n, x, i - 1, H, m, s, v);
            orc_spmv(n, rowptr, colind, val, x, t);
            for (int64_t l = 0; l < n; ++l) t[l] = b[l] - t[l];
            PC_APPLY(r, t);
        }
        beta = sqrt(plain_dot(n, r, r));
        resid = beta / normb;
        if (resid <= *tol) {
            *tol = resid;
            *max_iter = j;
            ret = ORC_OK;
            goto out;
        }
    }
    *tol = resid;
#undef PC_APPLY
out:
    for (int i = 0; i < m; ++i) free(v[i]);
    free(v); free(H); free(s); free(cs); free(sn); free(w); free(r); free(z); free(t);
    return ret;
}
