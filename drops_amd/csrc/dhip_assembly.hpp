// dhip_assembly.hpp -- P1 trace-FEM assembly on gfx950.
//
// Pipeline (one context stream, all data device resident):
//   k_classify   one lane per brick cell: reads the 27 half-grid P2 values of
//                the cell, builds 27-bit sign masks and tests the 6 Kuhn tets
//                (equalSigns, geom/subtriangulation.tpp:32-36); per-block counts.
//   k_compact    block scan of cut counts -> cut tets in mesh order and the
//                per-cell word (first cut index << 6 | 6-bit cut mask).
//   k_local_p1   one lane per cut tet: interface patch on the order-2 lattice
//                (SPatchCL::make_patch<MergeCutPolicyCL>), facet QR, composite
//                Quad2/Quad5 rules and all local matrices, facet by facet,
//                in the reference's summation order.  Output is SoA.
//   k_qualify / k_first / k_dofs
//                first-encounter interface numbering without atomics: a vertex
//                belongs to the first (tet, local vertex) in mesh order whose
//                phi_i^2 integral passes the limit; owners are ranked by a scan.
//   k_rows_w<false> / k_rows_w<true> (rowcount / rowfill)
//                one wave per DOF row: lists the <= 24 incident tets of the row
//                vertex in ascending tet id, loads their local entries together
//                and sums them in that order (the order of the reference's
//                hash-map +=), ranks the <= 15 columns and writes CSR rows, comb
//                values and rhs.  rowcount_one / rowfill_one are the same rows
//                for one lane each (the batched kernel's workgroup loops).
// Compiled with -ffp-contract=off: every product and sum is rounded as in the
// reference's evaluation order.
#pragma once

#include "dhip_common.hpp"

namespace dhip {

__device__ __forceinline__ int sgn(double d) { return d > 0. ? 1 : (d < 0. ? -1 : 0); }

__device__ __forceinline__ void vertex_coord(const MeshP &m, int64_t i1, int64_t i2, int64_t i3,
                                             double x[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
        x[c] = ((m.orig[c] + (double)i1 * m.off[0][c]) + (double)i2 * m.off[1][c]) +
               (double)i3 * m.off[2][c];
}

__device__ __forceinline__ void cell_of(const MeshP &m, int64_t cell, int64_t &i1, int64_t &i2,
                                        int64_t &i3) {
    i1 = cell % m.n[0];
    const int64_t r = cell / m.n[0];
    i2 = r % m.n[1];
    i3 = r / m.n[1];
}

// world point of barycentric b (Bary2WorldCoordCL: 0.0 + sum_i X_i b_i)
__device__ __forceinline__ void b2w(const double X[4][3], const double b[4], double w[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) s += X[i][r] * b[i];
        w[r] = s;
    }
}

__device__ __forceinline__ double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// P1 gradients and det as P1DiscCL::GetGradients (num/discretize.h:1724-1746)
__device__ inline double p1_grad(const double X[4][3], double H[4][3]) {
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) M[j][i] = X[i + 1][j] - X[0][j];
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    H[1][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    H[1][1] = (M[2][1] * M[0][2] - M[0][1] * M[2][2]) / det;
    H[1][2] = (M[0][1] * M[1][2] - M[1][1] * M[0][2]) / det;
    H[2][0] = (M[2][0] * M[1][2] - M[1][0] * M[2][2]) / det;
    H[2][1] = (M[0][0] * M[2][2] - M[2][0] * M[0][2]) / det;
    H[2][2] = (M[1][0] * M[0][2] - M[0][0] * M[1][2]) / det;
    H[3][0] = (M[1][0] * M[2][1] - M[2][0] * M[1][1]) / det;
    H[3][1] = (M[2][0] * M[0][1] - M[0][0] * M[2][1]) / det;
    H[3][2] = (M[0][0] * M[1][1] - M[1][0] * M[0][1]) / det;
#pragma unroll
    for (int c = 0; c < 3; ++c) H[0][c] = ((-H[1][c]) - H[2][c]) - H[3][c];
    return det;
}

__device__ inline double tet_volume(const double X[4][3]) {
    double v[3][3];
#pragma unroll
    for (int i = 1; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[i - 1][c] = X[i][c] - X[0][c];
    return fabs(v[0][0] * (v[1][1] * v[2][2] - v[1][2] * v[2][1]) -
                v[0][1] * (v[1][0] * v[2][2] - v[1][2] * v[2][0]) +
                v[0][2] * (v[1][0] * v[2][1] - v[1][1] * v[2][0])) / 6.0;
}

// P2 basis sum (FE_P2CL::H0..H9 + LinearCombinationCL) for the squared hat of vertex i:
// nodal values 1 at vertex i, 0.25 at its three edges, 0 elsewhere.
__device__ inline double hat_sq_eval(int i, const double p[4]) {
    const double h[10] = {p[0] * (2. * p[0] - 1.), p[1] * (2. * p[1] - 1.), p[2] * (2. * p[2] - 1.),
                          p[3] * (2. * p[3] - 1.), 4. * p[0] * p[1],         4. * p[0] * p[2],
                          4. * p[1] * p[2],         4. * p[0] * p[3],         4. * p[1] * p[3],
                          4. * p[2] * p[3]};
    // edges (0,1),(0,2),(1,2),(0,3),(1,3),(2,3)
    const int ea[6] = {0, 0, 1, 0, 1, 2}, eb[6] = {1, 2, 2, 3, 3, 3};
    double s = h[0] * (i == 0 ? 1. : 0.);
#pragma unroll
    for (int d = 1; d < 4; ++d) s += h[d] * (i == d ? 1. : 0.);
#pragma unroll
    for (int e = 0; e < 6; ++e) s += h[4 + e] * ((ea[e] == i || eb[e] == i) ? 0.25 : 0.);
    return s;
}

__device__ inline double rhs_fun(int kind, double p0, const double x[3]) {
    if (kind == DHIP_RHS_3XY_MINUS_X3) return 3.0 * x[0] * x[1] - x[0] * x[0] * x[0];
    if (kind == DHIP_RHS_CONST) return p0;
    return 0.;
}

// ---------------------------------------------------------------------------
// classification + compaction
// ---------------------------------------------------------------------------
constexpr int CLS_BLOCK = 256;

struct ClassArgs {
    MeshP m;
    const double *phi;
    uint8_t *cellmask;
    int32_t *blockcnt;
    int32_t *err;
    int zero_throw;
    double band = -1.0;   // >= 0: narrow band of that width
};

// 6-bit cut mask of brick cell `cell` (equalSigns on each Kuhn tet's 10 values); `zero` reports an
// exact-zero node in a cut tet.  With band >= 0 the mask also holds the uncut tets whose smallest
// |value| is at most band: distance() <= band (geom/subtriangulation.tpp:38-46,
// NarrowBandCommonDataP1CL::visit, surfactant/ifacetransp.h:688-700).
__device__ inline int cell_cut_mask(const MeshP &m, const double *phi, int64_t cell, bool &zero,
                                    double band = -1.0) {
    int mask = 0;
    zero = false;
    {
        int64_t i1, i2, i3;
        cell_of(m, cell, i1, i2, i3);
        const int64_t h1 = 2 * (int64_t)m.n[0] + 1, h2 = 2 * (int64_t)m.n[1] + 1;
        const double *base = phi + ((2 * i3) * h2 + 2 * i2) * h1 + 2 * i1;
        uint32_t pos = 0, neg = 0;
#pragma unroll
        for (int z = 0; z < 3; ++z)
#pragma unroll
            for (int y = 0; y < 3; ++y)
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const double v = base[(z * h2 + y) * h1 + x];
                    const int b = z * 9 + y * 3 + x;
                    pos |= (v > 0.) ? (1u << b) : 0u;
                    neg |= (v < 0.) ? (1u << b) : 0u;
                }
        if (pos != 0x7FFFFFFu && neg != 0x7FFFFFFu) {
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const uint32_t tm = c_tab.tetmask27[k];
                if ((pos & tm) != tm && (neg & tm) != tm) {
                    mask |= 1 << k;
                    zero |= ((pos | neg) & tm) != tm;
                }
            }
        }
        if (band >= 0. && mask != 63) {
            double mn[6] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
#pragma unroll
            for (int b = 0; b < 27; ++b) {
                const double av = fabs(base[((b / 9) * h2 + (b / 3) % 3) * h1 + b % 3]);
#pragma unroll
                for (int k = 0; k < 6; ++k)
                    if ((c_tab.tetmask27[k] >> b) & 1u) mn[k] = av < mn[k] ? av : mn[k];
            }
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (mn[k] <= band) mask |= 1 << k;
        }
    }
    return mask;
}

__global__ void __launch_bounds__(CLS_BLOCK) k_classify(ClassArgs a) {
    const int64_t cell = (int64_t)blockIdx.x * CLS_BLOCK + threadIdx.x;
    int mask = 0;
    if (cell < a.m.ncell) {
        bool zero;
        mask = cell_cut_mask(a.m, a.phi, cell, zero, a.band);
        if (zero && a.zero_throw) atomicOr(a.err, 1);
        a.cellmask[cell] = (uint8_t)mask;
    }
    int tot;
    block_excl_scan<CLS_BLOCK>(__popc(mask), &tot);
    if (threadIdx.x == 0) a.blockcnt[blockIdx.x] = tot;
}

struct CompactArgs {
    int64_t ncell;
    const uint8_t *cellmask;
    const int32_t *blockoff;
    uint32_t *cellinfo;
    int64_t *cut;
};

__global__ void __launch_bounds__(CLS_BLOCK) k_compact(CompactArgs a) {
    const int64_t cell = (int64_t)blockIdx.x * CLS_BLOCK + threadIdx.x;
    const int mask = cell < a.ncell ? a.cellmask[cell] : 0;
    const int off = block_excl_scan<CLS_BLOCK>(__popc(mask), nullptr);
    if (cell >= a.ncell) return;
    int64_t base = (int64_t)a.blockoff[blockIdx.x] + off;
    a.cellinfo[cell] = ((uint32_t)base << 6) | (uint32_t)mask;
    for (int k = 0; k < 6; ++k)
        if (mask & (1 << k)) a.cut[base++] = cell * 6 + k;
}

// ---------------------------------------------------------------------------
// per cut tet: patch + local matrices
// ---------------------------------------------------------------------------
// Vertex coordinates and the 10 P2 level-set values (lattice-2 order) of tet t.
__device__ inline void load_tet_geom(const MeshP &m, const double *phi, int64_t t, double X[4][3], double ls[10]) {
    const int k = (int)(t % 6);
    int64_t i1, i2, i3;
    cell_of(m, t / 6, i1, i2, i3);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        vertex_coord(m, i1 + c_tab.kuhn[k][q][0], i2 + c_tab.kuhn[k][q][1], i3 + c_tab.kuhn[k][q][2], X[q]);
    const int64_t h1 = 2 * (int64_t)m.n[0] + 1, h2 = 2 * (int64_t)m.n[1] + 1;
    const double *base = phi + ((2 * i3) * h2 + 2 * i2) * h1 + 2 * i1;
#pragma unroll
    for (int l = 0; l < 10; ++l) {
        const int nd = c_tab.tet_node27[k][l];
        ls[l] = base[((nd / 9) * h2 + (nd / 3) % 3) * h1 + nd % 3];
    }
}

// Barycentric vertices of triangle `tr` of sub-tet `lt` with sign pattern `pid`: zero
// vertices or edge cuts (1-a) v_a + a v_b, a = ls_a / (ls_a - ls_b) (ConvexComb,
// misc/container.h:403-410), merged per lattice edge.
__device__ inline void facet_vertices(const double ls[10], int lt, int pid, int tr, double vb[3][4]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int rep = c_tab.rp_tri[pid][tr][j];
        if (rep < 4) {
            const int lv = c_tab.lat_t[lt][rep];
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) vb[j][cc] = c_tab.lat_b[lv][cc];
        } else {
            const int E = c_tab.sub_edge[lt][rep - 4];
            const int la = c_tab.edge_ab[E][0], lb = c_tab.edge_ab[E][1];
            const double alpha = ls[la] / (ls[la] - ls[lb]);
#pragma unroll
            for (int cc = 0; cc < 4; ++cc)
                vb[j][cc] = (1.0 - alpha) * c_tab.lat_b[la][cc] + alpha * c_tab.lat_b[lb][cc];
        }
    }
}

// |det R| and unit normal Q e3 of the Householder QR of [w1-w0, w2-w0]
// (QRDecompCL<3,2>, misc/container.h:1270-1296); rank deficient -> 0.
__device__ inline void facet_qr(const double X[4][3], const double vb[3][4], double &absdet, double nrm[3]) {
    double w[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) b2w(X, vb[j], w[j]);
    double A[3][2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        A[r][0] = w[1][r] - w[0][r];
        A[r][1] = w[2][r] - w[0][r];
    }
    absdet = 0.;
    nrm[0] = nrm[1] = nrm[2] = 0.;
    double d0, d1, b0, b1;
    double sig = 0.;
    for (int i = 0; i < 3; ++i) sig += A[i][0] * A[i][0];
    bool def = sig == 0.;
    if (def) return;
    d0 = (A[0][0] < 0 ? 1. : -1.) * sqrt(sig);
    b0 = 1. / (d0 * A[0][0] - sig);
    A[0][0] -= d0;
    double sp = 0.;
    for (int i = 0; i < 3; ++i) sp += A[i][0] * A[i][1];
    sp *= b0;
    for (int i = 0; i < 3; ++i) A[i][1] += A[i][0] * sp;
    sig = 0.;
    for (int i = 1; i < 3; ++i) sig += A[i][1] * A[i][1];
    def = sig == 0.;
    if (def) return;
    d1 = (A[1][1] < 0 ? 1. : -1.) * sqrt(sig);
    b1 = 1. / (d1 * A[1][1] - sig);
    A[1][1] -= d1;
    double dd = 1.0;
    dd *= d0;
    dd *= d1;
    absdet = fabs(dd);
    double tq[3] = {0., 0., 1.};
    sp = 0.;
    for (int i = 1; i < 3; ++i) sp += A[i][1] * tq[i];
    sp *= b1;
    for (int i = 1; i < 3; ++i) tq[i] += A[i][1] * sp;
    sp = 0.;
    for (int i = 0; i < 3; ++i) sp += A[i][0] * tq[i];
    sp *= b0;
    for (int i = 0; i < 3; ++i) tq[i] += A[i][0] * sp;
    nrm[0] = tq[0];
    nrm[1] = tq[1];
    nrm[2] = tq[2];
}

// Normalised gradient of the P2 level set (P2 dof order v[10]) at barycentric b:
// grad phi_h = sum_k G_k grad(lambda_k), G_k = v_k (4 b_k - 1) + 4 sum_{e=(k,l)} v_e b_l.
__device__ inline void p2_unit_normal(const double v[10], const double b[4], const double G1[4][3], double n3[3]) {
    double G[4];
    const int ea[6] = {0, 0, 1, 0, 1, 2}, eb[6] = {1, 2, 2, 3, 3, 3};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) G[kk] = v[kk] * (4.0 * b[kk] - 1.0);
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        G[ea[e]] += 4.0 * (v[4 + e] * b[eb[e]]);
        G[eb[e]] += 4.0 * (v[4 + e] * b[ea[e]]);
    }
    double g[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] = ((G[0] * G1[0][d] + G[1] * G1[1][d]) + G[2] * G1[2][d]) + G[3] * G1[3][d];
    const double nn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    n3[0] = g[0] / nn;
    n3[1] = g[1] / nn;
    n3[2] = g[2] / nn;
}

constexpr int LOC_N = 57;  // 16 mass, 16 lb, 16 stab, 4 rhs, 4 hat_sq, 1 area
constexpr int LOC_MASS = 0, LOC_LB = 16, LOC_STAB = 32, LOC_RHS = 48, LOC_HAT = 52, LOC_AREA = 56;

struct LocalArgs {
    MeshP m;
    const double *phi;
    const int64_t *cut;
    int64_t ncut;
    double coef0, coef1, coef2;
    int rhs_kind;
    double rhs_p0;
    double *loc;      // SoA [LOC_N][ncut]
    uint8_t *flags;   // bit0: contributes (non-empty patch, or a narrow-band tet); bit5: band tet without patch
    int32_t *err;
    int band = 0;     // narrow band: listed tets without a patch still carry the stabilisation
    const double *rhs_vals = nullptr;   // DHIP_RHS_VALUES: f at every patch quadrature point
    const int32_t *rhs_off = nullptr;   //   first point of each cut tet
};

constexpr int LOCAL_BLOCK = 64;

// normal-gradient stabilisation on the Quad5 volume rule, normal = grad(phi_h)/|grad(phi_h)|
// (LocalNormalLaplaceBulkP1CL::setup, surfactant/ifacetransp.h:2005-2051)
__device__ inline void stab_p1(const LocalArgs &a, const double ls[10], const double G1[4][3], double det,
                               int64_t c) {
    const int64_t N = a.ncut;
    double *L = a.loc;
    {
        double v[10];
#pragma unroll
        for (int d = 0; d < 10; ++d) v[d] = ls[c_tab.lat_of_p2[d]];
        double R[10], S[10];
        for (int qn = 0; qn < 15; ++qn) {
            double n3[3];
            p2_unit_normal(v, c_tab.v5n[qn], G1, n3);
            double U[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) U[i] = dot3(n3, G1[i]);
            int o = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j <= i; ++j, ++o) {
                    const double f = U[i] * U[j];
                    if (qn == 0) R[o] = 8. / 405. * f;
                    else if (qn == 1 || qn == 5 || qn == 9) S[o] = f;
                    else S[o] += f;
                    if (qn == 4) R[o] = R[o] + c_tab.v5c1 * S[o];
                    else if (qn == 8) R[o] = R[o] + c_tab.v5c2 * S[o];
                    else if (qn == 14) R[o] = R[o] + 5. / 567. * S[o];
                }
        }
        const double ad = fabs(det) / 6;
        int o = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                const double val = a.coef2 * (ad * R[o++]);
                L[(LOC_STAB + i * 4 + j) * N + c] = val;
                L[(LOC_STAB + j * 4 + i) * N + c] = val;
            }
    }
}

// The local matrices of one cut tet.  ROLE < 0: every output; otherwise one group, so that four
// waves can share a tet: 0 mass + numbering integrals, 1 Laplace-Beltrami, 2 load vector, area and
// the flags, 3 stabilisation.  Each output is summed by the same statements in the same order in
// every form, so the parts equal the whole bit for bit.
template <int ROLE>
__device__ inline void local_p1_part(const LocalArgs &a, int64_t c) {
    constexpr bool ALL = ROLE < 0;
    constexpr bool MASS = ALL || ROLE == 0, LB = ALL || ROLE == 1, LOAD = ALL || ROLE == 2, STAB = ALL || ROLE == 3;
    constexpr bool FACETS = MASS || LB || LOAD;
    double X[4][3], ls[10];
    load_tet_geom(a.m, a.phi, a.cut[c], X, ls);
    double G1[4][3];
    const double det = p1_grad(X, G1);

    // accumulators (facet-major, node-minor, as the reference's flattened quadrature sums)
    double mdiag[4] = {0., 0., 0., 0.}, moff[6] = {0., 0., 0., 0., 0., 0.};
    double lbs[10] = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
    double hat[4] = {0., 0., 0., 0.}, rhs[4] = {0., 0., 0., 0.};
    double area = 0.;
    int nfacet = 0;
    const double *fvals = (LOAD && a.rhs_kind == DHIP_RHS_VALUES) ? a.rhs_vals + a.rhs_off[c] : nullptr;

    for (int lt = 0; lt < 8; ++lt) {
        int s[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] = sgn(ls[c_tab.lat_t[lt][q]]);
        if (!s[0] && !s[1] && !s[2] && !s[3]) {
            if (LOAD) {
                atomicOr(a.err, 2);
                a.flags[c] = 0;   // defined for the steps that run before the error is read
            }
            return;
        }
        const int pid = s[0] + 3 * s[1] + 9 * s[2] + 27 * s[3] + 40;
        const int ntri = c_tab.rp_size[pid];
        if constexpr (!FACETS) {
            nfacet += ntri;
            continue;
        }
        const bool bnd = c_tab.rp_bnd[pid] != 0;
        for (int tr = 0; tr < ntri; ++tr) {
            double vb[3][4];
            facet_vertices(ls, lt, pid, tr, vb);
            ++nfacet;
            double absdet, nrm[3];
            facet_qr(X, vb, absdet, nrm);
            double af = bnd ? 0.5 : 1.;
            af *= absdet;
            // Quad2 composite: mass + numbering integrals
            if (MASS) {
#pragma unroll
                for (int qn = 0; qn < 4; ++qn) {
                    double nb[4];
#pragma unroll
                    for (int cc = 0; cc < 4; ++cc) {
                        double sv = vb[0][cc] * c_tab.q2n[qn][0];
                        sv += vb[1][cc] * c_tab.q2n[qn][1];
                        sv += vb[2][cc] * c_tab.q2n[qn][2];
                        nb[cc] = sv;
                    }
                    const double wq = af * c_tab.q2w[qn];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        mdiag[i] += wq * (nb[i] * nb[i]);
                        hat[i] += wq * hat_sq_eval(i, nb);
                    }
                    int o = 0;
#pragma unroll
                    for (int i = 1; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < i; ++j) moff[o++] += wq * (nb[j] * nb[i]);
                }
            }
            // Laplace-Beltrami: projected gradients per facet
            if (LB) {
                double q[4][3];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double d = dot3(G1[i], nrm);
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) q[i][cc] = G1[i][cc] - d * nrm[cc];
                }
                int o = 0;
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j <= i; ++j) lbs[o++] += dot3(q[i], q[j]) * absdet;
            }
            // Quad5 composite: load vector and area
            if (LOAD) {
#pragma unroll
                for (int qn = 0; qn < 7; ++qn) {
                    double nb[4];
#pragma unroll
                    for (int cc = 0; cc < 4; ++cc) {
                        double sv = vb[0][cc] * c_tab.q5n[qn][0];
                        sv += vb[1][cc] * c_tab.q5n[qn][1];
                        sv += vb[2][cc] * c_tab.q5n[qn][2];
                        nb[cc] = sv;
                    }
                    const double wq = af * c_tab.q5w[qn];
                    area += wq * 1.;
                    if (a.rhs_kind != DHIP_RHS_NONE) {
                        double f;
                        if (fvals) {
                            f = fvals[7 * (nfacet - 1) + qn];
                        } else {
                            double x[3];
                            b2w(X, nb, x);
                            f = rhs_fun(a.rhs_kind, a.rhs_p0, x);
                        }
#pragma unroll
                        for (int i = 0; i < 4; ++i) rhs[i] += wq * (f * nb[i]);
                    }
                }
            }
        }
    }

    const int64_t N = a.ncut;
    double *L = a.loc;
    if (LOAD) a.flags[c] = nfacet > 0 ? 1 : (a.band ? 0x21 : 0);
    if (nfacet == 0) {
        auto zero = [&](int e0, int cnt) {
            for (int e = e0; e < e0 + cnt; ++e) L[e * N + c] = 0.;
        };
        if (MASS) {
            zero(LOC_MASS, 16);
            zero(LOC_HAT, 4);
        }
        if (LB) zero(LOC_LB, 16);
        if (LOAD) {
            zero(LOC_RHS, 4);
            zero(LOC_AREA, 1);
        }
        if (STAB) {
            zero(LOC_STAB, 16);
            // NarrowBandMatrixAccuP1CL: the stabilisation of a band tet does not need the patch
            if (a.band) stab_p1(a, ls, G1, det, c);
        }
        return;
    }
    // mass: diagonal without alpha, off-diagonal alpha * integral
    if (MASS) {
        int o = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            L[(LOC_MASS + i * 4 + i) * N + c] = mdiag[i];
#pragma unroll
            for (int j = 0; j < i; ++j) {
                const double v = a.coef0 * moff[o++];
                L[(LOC_MASS + i * 4 + j) * N + c] = v;
                L[(LOC_MASS + j * 4 + i) * N + c] = v;
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) L[(LOC_HAT + i) * N + c] = hat[i];
    }
    if (LB) {
        const double fac = a.coef1 * 0.5;
        int o = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                const double v = fac * lbs[o++];
                L[(LOC_LB + i * 4 + j) * N + c] = v;
                L[(LOC_LB + j * 4 + i) * N + c] = v;
            }
    }
    if (LOAD) {
#pragma unroll
        for (int i = 0; i < 4; ++i) L[(LOC_RHS + i) * N + c] = rhs[i];
        L[LOC_AREA * N + c] = area;
    }
    if (STAB) stab_p1(a, ls, G1, det, c);
}

__device__ inline void local_p1_one(const LocalArgs &a, int64_t c) { local_p1_part<-1>(a, c); }

// One lane per cut tet.  Splitting the outputs over four waves or four launches (local_p1_part<0..3>)
// was measured slower: the facet geometry, which every group needs, dominates the cost.
__global__ void __launch_bounds__(LOCAL_BLOCK) k_local_p1(LocalArgs a) {
    const int64_t c = (int64_t)blockIdx.x * LOCAL_BLOCK + threadIdx.x;
    if (c >= a.ncut) return;
    local_p1_one(a, c);
}

// The composite-Quad5 points of every cut tet's patch in the load vector's order (facets in
// make_patch order, 7 nodes each): COUNT writes 7 x facets per tet, else the world coordinates from
// off[c] on.  A sub-tet lying in the zero level set gives no points (the local pass reports it).
template <bool COUNT>
__global__ void __launch_bounds__(LOCAL_BLOCK) k_quad_points(LocalArgs a, int32_t *cnt, const int32_t *off,
                                                             double *xyz) {
    const int64_t c = (int64_t)blockIdx.x * LOCAL_BLOCK + threadIdx.x;
    if (c >= a.ncut) return;
    double X[4][3], ls[10];
    load_tet_geom(a.m, a.phi, a.cut[c], X, ls);
    int pids[8], nfacet = 0;
    for (int lt = 0; lt < 8; ++lt) {
        int s[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] = sgn(ls[c_tab.lat_t[lt][q]]);
        if (!s[0] && !s[1] && !s[2] && !s[3]) {
            if (COUNT) cnt[c] = 0;   // no points: the local pass raises the error for this tet
            return;
        }
        pids[lt] = s[0] + 3 * s[1] + 9 * s[2] + 27 * s[3] + 40;
        nfacet += c_tab.rp_size[pids[lt]];
    }
    if (COUNT) {
        cnt[c] = 7 * nfacet;
        return;
    }
    int64_t k = off[c];
    for (int lt = 0; lt < 8; ++lt) {
        const int pid = pids[lt];
        const int ntri = c_tab.rp_size[pid];
        for (int tr = 0; tr < ntri; ++tr) {
            double vb[3][4];
            facet_vertices(ls, lt, pid, tr, vb);
#pragma unroll
            for (int qn = 0; qn < 7; ++qn) {
                double nb[4];
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    double sv = vb[0][cc] * c_tab.q5n[qn][0];
                    sv += vb[1][cc] * c_tab.q5n[qn][1];
                    sv += vb[2][cc] * c_tab.q5n[qn][2];
                    nb[cc] = sv;
                }
                double x[3];
                b2w(X, nb, x);
                xyz[3 * k + 0] = x[0];
                xyz[3 * k + 1] = x[1];
                xyz[3 * k + 2] = x[2];
                ++k;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// interface numbering
// ---------------------------------------------------------------------------
struct QualArgs {
    MeshP m;
    const int64_t *cut;
    int64_t ncut;
    const double *loc;
    uint8_t *flags;
    double omit_bound;
    int band = 0;   // narrow band: every vertex of a listed tet (CreateNumbNearInterfaceVertex)
};

__device__ inline void qualify_one(const QualArgs &a, int64_t c) {
    int f = a.flags[c] & 1;
    const int keep = a.flags[c] & 0x20;
    if (f && a.band) {
        f |= 0x1E;
    } else if (f) {
        double limit = 0.;
        if (a.omit_bound >= 0.) {
            const int64_t t = a.cut[c];
            const int k = (int)(t % 6);
            int64_t i1, i2, i3;
            cell_of(a.m, t / 6, i1, i2, i3);
            double X[4][3];
            for (int q = 0; q < 4; ++q)
                vertex_coord(a.m, i1 + c_tab.kuhn[k][q][0], i2 + c_tab.kuhn[k][q][1],
                             i3 + c_tab.kuhn[k][q][2], X[q]);
            limit = a.omit_bound * pow(tet_volume(X) * 6, 4. / 3.);
        }
        for (int i = 0; i < 4; ++i)
            if (a.loc[(int64_t)(LOC_HAT + i) * a.ncut + c] > limit) f |= 2 << i;
    }
    a.flags[c] = (uint8_t)(f | keep);
}

__global__ void k_qualify(QualArgs a) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.ncut) return;
    qualify_one(a, c);
}

// Visit the tets containing grid vertex (x1,x2,x3) in ascending tet id; the
// functor gets (cut index, tet type k, local index of the vertex) for cut tets
// and returns true to stop.
template <class F>
__device__ inline void for_incident_cut_tets(const MeshP &m, const uint32_t *cellinfo, int64_t x1,
                                             int64_t x2, int64_t x3, F &&f) {
    for (int d = 7; d >= 0; --d) {
        const int64_t c1 = x1 - (d & 1), c2 = x2 - ((d >> 1) & 1), c3 = x3 - ((d >> 2) & 1);
        if (c1 < 0 || c2 < 0 || c3 < 0 || c1 >= m.n[0] || c2 >= m.n[1] || c3 >= m.n[2]) continue;
        const uint32_t info = cellinfo[(c3 * m.n[1] + c2) * m.n[0] + c1];
        const uint32_t mask = info & 63u;
        if (!mask) continue;
        const uint32_t base = info >> 6;
        for (int k = 0; k < 6; ++k) {
            const int li = c_tab.inc[d][k];
            if (li < 0 || !((mask >> k) & 1u)) continue;
            const int64_t ci = base + __popc(mask & ((1u << k) - 1u));
            if (f(ci, k, li)) return;
        }
    }
}

__device__ inline void slot_vertex(const MeshP &m, const int64_t *cut, int64_t s, int64_t &x1, int64_t &x2,
                                   int64_t &x3) {
    const int64_t t = cut[s >> 2];
    const int k = (int)(t % 6), i = (int)(s & 3);
    int64_t i1, i2, i3;
    cell_of(m, t / 6, i1, i2, i3);
    x1 = i1 + c_tab.kuhn[k][i][0];
    x2 = i2 + c_tab.kuhn[k][i][1];
    x3 = i3 + c_tab.kuhn[k][i][2];
}

struct FirstArgs {
    MeshP m;
    const int64_t *cut;
    int64_t ncut;
    const uint32_t *cellinfo;
    const uint8_t *flags;
    int32_t *first;
    int32_t *isown;
};

__device__ inline void first_one(const FirstArgs &a, int64_t s) {
    int64_t x1, x2, x3;
    slot_vertex(a.m, a.cut, s, x1, x2, x3);
    int64_t found = -1;
    for_incident_cut_tets(a.m, a.cellinfo, x1, x2, x3, [&](int64_t ci, int, int li) {
        if ((a.flags[ci] >> (1 + li)) & 1) {
            found = ci * 4 + li;
            return true;
        }
        return false;
    });
    a.first[s] = (int32_t)found;
    a.isown[s] = found == s ? 1 : 0;
}

__global__ void k_first(FirstArgs a) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= 4 * a.ncut) return;
    first_one(a, s);
}

struct DofArgs {
    int64_t nslot;
    const int32_t *first;
    const int32_t *ownrank;
    int32_t *dofs;
    int32_t *rowowner;
};

__global__ void k_dofs(DofArgs a) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nslot) return;
    const int32_t f = a.first[s];
    a.dofs[s] = f >= 0 ? a.ownrank[f] : -1;
    if (f == s) a.rowowner[a.ownrank[s]] = (int32_t)s;
}

// ---------------------------------------------------------------------------
// CSR rows
// ---------------------------------------------------------------------------
struct RowArgs {
    MeshP m;
    const int64_t *cut;
    int64_t ncut;
    const uint32_t *cellinfo;
    const uint8_t *flags;
    const int32_t *dofs;
    const int32_t *rowowner;
    int64_t ndof;
    const double *loc;
    int32_t *rowlen;
    const int32_t *rowptr;
    int32_t *col;
    double *vm, *vl, *vs, *vc, *rhs, *diag;
    double lc0, lc1, lc2;
};

__device__ inline void rowcount_one(const RowArgs &a, int64_t r) {
    int64_t x1, x2, x3;
    slot_vertex(a.m, a.cut, a.rowowner[r], x1, x2, x3);
    uint32_t present = 0;
    for_incident_cut_tets(a.m, a.cellinfo, x1, x2, x3, [&](int64_t ci, int k, int li) {
        if (!(a.flags[ci] & 1)) return false;
        for (int j = 0; j < 4; ++j)
            if (a.dofs[ci * 4 + j] >= 0) present |= 1u << c_tab.slot[k][li][j];
        return false;
    });
    a.rowlen[r] = __popc(present);
}

__device__ inline void rowfill_one(const RowArgs &a, int64_t r) {
    int64_t x1, x2, x3;
    slot_vertex(a.m, a.cut, a.rowowner[r], x1, x2, x3);
    uint32_t present = 0;
    int32_t cols[15];
    double am[15], al[15], as[15];
#pragma unroll
    for (int s = 0; s < 15; ++s) {
        cols[s] = 0;
        am[s] = al[s] = as[s] = 0.0;
    }
    double rh = 0.0;
    const int64_t N = a.ncut;
    for_incident_cut_tets(a.m, a.cellinfo, x1, x2, x3, [&](int64_t ci, int k, int li) {
        if (!(a.flags[ci] & 1)) return false;
        for (int j = 0; j < 4; ++j) {
            const int32_t dj = a.dofs[ci * 4 + j];
            if (dj < 0) continue;
            const int sl = c_tab.slot[k][li][j];
            const int e = li * 4 + j;
            const double m = a.loc[(int64_t)(LOC_MASS + e) * N + ci];
            const double l = a.loc[(int64_t)(LOC_LB + e) * N + ci];
            const double st = a.loc[(int64_t)(LOC_STAB + e) * N + ci];
#pragma unroll
            for (int s = 0; s < 15; ++s)
                if (s == sl) {
                    cols[s] = dj;
                    am[s] += m;
                    al[s] += l;
                    as[s] += st;
                }
            present |= 1u << sl;
        }
        rh += a.loc[(int64_t)(LOC_RHS + li) * N + ci];
        return false;
    });
    const int64_t base = a.rowptr[r];
#pragma unroll
    for (int s = 0; s < 15; ++s) {
        if (!((present >> s) & 1u)) continue;
        int rank = 0;
#pragma unroll
        for (int u = 0; u < 15; ++u)
            rank += (((present >> u) & 1u) && cols[u] < cols[s]) ? 1 : 0;
        const int64_t dst = base + rank;
        a.col[dst] = cols[s];
        a.vm[dst] = am[s];
        a.vl[dst] = al[s];
        a.vs[dst] = as[s];
        const double cv = 1.0 * (a.lc0 * am[s] + a.lc1 * al[s]) + a.lc2 * as[s];
        a.vc[dst] = cv;
        if (cols[s] == r) a.diag[r] = cv;
    }
    a.rhs[r] = rh;
}

// One row per wave (rowcount_one / rowfill_one's rows with the tets' loads side by side): lanes 0-47 look at
// the (cell, Kuhn tet) candidates of the row vertex at once, lane order being for_incident_cut_tets' walk
// order (d = 7 .. 0, then k), and list the flagged cut tets; sixteen tets at a time, lane 4 t + j loads
// column vertex j of tet t, and the sums into the 15 neighbour slots go tet by tet, so every slot is summed
// in walk order as rowfill_one sums it.
constexpr int ROWW_WAVES = 4;

struct RowWShared {
    double am[16], al[16], as[16];
    int32_t cols[16];
    int32_t pres[16];
    int32_t tci[48];
    int32_t tinf[48];   // k | li << 3
};

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <bool FILL>
__global__ void __launch_bounds__(64 * ROWW_WAVES) k_rows_w(RowArgs a) {
    __shared__ RowWShared shm[ROWW_WAVES];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // each XCD takes a contiguous run of rows, whose tets its L2 then shares
    const int64_t r = (int64_t)xcd_position(blockIdx.x, gridDim.x) * ROWW_WAVES + w;
    if (r >= a.ndof) return;
    RowWShared &sh = shm[w];
    if (lane < 16) {
        sh.pres[lane] = 0;
        sh.cols[lane] = 0;
        sh.am[lane] = sh.al[lane] = sh.as[lane] = 0.0;
    }
    int64_t x1, x2, x3;
    slot_vertex(a.m, a.cut, a.rowowner[r], x1, x2, x3);
    int32_t tci = -1;
    int tinf = 0;
    if (lane < 48) {
        const int d = 7 - lane / 6, k = lane % 6;
        const int64_t c1 = x1 - (d & 1), c2 = x2 - ((d >> 1) & 1), c3 = x3 - ((d >> 2) & 1);
        const int li = c_tab.inc[d][k];
        if (li >= 0 && c1 >= 0 && c2 >= 0 && c3 >= 0 && c1 < a.m.n[0] && c2 < a.m.n[1] && c3 < a.m.n[2]) {
            const uint32_t info = a.cellinfo[(c3 * a.m.n[1] + c2) * a.m.n[0] + c1];
            const uint32_t mask = info & 63u;
            if ((mask >> k) & 1u) {
                const int32_t t = (int32_t)((info >> 6) + __popc(mask & ((1u << k) - 1u)));
                if (a.flags[t] & 1) {
                    tci = t;
                    tinf = k | (li << 3);
                }
            }
        }
    }
    const uint64_t bal = __ballot(tci >= 0);
    const int ntet = __popcll(bal);
    if (tci >= 0) {
        const int pos = __popcll(bal & ((1ull << lane) - 1ull));
        sh.tci[pos] = tci;
        sh.tinf[pos] = tinf;
    }
    wave_sync_lds();
    const int64_t N = a.ncut;
    double rv = 0.0;
    if (FILL && lane < ntet) rv = a.loc[(int64_t)(LOC_RHS + (sh.tinf[lane] >> 3)) * N + sh.tci[lane]];
    const int tl = lane >> 2, j = lane & 3;
    for (int b = 0; b < ntet; b += 16) {
        const int t = b + tl;
        const bool act = t < ntet;
        int32_t dj = -1;
        int sl = 0;
        double vm = 0.0, vl = 0.0, vs = 0.0;
        if (act) {
            const int32_t ci = sh.tci[t];
            const int inf = sh.tinf[t], k = inf & 7, li = inf >> 3;
            dj = a.dofs[(int64_t)ci * 4 + j];
            sl = c_tab.slot[k][li][j];
            if (FILL) {
                const int e = li * 4 + j;
                vm = a.loc[(int64_t)(LOC_MASS + e) * N + ci];
                vl = a.loc[(int64_t)(LOC_LB + e) * N + ci];
                vs = a.loc[(int64_t)(LOC_STAB + e) * N + ci];
            }
        }
        if (!FILL) {
            if (act && dj >= 0) sh.pres[sl] = 1;
            continue;
        }
        const int nb = ntet - b < 16 ? ntet - b : 16;
        for (int u = 0; u < nb; ++u) {
            if (act && tl == u && dj >= 0) {
                sh.pres[sl] = 1;
                sh.cols[sl] = dj;
                sh.am[sl] += vm;
                sh.al[sl] += vl;
                sh.as[sl] += vs;
            }
            wave_sync_lds();
        }
    }
    wave_sync_lds();
    const bool pr = lane < 15 && sh.pres[lane] != 0;
    const uint64_t pm = __ballot(pr);
    if (!FILL) {
        if (lane == 0) a.rowlen[r] = __popcll(pm);
        return;
    }
    double rh = 0.0;
    for (int t = 0; t < ntet; ++t) rh += __shfl(rv, t, 64);
    if (pr) {
        const int32_t cs = sh.cols[lane];
        int rank = 0;
        for (int u = 0; u < 15; ++u) rank += (((pm >> u) & 1ull) && sh.cols[u] < cs) ? 1 : 0;
        const int64_t dst = a.rowptr[r] + rank;
        a.col[dst] = cs;
        a.vm[dst] = sh.am[lane];
        a.vl[dst] = sh.al[lane];
        a.vs[dst] = sh.as[lane];
        const double cv = 1.0 * (a.lc0 * sh.am[lane] + a.lc1 * sh.al[lane]) + a.lc2 * sh.as[lane];
        a.vc[dst] = cv;
        if (cs == r) a.diag[r] = cv;
    }
    if (lane == 0) a.rhs[r] = rh;
}

}  // namespace dhip
