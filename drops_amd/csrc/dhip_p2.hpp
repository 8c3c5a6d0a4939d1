// dhip_p2.hpp -- P2 interface space on gfx950: local forms, numbering, CSR rows.
//
// The P2 trace space lives on the same cut tets and the same planar patch as the P1
// space; only the basis changes.  Pipeline after k_classify / k_compact:
//   k_local_p2    one wave per cut tet.  Lanes 0-15 build the patch facets (sub-tet
//                 pattern, facet QR) into LDS, then the composite Quad5 points in
//                 chunks of 64 (one point per lane: P2 basis values, surface
//                 gradients projected with the P2 level-set normal, load value), then
//                 lane p < 55 accumulates pair p = (i >= j) of the 10x10 mass and
//                 Laplace-Beltrami matrices point by point in the composite order;
//                 the normal-gradient stabilisation uses the 15-point Quad5 volume
//                 rule.  Output is tet-major [ncut][LOC2_N].
//   k_qualify_p2 / k_first_p2 / k_dofs
//                 first-encounter numbering over 10 slots per cut tet
//                 (CreateNumbOnInterfaceP2, misc/problem.cpp:571-636): a half-grid node
//                 belongs to the first (tet, local node) in mesh order whose squared
//                 basis function has a large enough interface integral.
//   k_rowcount_p2 / k_rowfill_p2
//                 one wave per DOF row: lanes list the cut tets that contain the row
//                 node (ascending tet id) at once, load six tets' rows together and
//                 sum them tet by tet into 125 neighbour slots in LDS (the 5x5x5
//                 half-grid block around the node), so each entry is summed in the
//                 order of the reference's row map; then the present slots are
//                 compacted, ranked by column and written out.
// Compiled with -ffp-contract=off like the P1 path: bit-identical to the oracle.
#pragma once

#include "dhip_assembly.hpp"

namespace dhip {

constexpr int LOC2_N = 320;  // 100 mass, 100 lb, 100 stab, 10 rhs, 10 hat_sq
constexpr int LOC2_MASS = 0, LOC2_LB = 100, LOC2_STAB = 200, LOC2_RHS = 300, LOC2_HAT = 310;
// element e of cut tet c in the local array (tet-major: a tet's 320 values are contiguous, so the
// row builder's lanes read a tet's row of one operator as consecutive doubles)
__device__ __forceinline__ int64_t loc2(int e, int64_t c) { return c * LOC2_N + e; }
constexpr int P2_MAXF = 16;
constexpr int P2_CHUNK = 64;

// FE_P2CL::H0..H9 (num/fe.h:287-296): vertices l(2l-1), edges 4 l_a l_b
__device__ __forceinline__ void p2_basis(const double p[4], double h[10]) {
    const int ea[6] = {0, 0, 1, 0, 1, 2}, eb[6] = {1, 2, 2, 3, 3, 3};
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = p[k] * (2. * p[k] - 1.);
#pragma unroll
    for (int e = 0; e < 6; ++e) h[4 + e] = 4. * p[ea[e]] * p[eb[e]];
}

// gradient of P2 basis function d at p: (4 l_k - 1) grad l_k, or 4 l_b grad l_a + 4 l_a grad l_b
__device__ __forceinline__ void p2_grad_one(int d, const double p[4], const double G1[4][3], double g[3]) {
    const int ea[6] = {0, 0, 1, 0, 1, 2}, eb[6] = {1, 2, 2, 3, 3, 3};
    if (d < 4) {
        const double f = 4.0 * p[d] - 1.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) g[c] = f * G1[d][c];
    } else {
        const int a = ea[d - 4], b = eb[d - 4];
#pragma unroll
        for (int c = 0; c < 3; ++c) g[c] = 4.0 * p[b] * G1[a][c] + 4.0 * p[a] * G1[b][c];
    }
}

__device__ __forceinline__ void pair_of(int p, int &i, int &j) {
    i = 0;
    while ((i + 1) * (i + 2) / 2 <= p) ++i;
    j = p - i * (i + 1) / 2;
}

struct LocalP2Args {
    MeshP m;
    const double *phi;
    const int64_t *cut;
    int64_t ncut;
    double coef0, coef1, coef2;
    int rhs_kind;
    double rhs_p0;
    double *loc;      // tet-major [ncut][LOC2_N]
    uint8_t *flags;   // bit0: non-empty patch
    int32_t *err;
    const double *rhs_vals = nullptr;   // DHIP_RHS_VALUES: f at every patch quadrature point
    const int32_t *rhs_off = nullptr;   //   first point of each cut tet
};

struct P2Shared {
    double vb[P2_MAXF][3][4];
    double af[P2_MAXF];
    double h[P2_CHUNK][10];
    double sg[P2_CHUNK][10][3];
    double w[P2_CHUNK];
    double f[P2_CHUNK];
    double U[15][10];
    double n5[15][3];   // P2 level-set unit normal at the 15 volume points
    int nf;
};

// Patch facets of the tet (sub-tet lt, triangle tr, in the order of make_patch) into sh.vb / sh.af,
// lanes 0..nf-1 one facet each.  Returns the facet count, or -1 when a sub-tet lies entirely in
// the zero level set (the reference's "full body" error).  Uniform over the wave.
template <class Shared>
__device__ inline int p2_patch_facets(Shared &sh, const double X[4][3], const double ls[10], int lane) {
    int ntri[8], nf = 0;
    bool zero_body = false;
#pragma unroll
    for (int lt = 0; lt < 8; ++lt) {
        int s[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] = sgn(ls[c_tab.lat_t[lt][q]]);
        if (!s[0] && !s[1] && !s[2] && !s[3]) zero_body = true;
        ntri[lt] = c_tab.rp_size[s[0] + 3 * s[1] + 9 * s[2] + 27 * s[3] + 40];
        nf += ntri[lt];
    }
    if (zero_body) return -1;
    if (lane < nf) {
        int lt = 0, tr = lane;
        while (tr >= ntri[lt]) tr -= ntri[lt++];
        int s[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] = sgn(ls[c_tab.lat_t[lt][q]]);
        const int pid = s[0] + 3 * s[1] + 9 * s[2] + 27 * s[3] + 40;
        double vb[3][4];
        facet_vertices(ls, lt, pid, tr, vb);
        double absdet, nrm[3];
        facet_qr(X, vb, absdet, nrm);
        double af = c_tab.rp_bnd[pid] ? 0.5 : 1.;
        af *= absdet;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) sh.vb[lane][j][cc] = vb[j][cc];
        sh.af[lane] = af;
    }
    return nf;
}

// Stage composite-Quad5 point base + lane (< nq): weight, P2 basis values, surface gradients
// projected with the P2 level-set normal and (rhs_kind != NONE) the load value -- read from fvals
// (the tet's given values, DHIP_RHS_VALUES) or evaluated.
__device__ inline void p2_stage_points(P2Shared &sh, int base, int nq, int lane, const double X[4][3],
                                       const double G1[4][3], const double v[10], int rhs_kind, double rhs_p0,
                                       const double *fvals = nullptr) {
    const int q = base + lane;
    if (q >= nq) return;
    const int f = q / 7, i = q - 7 * (q / 7);
    double nb[4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
        double sv = sh.vb[f][0][cc] * c_tab.q5n[i][0];
        sv += sh.vb[f][1][cc] * c_tab.q5n[i][1];
        sv += sh.vb[f][2][cc] * c_tab.q5n[i][2];
        nb[cc] = sv;
    }
    sh.w[lane] = sh.af[f] * c_tab.q5w[i];
    double h[10];
    p2_basis(nb, h);
    double n3[3];
    p2_unit_normal(v, nb, G1, n3);
#pragma unroll
    for (int d = 0; d < 10; ++d) {
        double g[3];
        p2_grad_one(d, nb, G1, g);
        const double dn = dot3(g, n3);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) sh.sg[lane][d][cc] = g[cc] - dn * n3[cc];
        sh.h[lane][d] = h[d];
    }
    double fv = 0.;
    if (rhs_kind == DHIP_RHS_VALUES) {
        fv = fvals[q];
    } else if (rhs_kind != DHIP_RHS_NONE) {
        double x[3];
        b2w(X, nb, x);
        fv = rhs_fun(rhs_kind, rhs_p0, x);
    }
    sh.f[lane] = fv;
}

__global__ void __launch_bounds__(64) k_local_p2(LocalP2Args a) {
    __shared__ P2Shared sh;
    const int64_t c = blockIdx.x;
    const int lane = threadIdx.x;
    if (c >= a.ncut) return;
    double X[4][3], ls[10];
    load_tet_geom(a.m, a.phi, a.cut[c], X, ls);
    double G1[4][3];
    const double det = p1_grad(X, G1);
    double v[10];
#pragma unroll
    for (int d = 0; d < 10; ++d) v[d] = ls[c_tab.lat_of_p2[d]];

    const int nf = p2_patch_facets(sh, X, ls, lane);
    if (nf < 0) {
        if (lane == 0) atomicOr(a.err, 2);
        return;
    }
    double *L = a.loc;
    if (nf == 0) {
        for (int e = lane; e < LOC2_N; e += 64) L[loc2(e, c)] = 0.;
        if (lane == 0) a.flags[c] = 0;
        return;
    }
    __syncthreads();

    // ---- composite Quad5 surface integrals ----
    int pi = 0, pj = 0;
    pair_of(lane < 55 ? lane : 0, pi, pj);
    double am = 0., al = 0., hat = 0., rh = 0.;
    const int nq = 7 * nf;
    for (int base = 0; base < nq; base += P2_CHUNK) {
        p2_stage_points(sh, base, nq, lane, X, G1, v, a.rhs_kind, a.rhs_p0,
                        a.rhs_kind == DHIP_RHS_VALUES ? a.rhs_vals + a.rhs_off[c] : nullptr);
        __syncthreads();
        const int cnt = nq - base < P2_CHUNK ? nq - base : P2_CHUNK;
        if (lane < 55) {
            for (int k = 0; k < cnt; ++k) {
                const double w = sh.w[k];
                am += w * (sh.h[k][pj] * sh.h[k][pi]);
                al += w * dot3(sh.sg[k][pj], sh.sg[k][pi]);
            }
        }
        if (lane < 10) {
            for (int k = 0; k < cnt; ++k) {
                const double w = sh.w[k];
                hat += w * (sh.h[k][lane] * sh.h[k][lane]);
                if (a.rhs_kind != DHIP_RHS_NONE) rh += w * (sh.f[k] * sh.h[k][lane]);
            }
        }
        __syncthreads();
    }

    // ---- normal-gradient stabilisation: U[q][d] = n(x_q) . grad h_d(x_q), Quad5 volume rule ----
    // (the normal once per point, then one lane per (point, basis function))
    if (lane < 15) {
        double n3[3];
        p2_unit_normal(v, c_tab.v5n[lane], G1, n3);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) sh.n5[lane][cc] = n3[cc];
    }
    __syncthreads();
    for (int t = lane; t < 150; t += 64) {
        const int q = t / 10, d = t - 10 * (t / 10);
        double g[3];
        p2_grad_one(d, c_tab.v5n[q], G1, g);
        sh.U[q][d] = dot3(sh.n5[q], g);
    }
    __syncthreads();
    if (lane < 55) {
        double f[15];
#pragma unroll
        for (int q = 0; q < 15; ++q) f[q] = sh.U[q][pi] * sh.U[q][pj];
        const double ad = fabs(det);
        const double val = ad * (((8. / 405. * f[0] + c_tab.v5c1 * (((f[1] + f[2]) + f[3]) + f[4])) +
                                  c_tab.v5c2 * (((f[5] + f[6]) + f[7]) + f[8])) +
                                 5. / 567. * (((((f[9] + f[10]) + f[11]) + f[12]) + f[13]) + f[14]));
        const double st = a.coef2 * val;
        const double mv = a.coef0 * am, lv = a.coef1 * al;
        const int e1 = pi * 10 + pj, e2 = pj * 10 + pi;
        L[loc2(LOC2_MASS + e1, c)] = mv;
        L[loc2(LOC2_MASS + e2, c)] = mv;
        L[loc2(LOC2_LB + e1, c)] = lv;
        L[loc2(LOC2_LB + e2, c)] = lv;
        L[loc2(LOC2_STAB + e1, c)] = st;
        L[loc2(LOC2_STAB + e2, c)] = st;
    }
    if (lane < 10) {
        L[loc2(LOC2_RHS + lane, c)] = rh;
        L[loc2(LOC2_HAT + lane, c)] = hat;
    }
    if (lane == 0) a.flags[c] = 1;
}

// The numbering's part of k_local_p2 alone: the composite-Quad5 integrals of the squared basis functions
// (the hat slots) and the non-empty flag, summed as k_local_p2 sums them.  For passes whose forms come
// from elsewhere (the mapped surface).  Four tets per wave, 16 lanes each: lane j < nf builds facet j, the
// points go through LDS sixteen at a time and lanes j < 10 sum them in the composite order.
constexpr int HAT_TPW = 4;

struct HatShared {
    double vb[P2_MAXF][3][4];
    double af[P2_MAXF];
    double h[16][10];
    double w[16];
    int nq;
};

__global__ void __launch_bounds__(64) k_hat_p2(LocalP2Args a) {
    __shared__ HatShared sh[HAT_TPW];
    const int lane = threadIdx.x, t = lane >> 4, j = lane & 15;
    const int64_t c = (int64_t)blockIdx.x * HAT_TPW + t;
    HatShared &s = sh[t];
    double *L = a.loc;
    int nf = 0;
    if (c < a.ncut) {
        double X[4][3], ls[10];
        load_tet_geom(a.m, a.phi, a.cut[c], X, ls);
        nf = p2_patch_facets(s, X, ls, j);
        if (nf < 0) {
            if (j == 0) atomicOr(a.err, 2);
        } else if (nf == 0) {
            if (j < 10) L[loc2(LOC2_HAT + j, c)] = 0.;
            if (j == 0) a.flags[c] = 0;
        }
    }
    const int nq = nf > 0 ? 7 * nf : 0;
    if (j == 0) s.nq = nq;
    __syncthreads();
    const int mq = max(max(sh[0].nq, sh[1].nq), max(sh[2].nq, sh[3].nq));
    double hat = 0.;
    for (int base = 0; base < mq; base += 16) {
        const int q = base + j;
        if (q < nq) {
            const int f = q / 7, i = q - 7 * (q / 7);
            double nb[4];
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                double sv = s.vb[f][0][cc] * c_tab.q5n[i][0];
                sv += s.vb[f][1][cc] * c_tab.q5n[i][1];
                sv += s.vb[f][2][cc] * c_tab.q5n[i][2];
                nb[cc] = sv;
            }
            s.w[j] = s.af[f] * c_tab.q5w[i];
            double h[10];
            p2_basis(nb, h);
#pragma unroll
            for (int d = 0; d < 10; ++d) s.h[j][d] = h[d];
        }
        __syncthreads();
        const int cnt = nq - base < 16 ? nq - base : 16;
        if (j < 10)
            for (int k = 0; k < cnt; ++k) hat += s.w[k] * (s.h[k][j] * s.h[k][j]);
        __syncthreads();
    }
    if (nq > 0) {
        if (j < 10) L[loc2(LOC2_HAT + j, c)] = hat;
        if (j == 0) a.flags[c] = 1;
    }
}

// ---------------------------------------------------------------------------
// numbering
// ---------------------------------------------------------------------------
struct QualP2Args {
    MeshP m;
    const int64_t *cut;
    int64_t ncut;
    const double *loc;
    const uint8_t *flags;
    uint16_t *qual;   // bit d: P2 dof d of the tet qualifies
    double omit_bound;
};

__global__ void k_qualify_p2(QualP2Args a) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.ncut) return;
    uint32_t q = 0;
    if (a.flags[c] & 1) {
        double limit = 0.;
        if (a.omit_bound >= 0.) {
            const int64_t t = a.cut[c];
            const int k = (int)(t % 6);
            int64_t i1, i2, i3;
            cell_of(a.m, t / 6, i1, i2, i3);
            double X[4][3];
            for (int p = 0; p < 4; ++p)
                vertex_coord(a.m, i1 + c_tab.kuhn[k][p][0], i2 + c_tab.kuhn[k][p][1], i3 + c_tab.kuhn[k][p][2],
                             X[p]);
            limit = a.omit_bound * pow(tet_volume(X) * 6, 4. / 3.);
        }
        for (int d = 0; d < 10; ++d)
            if (a.loc[loc2(LOC2_HAT + d, c)] > limit) q |= 1u << d;
    }
    a.qual[c] = (uint16_t)q;
}

// doubled (half-grid) coordinates of P2 dof d of cut tet t
__device__ inline void p2_node_of(const MeshP &m, int64_t t, int d, int64_t J[3]) {
    const int k = (int)(t % 6);
    int64_t i1, i2, i3;
    cell_of(m, t / 6, i1, i2, i3);
    const int n27 = c_tab.p2_node27[k][d];
    J[0] = 2 * i1 + n27 % 3;
    J[1] = 2 * i2 + (n27 / 3) % 3;
    J[2] = 2 * i3 + n27 / 9;
}

// Visit the cut tets containing half-grid node J in ascending tet id; the functor gets
// (cut index, tet type k, local P2 dof of the node, cell-relative node index 0..26) and
// returns true to stop.
template <class F>
__device__ inline void for_cut_tets_at_node(const MeshP &m, const uint32_t *cellinfo, const int64_t J[3], F &&f) {
    int64_t lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (J[d] & 1) {
            lo[d] = hi[d] = J[d] >> 1;
        } else {
            lo[d] = (J[d] >> 1) - 1;
            hi[d] = J[d] >> 1;
        }
        if (lo[d] < 0) lo[d] = 0;
        if (hi[d] > m.n[d] - 1) hi[d] = m.n[d] - 1;
    }
    for (int64_t c3 = lo[2]; c3 <= hi[2]; ++c3)
        for (int64_t c2 = lo[1]; c2 <= hi[1]; ++c2)
            for (int64_t c1 = lo[0]; c1 <= hi[0]; ++c1) {
                const uint32_t info = cellinfo[(c3 * m.n[1] + c2) * m.n[0] + c1];
                const uint32_t mask = info & 63u;
                if (!mask) continue;
                const uint32_t base = info >> 6;
                const int rel = (int)((J[0] - 2 * c1) + 3 * (J[1] - 2 * c2) + 9 * (J[2] - 2 * c3));
                for (int k = 0; k < 6; ++k) {
                    if (!((mask >> k) & 1u)) continue;
                    const int li = c_tab.p2_of27[k][rel];
                    if (li < 0) continue;
                    const int64_t ci = base + __popc(mask & ((1u << k) - 1u));
                    if (f(ci, k, li, rel)) return;
                }
            }
}

struct FirstP2Args {
    MeshP m;
    const int64_t *cut;
    int64_t ncut;
    const uint32_t *cellinfo;
    const uint8_t *flags;
    const uint16_t *qual;
    int32_t *first;
    int32_t *isown;
};

__global__ void k_first_p2(FirstP2Args a) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= 10 * a.ncut) return;
    int64_t J[3];
    p2_node_of(a.m, a.cut[s / 10], (int)(s % 10), J);
    int64_t found = -1;
    for_cut_tets_at_node(a.m, a.cellinfo, J, [&](int64_t ci, int, int li, int) {
        if ((a.flags[ci] & 1) && ((a.qual[ci] >> li) & 1)) {
            found = ci * 10 + li;
            return true;
        }
        return false;
    });
    a.first[s] = (int32_t)found;
    a.isown[s] = found == s ? 1 : 0;
}

// ---------------------------------------------------------------------------
// CSR rows: one wave per row, 125 neighbour slots (offsets -2..2 in each half-grid direction)
// ---------------------------------------------------------------------------
struct RowP2Args {
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

constexpr int ROWP2_WAVES = 4;

struct RowP2Shared {
    double am[128], al[128], as[128];
    int32_t cols[128];
    int32_t pres[128];
    int32_t ccol[128];   // the present columns, compacted in slot order
    int32_t tci[48];     // the row node's flagged cut tets, in for_cut_tets_at_node's order
    int32_t tinf[48];    //   k | li << 3 | rel << 8
};

// One row.  Lanes 0-47 look at the (cell, Kuhn tet) candidates around the node at once, lane order being
// for_cut_tets_at_node's walk order (cells z, y, x ascending, then k), and list the flagged cut tets holding
// the node.  Six tets at a time, lanes 10 t + j load node j's DOF and local entries of tet t together; the
// sums into the 125 neighbour slots then go tet by tet, so each slot is summed in walk order.
template <bool FILL>
__device__ inline void p2_row(const RowP2Args &a, RowP2Shared &sh, int64_t r, int lane) {
    for (int s = lane; s < 128; s += 64) {
        sh.pres[s] = 0;
        sh.cols[s] = 0;
        sh.am[s] = sh.al[s] = sh.as[s] = 0.0;
    }
    const int32_t own = a.rowowner[r];
    int64_t J[3];
    p2_node_of(a.m, a.cut[own / 10], own % 10, J);
    int32_t tci = -1;
    int tinf = 0;
    if (lane < 48) {
        const int q = lane / 6, k = lane - 6 * q;
        int64_t cc[3];
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int64_t lo = (J[d] & 1) ? (J[d] >> 1) : (J[d] >> 1) - 1, hi = J[d] >> 1;
            if (lo < 0) lo = 0;
            if (hi > a.m.n[d] - 1) hi = a.m.n[d] - 1;
            cc[d] = lo + ((q >> d) & 1);
            ok = ok && cc[d] <= hi;
        }
        if (ok) {
            const uint32_t info = a.cellinfo[(cc[2] * a.m.n[1] + cc[1]) * a.m.n[0] + cc[0]];
            const uint32_t mask = info & 63u;
            if ((mask >> k) & 1u) {
                const int rel = (int)((J[0] - 2 * cc[0]) + 3 * (J[1] - 2 * cc[1]) + 9 * (J[2] - 2 * cc[2]));
                const int li = c_tab.p2_of27[k][rel];
                if (li >= 0) {
                    const int32_t t = (int32_t)((info >> 6) + __popc(mask & ((1u << k) - 1u)));
                    if (a.flags[t] & 1) {
                        tci = t;
                        tinf = k | (li << 3) | (rel << 8);
                    }
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
    double rv = 0.0;
    if (FILL && lane < ntet) rv = a.loc[loc2(LOC2_RHS + ((sh.tinf[lane] >> 3) & 15), sh.tci[lane])];
    const int tl = lane / 10, j = lane - 10 * tl;
    for (int b = 0; b < ntet; b += 6) {
        const int t = b + tl;
        const bool act = lane < 60 && t < ntet;
        int32_t dj = -1;
        int sl = 0;
        double vm = 0.0, vl = 0.0, vs = 0.0;
        if (act) {
            const int32_t c = sh.tci[t];
            const int inf = sh.tinf[t];
            const int k = inf & 7, li = (inf >> 3) & 15, rel = inf >> 8;
            dj = a.dofs[(int64_t)c * 10 + j];
            if (FILL) {
                const int e = li * 10 + j;
                vm = a.loc[loc2(LOC2_MASS + e, c)];
                vl = a.loc[loc2(LOC2_LB + e, c)];
                vs = a.loc[loc2(LOC2_STAB + e, c)];
            }
            const int nj = c_tab.p2_node27[k][j];
            sl = (nj % 3 - rel % 3 + 2) + 5 * ((nj / 3) % 3 - (rel / 3) % 3 + 2) + 25 * (nj / 9 - rel / 9 + 2);
        }
        const int nb = ntet - b < 6 ? ntet - b : 6;
        for (int u = 0; u < nb; ++u) {
            if (act && tl == u && dj >= 0) {
                sh.pres[sl] = 1;
                if (FILL) {
                    sh.cols[sl] = dj;
                    sh.am[sl] += vm;
                    sh.al[sl] += vl;
                    sh.as[sl] += vs;
                }
            }
            wave_sync_lds();
        }
    }
    const bool p0 = sh.pres[lane] != 0, p1 = sh.pres[lane + 64] != 0;
    const uint64_t m0 = __ballot(p0), m1 = __ballot(p1);
    const int cnt = __popcll(m0) + __popcll(m1);
    if (!FILL) {
        if (lane == 0) a.rowlen[r] = cnt;
        return;
    }
    double rh = 0.0;
    for (int t = 0; t < ntet; ++t) rh += __shfl(rv, t);
    const uint64_t lt = (1ull << lane) - 1ull;
    const int32_t c0 = sh.cols[lane], c1 = sh.cols[lane + 64];
    if (p0) sh.ccol[__popcll(m0 & lt)] = c0;
    if (p1) sh.ccol[__popcll(m0) + __popcll(m1 & lt)] = c1;
    wave_sync_lds();
    int rk0 = 0, rk1 = 0;
    for (int u = 0; u < cnt; ++u) {
        const int32_t cu = sh.ccol[u];
        rk0 += cu < c0 ? 1 : 0;
        rk1 += cu < c1 ? 1 : 0;
    }
    const int64_t base = a.rowptr[r];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (!(h ? p1 : p0)) continue;
        const int s = lane + 64 * h;
        const int32_t cs = h ? c1 : c0;
        const int64_t dst = base + (h ? rk1 : rk0);
        a.col[dst] = cs;
        a.vm[dst] = sh.am[s];
        a.vl[dst] = sh.al[s];
        a.vs[dst] = sh.as[s];
        const double cv = 1.0 * (a.lc0 * sh.am[s] + a.lc1 * sh.al[s]) + a.lc2 * sh.as[s];
        a.vc[dst] = cv;
        if (cs == r) a.diag[r] = cv;
    }
    if (lane == 0) a.rhs[r] = rh;
}

template <bool FILL>
__global__ void __launch_bounds__(64 * ROWP2_WAVES) k_rows_p2(RowP2Args a) {
    __shared__ RowP2Shared sh[ROWP2_WAVES];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t r = (int64_t)xcd_position(blockIdx.x, gridDim.x) * ROWP2_WAVES + w;   // a run of rows per XCD
    if (r >= a.ndof) return;
    p2_row<FILL>(a, sh[w], r, lane);
}

}  // namespace dhip
