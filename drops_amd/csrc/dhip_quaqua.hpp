// dhip_quaqua.hpp -- P2 interface forms on the QuaQua-mapped surface (gfx950).
//
// InterfaceCommonDataP2CL with a QuaQuaMapperCL (surfactant/ifacetransp.h:200-276,
// levelset/levelsetmapper.cpp:551-638): the patch is the tet's lattice-1 patch (the linear
// interpolant of the level set at its four vertices: one triangle or one planar quadrilateral), its
// composite Quad5 points x are mapped to base points y on the zero level of the P2 level set by
// damped Newton on (x - y - s g_h(y), -phi_h(y)) = 0 with the recovered gradient g_h, and the forms
// carry |det| of the map's Jacobian and, for the Laplace-Beltrami form, its gradient transformation
// W^-1 (LocalMassP2CL, LocalLaplaceBeltramiP2CL, LocalVectorP2CL, ifacetransp.h:2302-2403).
//   k_qq_first / k_qq_node_gradient  the recovered gradient of each P2 node of the cut tets once, at the
//                     first cut tet holding it: the running mean over every tet of the mesh that holds the
//                     node, in tet id order (averaging_P2_gradient_recovery through OswaldProjectionP2AccuCL,
//                     num/gradient_recovery.cpp:62-89), one thread per distinct node;
//   k_local_p2qq  four cut tets per wave, 16 lanes each:
//     lanes j < 10    the recovered gradient at the tet's ten P2 nodes from the node kernel;
//     all 16          the tet's P2 gradient fields T DH_iRef at its vertices into LDS;
//     lanes j < 7 nf  one quadrature point each: base point, Jacobian, |det|, W^-1, the ten mapped
//                     surface gradients and basis values, the load value at the base point;
//     all 64          the 4 x (55 pairs (i >= j) of the mass and Laplace-Beltrami matrices + 10 load
//                     entries), each summed over the points in the composite order.
//   Output: the mass, Laplace-Beltrami and load slots of the P2 local array (the stabilisation slot 0),
//   flags 1 only for tets with a lattice-1 patch, so the P2 row builder adds exactly their couplings.
// Every expression keeps the oracle's operation order (-ffp-contract=off): bit-identical except where
// cbrt of the tet's |det| (the Newton damping length) rounds differently.
#pragma once

#include "dhip_p2.hpp"

namespace dhip {

struct QQParams {
    int maxiter;
    double tol, armijo_c;
    int max_damping;
};

// Householder QR of a small row-major matrix held in registers (QRDecompCL<R,C>, misc/container.h:
// 1210-1300; std::pow(x, 2) is the correctly rounded x * x).
template <int R, int C>
struct SmallQR {
    double a[R * C], d[C], beta[C];
    __device__ __forceinline__ double &at(int i, int j) { return a[i * C + j]; }
    __device__ __forceinline__ double at(int i, int j) const { return a[i * C + j]; }
    // false: rank deficient (the reference throws)
    __device__ inline bool prepare() {
#pragma unroll
        for (int j = 0; j < C; ++j) {
            double sigma = 0.;
#pragma unroll
            for (int i = j; i < R; ++i) sigma += at(i, j) * at(i, j);
            if (sigma == 0.) return false;
            d[j] = (at(j, j) < 0 ? 1. : -1.) * sqrt(sigma);
            beta[j] = 1. / (d[j] * at(j, j) - sigma);
            at(j, j) -= d[j];
#pragma unroll
            for (int k = j + 1; k < C; ++k) {
                double sp = 0.;
#pragma unroll
                for (int i = j; i < R; ++i) sp += at(i, j) * at(i, k);
                sp *= beta[j];
#pragma unroll
                for (int i = j; i < R; ++i) at(i, k) += at(i, j) * sp;
            }
        }
        return true;
    }
    __device__ __forceinline__ void reflect(int j, double *b) const {
        double sp = 0.;
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (i >= j) sp += at(i, j) * b[i];
        sp *= beta[j];
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (i >= j) b[i] += at(i, j) * sp;
    }
    __device__ inline void apply_Q(double *b) const {
#pragma unroll
        for (int j = C - 1; j >= 0; --j) reflect(j, b);
    }
    __device__ inline void solve(double *b) const {
#pragma unroll
        for (int j = 0; j < C; ++j) reflect(j, b);
#pragma unroll
        for (int i = C - 1; i >= 0; --i) {
#pragma unroll
            for (int j = i + 1; j < C; ++j) b[i] -= at(i, j) * b[j];
            b[i] /= d[i];
        }
    }
};

// GetTrafoTr, geom/multigrid.h:627-647
__device__ inline double qq_trafo(const double X[4][3], double T[3][3]) {
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) M[j][i] = X[i + 1][j] - X[0][j];
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    T[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    T[0][1] = (M[2][0] * M[1][2] - M[1][0] * M[2][2]) / det;
    T[0][2] = (M[1][0] * M[2][1] - M[2][0] * M[1][1]) / det;
    T[1][0] = (M[2][1] * M[0][2] - M[0][1] * M[2][2]) / det;
    T[1][1] = (M[0][0] * M[2][2] - M[2][0] * M[0][2]) / det;
    T[1][2] = (M[2][0] * M[0][1] - M[0][0] * M[2][1]) / det;
    T[2][0] = (M[0][1] * M[1][2] - M[1][1] * M[0][2]) / det;
    T[2][1] = (M[1][0] * M[0][2] - M[0][0] * M[1][2]) / det;
    T[2][2] = (M[0][0] * M[1][1] - M[1][0] * M[0][1]) / det;
    return det;
}

// FE_P2CL::DH0Ref .. DH9Ref (num/fe.tpp:122-190) at reference vertex k: (0,0,0), e1, e2, e3
__device__ inline void qq_dhref(int i, int k, double g[3]) {
    const double v1 = k == 1 ? 1. : 0., v2 = k == 2 ? 1. : 0., v3 = k == 3 ? 1. : 0.;
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

__device__ __forceinline__ void qq_mv3(const double A[3][3], const double v[3], double r[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = A[i][0] * v[0] + A[i][1] * v[1] + A[i][2] * v[2];
}

__device__ __forceinline__ double qq_norm3(const double a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// T DH_iRef at vertex k of the tet (P2DiscCL::GetGradients, num/discretize.h:1430-1431)
__device__ __forceinline__ void qq_grad_at_vertex(const double T[3][3], int i, int k, double r[3]) {
    double g[3];
    qq_dhref(i, k, g);
    qq_mv3(T, g, r);
}

struct QQShared {
    double X[4][3];
    double T[3][3];
    double G[10][4][3];    // T DH_iRef at the vertices
    double v[10];          // level set, P2 dof order
    double gh[10][3];      // recovered gradient at the P2 nodes
    double w2b[16], w2bd[4], w2bb[4];
    double h;
    double vb[2][3][4];    // lattice-1 facets: barycentric vertices
    double af[2];          // (boundary ? 0.5 : 1) |det| of the facet
    double nrm[2][3];      // facet normals Q e3
    double U[2][3][2];     // Q e1, Q e2 of the facet QR
    double w[14], ad[14], f[14], H[14][10], Gq[14][10][3];
    int nf;
};

// LocalP1CL<Point3DCL> at b (LinearCombinationCL, num/fe.h:44-53)
__device__ __forceinline__ void qq_p1v(const QQShared &s, int i, const double b[4], double r[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
        r[c] = b[0] * s.G[i][0][c] + b[1] * s.G[i][1][c] + b[2] * s.G[i][2][c] + b[3] * s.G[i][3][c];
}

// LocalP2CL<Point3DCL> of the recovered gradient at b
__device__ __forceinline__ void qq_p2v(const QQShared &s, const double b[4], double r[3]) {
    double h[10];
    p2_basis(b, h);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double t = h[0] * s.gh[0][c];
#pragma unroll
        for (int d = 1; d < 10; ++d) t += h[d] * s.gh[d][c];
        r[c] = t;
    }
}

// LocalP2CL<double> of the level set at b
__device__ __forceinline__ double qq_p2s(const QQShared &s, const double b[4]) {
    double h[10];
    p2_basis(b, h);
    double t = h[0] * s.v[0];
#pragma unroll
    for (int d = 1; d < 10; ++d) t += h[d] * s.v[d];
    return t;
}

// World2BaryCoordCL (geom/simplex.cpp:479-484): the factorisation of [X; 1] in LDS
__device__ inline void qq_w2b(const QQShared &s, const double x[3], double b[4]) {
    SmallQR<4, 4> q;
#pragma unroll
    for (int e = 0; e < 16; ++e) q.a[e] = s.w2b[e];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        q.d[e] = s.w2bd[e];
        q.beta[e] = s.w2bb[e];
    }
    b[0] = x[0];
    b[1] = x[1];
    b[2] = x[2];
    b[3] = 1.;
    q.solve(b);
}

// QuaQuaMapperFunctionCL::set_point / compute_F: the tet is kept (locate_new_point never leaves it
// when neighbourhoods are given), the level set and recovered gradient are the tet's polynomials.
struct QQFun {
    double xcur[4], bxb[4], gh[3], F[4], x0[3];
};

__device__ inline void qq_set_point(const QQShared &s, QQFun &f, const double x[4]) {
    if (f.xcur[0] == x[0] && f.xcur[1] == x[1] && f.xcur[2] == x[2] && f.xcur[3] == x[3]) return;
    double b[4];
    qq_w2b(s, x, b);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f.xcur[i] = x[i];
        f.bxb[i] = b[i];
    }
    qq_p2v(s, f.bxb, f.gh);
#pragma unroll
    for (int i = 0; i < 3; ++i) f.F[i] = f.x0[i] - f.xcur[i] - f.xcur[3] * f.gh[i];
    f.F[3] = -qq_p2s(s, f.bxb);
}

__device__ __forceinline__ double qq_dot4(const double a[4], const double b[4]) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) t += a[i] * b[i];
    return t;
}

// base_point_newton (levelsetmapper.cpp:551-566) with newton_solve (num/newton.h:64-115); false on
// a singular Jacobian
__device__ inline bool qq_base_point(const QQShared &s, const QQParams &P, const double xb[4], double bxb[4],
                                     double &dh) {
    QQFun f;
    b2w(s.X, xb, f.x0);
#pragma unroll
    for (int i = 0; i < 4; ++i) f.bxb[i] = xb[i];
    f.xcur[0] = f.x0[0];
    f.xcur[1] = f.x0[1];
    f.xcur[2] = f.x0[2];
    f.xcur[3] = 0.;
    qq_p2v(s, f.bxb, f.gh);
    f.F[0] = f.F[1] = f.F[2] = 0.;
    f.F[3] = -qq_p2s(s, f.bxb);
    double x[4] = {f.xcur[0], f.xcur[1], f.xcur[2], f.xcur[3]};
    const double min_step_length = 1e-7;
    for (int iter = 0;; ++iter) {
        qq_set_point(s, f, x);
        const double normF = sqrt(qq_dot4(f.F, f.F));
        if (iter >= P.maxiter || normF < P.tol) break;
        // compute_dF (levelsetmapper.cpp:222-249) and its QR
        double g_ls[3] = {0., 0., 0.}, dgh[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
#pragma unroll
        for (int i = 0; i < 10; ++i) {
            double t[3];
            qq_p1v(s, i, f.bxb, t);
#pragma unroll
            for (int c = 0; c < 3; ++c) g_ls[c] += s.v[i] * t[c];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) dgh[r][c] += s.gh[i][r] * t[c];
        }
        SmallQR<4, 4> q;
        double dF[4][4];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) dF[i][j] = -f.xcur[3] * dgh[i][j];
            dF[i][i] -= 1.;
            dF[i][3] = -f.gh[i];
            dF[3][i] = -g_ls[i];
        }
        dF[3][3] = 0.;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) q.at(i, j) = dF[i][j];
        if (!q.prepare()) return false;
        double dx[4] = {f.F[0], f.F[1], f.F[2], f.F[3]}, Ddx[4];
        q.solve(dx);
#pragma unroll
        for (int i = 0; i < 4; ++i) Ddx[i] = dF[i][0] * dx[0] + dF[i][1] * dx[1] + dF[i][2] * dx[2] + dF[i][3] * dx[3];
        const double armijo_slope = P.armijo_c * qq_dot4(f.F, Ddx) / normF;
        const double dxn = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        double l = 0.5 * s.h / dxn;
        if (!(l < 1.)) l = 1.;
        for (int j = 0; j < P.max_damping && l >= min_step_length; ++j, l *= 0.5) {
            double xt[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) xt[i] = x[i] - l * dx[i];
            qq_set_point(s, f, xt);
            if (sqrt(qq_dot4(f.F, f.F)) < normF + armijo_slope * l) break;
        }
        if (l < min_step_length) break;
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] -= l * dx[i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) bxb[i] = f.bxb[i];
    dh = x[3] * qq_norm3(f.gh);
    return true;
}

// QuaQuaMapperCL::jacobian (levelsetmapper.cpp:590-638); also the unit level-set gradient n at the
// base point (local_ls_grad); false on a singular matrix
__device__ inline bool qq_jacobian(const QQShared &s, const double bxb[4], double dh, double dph[3][3], double n[3]) {
    double gh[3], dgh[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    qq_p2v(s, bxb, gh);
    const double ghn = qq_norm3(gh);
    double qn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) qn[c] = gh[c] / ghn;
    n[0] = n[1] = n[2] = 0.;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        double gp[3];
        qq_p1v(s, i, bxb, gp);
#pragma unroll
        for (int c = 0; c < 3; ++c) n[c] += s.v[i] * gp[c];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) dgh[r][c] += s.gh[i][r] * gp[c];
    }
    const double nn = qq_norm3(n);
#pragma unroll
    for (int c = 0; c < 3; ++c) n[c] /= nn;
    double tq[3];
#pragma unroll
    for (int col = 0; col < 3; ++col) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) t += dgh[i][col] * qn[i];
        tq[col] = t;
    }
    double dn[3][3], Qm[3][3];
    const double ig = 1. / ghn;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) dn[r][c] = ig * (dgh[r][c] - qn[r] * tq[c]);
    const double qnn = qn[0] * n[0] + qn[1] * n[1] + qn[2] * n[2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Qm[r][c] = (r == c ? 1. : 0.) - (qn[r] / qnn) * n[c];
    SmallQR<3, 3> A;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i) t += (dh * Qm[r][i]) * dn[i][c];
            A.at(r, c) = (r == c ? 1. : 0.) + t;
        }
    if (!A.prepare()) return false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double t[3] = {Qm[0][i], Qm[1][i], Qm[2][i]};
        A.solve(t);
#pragma unroll
        for (int r = 0; r < 3; ++r) dph[r][i] = t[r];
    }
    return true;
}

struct LocalP2QQArgs {
    MeshP m;
    const double *phi;
    const int64_t *cut;
    int64_t ncut;
    int rhs_kind;
    double rhs_p0;
    QQParams P;
    double *loc;      // tet-major [ncut][LOC2_N]
    uint8_t *flags;   // bit0: the tet adds to the mapped matrices
    int32_t *err;     // bit 1: full body in the zero level set, bit 2: singular matrix in the mapping
    const int32_t *first;   // [10 ncut] the first cut slot holding each slot's P2 node (k_qq_node_gradient)
    const double *gh;       // [10 ncut][3] the recovered gradient, at first slots
    double *pts = nullptr;          // points mode: the base points (world) out, 3 per point; no forms
    const double *vals = nullptr;   // DHIP_RHS_VALUES: f at the base points
    const int32_t *off = nullptr;   //   first point of each cut tet (pts, vals)
};

// the mapped quadrature points of each cut tet: 7 per facet of its lattice-1 patch
__global__ void __launch_bounds__(128) k_qq_count(MeshP m, const double *phi, const int64_t *cut, int64_t ncut,
                                                   int32_t *cnt) {
    const int64_t c = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (c >= ncut) return;
    double X[4][3], ls[10];
    load_tet_geom(m, phi, cut[c], X, ls);
    int sg[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) sg[q] = sgn(ls[c_tab.lat_of_p2[q]]);
    const bool zero = !sg[0] && !sg[1] && !sg[2] && !sg[3];
    cnt[c] = zero ? 0 : 7 * c_tab.rp_size[sg[0] + 3 * sg[1] + 9 * sg[2] + 27 * sg[3] + 40];
}

// LocalP2GradientCL::set_tetra (num/fe.h): the P1 gradient field of the P2 level set at the vertices (summed
// over the ten basis functions), at an edge node the mean of its two vertices; tet `cell * 6 + k`, P2 node dd
__device__ inline void qq_tet_node_gradient(const MeshP &m, const double *phi, int64_t cell, int k, int dd,
                                            double loc[3]) {
    double X[4][3], ls[10], T[3][3];
    load_tet_geom(m, phi, cell * 6 + k, X, ls);
    qq_trafo(X, T);
    double vv[10];
#pragma unroll
    for (int d = 0; d < 10; ++d) vv[d] = ls[c_tab.lat_of_p2[d]];
    const int ea[6] = {0, 0, 1, 0, 1, 2}, eb[6] = {1, 2, 2, 3, 3, 3};
    const int va = dd < 4 ? dd : ea[dd - 4], vb = dd < 4 ? dd : eb[dd - 4];
    double pa[3] = {0., 0., 0.}, pb[3] = {0., 0., 0.};
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        double g[3];
        qq_grad_at_vertex(T, i, va, g);
#pragma unroll
        for (int c = 0; c < 3; ++c) pa[c] += vv[i] * g[c];
        if (dd >= 4) {
            qq_grad_at_vertex(T, i, vb, g);
#pragma unroll
            for (int c = 0; c < 3; ++c) pb[c] += vv[i] * g[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) loc[c] = dd < 4 ? pa[c] : 0.5 * (pa[c] + pb[c]);
}

// the cells around half-grid node J (those whose closure holds it)
__device__ __forceinline__ void qq_cell_range(const MeshP &m, const int64_t J[3], int64_t lo[3], int64_t hi[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        lo[c] = J[c] >= 2 ? (J[c] - 1) >> 1 : 0;
        hi[c] = J[c] >> 1;
        if (hi[c] > m.n[c] - 1) hi[c] = m.n[c] - 1;
    }
}

// the running mean of averaging_P2_gradient_recovery: ((n-1)/n) avg + (1/n) val.  A half-grid node lies in at
// most 24 tets; for those n the two quotients come from tables the compiler divides out (IEEE division is
// correctly rounded, so they are the run-time quotients bit for bit)
#define DHIP_QQ_W(n) ((n - 1.) / n)
#define DHIP_QQ_R(n) (1. / n)
__constant__ const double qq_mean_a[25] = {0., DHIP_QQ_W(1.), DHIP_QQ_W(2.), DHIP_QQ_W(3.), DHIP_QQ_W(4.),
    DHIP_QQ_W(5.), DHIP_QQ_W(6.), DHIP_QQ_W(7.), DHIP_QQ_W(8.), DHIP_QQ_W(9.), DHIP_QQ_W(10.), DHIP_QQ_W(11.),
    DHIP_QQ_W(12.), DHIP_QQ_W(13.), DHIP_QQ_W(14.), DHIP_QQ_W(15.), DHIP_QQ_W(16.), DHIP_QQ_W(17.),
    DHIP_QQ_W(18.), DHIP_QQ_W(19.), DHIP_QQ_W(20.), DHIP_QQ_W(21.), DHIP_QQ_W(22.), DHIP_QQ_W(23.),
    DHIP_QQ_W(24.)};
__constant__ const double qq_mean_b[25] = {0., DHIP_QQ_R(1.), DHIP_QQ_R(2.), DHIP_QQ_R(3.), DHIP_QQ_R(4.),
    DHIP_QQ_R(5.), DHIP_QQ_R(6.), DHIP_QQ_R(7.), DHIP_QQ_R(8.), DHIP_QQ_R(9.), DHIP_QQ_R(10.), DHIP_QQ_R(11.),
    DHIP_QQ_R(12.), DHIP_QQ_R(13.), DHIP_QQ_R(14.), DHIP_QQ_R(15.), DHIP_QQ_R(16.), DHIP_QQ_R(17.),
    DHIP_QQ_R(18.), DHIP_QQ_R(19.), DHIP_QQ_R(20.), DHIP_QQ_R(21.), DHIP_QQ_R(22.), DHIP_QQ_R(23.),
    DHIP_QQ_R(24.)};
#undef DHIP_QQ_W
#undef DHIP_QQ_R

__device__ __forceinline__ void qq_mean_step(int &n, double avg[3], const double val[3]) {
    ++n;
    const double dn = (double)n;
    const double wa = n <= 24 ? qq_mean_a[n] : (dn - 1.) / dn, wb = n <= 24 ? qq_mean_b[n] : 1. / dn;
#pragma unroll
    for (int c = 0; c < 3; ++c) avg[c] = wa * avg[c] + wb * val[c];
}

// recovered gradient at half-grid node J: running mean over the tets holding J in ascending id
__device__ inline void qq_recovered(const MeshP &m, const double *phi, const int64_t J[3], double out[3]) {
    int64_t lo[3], hi[3];
    qq_cell_range(m, J, lo, hi);
    int n = 0;
    double avg[3] = {0., 0., 0.};
    for (int64_t i3 = lo[2]; i3 <= hi[2]; ++i3)
        for (int64_t i2 = lo[1]; i2 <= hi[1]; ++i2)
            for (int64_t i1 = lo[0]; i1 <= hi[0]; ++i1) {
                const int rel = (int)((J[2] - 2 * i3) * 9 + (J[1] - 2 * i2) * 3 + (J[0] - 2 * i1));
                const int64_t cell = (i3 * m.n[1] + i2) * m.n[0] + i1;
                for (int k = 0; k < 6; ++k) {
                    const int dd = c_tab.p2_of27[k][rel];
                    if (dd < 0) continue;
                    double loc[3];
                    qq_tet_node_gradient(m, phi, cell, k, dd, loc);
                    qq_mean_step(n, avg, loc);
                }
            }
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = avg[c];
}

// The recovered gradient of the P2 nodes of the cut tets, once per node.  k_qq_first: slot s = 10 c + d
// (P2 dof d of cut tet c) finds the first cut tet in id order that holds its node and records it; the
// slots that are their node's first append themselves to a list (one atomic per block).
struct QQGradArgs {
    MeshP m;
    const double *phi;
    const int64_t *cut;
    int64_t ncut;
    const uint32_t *cellinfo;
    int32_t *first;
    int32_t *list;    // owner slots, in no particular order
    int32_t *count;   // their number (zeroed before k_qq_first)
    double *gh;
};

constexpr int QQ_FIRST_THREADS = 1024;

__global__ void __launch_bounds__(QQ_FIRST_THREADS) k_qq_first(QQGradArgs a) {
    __shared__ int wbase[QQ_FIRST_THREADS / 64];
    __shared__ int bbase;
    const int64_t s = (int64_t)blockIdx.x * QQ_FIRST_THREADS + threadIdx.x;
    bool own = false;
    if (s < 10 * a.ncut) {
        int64_t J[3];
        p2_node_of(a.m, a.cut[s / 10], (int)(s % 10), J);
        int64_t found = s;
        for_cut_tets_at_node(a.m, a.cellinfo, J, [&](int64_t ci, int, int li, int) {
            found = ci * 10 + li;
            return true;
        });
        a.first[s] = (int32_t)found;
        own = found == s;
    }
    // one atomic per block: the owners of a block take consecutive list entries, wave by wave
    const uint64_t m = __ballot(own);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) wbase[w] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int i = 0; i < QQ_FIRST_THREADS / 64; ++i) {
            const int n = wbase[i];
            wbase[i] = t;
            t += n;
        }
        bbase = t ? atomicAdd(a.count, t) : 0;
    }
    __syncthreads();
    if (own) a.list[bbase + wbase[w] + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)s;
}

// k_qq_node_gradient: a wave takes QQN_NODES listed nodes at a time and lists their (node, tet) pairs -- node
// by node, each node's candidates (cell, Kuhn tet) in qq_recovered's walk order, cells z, y, x ascending, then
// k -- compacted in LDS; lanes evaluate the pairs' tet values 64 at a time, then lane b forms node b's running
// mean over its pairs in order.  A half-grid node lies in at most 24 tets (a mesh vertex: 6 in each of the two
// cells whose diagonal it ends, 2 in each of the other six), so QQN_NODES * 24 pairs bound a batch.
constexpr int QQN_NODES = 8;
constexpr int QQN_MAXP = QQN_NODES * 24;
constexpr int QQN_BLOCKS = 8192;   // one wave each, striding over the batches (the count is on the device)

struct QQNodeShared {
    int64_t J[QQN_NODES][3];
    int64_t cell[QQN_MAXP];
    int32_t kd[QQN_MAXP];   // k | dd << 3
    double val[QQN_MAXP][3];
};

__global__ void __launch_bounds__(64) k_qq_node_gradient(QQGradArgs a) {
    __shared__ QQNodeShared sh;
    const int lane = threadIdx.x;
    const int64_t cnt = *a.count;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int64_t i0 = (int64_t)blockIdx.x * QQN_NODES; i0 < cnt; i0 += (int64_t)gridDim.x * QQN_NODES) {
        const int nn = cnt - i0 < QQN_NODES ? (int)(cnt - i0) : QQN_NODES;
        int64_t s = 0;
        if (lane < nn) {
            s = a.list[i0 + lane];
            int64_t J[3];
            p2_node_of(a.m, a.cut[s / 10], (int)(s % 10), J);
#pragma unroll
            for (int c = 0; c < 3; ++c) sh.J[lane][c] = J[c];
        }
        wave_sync_lds();
        int np = 0, start = 0;   // lane b <= nn: the first pair of node b
#pragma unroll 1
        for (int r = 0; r < QQN_NODES * 48 / 64; ++r) {
            const int cand = r * 64 + lane, b = cand / 48, q6 = cand - 48 * b;
            bool valid = false;
            int64_t cell = 0;
            int kd = 0;
            if (b < nn) {
                const int q = q6 / 6, k = q6 - 6 * q;
                const int64_t J[3] = {sh.J[b][0], sh.J[b][1], sh.J[b][2]};
                int64_t lo[3], hi[3], cc[3];
                qq_cell_range(a.m, J, lo, hi);
                bool ok = true;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    cc[d] = lo[d] + ((q >> d) & 1);
                    ok = ok && cc[d] <= hi[d];
                }
                if (ok) {
                    const int rel = (int)((J[2] - 2 * cc[2]) * 9 + (J[1] - 2 * cc[1]) * 3 + (J[0] - 2 * cc[0]));
                    const int dd = c_tab.p2_of27[k][rel];
                    if (dd >= 0) {
                        valid = true;
                        cell = (cc[2] * a.m.n[1] + cc[1]) * a.m.n[0] + cc[0];
                        kd = k | (dd << 3);
                    }
                }
            }
            const uint64_t bal = __ballot(valid);
            const int p = np + __popcll(bal & lt);
            if (valid && p < QQN_MAXP) {
                sh.cell[p] = cell;
                sh.kd[p] = kd;
            }
            np += __popcll(bal);
            const int lim = 48 * lane - 64 * r;   // this round's candidates of the nodes before node `lane`
            const uint64_t low = lim <= 0 ? 0ull : lim >= 64 ? ~0ull : ((1ull << lim) - 1ull);
            start += __popcll(bal & low);
        }
        if (np > QQN_MAXP) np = QQN_MAXP;
        wave_sync_lds();
#pragma unroll 1
        for (int p = lane; p < np; p += 64) {
            const int kd = sh.kd[p];
            double loc[3];
            qq_tet_node_gradient(a.m, a.phi, sh.cell[p], kd & 7, kd >> 3, loc);
#pragma unroll
            for (int c = 0; c < 3; ++c) sh.val[p][c] = loc[c];
        }
        wave_sync_lds();
        const int end = __shfl(start, (lane + 1) & 63, 64);
        if (lane < nn) {
            int n = 0;
            double avg[3] = {0., 0., 0.};
            for (int p = start; p < end && p < np; ++p) {
                const double v[3] = {sh.val[p][0], sh.val[p][1], sh.val[p][2]};
                qq_mean_step(n, avg, v);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) a.gh[3 * s + c] = avg[c];
        }
        wave_sync_lds();
    }
}

// averaging_P2_gradient_recovery over every half-grid node: out[3 j + c]
struct RecGradArgs {
    MeshP m;
    const double *phi;
    int64_t nnodes;
    double *out;
};

__global__ void __launch_bounds__(128) k_recovered_gradient(RecGradArgs a) {
    const int64_t j = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (j >= a.nnodes) return;
    const int64_t h1 = 2 * (int64_t)a.m.n[0] + 1, h2 = 2 * (int64_t)a.m.n[1] + 1;
    const int64_t J[3] = {j % h1, (j / h1) % h2, j / (h1 * h2)};
    double g[3];
    qq_recovered(a.m, a.phi, J, g);
#pragma unroll
    for (int c = 0; c < 3; ++c) a.out[3 * j + c] = g[c];
}

// mass and Laplace-Beltrami sums of pair (pi, pj) over points k0 .. k0 + 6, added to am, al in point order
__device__ __forceinline__ void qq_pair7(const QQShared &S, int k0, int pi, int pj, double &am, double &al) {
    double w[7], ad[7], hj[7], hi[7], gj[7][3], gi[7][3];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        w[k] = S.w[k0 + k];
        ad[k] = S.ad[k0 + k];
        hj[k] = S.H[k0 + k][pj];
        hi[k] = S.H[k0 + k][pi];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gj[k][c] = S.Gq[k0 + k][pj][c];
            gi[k][c] = S.Gq[k0 + k][pi][c];
        }
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        am += w[k] * ((ad[k] * hj[k]) * hi[k]);
        const double dq = gj[k][0] * gi[k][0] + gj[k][1] * gi[k][1] + gj[k][2] * gi[k][2];
        al += w[k] * (ad[k] * dq);
    }
}

constexpr int QQ_TPW = 4;   // cut tets per wave (16 lanes each: a tet has at most 14 quadrature points)

__global__ void __launch_bounds__(64) k_local_p2qq(LocalP2QQArgs a) {
    __shared__ QQShared sh[QQ_TPW];
    const int lane = threadIdx.x, t = lane >> 4, j = lane & 15;
    const int64_t c = (int64_t)blockIdx.x * QQ_TPW + t;
    const bool live = c < a.ncut;
    QQShared &s = sh[t];
    const int64_t N = a.ncut;
    double *L = a.loc;
    // ---- per tet: level set, the lattice-1 patch's sign pattern ----
    double X[4][3], v[10];
    int nf = 0, pid = 40;
    if (live) {
        double ls[10];
        load_tet_geom(a.m, a.phi, a.cut[c], X, ls);
#pragma unroll
        for (int d = 0; d < 10; ++d) v[d] = ls[c_tab.lat_of_p2[d]];
        int sg[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) sg[q] = sgn(v[q]);
        if (!sg[0] && !sg[1] && !sg[2] && !sg[3]) {
            nf = -1;
        } else {
            pid = sg[0] + 3 * sg[1] + 9 * sg[2] + 27 * sg[3] + 40;
            nf = c_tab.rp_size[pid];
        }
    }
    if (live && nf < 0 && j == 0) atomicOr(a.err, 2);
    if (live && nf == 0 && !a.pts) {
        for (int e = j; e < LOC2_HAT; e += 16) L[loc2(e, c)] = 0.;
        if (j == 0) a.flags[c] = 0;
    }
    const bool work = live && nf > 0;
    if (j == 0) s.nf = work ? nf : 0;
    if (work && j == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) s.X[q][cc] = X[q][cc];
        double T[3][3];
        const double det = qq_trafo(X, T);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) s.T[r][cc] = T[r][cc];
        s.h = cbrt(fabs(det));
        SmallQR<4, 4> q;
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) {
#pragma unroll
            for (int i = 0; i < 3; ++i) q.at(i, vv) = X[vv][i];
            q.at(3, vv) = 1.;
        }
        q.prepare();   // a tet of positive volume
#pragma unroll
        for (int e = 0; e < 16; ++e) s.w2b[e] = q.a[e];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            s.w2bd[e] = q.d[e];
            s.w2bb[e] = q.beta[e];
        }
    }
    if (work && j < 10) {
        s.v[j] = v[j];
        const int64_t f = a.first[10 * c + j];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) s.gh[j][cc] = a.gh[3 * f + cc];
    }
    bool bad = false;
    if (work && j < nf) {
        // facet j of the lattice-1 patch: zero vertices or edge cuts (1 - a) e_a + a e_b
        const int ea[6] = {0, 0, 1, 0, 1, 2}, eb[6] = {1, 2, 2, 3, 3, 3};
        double vb[3][4];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const int rep = c_tab.rp_tri[pid][j][jj];
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) vb[jj][cc] = 0.;
            if (rep < 4) {
                vb[jj][rep] = 1.;
            } else {
                const int ia = ea[rep - 4], ib = eb[rep - 4];
                const double alpha = v[ia] / (v[ia] - v[ib]);
#pragma unroll
                for (int cc = 0; cc < 4; ++cc)
                    vb[jj][cc] = (1.0 - alpha) * (cc == ia ? 1. : 0.) + alpha * (cc == ib ? 1. : 0.);
            }
        }
        double w[3][3];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) b2w(X, vb[jj], w[jj]);
        SmallQR<3, 2> q;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            q.at(r, 0) = w[1][r] - w[0][r];
            q.at(r, 1) = w[2][r] - w[0][r];
        }
        if (!q.prepare()) {
            bad = true;
        } else {
            double dd = 1.0;
            dd *= q.d[0];
            dd *= q.d[1];
            double af = c_tab.rp_bnd[pid] ? 0.5 : 1.;
            af *= fabs(dd);
            s.af[j] = af;
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
                double e3[3] = {0., 0., 0.};
                e3[jj] = 1.;
                q.apply_Q(e3);
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    if (jj < 2) s.U[j][r][jj] = e3[r];
                    else s.nrm[j][r] = e3[r];
                }
            }
#pragma unroll
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) s.vb[j][jj][cc] = vb[jj][cc];
        }
    }
    if (work) {
        double T[3][3];
        qq_trafo(X, T);   // the lane's own copy (lane 0 of the group may still be writing s.T)
        for (int e = j; e < 40; e += 16) {
            const int i = e / 4, k = e % 4;
            double g[3];
            qq_grad_at_vertex(T, i, k, g);
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) s.G[i][k][cc] = g[cc];
        }
    }
    if (__syncthreads_or(bad)) {
        if (lane == 0) atomicOr(a.err, 4);
        return;
    }
    // ---- one quadrature point per lane of the tet's group ----
    const int nq = work ? 7 * nf : 0;
    if (j < nq) {
        const int f = j / 7, qi = j - 7 * (j / 7);
        double nb[4];
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            double sv = s.vb[f][0][cc] * c_tab.q5n[qi][0];
            sv += s.vb[f][1][cc] * c_tab.q5n[qi][1];
            sv += s.vb[f][2][cc] * c_tab.q5n[qi][2];
            nb[cc] = sv;
        }
        s.w[j] = s.af[f] * c_tab.q5w[qi];
        double bxb[4], dh, dph[3][3], n[3];
        if (!qq_base_point(s, a.P, nb, bxb, dh) || !qq_jacobian(s, bxb, dh, dph, n)) {
            bad = true;
        } else {
            // |det| = sqrt(det Gram(dph U)) (ProjectedQuadDomain2DCL::assign, ifacetransp.cpp:850-892)
            double DU[3][2];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    double t2 = 0.0;
#pragma unroll
                    for (int i = 0; i < 3; ++i) t2 += dph[r][i] * s.U[f][i][cc];
                    DU[r][cc] = t2;
                }
            double g00 = 0.0, g10 = 0.0, g11 = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i) g00 += DU[i][0] * DU[i][0];
#pragma unroll
            for (int i = 0; i < 3; ++i) g10 += DU[i][1] * DU[i][0];
#pragma unroll
            for (int i = 0; i < 3; ++i) g11 += DU[i][1] * DU[i][1];
            s.ad[j] = sqrt(g00 * g11 - g10 * g10);
            // W = dph^T + nl (n/(nl.n) - dph nl), nl the first facet's normal (gradient_trafo,
            // ifacetransp.cpp:894-918)
            double nl[3] = {s.nrm[0][0], s.nrm[0][1], s.nrm[0][2]};
            double dnl[3], tv[3];
            qq_mv3(dph, nl, dnl);
            const double nln = nl[0] * n[0] + nl[1] * n[1] + nl[2] * n[2];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) tv[cc] = n[cc] / nln - dnl[cc];
            SmallQR<3, 3> W;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) W.at(r, cc) = dph[cc][r] + nl[r] * tv[cc];
            if (!W.prepare()) {
                bad = true;
            } else {
                double Wi[3][3];
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) {
                    double e[3] = {0., 0., 0.};
                    e[jj] = 1.;
                    W.solve(e);
#pragma unroll
                    for (int r = 0; r < 3; ++r) Wi[r][jj] = e[r];
                }
                const double nf3[3] = {s.nrm[f][0], s.nrm[f][1], s.nrm[f][2]};
#pragma unroll
                for (int i = 0; i < 10; ++i) {
                    double g[3], tt[3], r[3];
                    qq_p1v(s, i, nb, g);
                    const double d = nf3[0] * g[0] + nf3[1] * g[1] + nf3[2] * g[2];
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) tt[cc] = g[cc] - d * nf3[cc];
                    qq_mv3(Wi, tt, r);
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) s.Gq[j][i][cc] = r[cc];
                }
                double h[10];
                p2_basis(nb, h);
#pragma unroll
                for (int i = 0; i < 10; ++i) s.H[j][i] = h[i];
                double fv = 0.;
                if (a.pts) {
                    double y[3];
                    b2w(s.X, bxb, y);
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) a.pts[3 * ((int64_t)a.off[c] + j) + cc] = y[cc];
                } else if (a.rhs_kind == DHIP_RHS_VALUES) {
                    fv = a.vals[(int64_t)a.off[c] + j];
                } else if (a.rhs_kind != DHIP_RHS_NONE) {
                    double y[3];
                    b2w(s.X, bxb, y);
                    fv = rhs_fun(a.rhs_kind, a.rhs_p0, y);
                }
                s.f[j] = fv;
            }
        }
    }
    if (__syncthreads_or(bad)) {
        if (lane == 0) atomicOr(a.err, 4);
        return;
    }
    if (a.pts) return;   // points mode: the base points only
    // ---- contractions in the composite order: per tet 55 pairs (i >= j) and 10 load entries ----
    for (int task = lane; task < QQ_TPW * 65; task += 64) {
        const int tt = task / 65, r = task - 65 * (task / 65);
        const int64_t ct = (int64_t)blockIdx.x * QQ_TPW + tt;
        const QQShared &S = sh[tt];
        if (ct >= N || S.nf <= 0) continue;
        const int nqt = 7 * S.nf;
        if (r < 55) {
            int pi = 0, pj = 0;
            pair_of(r, pi, pj);
            double am = 0.0, al = 0.0;
            if (nqt % 7 == 0) {
                // a facet's seven points at a time: every LDS read of the chunk is issued before the
                // first sum, which still runs point by point in the composite order
                for (int k0 = 0; k0 < nqt; k0 += 7) qq_pair7(S, k0, pi, pj, am, al);
            } else {
                for (int k = 0; k < nqt; ++k) {
                    const double w = S.w[k], ad = S.ad[k];
                    am += w * ((ad * S.H[k][pj]) * S.H[k][pi]);
                    const double dq = S.Gq[k][pj][0] * S.Gq[k][pi][0] + S.Gq[k][pj][1] * S.Gq[k][pi][1] +
                                      S.Gq[k][pj][2] * S.Gq[k][pi][2];
                    al += w * (ad * dq);
                }
            }
            const int e1 = pi * 10 + pj, e2 = pj * 10 + pi;
            L[loc2(LOC2_MASS + e1, ct)] = am;
            L[loc2(LOC2_MASS + e2, ct)] = am;
            L[loc2(LOC2_LB + e1, ct)] = al;
            L[loc2(LOC2_LB + e2, ct)] = al;
            L[loc2(LOC2_STAB + e1, ct)] = 0.;
            L[loc2(LOC2_STAB + e2, ct)] = 0.;
        } else {
            const int i = r - 55;
            double rh = 0.0;
            if (a.rhs_kind != DHIP_RHS_NONE)
                for (int k = 0; k < nqt; ++k) rh += S.w[k] * ((S.ad[k] * S.f[k]) * S.H[k][i]);
            L[loc2(LOC2_RHS + i, ct)] = rh;
        }
    }
    if (work && j == 0) a.flags[c] = 1;
}

}  // namespace dhip
