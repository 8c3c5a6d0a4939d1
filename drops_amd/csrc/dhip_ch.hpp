// dhip_ch.hpp -- one implicit-Euler step of the P1 trace-FEM Cahn-Hilliard system on gfx950.
//
// CahnHilliardcGP1CL (surfactant/ifacetransp.cpp:1526-1869) with theta = 1 on a resting
// interface (no wind: the convection and mass-divergence matrices vanish) and no sources.
// Unknowns [mu; c] on the P1 interface space of the context's assembly; per step
//   A_ = sigma dt LM(c_old) + rho dt C         B_ = M
//   C_ = -eps M                               D_ = eps^2 L + S M + rho eps^2 C
//   rhs3 = M c_old                            rhs4 = -M f'(c_old) + S M c_old
// LM = LocalLaplaceMobilityP1CL (surfactant/ifacetransp.h:1731-1766), the local products
// M c and M f'(c) are LocalMatVecP1CL accumulations (:800-828).  The block matrix has the
// layout of the SparseMatBaseCL(A, B, C, D) constructor (num/spmat.h:719-751).
//
//   k_ch_geom   once per system, one lane per cut tet: patch, Quad5 composite rule, projected
//               gradients at every quadrature point -> the 40 geometric coefficients of the
//               mobility matrix (linear in the vertex mobilities).
//   k_ch_tet    per step, one lane per cut tet: the 4x4 local mobility matrix from the vertex
//               concentrations and the local products M c, M f'(c).  SoA output.
//   k_ch_rows   one lane per DOF row: sums the local rows of the incident cut tets in
//               ascending tet id (the order of the reference's builder +=), writes the A_
//               block, the Precond3 diagonal and both right-hand sides.
//   k_ch_block  once per system: block CSR pattern, the constant B_, C_, D_ blocks and the
//               Precond4 diagonal.
#pragma once

#include "dhip_assembly.hpp"
#include "dhip_krylov.hpp"

namespace dhip {

// Mobility_function, surfactant/ifacetransp.cpp:3334-3339
__device__ __forceinline__ double mobility_fun(double x) {
    const double v = (1. - x) * x;
    return v > 0. ? v : 0.;
}

// Potential_prime_function, surfactant/ifacetransp.cpp:3360-3364
__device__ __forceinline__ double potential_prime(double x) { return (x - 1.) * x * (x - 0.5); }

constexpr int CHT_N = 28;  // 16 mobility, 4 M c, 4 M f'(c), 4 M c_r (BDF2 steps only)
constexpr int CHT_MOB = 0, CHT_TD = 16, CHT_WL = 20, CHT_TR = 24;
constexpr int CHG_N = 40;  // T[k][ij]: 4 vertices x 10 lower-triangular (i, j) pairs

struct ChGeomArgs {
    MeshP m;
    const double *phi;
    const int64_t *cut;
    int64_t ncut;
    const uint8_t *flags;
    double *T;               // SoA [CHG_N][ncut]
};

// Geometry of LocalLaplaceMobilityP1CL (surfactant/ifacetransp.h:1731-1766), computed once: the
// mobility enters through its P1 interpolant sum_k Mob(c_k) lambda_k, so
//   LM_ij = sum_k Mob(c_k) T_k,ij,   T_k,ij = sum_q w_q lambda_k(x_q) (P grad phi_i . P grad phi_j)(x_q)
// over the Quad5 composite rule of the patch, P = I - n n^T with n the normalised gradient of the
// P2 level set at x_q.  Per step only the contraction with the four vertex mobilities remains.
__global__ void __launch_bounds__(LOCAL_BLOCK) k_ch_geom(ChGeomArgs a) {
    const int64_t c = (int64_t)blockIdx.x * LOCAL_BLOCK + threadIdx.x;
    if (c >= a.ncut) return;
    const int64_t N = a.ncut;
    double acc[CHG_N];
#pragma unroll
    for (int o = 0; o < CHG_N; ++o) acc[o] = 0.;
    if (a.flags[c] & 1) {
        double X[4][3], ls[10];
        load_tet_geom(a.m, a.phi, a.cut[c], X, ls);
        double G1[4][3];
        p1_grad(X, G1);
        double v[10];
#pragma unroll
        for (int d = 0; d < 10; ++d) v[d] = ls[c_tab.lat_of_p2[d]];
        for (int lt = 0; lt < 8; ++lt) {
            int sg[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) sg[q] = sgn(ls[c_tab.lat_t[lt][q]]);
            const int pid = sg[0] + 3 * sg[1] + 9 * sg[2] + 27 * sg[3] + 40;
            const int ntri = c_tab.rp_size[pid];
            const bool bnd = c_tab.rp_bnd[pid] != 0;
            for (int tr = 0; tr < ntri; ++tr) {
                double vb[3][4];
                facet_vertices(ls, lt, pid, tr, vb);
                double absdet, nrm[3];
                facet_qr(X, vb, absdet, nrm);
                double af = bnd ? 0.5 : 1.;
                af *= absdet;
                for (int qn = 0; qn < 7; ++qn) {
                    double b[4];
#pragma unroll
                    for (int cc = 0; cc < 4; ++cc) {
                        double sv = vb[0][cc] * c_tab.q5n[qn][0];
                        sv += vb[1][cc] * c_tab.q5n[qn][1];
                        sv += vb[2][cc] * c_tab.q5n[qn][2];
                        b[cc] = sv;
                    }
                    const double wq = af * c_tab.q5w[qn];
                    double n3[3];
                    p2_unit_normal(v, b, G1, n3);
                    double q[4][3];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const double d = dot3(G1[i], n3);
#pragma unroll
                        for (int cc = 0; cc < 3; ++cc) q[i][cc] = G1[i][cc] - d * n3[cc];
                    }
                    int o = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j <= i; ++j, ++o) {
                            const double g = wq * dot3(q[i], q[j]);
#pragma unroll
                            for (int k = 0; k < 4; ++k) acc[k * 10 + o] += b[k] * g;
                        }
                }
            }
        }
    }
#pragma unroll
    for (int o = 0; o < CHG_N; ++o) a.T[(int64_t)o * N + c] = acc[o];
}

struct ChTetArgs {
    int64_t ncut;
    const uint8_t *flags;
    const int32_t *dofs;     // [ncut][4], -1 where the vertex has no DOF
    const double *loc;       // local mass (alpha = 1) from k_local_p1
    const double *T;         // k_ch_geom
    const double *conc;      // c on the DOFs (the extrapolated c_e of a BDF2 step)
    const double *conc_r;    // BDF2: the c_r of the time-derivative right-hand side; nullptr otherwise
    double *out;             // SoA [CHT_N][ncut]
};

// Per step: the local mobility matrix from the vertex concentrations and the local products
// M c, M f'(c) (LocalMatVecP1CL<LocalInterfaceMassP1CL>, surfactant/ifacetransp.h:800-828);
// a BDF2 step adds M c_r.
__global__ void k_ch_tet(ChTetArgs a) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.ncut) return;
    const int64_t N = a.ncut;
    double *O = a.out;
    if (!(a.flags[c] & 1)) {
        for (int e = 0; e < CHT_N; ++e) O[e * N + c] = 0.;
        return;
    }
    const bool bdf = a.conc_r != nullptr;
    double xc[4], mv[4], xr[4];
    int32_t nb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        nb[i] = a.dofs[c * 4 + i];
        xc[i] = nb[i] >= 0 ? a.conc[nb[i]] : 0.;
        xr[i] = nb[i] >= 0 && bdf ? a.conc_r[nb[i]] : 0.;
        mv[i] = mobility_fun(xc[i]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double vt = 0., vw = 0., vr = 0.;
        if (nb[i] >= 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (nb[j] < 0) continue;
                const double mij = a.loc[(int64_t)(LOC_MASS + i * 4 + j) * N + c];
                vt += mij * xc[j];
                vw += mij * potential_prime(xc[j]);
                vr += mij * xr[j];
            }
        }
        O[(CHT_TD + i) * N + c] = vt;
        O[(CHT_WL + i) * N + c] = vw;
        if (bdf) O[(CHT_TR + i) * N + c] = vr;
    }
    int o = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j, ++o) {
            double lm = 0.;
#pragma unroll
            for (int k = 0; k < 4; ++k) lm += mv[k] * a.T[(int64_t)(k * 10 + o) * N + c];
            O[(CHT_MOB + i * 4 + j) * N + c] = lm;
            O[(CHT_MOB + j * 4 + i) * N + c] = lm;
        }
}

struct ChCoef {
    double sdt;        // sigma * dt
    double rdt;        // rho * dt
    double dtr;        // dt * rho
    double dt;
    double S;
    double e2;         // eps * eps
    double re2;        // rho * eps * eps
    double e2r;        // eps * eps * rho
    double meps;       // -eps
};

struct ChRowArgs {
    MeshP m;
    const int64_t *cut;
    const uint32_t *cellinfo;
    const uint8_t *flags;
    const int32_t *dofs;
    const int32_t *rowowner;
    int64_t ndof, ncut;
    const double *tet;      // k_ch_tet output
    const int32_t *rowptr;  // P1 pattern
    const double *mass, *stab;
    double *bval;           // block values
    double *pcdiag;         // [2 ndof]
    double *rhs;            // [2 ndof]
    ChCoef k;
    int32_t bdf;            // rhs3 from the M c_r products of k_ch_tet
};

__global__ void k_ch_rows(ChRowArgs a) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.ndof) return;
    int64_t x1, x2, x3;
    slot_vertex(a.m, a.cut, a.rowowner[r], x1, x2, x3);
    uint32_t present = 0;
    int32_t cols[15];
    double lm[15];
#pragma unroll
    for (int s = 0; s < 15; ++s) {
        cols[s] = 0;
        lm[s] = 0.0;
    }
    double td = 0.0, wl = 0.0, tr = 0.0;
    const int64_t N = a.ncut;
    for_incident_cut_tets(a.m, a.cellinfo, x1, x2, x3, [&](int64_t ci, int k, int li) {
        if (!(a.flags[ci] & 1)) return false;
        for (int j = 0; j < 4; ++j) {
            const int32_t dj = a.dofs[ci * 4 + j];
            if (dj < 0) continue;
            const int sl = c_tab.slot[k][li][j];
            const double v = a.tet[(int64_t)(CHT_MOB + li * 4 + j) * N + ci];
#pragma unroll
            for (int s = 0; s < 15; ++s)
                if (s == sl) {
                    cols[s] = dj;
                    lm[s] += v;
                }
            present |= 1u << sl;
        }
        td += a.tet[(int64_t)(CHT_TD + li) * N + ci];
        wl += a.tet[(int64_t)(CHT_WL + li) * N + ci];
        if (a.bdf) tr += a.tet[(int64_t)(CHT_TR + li) * N + ci];
        return false;
    });
    const int64_t e0 = a.rowptr[r];
#pragma unroll
    for (int s = 0; s < 15; ++s) {
        if (!((present >> s) & 1u)) continue;
        int rank = 0;
#pragma unroll
        for (int u = 0; u < 15; ++u)
            rank += (((present >> u) & 1u) && cols[u] < cols[s]) ? 1 : 0;
        const int64_t e = e0 + rank;
        const double st = a.stab[e];
        // A_.LinComb(sigma dt, LaplaceM, rho dt, Volume_stab)
        a.bval[2 * e0 + rank] = a.k.sdt * lm[s] + a.k.rdt * st;
        // Precond3.LinComb(1, Mass, sigma dt, LaplaceM, dt rho, Volume_stab)
        if (cols[s] == r) a.pcdiag[r] = 1.0 * (1.0 * a.mass[e] + a.k.sdt * lm[s]) + a.k.dtr * st;
    }
    // rhs3 = theta (timeder + dt load3), rhs4 = -well + S timeder + load4 (no sources); a BDF2
    // step takes the time-derivative part of rhs3 from c_r
    a.rhs[r] = 1.0 * ((a.bdf ? tr : td) + a.k.dt * 0.0);
    a.rhs[a.ndof + r] = ((-1.0) * wl + a.k.S * td) + 0.0;
}

struct ChBlockArgs {
    int64_t ndof, nnz;
    const int32_t *rowptr, *col;
    const double *mass, *lb, *stab;
    int32_t *brp, *bcol;
    double *bval, *pcdiag;
    ChCoef k;
};

__global__ void k_ch_block(ChBlockArgs a) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > a.ndof) return;
    const int64_t n = a.ndof, nnz = a.nnz;
    if (r == n) {
        a.brp[2 * n] = (int32_t)(4 * nnz);
        return;
    }
    const int64_t e0 = a.rowptr[r], len = a.rowptr[r + 1] - e0;
    const int64_t top = 2 * e0, bot = 2 * nnz + 2 * e0;
    a.brp[r] = (int32_t)top;
    a.brp[n + r] = (int32_t)bot;
    for (int64_t k = 0; k < len; ++k) {
        const int64_t e = e0 + k;
        const int32_t cj = a.col[e];
        const double m = a.mass[e], l = a.lb[e], s = a.stab[e];
        a.bcol[top + k] = cj;                               // A_ (values per step)
        a.bcol[top + len + k] = (int32_t)(cj + n);          // B_.LinComb(1, M, 0 dt, Massd, dt, Conv)
        a.bval[top + len + k] = 1.0 * (1.0 * m + 0.0 * 0.0) + a.k.dt * 0.0;
        a.bcol[bot + k] = cj;                               // C_.LinComb(0, L, -eps, M)
        a.bval[bot + k] = 0.0 * l + a.k.meps * m;
        a.bcol[bot + len + k] = (int32_t)(cj + n);          // D_.LinComb(eps^2, L, S, M, rho eps^2, C)
        a.bval[bot + len + k] = 1.0 * (a.k.e2 * l + a.k.S * m) + a.k.re2 * s;
        // Precond4.LinComb(1, Mass, eps^2, Laplace, eps^2 rho, Volume_stab)
        if (cj == r) a.pcdiag[n + r] = 1.0 * (1.0 * m + a.k.e2 * l) + a.k.e2r * s;
    }
}

// BDF2 concentrations from c_n and c_{n-1} (surfphasesep/surfphasesep.cpp:451-456): the extrapolation
// c_e = w1 c_n + w2 c_{n-1} and the time-derivative vector c_r = w3 c_n + w4 c_{n-1} of the system scaled
// by dt_eff = 1/alpha.
__global__ void __launch_bounds__(KB) k_ch_bdf_combine(int64_t n, const double *cn, const double *cp, double w1,
                                                       double w2, double w3, double w4, double *ce, double *cr) {
    const int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (i >= n) return;
    const double a = cn[i], b = cp[i];
    ce[i] = w1 * a + w2 * b;
    cr[i] = w3 * a + w4 * b;
}

// Parts of the step error e^2 = (d, M d) / (c2, M c2), d = c2 - c1 (surfphasesep/surfphasesep.cpp:
// 477-479): per block the sums over its rows of d_r (M d)_r and c2_r (M c2)_r; k_reduce sums the parts.
__global__ void __launch_bounds__(KB) k_ch_mnorm(int64_t n, const int32_t *rp, const int32_t *col, const double *m,
                                                 const double *c1, const double *c2, double *part_d,
                                                 double *part_c) {
    double sd = 0., sc = 0.;
    for (int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x; r < n; r += (int64_t)gridDim.x * KB) {
        double md = 0., mc = 0.;
        for (int32_t q = rp[r]; q < rp[r + 1]; ++q) {
            const int32_t j = col[q];
            const double x2 = c2[j];
            md += m[q] * (x2 - c1[j]);
            mc += m[q] * x2;
        }
        const double x2 = c2[r];
        sd += (x2 - c1[r]) * md;
        sc += x2 * mc;
    }
    sd = block_sum<KB>(sd);
    sc = block_sum<KB>(sc);
    if (threadIdx.x == 0) {
        part_d[blockIdx.x] = sd;
        part_c[blockIdx.x] = sc;
    }
}

// |r_k|^2 of three residuals r_k = b - K x_k of candidate starts (each from k_spmv<true>: a row's
// products summed in column order) and |b|^2, per block: thread t of block g takes rows
// g * RES3_B + t + k * gridDim.x * RES3_B (block-order sums later)
constexpr int RES3_B = 256;
__global__ void __launch_bounds__(RES3_B) k_ch_three_norms(int64_t n2, const double *b, const double *r0,
                                                           const double *r1, const double *r2, double *part) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, sb = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RES3_B + threadIdx.x; i < n2; i += (int64_t)gridDim.x * RES3_B) {
        const double v0 = r0[i], v1 = r1[i], v2 = r2[i], bi = b[i];
        s0 += v0 * v0;
        s1 += v1 * v1;
        s2 += v2 * v2;
        sb += bi * bi;
    }
    s0 = block_sum<RES3_B>(s0);
    s1 = block_sum<RES3_B>(s1);
    s2 = block_sum<RES3_B>(s2);
    sb = block_sum<RES3_B>(sb);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = s0;
        part[4 * blockIdx.x + 1] = s1;
        part[4 * blockIdx.x + 2] = s2;
        part[4 * blockIdx.x + 3] = sb;
    }
}

// sum of the `stride`-strided partials k of nb blocks by one wave (lane-strided loads, then the
// wave's butterfly sum: the same order in every call)
__device__ inline double wave_sum_parts(const double *part, int nb, int stride, int k) {
    double v = 0.0;
    for (int b = threadIdx.x & 63; b < nb; b += 64) v += part[stride * b + k];
    return wave_sum(v);
}

// Candidates 0 = x_n, 1 = network, 2 = linear extrapolation.  The smaller residual of x_n and the
// extrapolation wins; the network's candidate replaces it when its residual is smaller still and
// the extrapolation is far from converged (residual above gate * |b|, the transient), where the
// residual is dominated by the change of the step and not by the solver's tolerance-level error.
__global__ void k_ch_pick(const double *part, int nb, double gate, int32_t *flag) {
    double s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = wave_sum_parts(part, nb, 4, q);
    if (threadIdx.x != 0) return;
    int best = s[2] < s[0] ? 2 : 0;
    if (s[1] < s[best] && s[2] > gate * gate * s[3]) best = 1;
    *flag = best;
}

// x (holding candidate 1) becomes candidate `flag`
__global__ void k_ch_keep(int64_t n2, const int32_t *flag, const double *x0, const double *x2, double *x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    const int f = *flag;
    if (f == 0) x[i] = x0[i];
    else if (f == 2) x[i] = x2[i];
}

}  // namespace dhip
