// dhip_ch_p2.hpp -- the Cahn-Hilliard step system on the P2 interface space.
//
// Same step as the P1 stepper (CahnHilliardcGP1CL with theta = 1, resting interface), with the
// P2 basis: the mobility enters through its P2 interpolant sum_d Mob(c_d) h_d, so
//   LM_ij = sum_d Mob(c_d) T_d,ij,   T_d,ij = sum_q w_q h_d(x_q) (P grad h_i . P grad h_j)(x_q)
// over the composite Quad5 rule of the patch.  T (10 x 55 per cut tet) is computed once.
//
//   k_ch_geom_p2        once, one wave per cut tet: T, tet-major [ncut][10][55].
//   k_ch_contrib_*      once: for every CSR entry the list of (cut tet, local pair) that add to
//                       it, and for every row the list of (cut tet, local dof), both sorted by
//                       tet -- the order of the reference's builder +=.  Counting pass with
//                       atomics, scan, fill pass, then each short list is sorted.
//   k_ch_tet_p2         per step, one wave per cut tet: the local 10x10 mobility matrix from T
//                       and the nodal mobilities, and the local products M c, M f'(c).
//   k_ch_entries_p2     per step, one lane per entry: sums its contributions in list order and
//                       writes the A_ block value and the Precond3 diagonal.
//   k_ch_rhs_p2         per step, one lane per row: both right-hand sides.
#pragma once

#include "dhip_p2.hpp"
#include "dhip_ch.hpp"

namespace dhip {

constexpr int CHG2_N = 550;   // T: 10 nodal mobilities x 55 lower-triangular pairs
constexpr int CHT2_N = 130;   // 100 local mobility entries, 10 M c, 10 M f'(c), 10 M c_r (BDF2)
constexpr int CHT2_TD = 100, CHT2_WL = 110, CHT2_TR = 120;

struct ChGeomP2Args {
    MeshP m;
    const double *phi;
    const int64_t *cut;
    int64_t ncut;
    const uint8_t *flags;
    double *T;   // [ncut][10][55]
    int32_t *err;
};

__global__ void __launch_bounds__(64) k_ch_geom_p2(ChGeomP2Args a) {
    __shared__ P2Shared sh;
    const int64_t c = blockIdx.x;
    const int lane = threadIdx.x;
    if (c >= a.ncut) return;
    double *Tc = a.T + c * CHG2_N;
    if (!(a.flags[c] & 1)) {
        for (int e = lane; e < CHG2_N; e += 64) Tc[e] = 0.;
        return;
    }
    double X[4][3], ls[10];
    load_tet_geom(a.m, a.phi, a.cut[c], X, ls);
    double G1[4][3];
    p1_grad(X, G1);
    double v[10];
#pragma unroll
    for (int d = 0; d < 10; ++d) v[d] = ls[c_tab.lat_of_p2[d]];
    const int nf = p2_patch_facets(sh, X, ls, lane);
    if (nf <= 0) {  // the assembly of the same tets succeeded: cannot happen
        if (lane == 0) atomicOr(a.err, 4);
        return;
    }
    __syncthreads();
    int pi = 0, pj = 0;
    pair_of(lane < 55 ? lane : 0, pi, pj);
    double acc[10];
#pragma unroll
    for (int d = 0; d < 10; ++d) acc[d] = 0.;
    const int nq = 7 * nf;
    for (int base = 0; base < nq; base += P2_CHUNK) {
        p2_stage_points(sh, base, nq, lane, X, G1, v, DHIP_RHS_NONE, 0.0);
        __syncthreads();
        const int cnt = nq - base < P2_CHUNK ? nq - base : P2_CHUNK;
        if (lane < 55) {
            for (int k = 0; k < cnt; ++k) {
                const double g = sh.w[k] * dot3(sh.sg[k][pi], sh.sg[k][pj]);
#pragma unroll
                for (int d = 0; d < 10; ++d) acc[d] += sh.h[k][d] * g;
            }
        }
        __syncthreads();
    }
    if (lane < 55) {
#pragma unroll
        for (int d = 0; d < 10; ++d) Tc[d * 55 + lane] = acc[d];
    }
}

// ---------------------------------------------------------------------------
// contribution lists (built once)
// ---------------------------------------------------------------------------
struct ContribArgs {
    int64_t ncut;
    const uint8_t *flags;
    const int32_t *dofs;     // [ncut][10]
    const int32_t *rowptr, *col;
    int32_t *ecnt, *rcnt;    // counting pass: per entry / per row counts
    const int32_t *eptr, *rptr;
    int32_t *econt, *rcont;  // fill pass: keys tet*100 + pair, tet*10 + local dof
    int32_t *err;
};

__device__ inline int64_t find_entry(const int32_t *rowptr, const int32_t *col, int32_t r, int32_t cj) {
    int32_t lo = rowptr[r], hi = rowptr[r + 1];
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (col[mid] < cj) lo = mid + 1;
        else hi = mid;
    }
    return (lo < rowptr[r + 1] && col[lo] == cj) ? lo : -1;
}

template <bool FILL>
__global__ void k_ch_contrib(ContribArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 100 * a.ncut) return;
    const int64_t c = t / 100;
    const int p = (int)(t - 100 * c), li = p / 10, lj = p - 10 * (p / 10);
    if (!(a.flags[c] & 1)) return;
    const int32_t ri = a.dofs[c * 10 + li], cj = a.dofs[c * 10 + lj];
    if (ri < 0 || cj < 0) return;
    const int64_t e = find_entry(a.rowptr, a.col, ri, cj);
    if (e < 0) {
        atomicOr(a.err, 8);
        return;
    }
    if (!FILL) {
        atomicAdd(&a.ecnt[e], 1);
        if (li == lj) atomicAdd(&a.rcnt[ri], 1);
    } else {
        a.econt[a.eptr[e] + atomicAdd(&a.ecnt[e], 1)] = (int32_t)(c * 100 + p);
        if (li == lj) a.rcont[a.rptr[ri] + atomicAdd(&a.rcnt[ri], 1)] = (int32_t)(c * 10 + li);
    }
}

// sort every list ptr[i]..ptr[i+1] ascending (insertion sort: lists hold at most 24 keys)
__global__ void k_sort_lists(int64_t n, const int32_t *ptr, int32_t *data) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = ptr[i], e = ptr[i + 1];
    for (int32_t k = b + 1; k < e; ++k) {
        const int32_t x = data[k];
        int32_t j = k - 1;
        while (j >= b && data[j] > x) {
            data[j + 1] = data[j];
            --j;
        }
        data[j + 1] = x;
    }
}

__global__ void k_entry_rows(int64_t n, const int32_t *rowptr, int32_t *erow) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int32_t e = rowptr[r]; e < rowptr[r + 1]; ++e) erow[e] = (int32_t)r;
}

// ---------------------------------------------------------------------------
// per step
// ---------------------------------------------------------------------------
struct ChTetP2Args {
    int64_t ncut;
    const uint8_t *flags;
    const int32_t *dofs;     // [ncut][10]
    const double *massloc;   // P2 local mass, [ncut][100]
    const double *T;         // [ncut][10][55]
    const double *conc;      // c on the DOFs (the extrapolated c_e of a BDF2 step)
    const double *conc_r;    // BDF2: the c_r of the time-derivative right-hand side; nullptr otherwise
    double *out;             // [ncut][CHT2_N]
};

constexpr int CHTET2_WAVES = 4;

__global__ void __launch_bounds__(64 * CHTET2_WAVES) k_ch_tet_p2(ChTetP2Args a) {
    const int lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * CHTET2_WAVES + (threadIdx.x >> 6);
    if (c >= a.ncut) return;
    double *O = a.out + c * CHT2_N;
    if (!(a.flags[c] & 1)) {
        for (int e = lane; e < CHT2_N; e += 64) O[e] = 0.;
        return;
    }
    double xc[10], mv[10];
    int32_t nb[10];
#pragma unroll
    for (int d = 0; d < 10; ++d) {
        nb[d] = a.dofs[c * 10 + d];
        xc[d] = nb[d] >= 0 ? a.conc[nb[d]] : 0.;
        mv[d] = mobility_fun(xc[d]);
    }
    if (lane < 55) {
        int pi, pj;
        pair_of(lane, pi, pj);
        const double *Tc = a.T + c * CHG2_N;
        double lm = 0.;
#pragma unroll
        for (int d = 0; d < 10; ++d) lm += mv[d] * Tc[d * 55 + lane];
        O[pi * 10 + pj] = lm;
        O[pj * 10 + pi] = lm;
    }
    // LocalMatVecP1CL-style products M c and M f'(c) with the P2 local mass, lane i < 10; a BDF2
    // step adds M c_r
    const int i = lane;
    if (i < 10) {
        const bool bdf = a.conc_r != nullptr;
        double vt = 0., vw = 0., vr = 0.;
        if (nb[i] >= 0) {
#pragma unroll
            for (int j = 0; j < 10; ++j) {
                if (nb[j] < 0) continue;
                const double mij = a.massloc[(int64_t)c * 100 + i * 10 + j];
                vt += mij * xc[j];
                vw += mij * potential_prime(xc[j]);
                if (bdf) vr += mij * a.conc_r[nb[j]];
            }
        }
        O[CHT2_TD + i] = vt;
        O[CHT2_WL + i] = vw;
        if (bdf) O[CHT2_TR + i] = vr;
    }
}

struct ChEntryP2Args {
    int64_t nnz;
    const int32_t *erow, *rowptr, *col;
    const int32_t *eptr, *econt;
    const double *tet;       // k_ch_tet_p2 output
    const double *mass, *stab;
    double *bval, *pcdiag;
    ChCoef k;
};

__global__ void k_ch_entries_p2(ChEntryP2Args a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nnz) return;
    const int32_t r = a.erow[e];
    double lm = 0.0;
    for (int32_t q = a.eptr[e]; q < a.eptr[e + 1]; ++q) {
        const int32_t key = a.econt[q];
        const int32_t c = key / 100;
        lm += a.tet[(int64_t)c * CHT2_N + (key - 100 * c)];
    }
    const double st = a.stab[e];
    // A_.LinComb(sigma dt, LaplaceM, rho dt, Volume_stab); the top block row r starts at 2 rowptr[r]
    a.bval[e + a.rowptr[r]] = a.k.sdt * lm + a.k.rdt * st;
    // Precond3.LinComb(1, Mass, sigma dt, LaplaceM, dt rho, Volume_stab)
    if (a.col[e] == r) a.pcdiag[r] = 1.0 * (1.0 * a.mass[e] + a.k.sdt * lm) + a.k.dtr * st;
}

struct ChRhsP2Args {
    int64_t ndof;
    const int32_t *rptr, *rcont;
    const double *tet;
    double *rhs;
    ChCoef k;
    int32_t bdf;             // rhs3 from the M c_r products of k_ch_tet_p2
};

__global__ void k_ch_rhs_p2(ChRhsP2Args a) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.ndof) return;
    double td = 0.0, wl = 0.0, tr = 0.0;
    for (int32_t q = a.rptr[r]; q < a.rptr[r + 1]; ++q) {
        const int32_t key = a.rcont[q];
        const int32_t c = key / 10;
        const int li = key - 10 * c;
        td += a.tet[(int64_t)c * CHT2_N + CHT2_TD + li];
        wl += a.tet[(int64_t)c * CHT2_N + CHT2_WL + li];
        if (a.bdf) tr += a.tet[(int64_t)c * CHT2_N + CHT2_TR + li];
    }
    // rhs3 = theta (timeder + dt load3), rhs4 = -well + S timeder + load4 (no sources); a BDF2
    // step takes the time-derivative part of rhs3 from c_r
    a.rhs[r] = 1.0 * ((a.bdf ? tr : td) + a.k.dt * 0.0);
    a.rhs[a.ndof + r] = ((-1.0) * wl + a.k.S * td) + 0.0;
}

}  // namespace dhip
