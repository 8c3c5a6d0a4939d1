// dhip_krylov.hpp -- CSR SpMV, dot/axpy and the preconditioned Krylov solvers.
//
// Single-problem PCG is latency bound (a 30k-row system moves ~9 MB per
// iteration, ~1 us at HBM speed), so the whole PCG loop runs in ONE
// cooperative persistent kernel: each block owns a contiguous row range, the
// p-update of iteration i is fused into the SpMV of iteration i+1 (p is double
// buffered), and an iteration needs two grid barriers.  Grid reductions are
// deterministic: every block sums the per-block partials in the same order,
// so every block takes the same convergence decision.
#pragma once

#include "dhip_common.hpp"

namespace dhip {

constexpr int KB = 256;       // threads per block of the vector kernels
constexpr int SPMV_G = 8;     // lanes per CSR row (P1 trace-FEM rows hold ~11 entries)

// ---------------------------------------------------------------------------
// plain kernels
// ---------------------------------------------------------------------------
template <int G, class VT, class XT>
__device__ __forceinline__ double row_dot(const int32_t *rp, const int32_t *col, const VT *val, const XT *x,
                                          int64_t r, int lane) {
    double s = 0.0;
    const int32_t e = rp[r + 1];
    for (int32_t k = rp[r] + lane; k < e; k += G) s += (double)val[k] * (double)x[col[k]];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

__global__ void __launch_bounds__(KB) k_spmv(int64_t n, const int32_t *rp, const int32_t *col,
                                             const double *val, const double *x, double *y) {
    const int64_t gid = (int64_t)blockIdx.x * KB + threadIdx.x;
    const int64_t r = gid / SPMV_G;
    const int lane = (int)(gid % SPMV_G);
    double s = 0.0;
    if (r < n) {
        const int32_t e = rp[r + 1];
        for (int32_t k = rp[r] + lane; k < e; k += SPMV_G) s += val[k] * x[col[k]];
    }
#pragma unroll
    for (int o = SPMV_G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (r < n && lane == 0) y[r] = s;
}

__global__ void __launch_bounds__(KB) k_dot_partial(int64_t n, const double *x, const double *y, double *part) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x; i < n; i += (int64_t)gridDim.x * KB)
        s += x[i] * y[i];
    s = block_sum<KB>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(KB) k_reduce(const double *part, int np, double *out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += KB) s += part[i];
    s = block_sum<KB>(s);
    if (threadIdx.x == 0) *out = s;
}

__global__ void __launch_bounds__(KB) k_axpy(int64_t n, double a, const double *x, double *y) {
    const int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// diagonal located by a row scan (SolveGSstep<PB_JAC0>, num/precond.h:163-168)
__global__ void __launch_bounds__(KB) k_diag(int64_t n, const int32_t *rp, const int32_t *col, const double *val,
                                             double *diag, float *diag32) {
    const int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (i >= n) return;
    int32_t nz = rp[i];
    const int32_t e = rp[i + 1];
    while (nz < e && col[nz] != i) ++nz;
    const double d = nz < e ? val[nz] : 0.0;
    diag[i] = d;
    if (diag32) diag32[i] = (float)d;
}

__global__ void __launch_bounds__(KB) k_to_f32(int64_t n, const double *a, float *b) {
    const int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_zero_u32(uint32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// grid barrier for cooperative launches (all blocks co-resident)
// ---------------------------------------------------------------------------
struct GridBar {
    uint32_t count;
    uint32_t gen;
    uint32_t timeout;
    uint32_t pad;
};

constexpr uint32_t SPIN_LIMIT = 1u << 22;  // ~seconds; a healthy barrier takes microseconds

// Returns false (in every thread of the block) if the barrier timed out.
__device__ inline bool grid_sync(GridBar *bar, uint32_t &gen) {
    __shared__ int ok_s;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const uint32_t g = gen;
        const uint32_t arrived = atomicAdd(&bar->count, 1u);
        if (arrived == gridDim.x - 1) {
            atomicExch(&bar->count, 0u);
            __threadfence();
            __hip_atomic_fetch_add(&bar->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (++spins > SPIN_LIMIT) {
                    ok = 0;
                    atomicOr(&bar->timeout, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __threadfence();
        gen = g + 1;
        ok_s = ok;
    }
    __syncthreads();
    return ok_s != 0;
}

// Sum of the per-block partials part[0..gridDim.x) in a fixed order; the same
// value in every block and thread.
__device__ inline double grid_total(const double *part) {
    __shared__ double tot_s;
    if (threadIdx.x < 64) {
        double s = 0.0;
        for (int i = threadIdx.x; i < (int)gridDim.x; i += 64) s += part[i];
        s = wave_sum(s);
        if (threadIdx.x == 0) tot_s = s;
    }
    __syncthreads();
    const double t = tot_s;
    __syncthreads();
    return t;
}

// ---------------------------------------------------------------------------
// persistent PCG (FP64), num/krylovsolver.h:107-196 semantics
// ---------------------------------------------------------------------------
struct PcgArgs {
    int64_t n;
    const int32_t *rp, *col;
    const double *val, *diag;   // diag == nullptr: no preconditioner
    const double *b;
    double *x, *r, *z, *p0, *p1, *q;
    double *part;               // [4][gridDim.x]
    GridBar *bar;
    int max_iter;
    double tol;
    int rel;
    int *out_iter;
    double *out_resid;
    int *out_status;            // 0 converged, 1 not converged, 2 barrier timeout
};

__global__ void __launch_bounds__(KB) k_pcg64(PcgArgs a) {
    const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * per;
    const int64_t r1 = r0 + per < a.n ? r0 + per : a.n;
    const int NB = gridDim.x;
    double *P0 = a.part, *P1 = a.part + NB, *P2 = a.part + 2 * NB, *P3 = a.part + 3 * NB;
    const int lane = threadIdx.x % SPMV_G, grp = threadIdx.x / SPMV_G;
    constexpr int GPB = KB / SPMV_G;
    uint32_t gen = 0;
    int status = 1, it_done = a.max_iter;
    double resid = 0.0;

    // r = b - A x ; |r|, |b|
    double rr = 0.0, bb = 0.0;
    for (int64_t r = r0 + grp; r < r1; r += GPB) {
        const double s = row_dot<SPMV_G>(a.rp, a.col, a.val, a.x, r, lane);
        if (lane == 0) {
            const double rv = a.b[r] - s;
            a.r[r] = rv;
            rr += rv * rv;
            bb += a.b[r] * a.b[r];
        }
    }
    rr = block_sum<KB>(rr);
    bb = block_sum<KB>(bb);
    if (threadIdx.x == 0) {
        P1[blockIdx.x] = rr;
        P3[blockIdx.x] = bb;
    }
    if (!grid_sync(a.bar, gen)) { status = 2; goto done; }
    {
        resid = sqrt(grid_total(P1));
        double normb = sqrt(grid_total(P3));
        if (normb == 0.0 || !a.rel) normb = 1.0;
        resid = resid / normb;
        if (resid <= a.tol) {
            status = 0;
            it_done = 0;
            goto done;
        }
        // z = M r, rho = z.r, p_old = 0
        double zr = 0.0;
        for (int64_t r = r0 + threadIdx.x; r < r1; r += KB) {
            const double rv = a.r[r];
            const double zv = a.diag ? 1.0 * rv / a.diag[r] : rv;
            a.z[r] = zv;
            a.p0[r] = 0.0;
            zr += zv * rv;
        }
        zr = block_sum<KB>(zr);
        if (threadIdx.x == 0) P2[blockIdx.x] = zr;
        if (!grid_sync(a.bar, gen)) { status = 2; goto done; }
        double rho = grid_total(P2), beta = 0.0;

        for (int it = 1; it <= a.max_iter; ++it) {
            double *pn = (it & 1) ? a.p1 : a.p0;
            const double *po = (it & 1) ? a.p0 : a.p1;
            // q = A p with p = z + beta p_old formed on the fly; lambda = q.p
            double lam = 0.0;
            for (int64_t r = r0 + grp; r < r1; r += GPB) {
                double s = 0.0;
                const int32_t e = a.rp[r + 1];
                for (int32_t k = a.rp[r] + lane; k < e; k += SPMV_G) {
                    const int32_t cc = a.col[k];
                    s += a.val[k] * (a.z[cc] + beta * po[cc]);
                }
#pragma unroll
                for (int o = SPMV_G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
                if (lane == 0) {
                    const double pv = a.z[r] + beta * po[r];
                    pn[r] = pv;
                    a.q[r] = s;
                    lam += s * pv;
                }
            }
            lam = block_sum<KB>(lam);
            if (threadIdx.x == 0) P0[blockIdx.x] = lam;
            if (!grid_sync(a.bar, gen)) { status = 2; goto done; }
            const double alpha = rho / grid_total(P0);
            double rr2 = 0.0, zr2 = 0.0;
            for (int64_t r = r0 + threadIdx.x; r < r1; r += KB) {
                const double pv = pn[r];
                a.x[r] += alpha * pv;
                double rv = a.r[r];
                rv += (-alpha) * a.q[r];
                a.r[r] = rv;
                rr2 += rv * rv;
                const double zv = a.diag ? 1.0 * rv / a.diag[r] : rv;
                a.z[r] = zv;
                zr2 += zv * rv;
            }
            rr2 = block_sum<KB>(rr2);
            zr2 = block_sum<KB>(zr2);
            if (threadIdx.x == 0) {
                P1[blockIdx.x] = rr2;
                P2[blockIdx.x] = zr2;
            }
            if (!grid_sync(a.bar, gen)) { status = 2; goto done; }
            resid = sqrt(grid_total(P1)) / normb;
            if (resid <= a.tol) {
                status = 0;
                it_done = it;
                break;
            }
            const double rho1 = rho;
            rho = grid_total(P2);
            beta = rho / rho1;
        }
    }
done:
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.out_iter = it_done;
        *a.out_resid = resid;
        *a.out_status = status;
    }
}

// ---------------------------------------------------------------------------
// persistent PCG with FP32 inner iterations and FP64 iterative refinement
//   outer (FP64): r = b - A x (FP64 values), resid = |r|/|b|, stop at tol
//   inner (FP32 storage, FP64 dot accumulation): PCG on A d = r / |r| from d = 0
//         until |r_in| <= inner_tol or inner_max iterations
//   update (FP64): x += |r| d
// ---------------------------------------------------------------------------
struct PcgMixArgs {
    int64_t n;
    const int32_t *rp, *col;
    const double *val;          // FP64 values (residual)
    const float *val32, *diag32;
    const double *b;
    double *x, *r;              // FP64
    float *d, *rin, *z, *p0, *p1, *q;
    double *part;
    GridBar *bar;
    int max_iter;               // total inner iterations allowed
    double tol;
    int rel;
    double inner_tol;
    int inner_max;
    int *out_iter;
    double *out_resid;
    int *out_status;
    int *out_outer;
};

__global__ void __launch_bounds__(KB) k_pcg_mixed(PcgMixArgs a) {
    const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * per;
    const int64_t r1 = r0 + per < a.n ? r0 + per : a.n;
    const int NB = gridDim.x;
    double *P0 = a.part, *P1 = a.part + NB, *P2 = a.part + 2 * NB, *P3 = a.part + 3 * NB;
    const int lane = threadIdx.x % SPMV_G, grp = threadIdx.x / SPMV_G;
    constexpr int GPB = KB / SPMV_G;
    uint32_t gen = 0;
    int status = 1, total_it = 0, outer = 0;
    double resid = 0.0, normb = 1.0;
    bool first = true;

    for (;;) {
        // ---- FP64 residual ----
        double rr = 0.0, bb = 0.0;
        for (int64_t r = r0 + grp; r < r1; r += GPB) {
            const double s = row_dot<SPMV_G>(a.rp, a.col, a.val, a.x, r, lane);
            if (lane == 0) {
                const double rv = a.b[r] - s;
                a.r[r] = rv;
                rr += rv * rv;
                if (first) bb += a.b[r] * a.b[r];
            }
        }
        rr = block_sum<KB>(rr);
        if (first) bb = block_sum<KB>(bb);
        if (threadIdx.x == 0) {
            P1[blockIdx.x] = rr;
            if (first) P3[blockIdx.x] = bb;
        }
        if (!grid_sync(a.bar, gen)) { status = 2; break; }
        const double rnorm = sqrt(grid_total(P1));
        if (first) {
            normb = sqrt(grid_total(P3));
            if (normb == 0.0 || !a.rel) normb = 1.0;
            first = false;
        }
        resid = rnorm / normb;
        if (resid <= a.tol) { status = 0; break; }
        if (total_it >= a.max_iter) { status = 1; break; }
        ++outer;
        // ---- FP32 inner PCG on A d = r/|r| ----
        const double scale = 1.0 / rnorm;
        double zr = 0.0;
        for (int64_t r = r0 + threadIdx.x; r < r1; r += KB) {
            const float rv = (float)(a.r[r] * scale);
            a.rin[r] = rv;
            a.d[r] = 0.0f;
            a.p0[r] = 0.0f;
            const float zv = rv / a.diag32[r];
            a.z[r] = zv;
            zr += (double)zv * (double)rv;
        }
        zr = block_sum<KB>(zr);
        if (threadIdx.x == 0) P2[blockIdx.x] = zr;
        if (!grid_sync(a.bar, gen)) { status = 2; break; }
        double rho = grid_total(P2);
        float beta = 0.0f;
        bool fail = false;
        for (int it = 1; it <= a.inner_max && total_it < a.max_iter; ++it) {
            ++total_it;
            float *pn = (it & 1) ? a.p1 : a.p0;
            const float *po = (it & 1) ? a.p0 : a.p1;
            double lam = 0.0;
            for (int64_t r = r0 + grp; r < r1; r += GPB) {
                float s = 0.0f;
                const int32_t e = a.rp[r + 1];
                for (int32_t k = a.rp[r] + lane; k < e; k += SPMV_G) {
                    const int32_t cc = a.col[k];
                    s += a.val32[k] * (a.z[cc] + beta * po[cc]);
                }
#pragma unroll
                for (int o = SPMV_G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
                if (lane == 0) {
                    const float pv = a.z[r] + beta * po[r];
                    pn[r] = pv;
                    a.q[r] = s;
                    lam += (double)s * (double)pv;
                }
            }
            lam = block_sum<KB>(lam);
            if (threadIdx.x == 0) P0[blockIdx.x] = lam;
            if (!grid_sync(a.bar, gen)) { fail = true; break; }
            const float alpha = (float)(rho / grid_total(P0));
            double rr2 = 0.0, zr2 = 0.0;
            for (int64_t r = r0 + threadIdx.x; r < r1; r += KB) {
                a.d[r] += alpha * pn[r];
                float rv = a.rin[r];
                rv -= alpha * a.q[r];
                a.rin[r] = rv;
                rr2 += (double)rv * (double)rv;
                const float zv = rv / a.diag32[r];
                a.z[r] = zv;
                zr2 += (double)zv * (double)rv;
            }
            rr2 = block_sum<KB>(rr2);
            zr2 = block_sum<KB>(zr2);
            if (threadIdx.x == 0) {
                P1[blockIdx.x] = rr2;
                P2[blockIdx.x] = zr2;
            }
            if (!grid_sync(a.bar, gen)) { fail = true; break; }
            if (sqrt(grid_total(P1)) <= a.inner_tol) break;
            const double rho1 = rho;
            rho = grid_total(P2);
            beta = (float)(rho / rho1);
        }
        if (fail) { status = 2; break; }
        // ---- FP64 refinement update ----
        for (int64_t r = r0 + threadIdx.x; r < r1; r += KB) a.x[r] += rnorm * (double)a.d[r];
        if (!grid_sync(a.bar, gen)) { status = 2; break; }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.out_iter = total_it;
        *a.out_resid = resid;
        *a.out_status = status;
        *a.out_outer = outer;
    }
}

}  // namespace dhip
