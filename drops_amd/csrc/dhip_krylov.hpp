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

// The persistent solver kernels (k_pcg64, k_pcg64_sr, k_gmres, k_pcg_mixed) are defined in the
// translation unit that defines DHIP_SOLVER_KERNELS (dhip_solvers.hip, which also instantiates
// them); every other unit sees their declarations.

#include "dhip_common.hpp"

#include <type_traits>

namespace dhip {

constexpr int KB = 256;       // threads per block of the vector kernels

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

// CSR SpMV over tiles of rb consecutive rows.  A tile's entries are streamed in coalesced,
// non-temporal loads (thread t takes entries t, t + 256, ..., all of a thread's loads issued before
// the first use; the matrix is read once, the caches are left to x), their products val * x[col]
// staged in LDS, and each row summed by one thread in column order: the serial row loop's rounding
// bit for bit.  rb comes from the mean row length so that a tile normally fits the stage; a tile
// with more than SPMV_CAP entries sums its rows straight from global memory, one row per thread, in
// the same order.  The 16 KB stage keeps up to 8 blocks per CU resident (a 32 KB stage measured
// 3.5 TB/s, 16 KB 4.7 TB/s, 8 KB 4.5 TB/s on a 315 MB matrix of 11-entry rows).
// RESID: y = b - A x instead of y = A x (b - the serial row sum).
constexpr int SPMV_T = 256, SPMV_CAP = 2048, SPMV_PT = SPMV_CAP / SPMV_T;

#ifndef DHIP_SOLVER_KERNELS
template <bool RESID>
__global__ void __launch_bounds__(SPMV_T) k_spmv(int64_t n, int rb, const int32_t *rp, const int32_t *col,
                                                 const double *val, const double *x, const double *b, double *y) {
    __shared__ double prod[SPMV_CAP];
    const int64_t r0 = (int64_t)blockIdx.x * rb;
    if (r0 >= n) return;   // the same for every thread of the block
    const int64_t r1 = r0 + rb < n ? r0 + rb : n;
    const int32_t z0 = rp[r0], nz = rp[r1] - z0;
    if (nz <= SPMV_CAP) {
        int32_t cc[SPMV_PT];
        double vv[SPMV_PT];
#pragma unroll
        for (int i = 0; i < SPMV_PT; ++i) {
            const int k = threadIdx.x + i * SPMV_T;
            if (k < nz) {
                cc[i] = __builtin_nontemporal_load(col + z0 + k);
                vv[i] = __builtin_nontemporal_load(val + z0 + k);
            }
        }
#pragma unroll
        for (int i = 0; i < SPMV_PT; ++i) {
            const int k = threadIdx.x + i * SPMV_T;
            if (k < nz) prod[k] = vv[i] * x[cc[i]];
        }
        __syncthreads();
        const int64_t r = r0 + threadIdx.x;
        if (r < r1) {
            double s = 0.0;
            for (int32_t k = rp[r] - z0, e = rp[r + 1] - z0; k < e; ++k) s += prod[k];
            y[r] = RESID ? b[r] - s : s;
        }
    } else {
        for (int64_t r = r0 + threadIdx.x; r < r1; r += SPMV_T) {
            double s = 0.0;
            for (int32_t k = rp[r], e = rp[r + 1]; k < e; ++k) s += val[k] * x[col[k]];
            y[r] = RESID ? b[r] - s : s;
        }
    }
}
#endif

// rows per SpMV tile: the largest power of two <= 256 whose tile of mean-length rows, with 25 %
// headroom, fits the LDS stage
inline int spmv_tile_rows(int64_t n, int64_t nnz) {
    const double mean = n > 0 ? (double)nnz / (double)n : 0.0;
    int rb = SPMV_T;
    while (rb > 16 && rb * mean * 1.25 > SPMV_CAP) rb /= 2;
    return rb;
}

// The transposed pattern for y = A^T x (y_ATx, num/spmat.h:1729-1747): column counts, their scan,
// a fill pass, then each transposed row sorted by original row index -- the order in which y_ATx
// adds the products into y_j -- so that k_spmv over it gives y_ATx's sums bit for bit.
#ifndef DHIP_SOLVER_KERNELS
__global__ void k_tcount(int64_t nnz, int64_t ncols, const int32_t *col, int32_t *cnt, int32_t *err) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const int32_t j = col[e];
    if (j < 0 || j >= ncols) {
        atomicOr(err, 1);
        return;
    }
    atomicAdd(&cnt[j], 1);
}

__global__ void k_tfill(int64_t n, const int32_t *rp, const int32_t *col, const double *val, const int32_t *trp,
                        int32_t *fill, int32_t *tcol, double *tval) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
        const int32_t j = col[q];
        const int32_t p = trp[j] + atomicAdd(&fill[j], 1);
        tcol[p] = (int32_t)i;
        tval[p] = val[q];
    }
}

__global__ void k_tsort(int64_t m, const int32_t *trp, int32_t *tcol, double *tval) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int32_t b = trp[j], e = trp[j + 1];
    for (int32_t k = b + 1; k < e; ++k) {
        const int32_t c = tcol[k];
        const double v = tval[k];
        int32_t q = k - 1;
        while (q >= b && tcol[q] > c) {
            tcol[q + 1] = tcol[q];
            tval[q + 1] = tval[q];
            --q;
        }
        tcol[q + 1] = c;
        tval[q + 1] = v;
    }
}
#endif

#ifndef DHIP_SOLVER_KERNELS
__global__ void __launch_bounds__(KB) k_dot_partial(int64_t n, const double *x, const double *y, double *part) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x; i < n; i += (int64_t)gridDim.x * KB)
        s += x[i] * y[i];
    s = block_sum<KB>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
#endif

#ifndef DHIP_SOLVER_KERNELS
__global__ void __launch_bounds__(KB) k_reduce(const double *part, int np, double *out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += KB) s += part[i];
    s = block_sum<KB>(s);
    if (threadIdx.x == 0) *out = s;
}
#endif

#ifndef DHIP_SOLVER_KERNELS
__global__ void __launch_bounds__(KB) k_axpy(int64_t n, double a, const double *x, double *y) {
    const int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
#endif

// diagonal located by a row scan (SolveGSstep<PB_JAC0>, num/precond.h:163-168)
#ifndef DHIP_SOLVER_KERNELS
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
#endif

#ifndef DHIP_SOLVER_KERNELS
__global__ void __launch_bounds__(KB) k_to_f32(int64_t n, const double *a, float *b) {
    const int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
#endif

#ifndef DHIP_SOLVER_KERNELS
__global__ void k_zero_u32(uint32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
#endif

#ifndef DHIP_SOLVER_KERNELS
__global__ void k_copy_i32(const int32_t *src, int32_t *dst) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = *src;
}
#endif

// ---------------------------------------------------------------------------
// Grid synchronisation for cooperative (co-resident) launches.
//
// Every block owns one slot of 64-bit words.  A block (1) waits until its own
// global stores have reached L2 (s_waitcnt vmcnt(0) in every wave -- s_barrier
// alone does not wait for them), (2) one agent-scope release (L2 write-back, so
// other XCDs see the block's data), (3) publishes each partial sum as two
// 64-bit words {generation:32 | half of the double:32}.  64-bit atomic accesses
// are single-copy atomic, so a reader that sees the generation in both words
// holds the matching value: one polling round trip returns arrival and data.
// One wave per block polls every slot, then acquires (for the vector data) and
// sums the partials in lane order -- the same order in every block, so every
// block holds bit-identical totals and takes identical decisions.  No atomic
// read-modify-write, so arrival cost does not grow with a serialised counter.
// Words are double buffered by generation parity: a block can run at most one
// generation ahead of the slowest reader.
// ---------------------------------------------------------------------------
constexpr int SYNC_VALS = 3;

struct alignas(128) SyncSlot {
    uint64_t w[2][2 * SYNC_VALS];
    uint32_t timeout;
};

constexpr uint32_t SPIN_LIMIT = 1u << 22;  // polls; a healthy barrier takes microseconds
constexpr int POLL_CHUNKS = 4;             // slots polled per lane in one round (covers 256 blocks)

__device__ __forceinline__ void wait_own_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Vector values exchanged between the blocks of a persistent solver.  COH = false: plain stores and
// loads, made visible by the agent-scope release and acquire fences of the synchronisation, which
// write back and invalidate the L2 of every block's XCD on each synchronisation.  COH = true: each
// exchanged value is stored and loaded as a relaxed agent-scope atomic -- coherent across the XCDs
// by itself -- and the synchronisation runs without fences: a block publishes its slot only after
// s_waitcnt vmcnt(0) has seen all its stores complete, and a reader issues its loads only after it
// has seen every slot of the generation.
template <bool COH, class T>
__device__ __forceinline__ T xload(const T *p) {
    if (COH) return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}
template <bool COH, class T>
__device__ __forceinline__ void xstore(T *p, T v) {
    if (COH)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}

// Returns false in every thread of a block that saw a timeout.  FENCES = false: for solvers whose
// exchanged values are coherent atomics (xload / xstore with COH = true).
template <int BLOCK, bool FENCES = true>
__device__ inline bool sync_reduce(SyncSlot *slots, uint32_t g, double a0, double a1, double a2, double out[3]) {
    __shared__ double sh[SYNC_VALS];
    __shared__ int ok_sh;
    wait_own_stores();
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x, nb = gridDim.x, par = g & 1;
        const uint64_t tag = (uint64_t)g << 32;
        if (lane == 0) {
            if (FENCES) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            uint64_t *w = slots[blockIdx.x].w[par];
            const double in[SYNC_VALS] = {a0, a1, a2};
#pragma unroll
            for (int i = 0; i < SYNC_VALS; ++i) {
                const uint64_t bits = __double_as_longlong(in[i]);
                __hip_atomic_store(&w[2 * i], tag | (bits & 0xffffffffull), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&w[2 * i + 1], tag | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        int ok = 1;
        double v[SYNC_VALS] = {0.0, 0.0, 0.0};
        // lane l polls slots l, l + 64, l + 128, l + 192 with all loads of a round in flight
        // together (one polling round trip, not one per 64 blocks), then adds them in that order
        for (int base = 0; base < nb; base += POLL_CHUNKS * 64) {
            uint64_t x[POLL_CHUNKS][2 * SYNC_VALS];
            uint32_t spins = 0;
            for (;;) {
                bool ready = true;
#pragma unroll
                for (int j = 0; j < POLL_CHUNKS; ++j) {
                    const int l = base + j * 64 + lane;
                    if (l < nb) {
                        const uint64_t *w = slots[l].w[par];
#pragma unroll
                        for (int i = 0; i < 2 * SYNC_VALS; ++i) {
                            x[j][i] = __hip_atomic_load(&w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            ready &= (uint32_t)(x[j][i] >> 32) == g;
                        }
                    }
                }
                if (ready) break;
                if (++spins > SPIN_LIMIT) {
                    ok = 0;
                    break;
                }
            }
#pragma unroll
            for (int j = 0; j < POLL_CHUNKS; ++j)
                if (base + j * 64 + lane < nb) {
#pragma unroll
                    for (int i = 0; i < SYNC_VALS; ++i)
                        v[i] += __longlong_as_double(
                            (long long)((x[j][2 * i] & 0xffffffffull) | (x[j][2 * i + 1] << 32)));
                }
        }
        if (FENCES) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
        for (int i = 0; i < SYNC_VALS; ++i) v[i] = wave_sum(v[i]);
        const bool all_ok = __ballot(ok) == __ballot(1);
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < SYNC_VALS; ++i) sh[i] = v[i];
            ok_sh = all_ok ? 1 : 0;
            if (!all_ok) slots[blockIdx.x].timeout = 1;
        }
    }
    __syncthreads();
    out[0] = sh[0];
    out[1] = sh[1];
    out[2] = sh[2];
    const bool ok = ok_sh != 0;
    __syncthreads();
    return ok;
}

// ---------------------------------------------------------------------------
// persistent PCG (FP64), num/krylovsolver.h:107-196 semantics
//
// Block b owns rows [r0, r1) (one or RPT rows per thread, values in registers)
// and keeps its CSR slice in LDS when it fits.  Only z and p are exchanged
// through global memory; the p-update of iteration i is formed on the fly
// inside the SpMV of iteration i+1 (p double buffered), so an iteration costs
// two grid synchronisations: after q = A p (lambda) and after the vector
// update (|r|^2, z.r).
// ---------------------------------------------------------------------------

// Row range of this block among gridDim.x equal ranges: range index blockIdx.x, or with xcd_map the
// block's position above, so that the rows a block gathers from its neighbours were mostly written
// by blocks of its own XCD and are read from the L2 they share.
__device__ inline unsigned row_range_index(int xcd_map) {
    return xcd_map ? xcd_position(blockIdx.x, gridDim.x) : blockIdx.x;
}

constexpr int PB = 512;                   // threads per block of the persistent solvers
constexpr int LDS_BYTES = 64 * 1024;      // dynamic LDS per block for the CSR slice

struct PcgArgs {
    int64_t n;
    const int32_t *rp, *col;
    const double *val, *diag;   // diag == nullptr: no preconditioner
    const double *b;
    double *x, *z, *p0, *p1;
    SyncSlot *slots;
    int max_iter;
    double tol;
    int rel;
    int *out_iter;
    double *out_resid;
    int *out_status;            // 0 converged, 1 not converged, 2 synchronisation timeout
    int xcd_map = 0;            // 1: the blocks of one XCD own one contiguous run of rows
};

// CSR slice of a block: LDS copy when it fits, else the global arrays.
template <class VT>
struct Slice {
    const int32_t *rp;   // local row pointers (LDS) or global rp
    const int32_t *col;
    const VT *val;
    int32_t off;         // subtract from rp entries when reading the global arrays
    int64_t r0;
};

template <class VT>
__device__ inline Slice<VT> load_slice(unsigned char *lds, const int32_t *rp, const int32_t *col, const VT *val,
                                       int64_t r0, int64_t r1) {
    Slice<VT> S;
    S.r0 = r0;
    const int64_t nr = r1 > r0 ? r1 - r0 : 0;
    const int32_t nz0 = nr ? rp[r0] : 0;
    const int64_t nz = nr ? (int64_t)rp[r1] - nz0 : 0;
    const size_t need = (size_t)(nr + 1) * 4 + (size_t)nz * (sizeof(VT) + 4) + 16;
    if (nr && need <= (size_t)LDS_BYTES) {
        int32_t *lrp = reinterpret_cast<int32_t *>(lds);
        VT *lval = reinterpret_cast<VT *>(lds + (((nr + 1) * 4 + 15) & ~(int64_t)15));
        int32_t *lcol = reinterpret_cast<int32_t *>(lval + nz);
        for (int64_t i = threadIdx.x; i <= nr; i += blockDim.x) lrp[i] = rp[r0 + i] - nz0;
        for (int64_t i = threadIdx.x; i < nz; i += blockDim.x) {
            lval[i] = val[nz0 + i];
            lcol[i] = col[nz0 + i];
        }
        __syncthreads();
        S.rp = lrp;
        S.col = lcol;
        S.val = lval;
        S.off = 0;
    } else {
        S.rp = rp + r0;
        S.col = col + nz0;
        S.val = val + nz0;
        S.off = nz0;
    }
    return S;
}

// Row product sum_k val_k * xv(col_k) for local row l of a slice.  Rows of at
// most MAXR entries (P1 trace-FEM rows hold <= 15) are processed as one batch:
// indices and values first, then all gathers issued together, then the sum in
// column order -- the same rounding as the sequential loop, one memory latency
// instead of one per entry.
constexpr int MAXR = 16;

template <class ACC, int BATCH = MAXR, class VT, class F>
__device__ __forceinline__ ACC row_product(const int32_t *rp, const int32_t *col, const VT *val, int32_t off,
                                           int64_t l, F &&xv) {
    const int32_t b = rp[l] - off, e = rp[l + 1] - off;
    ACC s = ACC(0);
    // chunks of BATCH entries: all gathers of a chunk are issued before the first use,
    // the products are summed in column order (the order of a serial row loop)
    for (int32_t q0 = b; q0 < e; q0 += BATCH) {
        const int len = e - q0 < BATCH ? e - q0 : BATCH;
        int32_t cc[BATCH];
        VT vv[BATCH];
        ACC xx[BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; ++j)
            if (j < len) {
                cc[j] = col[q0 + j];
                vv[j] = val[q0 + j];
            }
#pragma unroll
        for (int j = 0; j < BATCH; ++j)
            if (j < len) xx[j] = xv(cc[j]);
#pragma unroll
        for (int j = 0; j < BATCH; ++j)
            if (j < len) s += (ACC)vv[j] * xx[j];
    }
    return s;
}

// The row product of row_product summed in two halves, columns [0, H) and [H, len) each in column
// order, then added: the summation of k_pcg64_sr, which k_pcg64_tx reproduces over its LDS halo.
template <int H, class F>
__device__ __forceinline__ double row_product_halves(const int32_t *rp, const int32_t *col, const double *val,
                                                     int32_t off, int64_t l, F &&xv) {
    const int32_t b = rp[l] - off, e = rp[l + 1] - off;
    const int len = e - b;
    double s0 = 0.0, s1 = 0.0;
    int32_t cc[H];
    double vv[H], xx[H];
#pragma unroll
    for (int j = 0; j < H; ++j)
        if (j < len) {
            cc[j] = col[b + j];
            vv[j] = val[b + j];
        }
#pragma unroll
    for (int j = 0; j < H; ++j)
        if (j < len) xx[j] = xv(cc[j]);
#pragma unroll
    for (int j = 0; j < H; ++j)
        if (j < len) s0 += vv[j] * xx[j];
    for (int32_t q0 = b + H; q0 < e; q0 += H) {
        const int cnt = e - q0 < H ? e - q0 : H;
#pragma unroll
        for (int j = 0; j < H; ++j)
            if (j < cnt) {
                cc[j] = col[q0 + j];
                vv[j] = val[q0 + j];
            }
#pragma unroll
        for (int j = 0; j < H; ++j)
            if (j < cnt) xx[j] = xv(cc[j]);
#pragma unroll
        for (int j = 0; j < H; ++j)
            if (j < cnt) s1 += vv[j] * xx[j];
    }
    return s0 + s1;
}

constexpr int TX_HALF = 8;   // columns of the first half of a row (row_product_halves)

template <int RPT>
__global__ void __launch_bounds__(PB) k_pcg64(PcgArgs a)
#ifndef DHIP_SOLVER_KERNELS
    ;
#else
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)row_range_index(a.xcd_map) * per;
    const int64_t r1 = r0 + per < a.n ? r0 + per : a.n;
    const Slice<double> S = load_slice<double>(lds, a.rp, a.col, a.val, r0, r1);
    uint32_t gen = 0;
    double red[3];
    int status = 1, it_done = a.max_iter;
    double resid = 0.0;

    double xr[RPT], rr[RPT], zr[RPT], pr[RPT], dr[RPT];
    bool own[RPT];
    int64_t lr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        lr[k] = threadIdx.x + (int64_t)k * PB;
        own[k] = r0 + lr[k] < r1;
    }
    auto row = [&](int64_t l, auto &&xv) { return row_product<double>(S.rp, S.col, S.val, S.off, l, xv); };

    // r = b - A x, |r|^2, |b|^2
    double p_rr = 0.0, p_bb = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        if (!own[k]) continue;
        const int64_t r = r0 + lr[k];
        const double ax = row(lr[k], [&](int32_t c) { return a.x[c]; });
        const double bv = a.b[r];
        xr[k] = a.x[r];
        rr[k] = bv - ax;
        dr[k] = a.diag ? a.diag[r] : 1.0;
        p_rr += rr[k] * rr[k];
        p_bb += bv * bv;
    }
    p_rr = block_sum<PB>(p_rr);
    p_bb = block_sum<PB>(p_bb);
    if (!sync_reduce<PB>(a.slots, ++gen, p_rr, p_bb, 0.0, red)) { status = 2; goto done; }
    {
        double normb = sqrt(red[1]);
        if (normb == 0.0 || !a.rel) normb = 1.0;
        resid = sqrt(red[0]) / normb;
        if (resid <= a.tol) {
            status = 0;
            it_done = 0;
            goto done;
        }
        double p_zr = 0.0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            const int64_t r = r0 + lr[k];
            zr[k] = a.diag ? 1.0 * rr[k] / dr[k] : rr[k];
            pr[k] = 0.0;
            a.z[r] = zr[k];
            a.p0[r] = 0.0;
            p_zr += zr[k] * rr[k];
        }
        p_zr = block_sum<PB>(p_zr);
        if (!sync_reduce<PB>(a.slots, ++gen, p_zr, 0.0, 0.0, red)) { status = 2; goto done; }
        double rho = red[0], beta = 0.0;

        for (int it = 1; it <= a.max_iter; ++it) {
            double *pn = (it & 1) ? a.p1 : a.p0;
            const double *po = (it & 1) ? a.p0 : a.p1;
            const double *zg = a.z;
            double qv[RPT], p_lam = 0.0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                const int64_t r = r0 + lr[k];
                pr[k] = zr[k] + beta * pr[k];
                pn[r] = pr[k];
                qv[k] = row(lr[k], [&](int32_t c) { return zg[c] + beta * po[c]; });
                p_lam += qv[k] * pr[k];
            }
            p_lam = block_sum<PB>(p_lam);
            if (!sync_reduce<PB>(a.slots, ++gen, p_lam, 0.0, 0.0, red)) { status = 2; goto done; }
            const double alpha = rho / red[0];
            double p_rr2 = 0.0, p_zr2 = 0.0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                const int64_t r = r0 + lr[k];
                xr[k] += alpha * pr[k];
                rr[k] += (-alpha) * qv[k];
                p_rr2 += rr[k] * rr[k];
                zr[k] = a.diag ? 1.0 * rr[k] / dr[k] : rr[k];
                a.z[r] = zr[k];
                p_zr2 += zr[k] * rr[k];
            }
            p_rr2 = block_sum<PB>(p_rr2);
            p_zr2 = block_sum<PB>(p_zr2);
            if (!sync_reduce<PB>(a.slots, ++gen, p_rr2, p_zr2, 0.0, red)) { status = 2; goto done; }
            resid = sqrt(red[0]) / normb;
            if (resid <= a.tol) {
                status = 0;
                it_done = it;
                break;
            }
            const double rho1 = rho;
            rho = red[1];
            beta = rho / rho1;
        }
    }
done:
    if (status != 2) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (own[k] && it_done > 0) a.x[r0 + lr[k]] = xr[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.out_iter = it_done;
        *a.out_resid = resid;
        *a.out_status = status;
    }
}
#endif

// ---------------------------------------------------------------------------
// persistent single-reduction PCG (FP64): the Chronopoulos-Gear arrangement of
// the same recurrences -- one fused reduction of (r,u), (w,u), (r,r) per
// iteration instead of two dependent ones:
//   p = u + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s,  u = M r,  w = A u
//   gamma = (r,u), delta = (w,u), beta' = gamma/gamma_old,
//   alpha' = gamma / (delta - beta' gamma / alpha)
// with u = M r the Jacobi step.  w = A u is formed from the neighbours' values
// of the previous synchronisation: u_c = (r_c - alpha (w_c + beta s_c)) / d_c,
// using the column-scaled slice a_rc / d_c kept in LDS.  r, w and s are double
// buffered by iteration parity (separate arrays: a layout with the three values of
// a row side by side measured slower, its synchronisations took longer).
// ---------------------------------------------------------------------------
struct PcgSrArgs {
    int64_t n;
    const int32_t *rp, *col;
    const double *val, *diag;   // diag == nullptr: no preconditioner
    const double *b;
    double *x;
    double *R[2], *W[2], *S[2];
    SyncSlot *slots;
    int max_iter;
    double tol;
    int rel;
    int *out_iter;
    double *out_resid;
    int *out_status;
    uint64_t *probe;            // optional [gridDim.x][4] wall-clock ticks: setup, compute, sync, total
    int xcd_map = 0;            // 1: the blocks of one XCD own one contiguous run of rows
};

// LDS copy of the slice with values divided by the diagonal of their column, held as VT (the
// quotient rounded once).  When it does not fit: the global arrays, with val set only for VT = double.
template <class VT = double>
__device__ inline Slice<VT> load_slice_scaled(unsigned char *lds, const int32_t *rp, const int32_t *col,
                                              const double *val, const double *diag, int64_t r0, int64_t r1,
                                              bool &scaled) {
    Slice<VT> S;
    S.r0 = r0;
    const int64_t nr = r1 > r0 ? r1 - r0 : 0;
    const int32_t nz0 = nr ? rp[r0] : 0;
    const int64_t nz = nr ? (int64_t)rp[r1] - nz0 : 0;
    const size_t need = (size_t)(nr + 1) * 4 + (size_t)nz * (4 + sizeof(VT)) + 16;
    scaled = false;
    if (nr && need <= (size_t)LDS_BYTES) {
        int32_t *lrp = reinterpret_cast<int32_t *>(lds);
        VT *lval = reinterpret_cast<VT *>(lds + (((nr + 1) * 4 + 15) & ~(int64_t)15));
        int32_t *lcol = reinterpret_cast<int32_t *>(lval + nz);
        for (int64_t i = threadIdx.x; i <= nr; i += blockDim.x) lrp[i] = rp[r0 + i] - nz0;
        for (int64_t i = threadIdx.x; i < nz; i += blockDim.x) {
            const int32_t c = col[nz0 + i];
            lval[i] = (VT)(diag ? val[nz0 + i] / diag[c] : val[nz0 + i]);
            lcol[i] = c;
        }
        __syncthreads();
        S.rp = lrp;
        S.col = lcol;
        S.val = lval;
        S.off = 0;
        scaled = true;
    } else {
        S.rp = rp + r0;
        S.col = col + nz0;
        if constexpr (sizeof(VT) == sizeof(double)) S.val = val + nz0;
        else S.val = nullptr;
        S.off = nz0;
    }
    return S;
}

template <int RPT, bool COH>
__global__ void __launch_bounds__(PB) k_pcg64_sr(PcgSrArgs a)
#ifndef DHIP_SOLVER_KERNELS
    ;
#else
{
    extern __shared__ __align__(16) unsigned char lds[];
    const uint64_t t_start = wall_clock64();
    uint64_t t_compute = 0, t_sync = 0, t_mark = 0;
    const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)row_range_index(a.xcd_map) * per;
    const int64_t r1 = r0 + per < a.n ? r0 + per : a.n;
    bool scaled;
    const Slice<double> S = load_slice_scaled(lds, a.rp, a.col, a.val, a.diag, r0, r1, scaled);
    const double *dg = a.diag;
    // sum_c (a_rc / d_c) * y_c over local row l, y gathered by yv
    auto scaled_row = [&](int64_t l, auto &&yv) {
        if (scaled || !dg) return row_product_halves<TX_HALF>(S.rp, S.col, S.val, S.off, l, yv);
        return row_product_halves<TX_HALF>(S.rp, S.col, S.val, S.off, l, [&](int32_t c) { return yv(c) / dg[c]; });
    };
    uint32_t gen = 0;
    double red[3];
    int status = 1, it_done = a.max_iter;
    double resid = 0.0;
    double xr[RPT], rr[RPT], ur[RPT], wr[RPT], pr[RPT], sr[RPT], dr[RPT];
    bool own[RPT];
    int64_t lr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        lr[k] = threadIdx.x + (int64_t)k * PB;
        own[k] = r0 + lr[k] < r1;
    }
    // r0 = b - A x0 (unscaled values from global memory)
    double p_rr = 0.0, p_bb = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        if (!own[k]) continue;
        const int64_t r = r0 + lr[k];
        const double ax = row_product<double>(a.rp, a.col, a.val, 0, r, [&](int32_t c) { return a.x[c]; });
        const double bv = a.b[r];
        xr[k] = a.x[r];
        rr[k] = bv - ax;
        dr[k] = dg ? dg[r] : 1.0;
        xstore<COH>(&a.R[0][r], rr[k]);
        p_rr += rr[k] * rr[k];
        p_bb += bv * bv;
    }
    {
        double unused = 0.0;
        block_sum3<PB>(p_rr, p_bb, unused);
    }
    if (!sync_reduce<PB, !COH>(a.slots, ++gen, p_rr, p_bb, 0.0, red)) { status = 2; goto done; }
    {
        double normb = sqrt(red[1]);
        if (normb == 0.0 || !a.rel) normb = 1.0;
        resid = sqrt(red[0]) / normb;
        if (resid <= a.tol) {
            status = 0;
            it_done = 0;
            goto done;
        }
        // w0 = A M r0, gamma0 = (r0, u0), delta0 = (w0, u0)
        double p_g = 0.0, p_d = 0.0;
        const double *R0 = a.R[0];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            const int64_t r = r0 + lr[k];
            ur[k] = dg ? 1.0 * rr[k] / dr[k] : rr[k];
            wr[k] = scaled_row(lr[k], [&](int32_t c) { return xload<COH>(&R0[c]); });
            sr[k] = 0.0;
            pr[k] = 0.0;
            xstore<COH>(&a.W[0][r], wr[k]);
            xstore<COH>(&a.S[0][r], 0.0);
            p_g += rr[k] * ur[k];
            p_d += wr[k] * ur[k];
        }
        {
            double unused = 0.0;
            block_sum3<PB>(p_g, p_d, unused);
        }
        if (!sync_reduce<PB, !COH>(a.slots, ++gen, p_g, p_d, 0.0, red)) { status = 2; goto done; }
        double gamma = red[0], alpha = red[0] / red[1], beta = 0.0;
        const uint64_t t_setup = wall_clock64() - t_start;
        if (a.probe && threadIdx.x == 0) a.probe[blockIdx.x * 4 + 0] = t_setup;

        for (int it = 0; it < a.max_iter; ++it) {
            if (a.probe) t_mark = wall_clock64();
            const int cur = it & 1, nxt = cur ^ 1;
            const double *Rc = a.R[cur], *Wc = a.W[cur], *Sc = a.S[cur];
            double p_g2 = 0.0, p_d2 = 0.0, p_rr2 = 0.0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                const int64_t r = r0 + lr[k];
                pr[k] = ur[k] + beta * pr[k];
                sr[k] = wr[k] + beta * sr[k];
                xr[k] += alpha * pr[k];
                rr[k] = rr[k] - alpha * sr[k];
                ur[k] = dg ? 1.0 * rr[k] / dr[k] : rr[k];
                wr[k] = scaled_row(lr[k], [&](int32_t c) {
                    return xload<COH>(&Rc[c]) - alpha * (xload<COH>(&Wc[c]) + beta * xload<COH>(&Sc[c]));
                });
                xstore<COH>(&a.R[nxt][r], rr[k]);
                xstore<COH>(&a.W[nxt][r], wr[k]);
                xstore<COH>(&a.S[nxt][r], sr[k]);
                p_g2 += rr[k] * ur[k];
                p_d2 += wr[k] * ur[k];
                p_rr2 += rr[k] * rr[k];
            }
            block_sum3<PB>(p_g2, p_d2, p_rr2);
            if (a.probe) {
                const uint64_t t1 = wall_clock64();
                t_compute += t1 - t_mark;
                t_mark = t1;
            }
            if (!sync_reduce<PB, !COH>(a.slots, ++gen, p_g2, p_d2, p_rr2, red)) { status = 2; goto done; }
            if (a.probe) t_sync += wall_clock64() - t_mark;
            resid = sqrt(red[2]) / normb;
            if (resid <= a.tol) {
                status = 0;
                it_done = it + 1;
                break;
            }
            const double gamma_new = red[0];
            beta = gamma_new / gamma;
            alpha = gamma_new / (red[1] - beta * gamma_new / alpha);
            gamma = gamma_new;
        }
    }
done:
    if (status != 2) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (own[k] && it_done > 0) a.x[r0 + lr[k]] = xr[k];
    }
    if (a.probe && threadIdx.x == 0) {
        a.probe[blockIdx.x * 4 + 1] = t_compute;
        a.probe[blockIdx.x * 4 + 2] = t_sync;
        a.probe[blockIdx.x * 4 + 3] = wall_clock64() - t_start;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.out_iter = it_done;
        *a.out_resid = resid;
        *a.out_status = status;
    }
}
#endif

// ---------------------------------------------------------------------------
// Tagged exchange: every exchanged double travels with the generation that wrote it, as one
// 16-byte record {lo32, gen, hi32, gen} written and read with device-coherent (sc1) 128-bit
// buffer accesses.  A reader takes a value only when both 64-bit halves carry the generation it
// expects (each half is single-copy atomic), and reloads it otherwise, so a block needs neither
// a store acknowledgement nor a flag before its neighbours may read its rows: the data are the
// flag.  The reductions still go through the slots, published without fences and collected
// while the gathers are in flight.  Records are double buffered by generation parity.
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct TxVec {
    __amdgpu_buffer_rsrc_t rsrc;   // [2][n] records of 16 bytes
    int64_t n;
};

__device__ __forceinline__ TxVec tx_vec(uint4 *base, int64_t n) {
    TxVec v;
    v.rsrc = __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)(2 * n * 16), 0x00020000);
    v.n = n;
    return v;
}

constexpr int TX_SC1 = 16;   // cache policy: device-coherent (sc1)

__device__ __forceinline__ void tx_store(const TxVec &v, uint32_t g, int64_t i, double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t off = (uint32_t)(((g & 1) * v.n + i) * 16);
    const u32x4 rec = {(unsigned)(b & 0xffffffffu), g, (unsigned)(b >> 32), g};
    __builtin_amdgcn_raw_buffer_store_b128(rec, v.rsrc, off, 0, TX_SC1);
}

__device__ __forceinline__ u32x4 tx_load_raw(const TxVec &v, uint32_t g, int64_t i) {
    const uint32_t off = (uint32_t)(((g & 1) * v.n + i) * 16);
    return __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, off, 0, TX_SC1);
}

__device__ __forceinline__ bool tx_ready(const u32x4 &r, uint32_t g) { return r[1] == g && r[3] == g; }

__device__ __forceinline__ double tx_value(const u32x4 &r) {
    return __longlong_as_double((long long)((uint64_t)r[0] | ((uint64_t)r[2] << 32)));
}

// Publish three block partials in this block's slot for generation g (no fence: the partials are
// the slot's own tagged words).  Called by one thread.
__device__ __forceinline__ void slot_publish3(SyncSlot *slots, uint32_t g, double a0, double a1, double a2) {
    uint64_t *w = slots[blockIdx.x].w[g & 1];
    const uint64_t tag = (uint64_t)g << 32;
    const double in[SYNC_VALS] = {a0, a1, a2};
#pragma unroll
    for (int i = 0; i < SYNC_VALS; ++i) {
        const uint64_t bits = __double_as_longlong(in[i]);
        __hip_atomic_store(&w[2 * i], tag | (bits & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&w[2 * i + 1], tag | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Collect generation g of every block's slot by one wave and sum the partials in the order of
// sync_reduce (lane l adds slots l, l + 64, l + 128, l + 192, then the wave sum), so that the totals
// are the ones sync_reduce gives.  Lane 0 returns them; false after a timeout.
__device__ inline bool slot_collect3(SyncSlot *slots, uint32_t g, double out[3]) {
    const int lane = threadIdx.x & 63, nb = gridDim.x, par = g & 1;
    int ok = 1;
    double v[SYNC_VALS] = {0.0, 0.0, 0.0};
    for (int base = 0; base < nb; base += POLL_CHUNKS * 64) {
        uint64_t x[POLL_CHUNKS][2 * SYNC_VALS];
        uint32_t spins = 0;
        for (;;) {
            bool ready = true;
#pragma unroll
            for (int j = 0; j < POLL_CHUNKS; ++j) {
                const int l = base + j * 64 + lane;
                if (l < nb) {
                    const uint64_t *w = slots[l].w[par];
#pragma unroll
                    for (int i = 0; i < 2 * SYNC_VALS; ++i) {
                        x[j][i] = __hip_atomic_load(&w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ready &= (uint32_t)(x[j][i] >> 32) == g;
                    }
                }
            }
            if (ready) break;
            if (++spins > SPIN_LIMIT) {
                ok = 0;
                break;
            }
        }
#pragma unroll
        for (int j = 0; j < POLL_CHUNKS; ++j)
            if (base + j * 64 + lane < nb) {
#pragma unroll
                for (int i = 0; i < SYNC_VALS; ++i)
                    v[i] += __longlong_as_double((long long)((x[j][2 * i] & 0xffffffffull) | (x[j][2 * i + 1] << 32)));
            }
    }
#pragma unroll
    for (int i = 0; i < SYNC_VALS; ++i) out[i] = wave_sum(v[i]);
    return __ballot(ok) == __ballot(1);
}

__device__ inline int block_min_int(int v) {
    __shared__ int ws[PB / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int t = __shfl_xor(v, o, 64);
        v = t < v ? t : v;
    }
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    int m = ws[0];
#pragma unroll
    for (int i = 1; i < PB / 64; ++i) m = ws[i] < m ? ws[i] : m;
    __syncthreads();
    return m;
}

// A block's CSR slice with block-local column indices in LDS: the distinct columns of rows
// [r0, r0 + nr) in an LDS bitmap over [cmn, cmn + range), ranked by a block scan into the list U
// (nu entries, ascending), the slice's columns rewritten as indices into U and its values divided by
// the diagonal of their column (dg != nullptr), and nvec halo arrays X[nvec][nu].  Layout:
// lrp[nr + 1] | lcol[nzb] | bits[nwords] | pre[nwords] | U[nu] | (16-aligned) lval[nzb] | X[nvec][nu].
// fits is the same in every thread; false: the LDS holds neither (the arrays are then unusable).
struct TxLocal {
    int32_t *lrp, *lcol, *pre, *U;
    uint32_t *bits;
    double *lval, *X;
    int cmn, nu;
    int64_t range;
    bool fits;
};

__device__ inline TxLocal tx_local_setup(unsigned char *lds, int64_t lds_bytes, const int32_t *rp, const int32_t *col,
                                         const double *val, const double *dg, int64_t r0, int64_t nr, int32_t nz0,
                                         int64_t nzb, int nvec) {
    TxLocal T{};
    int cmn = 0x7fffffff, cmx = -1;
    for (int64_t k = threadIdx.x; k < nzb; k += PB) {
        const int c = col[nz0 + k];
        cmn = c < cmn ? c : cmn;
        cmx = c > cmx ? c : cmx;
    }
    cmn = block_min_int(cmn);
    cmx = -block_min_int(-cmx);
    const int64_t range = nzb ? (int64_t)cmx - cmn + 1 : 0;
    const int64_t nwords = (range + 31) / 32;
    int32_t *lrp = reinterpret_cast<int32_t *>(lds);
    int32_t *lcol = lrp + nr + 1;
    uint32_t *bits = reinterpret_cast<uint32_t *>(lcol + nzb);
    int32_t *pre = reinterpret_cast<int32_t *>(bits + nwords);
    int32_t *U = pre + nwords;
    bool fits = 4 * (size_t)(nr + 1 + nzb + 2 * nwords) <= (size_t)lds_bytes;
    int nu = 0;
    double *lval = nullptr, *X = nullptr;
    if (fits && nzb > 0) {
        for (int64_t w = threadIdx.x; w < nwords; w += PB) bits[w] = 0u;
        __syncthreads();
        for (int64_t k = threadIdx.x; k < nzb; k += PB) {
            const int d = col[nz0 + k] - cmn;
            atomicOr(&bits[d >> 5], 1u << (d & 31));
        }
        __syncthreads();
        // rank of every set bit: each thread takes a run of words
        const int64_t wpt = (nwords + PB - 1) / PB;
        const int64_t w0 = threadIdx.x * wpt, w1 = w0 + wpt < nwords ? w0 + wpt : nwords;
        int cnt = 0;
        for (int64_t w = w0; w < w1; ++w) cnt += __popc(bits[w]);
        int tot = 0;
        int off = block_excl_scan<PB>(cnt, &tot);
        nu = tot;
        const size_t int_bytes = 4 * (size_t)(nr + 1 + nzb + 2 * nwords + nu);
        const size_t lval_off = (int_bytes + 15) & ~(size_t)15;
        // the same in every thread (the scan total)
        fits = lval_off + 8 * (size_t)(nzb + (int64_t)nvec * nu) <= (size_t)lds_bytes;
        if (fits) {
            lval = reinterpret_cast<double *>(lds + lval_off);
            X = lval + nzb;
            for (int64_t w = w0; w < w1; ++w) {
                pre[w] = off;
                uint32_t m = bits[w];
                while (m) {
                    const int bpos = __ffs(m) - 1;
                    U[off++] = cmn + (int)(32 * w) + bpos;
                    m &= m - 1;
                }
            }
            __syncthreads();
            for (int64_t i = threadIdx.x; i <= nr; i += PB) lrp[i] = rp[r0 + i] - nz0;
            for (int64_t k = threadIdx.x; k < nzb; k += PB) {
                const int c = col[nz0 + k], d = c - cmn;
                lcol[k] = pre[d >> 5] + __popc(bits[d >> 5] & ((1u << (d & 31)) - 1u));
                lval[k] = dg ? val[nz0 + k] / dg[c] : val[nz0 + k];
            }
            __syncthreads();
        }
    } else if (fits) {
        for (int64_t i = threadIdx.x; i <= nr; i += PB) lrp[i] = 0;   // rows without entries
        __syncthreads();
    }
    T.lrp = lrp;
    T.lcol = lcol;
    T.pre = pre;
    T.U = U;
    T.bits = bits;
    T.lval = lval;
    T.X = X;
    T.cmn = cmn;
    T.nu = nu;
    T.range = range;
    T.fits = fits;
    return T;
}

// Persistent FP64 PCG, the single-reduction arrangement of k_pcg64_sr with the tagged exchange and
// the block's halo in LDS.  At set-up each block lists the distinct columns of its rows (an LDS
// bitmap over their range, ranked by a block scan) and rewrites its CSR slice to indices into that
// list.  Per iteration the block loads the r, w, s records of generation it + 2 once per distinct
// column it does not own -- one tagged load per value instead of one per matrix entry -- into LDS,
// while one wave collects the block partials of the same generation; its own rows' values enter
// the LDS copy from registers.  The row products then read LDS only.  One row per thread as in
// k_pcg64_sr, rows summed in row_product_halves' order, block partials added as there: the
// iterates equal k_pcg64_sr's.  A block whose slice and halo do not fit the LDS makes every block
// leave with status 3 after the first reduction (the host then runs k_pcg64_sr).
struct PcgTxArgs {
    int64_t n;
    const int32_t *rp, *col;
    const double *val, *diag;   // diag == nullptr: no preconditioner
    const double *b;
    double *x;
    uint4 *R, *W, *S;           // [2][n] tagged records each
    SyncSlot *slots;
    int max_iter;
    double tol;
    int rel;
    int *out_iter;
    double *out_resid;
    int *out_status;            // 0 converged, 1 not converged, 2 timeout, 3 slice and halo over the LDS
    uint64_t *probe;            // optional [gridDim.x][4] wall-clock ticks: setup, compute, wait, total
    int xcd_map = 0;
    int64_t lds_bytes = 0;      // dynamic LDS of the launch
};


template <int RPT>
__global__ void __launch_bounds__(PB) k_pcg64_tx(PcgTxArgs a)
#ifndef DHIP_SOLVER_KERNELS
    ;
#else
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ double red_sh[SYNC_VALS];
    const uint64_t t_start = wall_clock64();
    uint64_t t_wait = 0;
    bool tbad = false;   // this thread saw a timeout
    const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)row_range_index(a.xcd_map) * per;
    const int64_t r1 = r0 + per < a.n ? r0 + per : a.n;
    const int64_t nr = r1 > r0 ? r1 - r0 : 0;
    const int32_t nz0 = nr ? a.rp[r0] : 0;
    const int64_t nzb = nr ? (int64_t)a.rp[r1] - nz0 : 0;
    const double *dg = a.diag;
    const TxVec TR = tx_vec(a.R, a.n), TW = tx_vec(a.W, a.n), TS = tx_vec(a.S, a.n);
    // ---- set-up A: the distinct columns of the block and the slice in local column indices ----
    const TxLocal TL = tx_local_setup(lds, a.lds_bytes, a.rp, a.col, a.val, dg, r0, nr, nz0, nzb, 3);
    int32_t *lrp = TL.lrp, *lcol = TL.lcol, *pre = TL.pre, *U = TL.U;
    uint32_t *bits = TL.bits;
    const int cmn = TL.cmn;
    const int64_t range = TL.range;
    const bool fits = TL.fits;
    const int nu = TL.nu;
    double *lval = TL.lval, *XR = TL.X, *XW = TL.X ? TL.X + nu : nullptr, *XS = TL.X ? TL.X + 2 * nu : nullptr;
    // local index of global column c (-1: not a column of the block)
    auto local_of = [&](int64_t c) -> int {
        const int64_t d = c - cmn;
        if (d < 0 || d >= range) return -1;
        const uint32_t w = bits[d >> 5], m = 1u << (d & 31);
        return (w & m) ? pre[d >> 5] + __popc(w & (m - 1u)) : -1;
    };
    // row sum over LDS in row_product_halves' order: columns [0, 8), then the rest
    auto lds_row = [&](int64_t l, auto &&yv) {
        const int32_t b0 = lrp[l], e0 = lrp[l + 1];
        const int32_t m0 = e0 - b0 < TX_HALF ? e0 : b0 + TX_HALF;
        double s0 = 0.0, s1 = 0.0;
        for (int32_t q = b0; q < m0; ++q) s0 += lval[q] * yv(lcol[q]);
        for (int32_t q = m0; q < e0; ++q) s1 += lval[q] * yv(lcol[q]);
        return s0 + s1;
    };
    // the halo of generation g: r (and w, s when three) of every distinct column not owned by the
    // block, loaded together and reloaded until written
    auto halo = [&](uint32_t g, bool three) {
        for (int t0 = 0; t0 < nu; t0 += PB) {
            const int t = t0 + threadIdx.x;
            const bool mine = t < nu && !(U[t] >= r0 && U[t] < r1);
            if (!mine) continue;
            const int64_t c = U[t];
            u32x4 vr = tx_load_raw(TR, g, c), vw, vs;
            if (three) {
                vw = tx_load_raw(TW, g, c);
                vs = tx_load_raw(TS, g, c);
            }
            uint32_t spins = 0;
            for (;;) {
                bool ok = tx_ready(vr, g);
                if (three) ok = ok && tx_ready(vw, g) && tx_ready(vs, g);
                if (ok) break;
                if (++spins > SPIN_LIMIT) {
                    tbad = true;
                    break;
                }
                // the buffer-load builtin reads memory the compiler assumes unchanged: without this
                // barrier it reuses the first load's result and the reload never happens
                __asm__ __volatile__("" ::: "memory");
                if (!tx_ready(vr, g)) vr = tx_load_raw(TR, g, c);
                if (three) {
                    if (!tx_ready(vw, g)) vw = tx_load_raw(TW, g, c);
                    if (!tx_ready(vs, g)) vs = tx_load_raw(TS, g, c);
                }
            }
            XR[t] = tx_value(vr);
            if (three) {
                XW[t] = tx_value(vw);
                XS[t] = tx_value(vs);
            }
        }
    };
    // the block partials of generation g into red_sh; false in every thread when any thread of the
    // block has seen a timeout
    auto collect = [&](uint32_t g) {
        const uint64_t t0 = a.probe ? wall_clock64() : 0;
        if (threadIdx.x < 64) {
            double o[3];
            const bool ok = slot_collect3(a.slots, g, o);
            if (threadIdx.x == 0) {
                red_sh[0] = o[0];
                red_sh[1] = o[1];
                red_sh[2] = o[2];
                if (!ok) {
                    tbad = true;
                    a.slots[blockIdx.x].timeout = 1;
                }
            }
        }
        const bool any = __syncthreads_or(tbad);
        if (a.probe) t_wait += wall_clock64() - t0;
        return !any;
    };
    int status = 1, it_done = a.max_iter;
    double resid = 0.0, normb = 1.0;
    double xr[RPT], rr[RPT], ur[RPT], wr[RPT], pr[RPT], sr[RPT], dr[RPT];
    int lo[RPT];
    bool own[RPT];
    int64_t lr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        lr[k] = threadIdx.x + (int64_t)k * PB;
        own[k] = lr[k] < nr;
        lo[k] = (own[k] && fits) ? local_of(r0 + lr[k]) : -1;
    }
    // ---- set-up B: r0 = b - A x0 (generation 1), |r|^2, |b|^2, and whether every block fits ----
    {
        double p_rr = 0.0, p_bb = 0.0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            const int64_t r = r0 + lr[k];
            const double ax = row_product<double>(a.rp, a.col, a.val, 0, r, [&](int32_t c) { return a.x[c]; });
            const double bv = a.b[r];
            xr[k] = a.x[r];
            rr[k] = bv - ax;
            dr[k] = dg ? dg[r] : 1.0;
            tx_store(TR, 1, r, rr[k]);
            if (lo[k] >= 0) XR[lo[k]] = rr[k];
            p_rr += rr[k] * rr[k];
            p_bb += bv * bv;
        }
        double unused = 0.0;
        block_sum3<PB>(p_rr, p_bb, unused);
        if (threadIdx.x == 0) slot_publish3(a.slots, 1, p_rr, p_bb, fits ? 0.0 : 1.0);
    }
    if (fits) halo(1, false);
    if (!collect(1)) { status = 2; goto done; }
    if (red_sh[2] != 0.0) { status = 3; goto done; }
    {
        normb = sqrt(red_sh[1]);
        if (normb == 0.0 || !a.rel) normb = 1.0;
        resid = sqrt(red_sh[0]) / normb;
        if (resid <= a.tol) {
            status = 0;
            it_done = 0;
            goto done;
        }
        // w0 = A M r0 (generation 2 of r, w, s), gamma0 = (r0, u0), delta0 = (w0, u0)
        double p_g = 0.0, p_d = 0.0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            const int64_t r = r0 + lr[k];
            ur[k] = dg ? 1.0 * rr[k] / dr[k] : rr[k];
            wr[k] = lds_row(lr[k], [&](int q) { return XR[q]; });
            sr[k] = 0.0;
            pr[k] = 0.0;
            tx_store(TR, 2, r, rr[k]);
            tx_store(TW, 2, r, wr[k]);
            tx_store(TS, 2, r, 0.0);
            p_g += rr[k] * ur[k];
            p_d += wr[k] * ur[k];
        }
        double unused = 0.0;
        block_sum3<PB>(p_g, p_d, unused);
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (lo[k] >= 0) {
                XW[lo[k]] = wr[k];
                XS[lo[k]] = 0.0;
            }
        if (threadIdx.x == 0) slot_publish3(a.slots, 2, p_g, p_d, 0.0);
    }
    {
        double gamma = 0.0, alpha = 0.0, beta = 0.0;
        if (a.probe && threadIdx.x == 0) a.probe[blockIdx.x * 4 + 0] = wall_clock64() - t_start;
        for (int it = 0;; ++it) {
            const uint32_t gd = (uint32_t)it + 2;
            halo(gd, true);
            if (!collect(gd)) { status = 2; break; }
            if (it == 0) {
                gamma = red_sh[0];
                alpha = red_sh[0] / red_sh[1];
                beta = 0.0;
            } else {
                resid = sqrt(red_sh[2]) / normb;
                if (resid <= a.tol) {
                    status = 0;
                    it_done = it;
                    break;
                }
                const double gamma_new = red_sh[0];
                beta = gamma_new / gamma;
                alpha = gamma_new / (red_sh[1] - beta * gamma_new / alpha);
                gamma = gamma_new;
            }
            if (it == a.max_iter) break;
            double p_g2 = 0.0, p_d2 = 0.0, p_rr2 = 0.0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                const int64_t r = r0 + lr[k];
                pr[k] = ur[k] + beta * pr[k];
                sr[k] = wr[k] + beta * sr[k];
                xr[k] += alpha * pr[k];
                rr[k] = rr[k] - alpha * sr[k];
                ur[k] = dg ? 1.0 * rr[k] / dr[k] : rr[k];
                wr[k] = lds_row(lr[k], [&](int q) { return XR[q] - alpha * (XW[q] + beta * XS[q]); });
                tx_store(TR, gd + 1, r, rr[k]);
                tx_store(TW, gd + 1, r, wr[k]);
                tx_store(TS, gd + 1, r, sr[k]);
                p_g2 += rr[k] * ur[k];
                p_d2 += wr[k] * ur[k];
                p_rr2 += rr[k] * rr[k];
            }
            block_sum3<PB>(p_g2, p_d2, p_rr2);
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (lo[k] >= 0) {
                    XR[lo[k]] = rr[k];
                    XW[lo[k]] = wr[k];
                    XS[lo[k]] = sr[k];
                }
            if (threadIdx.x == 0) slot_publish3(a.slots, gd + 1, p_g2, p_d2, p_rr2);
        }
    }
done:
    if (status == 0 || status == 1) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (own[k] && it_done > 0) a.x[r0 + lr[k]] = xr[k];
    }
    if (a.probe && threadIdx.x == 0) {
        const uint64_t total = wall_clock64() - t_start;
        a.probe[blockIdx.x * 4 + 1] = total - t_wait - a.probe[blockIdx.x * 4 + 0];
        a.probe[blockIdx.x * 4 + 2] = t_wait;
        a.probe[blockIdx.x * 4 + 3] = total;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.out_iter = it_done;
        *a.out_resid = resid;
        *a.out_status = status;
    }
}
#endif

// Persistent FP64 PCG in the pipelined arrangement (Ghysels and Vanroose, "Hiding global synchronization
// latency in the preconditioned Conjugate Gradient algorithm", Parallel Computing 40 (2014), Alg. 4, with
// Jacobi preconditioning M = diag(A)^-1): per iteration the product n = A M w needs only w, which every
// block publishes as tagged records at the end of the previous iteration together with its partials of
// gamma = (r, u), delta = (w, u) and |r|^2.  So each iteration loads the halo of w, runs the row
// products from LDS, and only then collects the partials, whose wait overlaps the halo and the products;
// one vector is exchanged instead of three.  The recurrences
//   z = n + beta z, q = M w + beta q, s = w + beta s, p = u + beta p,
//   x += alpha p, r -= alpha s, u -= alpha q, w -= alpha z
// are CG's in exact arithmetic (same alphas and betas); their rounding drifts from b - A x, so a
// residual that meets the tolerance is checked once as b - A x (x exchanged as records) and, if it
// does not meet it, the recurrences restart from that residual.  The LDS slice, the halo and the
// set-up follow k_pcg64_tx; a block whose slice and halo do not fit makes every block leave with
// status 3.
template <int RPT>
__global__ void __launch_bounds__(PB) k_pcg64_pipe(PcgTxArgs a)
#ifndef DHIP_SOLVER_KERNELS
    ;
#else
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ double red_sh[SYNC_VALS];
    const uint64_t t_start = wall_clock64();
    uint64_t t_wait = 0;
    bool tbad = false;
    const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)row_range_index(a.xcd_map) * per;
    const int64_t r1 = r0 + per < a.n ? r0 + per : a.n;
    const int64_t nr = r1 > r0 ? r1 - r0 : 0;
    const int32_t nz0 = nr ? a.rp[r0] : 0;
    const int64_t nzb = nr ? (int64_t)a.rp[r1] - nz0 : 0;
    const double *dg = a.diag;
    const TxVec TR = tx_vec(a.R, a.n), TW = tx_vec(a.W, a.n), TX = tx_vec(a.S, a.n);
    const TxLocal TL = tx_local_setup(lds, a.lds_bytes, a.rp, a.col, a.val, dg, r0, nr, nz0, nzb, 1);
    const bool fits = TL.fits;
    const int nu = TL.nu;
    double *XH = TL.X;   // the halo of the vector being exchanged (own rows from registers)
    auto local_of = [&](int64_t c) -> int {
        const int64_t d = c - TL.cmn;
        if (d < 0 || d >= TL.range) return -1;
        const uint32_t w = TL.bits[d >> 5], m = 1u << (d & 31);
        return (w & m) ? TL.pre[d >> 5] + __popc(w & (m - 1u)) : -1;
    };
    auto lds_row = [&](int64_t l) {
        const int32_t b0 = TL.lrp[l], e0 = TL.lrp[l + 1];
        const int32_t m0 = e0 - b0 < TX_HALF ? e0 : b0 + TX_HALF;
        double s0 = 0.0, s1 = 0.0;
        for (int32_t q = b0; q < m0; ++q) s0 += TL.lval[q] * XH[TL.lcol[q]];
        for (int32_t q = m0; q < e0; ++q) s1 += TL.lval[q] * XH[TL.lcol[q]];
        return s0 + s1;
    };
    // the halo of V at generation g into XH, columns the block does not own, reloaded until written
    auto halo = [&](const TxVec &V, uint32_t g) {
        for (int t0 = 0; t0 < nu; t0 += PB) {
            const int t = t0 + threadIdx.x;
            if (!(t < nu && !(TL.U[t] >= r0 && TL.U[t] < r1))) continue;
            const int64_t c = TL.U[t];
            u32x4 v = tx_load_raw(V, g, c);
            uint32_t spins = 0;
            while (!tx_ready(v, g)) {
                if (++spins > SPIN_LIMIT) {
                    tbad = true;
                    break;
                }
                // the buffer-load builtin reads memory the compiler assumes unchanged: without this
                // barrier it reuses the first load's result and the reload never happens
                __asm__ __volatile__("" ::: "memory");
                v = tx_load_raw(V, g, c);
            }
            XH[t] = tx_value(v);
        }
    };
    auto collect = [&](uint32_t g) {
        const uint64_t t0 = a.probe ? wall_clock64() : 0;
        if (threadIdx.x < 64) {
            double o[3];
            const bool ok = slot_collect3(a.slots, g, o);
            if (threadIdx.x == 0) {
                red_sh[0] = o[0];
                red_sh[1] = o[1];
                red_sh[2] = o[2];
                if (!ok) {
                    tbad = true;
                    a.slots[blockIdx.x].timeout = 1;
                }
            }
        }
        const bool any = __syncthreads_or(tbad);
        if (a.probe) t_wait += wall_clock64() - t0;
        return !any;
    };
    if (a.probe && threadIdx.x == 0) a.probe[blockIdx.x * 4 + 0] = 0;
    int status = 1, it = 0, restarts = 0;
    uint32_t gen = 0;
    double resid = 0.0, normb = 1.0, alpha = 0.0, gamma = 0.0;
    double xr[RPT], rr[RPT], ur[RPT], wr[RPT], zr[RPT], qr[RPT], sr[RPT], pr[RPT], dr[RPT], nrv[RPT];
    int lo[RPT];
    bool own[RPT];
    int64_t lr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        lr[k] = threadIdx.x + (int64_t)k * PB;
        own[k] = lr[k] < nr;
        lo[k] = (own[k] && fits) ? local_of(r0 + lr[k]) : -1;
        xr[k] = rr[k] = ur[k] = wr[k] = zr[k] = qr[k] = sr[k] = pr[k] = nrv[k] = 0.0;
        dr[k] = 1.0;
    }
    // ---- set-up: r0 = b - A x0 (records and slot of generation 1), |r|^2, |b|^2, whether every block fits
    {
        double p_rr = 0.0, p_bb = 0.0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            const int64_t r = r0 + lr[k];
            const double ax = row_product<double>(a.rp, a.col, a.val, 0, r, [&](int32_t c) { return a.x[c]; });
            const double bv = a.b[r];
            xr[k] = a.x[r];
            rr[k] = bv - ax;
            dr[k] = dg ? dg[r] : 1.0;
            tx_store(TR, 1, r, rr[k]);
            p_rr += rr[k] * rr[k];
            p_bb += bv * bv;
        }
        double unused = 0.0;
        block_sum3<PB>(p_rr, p_bb, unused);
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (lo[k] >= 0) XH[lo[k]] = rr[k];
        gen = 1;
        if (threadIdx.x == 0) slot_publish3(a.slots, gen, p_rr, p_bb, fits ? 0.0 : 1.0);
    }
    if (fits) halo(TR, gen);
    if (!collect(gen)) { status = 2; goto done; }
    if (red_sh[2] != 0.0) { status = 3; goto done; }
    normb = sqrt(red_sh[1]);
    if (normb == 0.0 || !a.rel) normb = 1.0;
    resid = sqrt(red_sh[0]) / normb;
    if (resid <= a.tol) {
        status = 0;
        goto done;
    }
    if (a.probe && threadIdx.x == 0) a.probe[blockIdx.x * 4 + 0] = wall_clock64() - t_start;
    for (;;) {
        // ---- (re)start from r (registers) with the halo of r in XH: u = M r, w = A M r ----
        {
            double p_g = 0.0, p_d = 0.0, p_r = 0.0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                ur[k] = dg ? 1.0 * rr[k] / dr[k] : rr[k];
                wr[k] = lds_row(lr[k]);
                zr[k] = qr[k] = sr[k] = pr[k] = 0.0;
                p_g += rr[k] * ur[k];
                p_d += wr[k] * ur[k];
                p_r += rr[k] * rr[k];
            }
            ++gen;
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (own[k]) tx_store(TW, gen, r0 + lr[k], wr[k]);
            block_sum3<PB>(p_g, p_d, p_r);
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (lo[k] >= 0) XH[lo[k]] = wr[k];
            if (threadIdx.x == 0) slot_publish3(a.slots, gen, p_g, p_d, p_r);
        }
        bool first = true, converged = false;
        for (;; ++it) {
            // n = A M w from the halo of w; then the partials of the same generation
            halo(TW, gen);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (own[k]) nrv[k] = lds_row(lr[k]);
            if (!collect(gen)) { status = 2; goto done; }
            resid = sqrt(red_sh[2]) / normb;
            if (resid <= a.tol) {
                converged = true;
                break;
            }
            if (it == a.max_iter) break;
            const double gamma_new = red_sh[0];
            double beta;
            if (first) {
                beta = 0.0;
                alpha = gamma_new / red_sh[1];
                first = false;
            } else {
                beta = gamma_new / gamma;
                alpha = gamma_new / (red_sh[1] - beta * gamma_new / alpha);
            }
            gamma = gamma_new;
            double p_g = 0.0, p_d = 0.0, p_r = 0.0;
            ++gen;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                const double m = dg ? 1.0 * wr[k] / dr[k] : wr[k];
                zr[k] = nrv[k] + beta * zr[k];
                qr[k] = m + beta * qr[k];
                sr[k] = wr[k] + beta * sr[k];
                pr[k] = ur[k] + beta * pr[k];
                xr[k] += alpha * pr[k];
                rr[k] = rr[k] - alpha * sr[k];
                ur[k] = ur[k] - alpha * qr[k];
                wr[k] = wr[k] - alpha * zr[k];
                tx_store(TW, gen, r0 + lr[k], wr[k]);
                p_g += rr[k] * ur[k];
                p_d += wr[k] * ur[k];
                p_r += rr[k] * rr[k];
            }
            block_sum3<PB>(p_g, p_d, p_r);
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (lo[k] >= 0) XH[lo[k]] = wr[k];
            if (threadIdx.x == 0) slot_publish3(a.slots, gen, p_g, p_d, p_r);
        }
        if (!converged) break;   // max_iter: status 1
        // ---- the true residual b - A x (x exchanged as records; its generation carries an empty slot) ----
        ++gen;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (own[k]) tx_store(TX, gen, r0 + lr[k], xr[k]);
            if (lo[k] >= 0) XH[lo[k]] = xr[k];
        }
        if (threadIdx.x == 0) slot_publish3(a.slots, gen, 0.0, 0.0, 0.0);
        halo(TX, gen);
        if (!collect(gen)) { status = 2; goto done; }
        {
            double p_r = 0.0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                const int64_t r = r0 + lr[k];
                const double ax = row_product<double>(a.rp, a.col, a.val, 0, r,
                                                      [&](int32_t c) { return XH[local_of(c)]; });
                rr[k] = a.b[r] - ax;
                p_r += rr[k] * rr[k];
            }
            double u0 = 0.0, u1 = 0.0;
            block_sum3<PB>(p_r, u0, u1);
            ++gen;
            if (threadIdx.x == 0) slot_publish3(a.slots, gen, p_r, 0.0, 0.0);
            if (!collect(gen)) { status = 2; goto done; }
            resid = sqrt(red_sh[0]) / normb;
            if (resid <= a.tol) {
                status = 0;
                break;
            }
        }
        if (it >= a.max_iter || ++restarts > 20) break;
        // ---- restart from the true residual: its halo (generation with an empty slot) ----
        ++gen;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (own[k]) tx_store(TR, gen, r0 + lr[k], rr[k]);
            if (lo[k] >= 0) XH[lo[k]] = rr[k];
        }
        if (threadIdx.x == 0) slot_publish3(a.slots, gen, 0.0, 0.0, 0.0);
        halo(TR, gen);
        if (!collect(gen)) { status = 2; goto done; }
    }
done:
    if (status == 0 || status == 1) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (own[k] && it > 0) a.x[r0 + lr[k]] = xr[k];
    }
    if (a.probe && threadIdx.x == 0) {
        const uint64_t total = wall_clock64() - t_start;
        a.probe[blockIdx.x * 4 + 1] = total - t_wait - a.probe[blockIdx.x * 4 + 0];
        a.probe[blockIdx.x * 4 + 2] = t_wait;
        a.probe[blockIdx.x * 4 + 3] = total;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.out_iter = status == 0 ? it : a.max_iter;
        *a.out_resid = resid;
        *a.out_status = status;
    }
}
#endif

// ---------------------------------------------------------------------------
// Vector-payload synchronisation: up to KMAX partial sums per block.  The
// payload is written with plain stores before the release; the generation is
// one 64-bit tagged word per slot.  After all generations are seen (acquire),
// thread k sums partial k over the blocks in block order -- the same order in
// every block.
// ---------------------------------------------------------------------------
constexpr int KMAX = 64;

template <int BLOCK, bool FENCES = true>
__device__ inline bool sync_reduce_vec(SyncSlot *slots, double *parts, uint32_t g, const double *mine, int K,
                                       double *out_sh) {
    constexpr bool COH = !FENCES;
    static_assert(BLOCK == 512, "the chunked reduction maps 8 waves to 8 chunks of blocks");
    __shared__ int ok_sh;
    __shared__ double chunk_sh[BLOCK / 64][KMAX];
    const int nb = gridDim.x, par = g & 1;
    double *P = parts + (size_t)par * nb * KMAX;
    if (threadIdx.x < K) xstore<COH>(&P[(size_t)blockIdx.x * KMAX + threadIdx.x], mine[threadIdx.x]);
    wait_own_stores();
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const uint64_t tag = (uint64_t)g << 32;
        if (lane == 0) {
            if (FENCES) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(&slots[blockIdx.x].w[par][0], tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        int ok = 1;
        for (int base = 0; base < nb; base += POLL_CHUNKS * 64) {
            uint32_t spins = 0;
            for (;;) {
                bool ready = true;
#pragma unroll
                for (int j = 0; j < POLL_CHUNKS; ++j) {
                    const int l = base + j * 64 + lane;
                    if (l < nb)
                        ready &= (uint32_t)(__hip_atomic_load(&slots[l].w[par][0], __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT) >> 32) == g;
                }
                if (ready) break;
                if (++spins > SPIN_LIMIT) {
                    ok = 0;
                    break;
                }
            }
        }
        if (FENCES) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const bool all_ok = __ballot(ok) == __ballot(1);
        if (lane == 0) {
            ok_sh = all_ok ? 1 : 0;
            if (!all_ok) slots[blockIdx.x].timeout = 1;
        }
    }
    __syncthreads();
    if (K > 0) {
        // wave w sums blocks w, w+8, w+16, ... of entry (lane); the gathers of a batch are all
        // issued before the first add; then the 8 chunk sums are added in chunk order.  Every
        // block performs the identical sequence, so all blocks hold bit-identical results.
        const int kk = threadIdx.x & 63, ch = threadIdx.x >> 6;
        constexpr int NCH = BLOCK / 64, BATCH = 32;
        if (kk < K) {
            double s = 0.0;
            for (int q0 = 0; ch + NCH * q0 < nb; q0 += BATCH) {
                double v[BATCH];
#pragma unroll
                for (int q = 0; q < BATCH; ++q) {
                    const int b = ch + NCH * (q0 + q);
                    v[q] = b < nb ? xload<COH>(&P[(size_t)b * KMAX + kk]) : 0.0;
                }
#pragma unroll
                for (int q = 0; q < BATCH; ++q)
                    if (ch + NCH * (q0 + q) < nb) s += v[q];
            }
            chunk_sh[ch][kk] = s;
        }
        __syncthreads();
        if (threadIdx.x < K) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < NCH; ++c) s += chunk_sh[c][threadIdx.x];
            out_sh[threadIdx.x] = s;
        }
        __syncthreads();
    }
    return ok_sh != 0;
}

// Eight wave sums at once by a butterfly transpose (10 shuffles instead of 48): on return lane l
// holds the wave sum of p[l >> 3], bit-identical in the 8 lanes of its group.
__device__ inline double wave_sum8(const double p[8]) {
    const int lane = threadIdx.x & 63;
    const bool h = (lane >> 5) & 1, g = (lane >> 4) & 1, f = (lane >> 3) & 1;
    double q[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const double keep = h ? p[4 + t] : p[t], send = h ? p[t] : p[4 + t];
        q[t] = keep + __shfl_xor(send, 32, 64);
    }
    double r2[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const double keep = g ? q[2 + u] : q[u], send = g ? q[u] : q[2 + u];
        r2[u] = keep + __shfl_xor(send, 16, 64);
    }
    double v = (f ? r2[1] : r2[0]) + __shfl_xor(f ? r2[0] : r2[1], 8, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 1, 64);
    return v;
}

// ---------------------------------------------------------------------------
// persistent restarted GMRES(m), num/krylovsolver.h:320-463 semantics
// (serial variant: m-1 Arnoldi steps per cycle, Givens rotations, right or left
// Jacobi preconditioning, relative or absolute tolerance).  Orthogonalisation
// is classical Gram-Schmidt with one re-orthogonalisation (CGS2): the i+1
// dependent reductions of modified Gram-Schmidt become two fused ones; the norm
// of the twice-projected vector is |w1|^2 - |h2|^2.  The small dense part
// (Hessenberg column, rotations, back substitution) is computed redundantly and
// identically by every block from the identical reduced values.
// ---------------------------------------------------------------------------
struct GmresArgs {
    int64_t n;
    const int32_t *rp, *col;
    const double *val, *diag;   // diag == nullptr: no preconditioner
    const double *b;
    double *x;                  // start vector in, solution out (also exchanged for residuals)
    double *V;                  // [m][n] Krylov basis
    SyncSlot *slots;
    double *parts;              // [2][gridDim.x][KMAX]
    int m;
    int max_iter;
    double tol;
    int rel;
    int left;
    int *out_iter;
    double *out_resid;
    int *out_status;
    uint64_t *probe;            // optional [gridDim.x][4] wall-clock ticks: setup, compute, sync, total
    const double *sval;         // right preconditioning: values a_rc / d_c precomputed (or nullptr)
    int64_t lds_bytes;          // dynamic LDS of the launch
    // SELL-64 copy of the matrix (values as the slice would hold them) for blocks whose CSR slice
    // does not fit in LDS; nullptr: such blocks read the CSR arrays
    const int32_t *sell_off, *sell_col;
    const double *sell_val;
    int sell_spb;               // slices per block
    int xcd_map = 0;            // 1: the blocks of one XCD own one contiguous run of rows
    const float *sell_val32 = nullptr;  // FP32 Arnoldi: the SELL-64 values rounded to float
};

// SELL-64 layout of a CSR matrix partitioned into blocks of `per` consecutive rows: the local rows
// of a block form slices of 64; entry k of local row l sits at off[slice] + 64 k + (l % 64), so the
// thread-per-row loads of entry k are coalesced over a wave.  width[s] = 64 * longest row.
#ifndef DHIP_SOLVER_KERNELS
__global__ void k_sell_width(int64_t n, int64_t per, int spb, int64_t nslices, const int32_t *rp, int32_t *width) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    const int64_t b = s / spb, j = s - b * spb;
    int32_t w = 0;
    for (int64_t l = 64 * j; l < 64 * j + 64 && l < per; ++l) {
        const int64_t r = b * per + l;
        if (r >= n) break;
        const int32_t len = rp[r + 1] - rp[r];
        w = len > w ? len : w;
    }
    width[s] = 64 * w;
}
#endif

#ifndef DHIP_SOLVER_KERNELS
__global__ void k_sell_fill(int64_t n, int64_t per, int spb, const int32_t *rp, const int32_t *col,
                            const double *val, const double *diag, const int32_t *off, int32_t *scol,
                            double *sval, int32_t *pos, float *sval32) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t b = r / per, l = r - b * per;
    const int64_t base = (int64_t)off[b * spb + (l >> 6)] + (l & 63);
    for (int32_t e = rp[r], k = 0; e < rp[r + 1]; ++e, ++k) {
        const int32_t c = col[e];
        const double v = diag ? val[e] / diag[c] : val[e];
        scol[base + 64 * (int64_t)k] = c;
        sval[base + 64 * (int64_t)k] = v;
        if (sval32) sval32[base + 64 * (int64_t)k] = (float)v;
        if (pos) pos[e] = (int32_t)(base + 64 * (int64_t)k);
    }
}
#endif

// new values on an unchanged pattern: one lane per CSR entry through the stored SELL slots
#ifndef DHIP_SOLVER_KERNELS
__global__ void k_sell_vals(int64_t nnz, const int32_t *pos, const int32_t *col, const double *val,
                            const double *diag, double *sval, float *sval32) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const double v = diag ? val[e] / diag[col[e]] : val[e];
    sval[pos[e]] = v;
    if (sval32) sval32[pos[e]] = (float)v;
}
#endif

// a_rc / d_c for right preconditioning (the same quotient the LDS slice loader forms)
#ifndef DHIP_SOLVER_KERNELS
__global__ void k_scale_vals(int64_t nnz, const int32_t *col, const double *val, const double *diag, double *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    out[e] = val[e] / diag[col[e]];
}
#endif

constexpr int KREG = 50;  // basis vectors kept in registers (restart m <= KREG, one row per thread)

// F32: the Arnoldi cycles store the Krylov basis and read the matrix values (LDS slice, SELL copy)
// as float, accumulating every product, dot and combination in double; the restart residual
// b - A x, the Hessenberg least squares and x stay FP64, and a cycle that meets the tolerance by
// its estimate is accepted only when the recomputed FP64 residual meets it too (iterative
// refinement with one GMRES cycle as the inner solve).
template <int RPT, int KR, bool COH, bool F32 = false>
__global__ void __launch_bounds__(PB) k_gmres(GmresArgs a)
#ifndef DHIP_SOLVER_KERNELS
    ;
#else
{
    using BT = typename std::conditional<F32, float, double>::type;
    extern __shared__ __align__(16) unsigned char lds[];
    const uint64_t t_start = a.probe ? wall_clock64() : 0;
    uint64_t t_sync = 0, t_spmv = 0;
    __shared__ double Hc[KMAX + 1];          // current Hessenberg column
    __shared__ double Hs[KMAX * KMAX / 2 + KMAX];  // rotated upper triangle, packed by columns
    __shared__ double cs[KMAX], sn[KMAX], sv[KMAX + 1], yv[KMAX];
    __shared__ double red_sh[KMAX + 2];
    __shared__ double msh[KMAX + 2];         // this block's partial sums
    const int64_t n = a.n;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const unsigned range = row_range_index(a.xcd_map);
    const int64_t r0 = (int64_t)range * per;
    const int64_t r1 = r0 + per < n ? r0 + per : n;
    const bool right = !a.left;
    bool in_lds;
    // right preconditioning: the slice holds a_rc / d_c (A M v); left: plain values (M A v)
    const bool pre = right && a.diag && a.sval;
    const Slice<BT> S = load_slice_scaled<BT>(lds, a.rp, a.col, pre ? a.sval : a.val,
                                              (right && !pre) ? a.diag : nullptr, r0, r1, in_lds);
    const bool scaled = in_lds || pre;
    const double *dg = a.diag;
    // the global CSR values of a slice that is not in LDS (FP64 in both forms)
    const double *gval = (pre ? a.sval : a.val) + S.off;
    // gathers in batches of 8 when the basis rows occupy registers
    constexpr int GB = KR > 0 ? 8 : MAXR;
    const BT *sell_val;
    if constexpr (F32) sell_val = a.sell_val32;
    else sell_val = a.sell_val;
    const bool sell = !in_lds && sell_val;
    const int32_t *sell_col = a.sell_col;
    const int32_t *sell_off = a.sell_off + (sell ? (int64_t)range * a.sell_spb : 0);
    // SELL batches: one wave per CU-quarter runs here, so a batch may hold a whole long row in
    // registers; each batch costs two dependent memory round trips (entries, then gathers)
    constexpr int GBS = KR > 0 ? 8 : (F32 ? 40 : 20);
    auto spmv_row = [&](int64_t l, auto &&yv_) {
        if (sell) {
            // the same sum as row_product, entries read from the coalesced SELL-64 copy
            using XT = decltype(yv_(0));
            const int64_t base = (int64_t)sell_off[l >> 6] + (l & 63);
            const int32_t len = a.rp[r0 + l + 1] - a.rp[r0 + l];
            double s_ = 0.0;
            for (int32_t q0 = 0; q0 < len; q0 += GBS) {
                const int cnt = len - q0 < GBS ? len - q0 : GBS;
                int32_t cc[GBS];
                BT vv[GBS];
                XT xx[GBS];
#pragma unroll
                for (int j = 0; j < GBS; ++j)
                    if (j < cnt) {
                        cc[j] = sell_col[base + 64 * (int64_t)(q0 + j)];
                        vv[j] = sell_val[base + 64 * (int64_t)(q0 + j)];
                    }
#pragma unroll
                for (int j = 0; j < GBS; ++j)
                    if (j < cnt) xx[j] = yv_(cc[j]);
#pragma unroll
                for (int j = 0; j < GBS; ++j)
                    if (j < cnt) s_ += (double)vv[j] * (double)xx[j];
            }
            return s_;
        }
        if (in_lds) return row_product<double, GB>(S.rp, S.col, S.val, S.off, l, yv_);
        if (scaled || !right || !dg) return row_product<double, GB>(S.rp, S.col, gval, S.off, l, yv_);
        return row_product<double, GB>(S.rp, S.col, gval, S.off, l, [&](int32_t c) { return yv_(c) / dg[c]; });
    };
    auto col_of = [](int k) { return k * (k + 1) / 2; };  // start of packed column k (rows 0..k)
    const int m = a.m;
    uint32_t gen = 0;
    auto syncv = [&](int K, double *out) {
        const uint64_t t0 = a.probe ? wall_clock64() : 0;
        const bool ok = sync_reduce_vec<PB, !COH>(a.slots, a.parts, ++gen, msh, K, out);
        if (a.probe) t_sync += wall_clock64() - t0;
        return ok;
    };
    int status = 1, j = 1;
    double resid = 0.0, normb = 1.0, beta = 0.0;
    bool own[RPT];
    int64_t lr[RPT];
    double xr[RPT], wr[RPT], dr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        lr[k] = threadIdx.x + (int64_t)k * PB;
        own[k] = r0 + lr[k] < r1;
        if (own[k]) {
            xr[k] = a.x[r0 + lr[k]];
            dr[k] = dg ? dg[r0 + lr[k]] : 1.0;
        }
    }
    auto prec = [&](int k, double v) { return dg ? 1.0 * v / dr[k] : v; };
    // One row per thread: the thread's entries of the Krylov basis also live in registers, so
    // the dot products and basis combinations read no memory; the global basis holds the
    // vectors the SpMV gathers from other blocks.
    static_assert(KR == 0 || RPT == 1, "register basis needs one row per thread");
    constexpr bool REGV = KR > 0;
    // Register basis: vector q of the current cycle sits in slot KR - nvec + q; a new vector
    // shifts the slots down by one and enters at the top, so every register index is static.
    double vreg[REGV ? KR : 1];
    int nvec = 0;
    // Without a register basis and with the CSR slice in global memory, the first KL basis
    // vectors of the block's rows are also kept in LDS: the dots and combinations read them there
    // (each thread reads back only its own rows), the SpMV gathers read the global copy.
    const int64_t nrb = r1 > r0 ? r1 - r0 : 0;
    int KL = 0;
    BT *vl = reinterpret_cast<BT *>(lds);
    BT *V = reinterpret_cast<BT *>(a.V);
    if constexpr (!REGV) {
        if (!in_lds && nrb > 0) {
            const int64_t fit = a.lds_bytes / ((int64_t)sizeof(BT) * nrb);
            KL = (int)(fit < a.m ? fit : a.m);
        }
    }
    auto vbasis = [&](int q, int64_t l) -> double {
        return q < KL ? vl[(size_t)q * nrb + l] : V[(size_t)q * n + r0 + l];
    };
    auto vstore = [&](int q, int64_t r, double v, bool mine) {
        const BT vb = (BT)v;
        if (mine) {
            xstore<COH>(&V[(size_t)q * n + r], vb);
            if (q < KL) vl[(size_t)q * nrb + (r - r0)] = vb;
        }
        if constexpr (REGV) {
#pragma unroll
            for (int t = 0; t < KR - 1; ++t) vreg[t] = vreg[t + 1];
            vreg[KR - 1] = (double)vb;
        }
    };
    // block sums of w . V[q] (q < K) and, with `self`, w . w as entry K
    __shared__ double wave_part[PB / 64][KMAX + 2];
    auto dots = [&](int K, bool self, double *out) {
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int KT = K + (self ? 1 : 0);
        if constexpr (REGV) {
            const int off = KR - nvec;  // slot of vector 0
#pragma unroll
            for (int c0 = 0; c0 < KR; c0 += 8) {
                if (c0 + 8 > off) {
                    double p[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int t = c0 + u, q = t - off;
                        p[u] = (t < KR && q >= 0 && q < K && own[0]) ? wr[0] * vreg[t < KR ? t : KR - 1] : 0.0;
                    }
                    const double v = wave_sum8(p);
                    const int q = c0 + (lane >> 3) - off;
                    if ((lane & 7) == 0 && q >= 0 && q < K) wave_part[wv][q] = v;
                }
            }
            if (self) {
                const double v = wave_sum(own[0] ? wr[0] * wr[0] : 0.0);
                if (lane == 0) wave_part[wv][K] = v;
            }
        } else {
            for (int c0 = 0; c0 < KT; c0 += 8) {
                double p[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int q = c0 + u;
                    double sacc = 0.0;
                    if (q < KT) {
#pragma unroll
                        for (int k = 0; k < RPT; ++k)
                            if (own[k]) sacc += wr[k] * (q < K ? vbasis(q, lr[k]) : wr[k]);
                    }
                    p[u] = sacc;
                }
                const double v = wave_sum8(p);
                if ((lane & 7) == 0 && c0 + (lane >> 3) < KT) wave_part[wv][c0 + (lane >> 3)] = v;
            }
        }
        __syncthreads();
        if (threadIdx.x < KT) {
            double sacc = 0.0;
#pragma unroll
            for (int ww = 0; ww < PB / 64; ++ww) sacc += wave_part[ww][threadIdx.x];
            out[threadIdx.x] = sacc;
        }
        __syncthreads();
    };
    // residual r = b - A x (left: M r) into wr; returns block partials |r|^2 (and |b|^2 or |M b|^2)
    auto residual = [&](bool with_b, double &p_rr, double &p_bb) {
        p_rr = 0.0;
        p_bb = 0.0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            const int64_t r = r0 + lr[k];
            const double ax =
                row_product<double>(a.rp, a.col, a.val, 0, r, [&](int32_t c) { return xload<COH>(&a.x[c]); });
            double rv = a.b[r] - ax;
            double bv = a.b[r];
            if (!right) {
                rv = prec(k, rv);
                bv = prec(k, bv);
            }
            wr[k] = rv;
            p_rr += rv * rv;
            if (with_b) p_bb += bv * bv;
        }
        p_rr = block_sum<PB>(p_rr);
        if (with_b) p_bb = block_sum<PB>(p_bb);
    };
    // x += (M) V y over columns 0..kk of the basis
    // w - sign * sum_{q < cnt} h_q V[q][r], subtracting in q order (registers, or gathers of 8
    // basis entries issued before their first use)
    auto combine_V = [&](double w, int64_t r, int cnt, const double *h, double sign) {
        if constexpr (REGV) {
            const int off = KR - nvec;
#pragma unroll
            for (int t = 0; t < KR; ++t) {
                const int q = t - off;
                if (q >= 0 && q < cnt) w -= sign * (h[q] * vreg[t]);
            }
        } else {
            for (int c0 = 0; c0 < cnt; c0 += 8) {
                double vv[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) vv[q] = c0 + q < cnt ? vbasis(c0 + q, r - r0) : 0.0;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (c0 + q < cnt) w -= sign * (h[c0 + q] * vv[q]);
            }
        }
        return w;
    };
    auto update_x = [&](int kk) {
        if (threadIdx.x == 0) {
            for (int i = kk; i >= 0; --i) yv[i] = sv[i];
            for (int i = kk; i >= 0; --i) {
                yv[i] /= Hs[col_of(i) + i];
                for (int jj = i - 1; jj >= 0; --jj) yv[jj] -= Hs[col_of(i) + jj] * yv[i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            const int64_t r = r0 + lr[k];
            const double z = combine_V(0.0, r, kk + 1, yv, -1.0);
            xr[k] += right ? prec(k, z) : z;
        }
    };

    {
        double p_rr, p_bb;
        residual(true, p_rr, p_bb);
        if (threadIdx.x == 0) {
            msh[0] = p_rr;
            msh[1] = p_bb;
        }
        if (!syncv(2, red_sh)) { status = 2; goto done; }
        beta = sqrt(red_sh[0]);
        normb = sqrt(red_sh[1]);
        if (normb == 0.0 || !a.rel) normb = 1.0;
        resid = beta / normb;
        if (resid <= a.tol) {
            status = 0;
            j = 0;
            goto done;
        }
    }
    while (j <= a.max_iter) {
        // v0 = r / beta
        {
            const double inv = 1.0 / beta;
            nvec = 0;
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                vstore(0, r0 + lr[k], own[k] ? wr[k] * inv : 0.0, own[k]);
            nvec = 1;
            if (threadIdx.x < KMAX + 1) sv[threadIdx.x] = 0.0;
            if (threadIdx.x == 0) sv[0] = beta;
            if (!syncv(0, red_sh)) { status = 2; goto done; }
        }
        int i;
        bool converged = false;
        for (i = 0; i < m - 1 && j <= a.max_iter; ++i, ++j) {
            const BT *vi = V + (size_t)i * n;
            // w = A M v_i (right) or M A v_i (left); h1 = V^T w
            const uint64_t t_sp = a.probe ? wall_clock64() : 0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                double w = spmv_row(lr[k], [&](int32_t c) { return xload<COH>(&vi[c]); });
                if (!right) w = prec(k, w);
                wr[k] = w;
            }
            if (a.probe) t_spmv += wall_clock64() - t_sp;
            dots(i + 1, false, msh);
            if (!syncv(i + 1, red_sh)) { status = 2; goto done; }
            if (threadIdx.x <= i) Hc[threadIdx.x] = red_sh[threadIdx.x];
            __syncthreads();
            // w1 = w - V h1; h2 = V^T w1; |w1|^2
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                const int64_t r = r0 + lr[k];
                wr[k] = combine_V(wr[k], r, i + 1, Hc, 1.0);
            }
            dots(i + 1, true, msh);
            if (!syncv(i + 2, red_sh)) { status = 2; goto done; }
            double hn2 = red_sh[i + 1];
            for (int kk = 0; kk <= i; ++kk) hn2 -= red_sh[kk] * red_sh[kk];
            const double hnorm = sqrt(hn2 > 0.0 ? hn2 : 0.0);
            // v_{i+1} = (w1 - V h2) / |.|
            {
                const double inv = 1.0 / hnorm;
                double vnew[RPT];
#pragma unroll
                for (int k = 0; k < RPT; ++k)
                    vnew[k] = own[k] ? combine_V(wr[k], r0 + lr[k], i + 1, red_sh, 1.0) * inv : 0.0;
#pragma unroll
                for (int k = 0; k < RPT; ++k) vstore(i + 1, r0 + lr[k], vnew[k], own[k]);
                nvec = i + 2;
            }
            // Hessenberg column i: h1 + h2 (one thread per entry), then the Givens rotations by one
            // thread with the rotated entry carried in a register (num/krylovsolver.h:258-280, 392-399)
            if (threadIdx.x <= (unsigned)i) Hc[threadIdx.x] += red_sh[threadIdx.x];
            __syncthreads();
            if (threadIdx.x == 0) {
                double x = Hc[0];
                for (int kk = 0; kk < i; ++kk) {
                    const double y = Hc[kk + 1], ck = cs[kk], sk = sn[kk];
                    Hs[col_of(i) + kk] = ck * x + sk * y;
                    x = -sk * x + ck * y;
                }
                const double dx = x, dy = hnorm;
                if (dy == 0.0) {
                    cs[i] = 1.0;
                    sn[i] = 0.0;
                } else {
                    const double rr = sqrt(dx * dx + dy * dy);
                    cs[i] = dx / rr;
                    sn[i] = dy / rr;
                }
                Hs[col_of(i) + i] = cs[i] * dx + sn[i] * dy;
                {
                    const double t = cs[i] * sv[i] + sn[i] * sv[i + 1];
                    sv[i + 1] = -sn[i] * sv[i] + cs[i] * sv[i + 1];
                    sv[i] = t;
                }
            }
            __syncthreads();
            resid = fabs(sv[i + 1]) / normb;
            if (resid <= a.tol) {
                update_x(i);
                converged = true;
                break;
            }
            // v_{i+1} visible to all blocks before the next SpMV
            if (!syncv(0, red_sh)) { status = 2; goto done; }
        }
        if (converged && !F32) {
            status = 0;
            break;
        }
        // FP32 cycles: a converged estimate is checked against the FP64 residual below
        if (!converged) update_x(i - 1);
        // restart: publish x, r = b - A x (left: M r)
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (own[k]) xstore<COH>(&a.x[r0 + lr[k]], xr[k]);
        if (!syncv(0, red_sh)) { status = 2; goto done; }
        {
            double p_rr, p_bb;
            residual(false, p_rr, p_bb);
            if (threadIdx.x == 0) msh[0] = p_rr;
            if (!syncv(1, red_sh)) { status = 2; goto done; }
            beta = sqrt(red_sh[0]);
            resid = beta / normb;
            if (resid <= a.tol) {
                status = 0;
                break;
            }
        }
    }
done:
    if (status != 2) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (own[k]) a.x[r0 + lr[k]] = xr[k];
    }
    if (a.probe && threadIdx.x == 0) {
        const uint64_t total = wall_clock64() - t_start;
        a.probe[blockIdx.x * 4 + 0] = t_spmv;   // part of the compute time spent in the SpMV
        a.probe[blockIdx.x * 4 + 1] = total - t_sync;
        a.probe[blockIdx.x * 4 + 2] = t_sync;
        a.probe[blockIdx.x * 4 + 3] = total;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.out_iter = status == 0 ? j : a.max_iter;
        *a.out_resid = resid;
        *a.out_status = status;
    }
}
#endif

// ---------------------------------------------------------------------------
// persistent PCG with FP32 inner iterations and FP64 iterative refinement
//   outer (FP64): r = b - A x with the FP64 values, resid = |r|/|b|, stop at tol
//   inner (FP32 values and vectors, FP64 dot accumulation): PCG on A d = r/|r|
//         from d = 0 until |r_in| <= inner_tol or inner_max iterations
//   update (FP64): x += |r| d
// ---------------------------------------------------------------------------
struct PcgMixArgs {
    int64_t n;
    const int32_t *rp, *col;
    const double *val, *diag;   // FP64 values (outer residual) and diagonal
    const float *val32;
    const double *b;
    double *x;                  // FP64 solution (exchanged for the outer residual)
    float *z, *p0, *p1;         // FP32 exchange vectors
    SyncSlot *slots;
    int max_iter;               // total inner iterations allowed
    double tol;
    int rel;
    double inner_tol;
    int inner_max;
    int *out_iter;
    double *out_resid;
    int *out_status;
    int *out_outer;
    int xcd_map = 0;            // 1: the blocks of one XCD own one contiguous run of rows
};

template <int RPT>
__global__ void __launch_bounds__(PB) k_pcg_mixed(PcgMixArgs a)
#ifndef DHIP_SOLVER_KERNELS
    ;
#else
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)row_range_index(a.xcd_map) * per;
    const int64_t r1 = r0 + per < a.n ? r0 + per : a.n;
    const Slice<float> S = load_slice<float>(lds, a.rp, a.col, a.val32, r0, r1);
    uint32_t gen = 0;
    double red[3];
    int status = 1, total_it = 0, outer = 0;
    double resid = 0.0, normb = 1.0;

    double xr[RPT], dd[RPT];
    float df[RPT], dv[RPT], rin[RPT], zv[RPT], pv[RPT];
    bool own[RPT];
    int64_t lr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        lr[k] = threadIdx.x + (int64_t)k * PB;
        own[k] = r0 + lr[k] < r1;
        if (own[k]) {
            xr[k] = a.x[r0 + lr[k]];
            dd[k] = a.diag[r0 + lr[k]];
            df[k] = (float)dd[k];
        }
    }
    bool first = true;
    for (;;) {
        // ---- FP64 residual ----
        double p_rr = 0.0, p_bb = 0.0, r64[RPT];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            const int64_t r = r0 + lr[k];
            const double s = row_product<double>(a.rp, a.col, a.val, 0, r, [&](int32_t c) { return a.x[c]; });
            const double bv = a.b[r];
            r64[k] = bv - s;
            p_rr += r64[k] * r64[k];
            p_bb += bv * bv;
        }
        p_rr = block_sum<PB>(p_rr);
        if (first) p_bb = block_sum<PB>(p_bb);
        if (!sync_reduce<PB>(a.slots, ++gen, p_rr, p_bb, 0.0, red)) { status = 2; break; }
        const double rnorm = sqrt(red[0]);
        if (first) {
            normb = sqrt(red[1]);
            if (normb == 0.0 || !a.rel) normb = 1.0;
            first = false;
        }
        resid = rnorm / normb;
        if (resid <= a.tol) { status = 0; break; }
        if (total_it >= a.max_iter) { status = 1; break; }
        ++outer;
        // ---- FP32 inner PCG on A d = r/|r| ----
        const double scale = 1.0 / rnorm;
        double p_zr = 0.0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            const int64_t r = r0 + lr[k];
            rin[k] = (float)(r64[k] * scale);
            dv[k] = 0.0f;
            pv[k] = 0.0f;
            zv[k] = rin[k] / df[k];
            a.z[r] = zv[k];
            a.p0[r] = 0.0f;
            p_zr += (double)zv[k] * (double)rin[k];
        }
        p_zr = block_sum<PB>(p_zr);
        if (!sync_reduce<PB>(a.slots, ++gen, p_zr, 0.0, 0.0, red)) { status = 2; break; }
        double rho = red[0];
        float beta = 0.0f;
        bool fail = false;
        for (int it = 1; it <= a.inner_max && total_it < a.max_iter; ++it) {
            ++total_it;
            float *pn = (it & 1) ? a.p1 : a.p0;
            const float *po = (it & 1) ? a.p0 : a.p1;
            const float *zg = a.z;
            float qv[RPT];
            double p_lam = 0.0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                const int64_t r = r0 + lr[k];
                pv[k] = zv[k] + beta * pv[k];
                pn[r] = pv[k];
                const float s = row_product<float>(S.rp, S.col, S.val, S.off, lr[k],
                                                   [&](int32_t c) { return zg[c] + beta * po[c]; });
                qv[k] = s;
                p_lam += (double)s * (double)pv[k];
            }
            p_lam = block_sum<PB>(p_lam);
            if (!sync_reduce<PB>(a.slots, ++gen, p_lam, 0.0, 0.0, red)) { fail = true; break; }
            const float alpha = (float)(rho / red[0]);
            double p_rr2 = 0.0, p_zr2 = 0.0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (!own[k]) continue;
                const int64_t r = r0 + lr[k];
                dv[k] += alpha * pv[k];
                rin[k] -= alpha * qv[k];
                p_rr2 += (double)rin[k] * (double)rin[k];
                zv[k] = rin[k] / df[k];
                a.z[r] = zv[k];
                p_zr2 += (double)zv[k] * (double)rin[k];
            }
            p_rr2 = block_sum<PB>(p_rr2);
            p_zr2 = block_sum<PB>(p_zr2);
            if (!sync_reduce<PB>(a.slots, ++gen, p_rr2, p_zr2, 0.0, red)) { fail = true; break; }
            if (sqrt(red[0]) <= a.inner_tol) break;
            const double rho1 = rho;
            rho = red[1];
            beta = (float)(rho / rho1);
        }
        if (fail) { status = 2; break; }
        // ---- FP64 refinement update, published for the next outer residual ----
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!own[k]) continue;
            xr[k] += rnorm * (double)dv[k];
            a.x[r0 + lr[k]] = xr[k];
        }
        if (!sync_reduce<PB>(a.slots, ++gen, 0.0, 0.0, 0.0, red)) { status = 2; break; }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.out_iter = total_it;
        *a.out_resid = resid;
        *a.out_status = status;
        *a.out_outer = outer;
    }
}
#endif

}  // namespace dhip
