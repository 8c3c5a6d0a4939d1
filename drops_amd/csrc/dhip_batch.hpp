// dhip_batch.hpp -- many independent P1 trace-FEM problems assembled and solved in one launch.
//
// Each problem has its own Kuhn brick, level set and sizes.  One workgroup owns one problem at a
// time and runs the whole path for it with workgroup barriers only (no grid synchronisation):
// classification + compaction, local matrices, first-encounter numbering, CSR rows, and a Jacobi
// PCG with FP32 inner iterations and FP64 refinement.  Workgroups take problems from a queue
// ordered largest first (longest-processing-time order), so uneven sizes balance over the grid.
// The per-item code is the single-problem path's (cell_cut_mask, local_p1_one, qualify_one,
// first_one, rowcount_one, rowfill_one), so every assembled array equals what the single-problem
// entry points produce.
//
//   k_batch_cells / k_batch_verts   planning: cut tets and cut-tet vertices per problem (sizes of
//                                   the per-problem scratch; the rows of a problem are at most its
//                                   cut-tet vertices, each row has at most 15 entries)
//   k_batch<BB>                     the persistent solve, BB threads per problem
//   k_batch_solve<BB>               with the tuning key batch_split: k_batch stops after assembly and
//                                   this second launch runs the PCGs, BB (up to 1024) threads per
//                                   problem, largest first; the solve is bandwidth-bound per problem,
//                                   so more threads than the assembly's workgroup shorten the span
#pragma once

#include "dhip_assembly.hpp"
#include "dhip_krylov.hpp"

namespace dhip {

struct BatchProb {
    MeshP m;
    int64_t phi_off;    // first level-set value of the problem
    int64_t cell_off;   // first cell word
    int64_t cut_off;    // first cut-tet slot (capacity ncut_cap)
    int64_t ncut_cap;
    int64_t row_off;    // first row slot (capacity row_cap; row pointers at row_off + problem index)
    int64_t row_cap;
    int64_t nz_off;     // first entry slot (capacity nz_cap)
    int64_t nz_cap;
    double lc[3];       // the solved operator combination
    int64_t sell_off;   // first SELL-64 slot (capacity 15 * 64 * ceil(row_cap / 64))
    int64_t slice_off;  // first slice offset (capacity ceil(row_cap / 64) + 1)
};

struct BatchStat {
    int32_t status;     // DHIP_OK, DHIP_E_NOT_CONVERGED, DHIP_E_ZERO_NODE, DHIP_E_ZERO_PATTERN, DHIP_E_CAPACITY
    int32_t ncut, ndof, nnz;
    int32_t iters, outer;
    double resid;
    uint64_t t0, t1;    // wall clock (100 MHz) at the start and end of the problem
    uint64_t t_asm;     // wall clock when its CSR system was complete
};

// ---------------------------------------------------------------------------
// planning
// ---------------------------------------------------------------------------
constexpr int BPLAN = 256;

__global__ void __launch_bounds__(BPLAN) k_batch_cells(const BatchProb *prob, const double *phi, uint8_t *mask,
                                                       int32_t *ncut, int32_t *zero) {
    const BatchProb &P = prob[blockIdx.y];
    const int64_t cell = (int64_t)blockIdx.x * BPLAN + threadIdx.x;
    int mk = 0;
    bool z = false;
    if (cell < P.m.ncell) {
        mk = cell_cut_mask(P.m, phi + P.phi_off, cell, z);
        mask[P.cell_off + cell] = (uint8_t)mk;
    }
    const double cnt = block_sum<BPLAN>((double)__popc(mk));
    const double zs = block_sum<BPLAN>(z ? 1.0 : 0.0);
    if (threadIdx.x == 0) {
        if (cnt > 0.0) atomicAdd(&ncut[blockIdx.y], (int32_t)cnt);
        if (zs > 0.0) atomicOr(&zero[blockIdx.y], 1);
    }
}

__global__ void __launch_bounds__(BPLAN) k_batch_verts(const BatchProb *prob, const uint8_t *mask, int32_t *nvert) {
    const BatchProb &P = prob[blockIdx.y];
    const MeshP &m = P.m;
    const int64_t nv = (int64_t)(m.n[0] + 1) * (m.n[1] + 1) * (m.n[2] + 1);
    const int64_t v = (int64_t)blockIdx.x * BPLAN + threadIdx.x;
    bool in = false;
    if (v < nv) {
        const int64_t x1 = v % (m.n[0] + 1), x2 = (v / (m.n[0] + 1)) % (m.n[1] + 1),
                      x3 = v / ((int64_t)(m.n[0] + 1) * (m.n[1] + 1));
        for (int d = 0; d < 8 && !in; ++d) {
            const int64_t c1 = x1 - (d & 1), c2 = x2 - ((d >> 1) & 1), c3 = x3 - ((d >> 2) & 1);
            if (c1 < 0 || c2 < 0 || c3 < 0 || c1 >= m.n[0] || c2 >= m.n[1] || c3 >= m.n[2]) continue;
            const int mk = mask[P.cell_off + (c3 * m.n[1] + c2) * m.n[0] + c1];
            for (int k = 0; k < 6; ++k)
                if (((mk >> k) & 1) && c_tab.inc[d][k] >= 0) in = true;
        }
    }
    const double cnt = block_sum<BPLAN>(in ? 1.0 : 0.0);
    if (threadIdx.x == 0 && cnt > 0.0) atomicAdd(&nvert[blockIdx.y], (int32_t)cnt);
}

// exclusive scan of a[0..n) in place by the whole workgroup; returns the total
template <int BB>
__device__ inline int64_t wg_scan_inplace(int32_t *a, int64_t n) {
    int64_t carry = 0;
    for (int64_t i0 = 0; i0 < n; i0 += BB) {
        const int64_t i = i0 + threadIdx.x;
        const int v = i < n ? a[i] : 0;
        int tot;
        const int off = block_excl_scan<BB>(v, &tot);
        if (i < n) a[i] = (int32_t)(carry + off);
        carry += tot;
    }
    __syncthreads();
    return carry;
}

// ---------------------------------------------------------------------------
// pre-pass (tuning key batch_prepass): classification, compaction and local matrices of every
// problem over the whole grid, so that a large problem's most expensive per-tet work does not
// lengthen its workgroup's span.  The same per-item functions produce the same arrays as the
// in-kernel path.  The largest problem's assembly leaves its span, but every problem then enters its
// solve at once and the solves slow down; which effect wins differs between runs, so the autotuner
// decides per problem set.
//   k_batch_cut_count   per (problem, block of cells): cell masks, the block's cut-tet count,
//                       the problem's exact-zero flag
//   k_batch_cut_scan    per problem: exclusive scan of its block counts
//   k_batch_cut_fill    per (problem, block of cells): cut list in cell order and cell words
//   k_batch_local       one thread per cut tet of any problem: local_p1_one
// ---------------------------------------------------------------------------
struct PrepassArgs {
    const BatchProb *prob;
    int32_t nprob;
    const double *phi;
    uint8_t *cellmask;          // per problem cell (cell_off layout)
    int32_t *blkcnt;            // [nprob][max_blocks]
    int32_t max_blocks;
    int32_t *zero;              // per problem: exact-zero node in a cut tet
    int32_t *perr;              // per problem: local error bits (2: all-zero sub-tet)
    uint32_t *cellinfo;
    int64_t *cut;
    int64_t cut_total;
    uint8_t *flags;
    double *loc;
    double coef0, coef1, coef2;
    int rhs_kind;
    double rhs_p0;
};

__global__ void __launch_bounds__(BPLAN) k_batch_cut_count(PrepassArgs a) {
    const int p = blockIdx.y;
    const BatchProb &P = a.prob[p];
    const int64_t cell = (int64_t)blockIdx.x * BPLAN + threadIdx.x;
    int mk = 0;
    bool z = false;
    if (cell < P.m.ncell) {
        mk = cell_cut_mask(P.m, a.phi + P.phi_off, cell, z);
        a.cellmask[P.cell_off + cell] = (uint8_t)mk;
    }
    int tot;
    block_excl_scan<BPLAN>(__popc(mk), &tot);
    const double zs = block_sum<BPLAN>(z ? 1.0 : 0.0);
    if (threadIdx.x == 0) {
        a.blkcnt[(int64_t)p * a.max_blocks + blockIdx.x] = tot;
        if (zs > 0.0) atomicOr(&a.zero[p], 1);
    }
}

__global__ void __launch_bounds__(BPLAN) k_batch_cut_scan(PrepassArgs a) {
    const int p = blockIdx.x;
    const int64_t nb = (a.prob[p].m.ncell + BPLAN - 1) / BPLAN;
    wg_scan_inplace<BPLAN>(a.blkcnt + (int64_t)p * a.max_blocks, nb);
}

__global__ void __launch_bounds__(BPLAN) k_batch_cut_fill(PrepassArgs a) {
    const int p = blockIdx.y;
    const BatchProb &P = a.prob[p];
    const int64_t cell = (int64_t)blockIdx.x * BPLAN + threadIdx.x;
    const int mk = cell < P.m.ncell ? a.cellmask[P.cell_off + cell] : 0;
    const int off = block_excl_scan<BPLAN>(__popc(mk), nullptr);
    if (cell >= P.m.ncell) return;
    int64_t b = (int64_t)a.blkcnt[(int64_t)p * a.max_blocks + blockIdx.x] + off;
    a.cellinfo[P.cell_off + cell] = ((uint32_t)b << 6) | (uint32_t)mk;
    for (int k = 0; k < 6; ++k)
        if (mk & (1 << k)) {
            if (b < P.ncut_cap) a.cut[P.cut_off + b] = cell * 6 + k;
            ++b;
        }
}

__global__ void __launch_bounds__(LOCAL_BLOCK) k_batch_local(PrepassArgs a) {
    const int64_t g = (int64_t)blockIdx.x * LOCAL_BLOCK + threadIdx.x;
    if (g >= a.cut_total) return;
    // the problem whose cut slots hold g: the last one with cut_off <= g
    int lo = 0, hi = a.nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.prob[mid].cut_off <= g) lo = mid;
        else hi = mid - 1;
    }
    const BatchProb &P = a.prob[lo];
    const int64_t c = g - P.cut_off;
    if (c < 0 || c >= P.ncut_cap) return;
    LocalArgs la{P.m, a.phi + P.phi_off, a.cut + P.cut_off, P.ncut_cap, a.coef0, a.coef1, a.coef2, a.rhs_kind,
                 a.rhs_p0, a.loc + (int64_t)LOC_N * P.cut_off, a.flags + P.cut_off, &a.perr[lo]};
    local_p1_one(la, c);
}

// ---------------------------------------------------------------------------
// the persistent solve
// ---------------------------------------------------------------------------
struct BatchArgs {
    const BatchProb *prob;
    const int32_t *order;       // problem ids, largest first
    int32_t nprob;
    int32_t *queue;             // next position in `order` (zeroed before the launch)
    const double *phi;
    uint32_t *cellinfo;
    int64_t *cut;
    uint8_t *flags;
    double *loc;                // SoA [LOC_N][ncut] per problem
    int32_t *first, *slot, *dofs;
    int32_t *rowowner, *rowlen, *rowptr;
    int32_t *col;
    double *vm, *vl, *vs, *vc;
    double *rhs, *diag, *x, *r64;
    float *d32, *r32, *z32, *p32, *q32;
    int32_t *scol, *soff;       // SELL-64 copy of the FP32 inner matrix
    float *sval;
    int64_t lds_floats;         // dynamic LDS of the launch: the search direction when it fits
    BatchStat *stat;
    double coef0, coef1, coef2;
    int rhs_kind;
    double rhs_p0;
    double omit_bound;
    int zero_throw;
    double tol;
    int max_iter;
    int rel;
    double inner_tol;
    int inner_max;
    int prepass = 0;            // 1: cut lists, cell words, flags and local data come from the pre-pass
    int split = 0;              // 1: k_batch assembles only, k_batch_solve runs the PCGs
    const int32_t *pre_zero = nullptr, *pre_err = nullptr;
};

// two workgroup sums with one LDS round
template <int BB>
__device__ inline void block_sum2(double &u, double &v) {
    __shared__ double ws[2][BB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    u = wave_sum(u);
    v = wave_sum(v);
    if (lane == 0) {
        ws[0][w] = u;
        ws[1][w] = v;
    }
    __syncthreads();
    double su = 0.0, sv = 0.0;
#pragma unroll
    for (int i = 0; i < BB / 64; ++i) {
        su += ws[0][i];
        sv += ws[1][i];
    }
    __syncthreads();
    u = su;
    v = sv;
}

// Row loops of the inner iteration in groups: a thread's rows r = tid + (g + u BB) are handled G at a
// time with every load of the group issued before the first use, so a group costs one memory round
// trip instead of one per row (the loops are latency-bound: a few dozen rows per thread).
constexpr int SELL_MAXW = 15;   // entries per row of a P1 interface matrix on a Kuhn mesh

template <int BB, int G>
__device__ __forceinline__ double sell_spmv_group(const int32_t *soff, const int32_t *scol, const float *sval,
                                                  const int32_t *rowptr, const float *pv, float *q32, int64_t nd) {
    double p_lam = 0.0;
    for (int64_t r0 = threadIdx.x; r0 < nd; r0 += (int64_t)G * BB) {
        int64_t base[G];
        int32_t len[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int64_t r = r0 + (int64_t)u * BB;
            len[u] = 0;
            base[u] = 0;
            if (r < nd) {
                base[u] = (int64_t)soff[r >> 6] + (r & 63);
                len[u] = rowptr[r + 1] - rowptr[r];
            }
        }
        int32_t cc[G][SELL_MAXW];
        float vv[G][SELL_MAXW];
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int k = 0; k < SELL_MAXW; ++k)
                if (k < len[u]) {
                    cc[u][k] = scol[base[u] + 64 * k];
                    vv[u][k] = sval[base[u] + 64 * k];
                }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int64_t r = r0 + (int64_t)u * BB;
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < SELL_MAXW; ++k)
                if (k < len[u]) s += vv[u][k] * pv[cc[u][k]];
            if (r < nd) {
                q32[r] = s;
                p_lam += (double)s * (double)pv[r];
            }
        }
    }
    return p_lam;
}

template <int BB>
__device__ void batch_pcg(const BatchArgs &a, int p, float *lds_p, BatchStat &st);

// Assembly of problem p, then (solve) its PCG; without solve the stat entry records the assembly
// and k_batch_solve runs the PCG.
template <int BB>
__device__ void batch_problem(const BatchArgs &a, int p, float *lds_p, bool solve) {
    __shared__ int sh_zero, sh_err;   // workgroup flags (LDS: no stale cache lines)
    const BatchProb P = a.prob[p];
    const MeshP m = P.m;
    const int tid = threadIdx.x;
    const double *phi = a.phi + P.phi_off;
    uint32_t *cellinfo = a.cellinfo + P.cell_off;
    int64_t *cut = a.cut + P.cut_off;
    uint8_t *flags = a.flags + P.cut_off;
    double *loc = a.loc + (int64_t)LOC_N * P.cut_off;
    int32_t *first = a.first + 4 * P.cut_off, *slot = a.slot + 4 * P.cut_off, *dofs = a.dofs + 4 * P.cut_off;
    int32_t *rowowner = a.rowowner + P.row_off, *rowlen = a.rowlen + P.row_off, *rowptr = a.rowptr + P.row_off + p;
    int32_t *col = a.col + P.nz_off;
    double *vm = a.vm + P.nz_off, *vl = a.vl + P.nz_off, *vs = a.vs + P.nz_off, *vc = a.vc + P.nz_off;
    double *rhs = a.rhs + P.row_off, *diag = a.diag + P.row_off, *x = a.x + P.row_off, *r64 = a.r64 + P.row_off;
    float *d32 = a.d32 + P.row_off, *r32 = a.r32 + P.row_off, *z32 = a.z32 + P.row_off;
    float *p32 = a.p32 + P.row_off, *q32 = a.q32 + P.row_off;
    BatchStat st{};
    st.t0 = wall_clock64();
    st.status = DHIP_OK;

    // ---- classification + compaction in cell order ----
    if (tid == 0) {
        sh_zero = 0;
        sh_err = 0;
    }
    __syncthreads();
    int64_t ncut = 0;
    bool over = false;
    if (a.prepass) {
        ncut = P.ncut_cap;   // the planned count is exact: same level set, same classification
        if (a.pre_zero[p]) sh_zero = 1;
        if (a.pre_err[p] & 2) sh_err = 2;
    }
    for (int64_t c0 = 0; c0 < (a.prepass ? 0 : m.ncell); c0 += BB) {
        const int64_t cell = c0 + tid;
        int mk = 0;
        if (cell < m.ncell) {
            bool z;
            mk = cell_cut_mask(m, phi, cell, z);
            if (z) sh_zero = 1;
        }
        int tot;
        const int off = block_excl_scan<BB>(__popc(mk), &tot);
        if (cell < m.ncell) {
            int64_t b = ncut + off;
            cellinfo[cell] = ((uint32_t)b << 6) | (uint32_t)mk;
            for (int k = 0; k < 6; ++k)
                if (mk & (1 << k)) {
                    if (b < P.ncut_cap) cut[b] = cell * 6 + k;
                    ++b;
                }
        }
        ncut += tot;
    }
    __syncthreads();
    st.ncut = (int32_t)ncut;
    if (ncut > P.ncut_cap) over = true;
    if (a.zero_throw && sh_zero) st.status = DHIP_E_ZERO_NODE;
    if (over) st.status = DHIP_E_CAPACITY;

    int64_t nd = 0, nnz = 0;
    if (st.status == DHIP_OK) {
        // ---- local matrices ----
        if (!a.prepass) {
            LocalArgs la{m, phi, cut, ncut, a.coef0, a.coef1, a.coef2, a.rhs_kind, a.rhs_p0, loc, flags, &sh_err};
            for (int64_t c = tid; c < ncut; c += BB) local_p1_one(la, c);
        }
        __syncthreads();
        if (sh_err & 2) st.status = DHIP_E_ZERO_PATTERN;
    }
    if (st.status == DHIP_OK) {
        // ---- first-encounter numbering ----
        QualArgs qa{m, cut, ncut, loc, flags, a.omit_bound};
        for (int64_t c = tid; c < ncut; c += BB) qualify_one(qa, c);
        __syncthreads();
        FirstArgs fa{m, cut, ncut, cellinfo, flags, first, slot};
        for (int64_t s = tid; s < 4 * ncut; s += BB) first_one(fa, s);
        __syncthreads();
        nd = wg_scan_inplace<BB>(slot, 4 * ncut);   // slot: owner flag -> owner rank
        if (nd > P.row_cap) st.status = DHIP_E_CAPACITY;
    }
    if (st.status == DHIP_OK) {
        for (int64_t s = tid; s < 4 * ncut; s += BB) {
            const int32_t f = first[s];
            dofs[s] = f >= 0 ? slot[f] : -1;
            if (f == s) rowowner[slot[s]] = (int32_t)s;
        }
        __syncthreads();
        // ---- CSR rows ----
        RowArgs ra{m, cut, ncut, cellinfo, flags, dofs, rowowner, nd, loc, rowlen, rowptr, col,
                   vm, vl, vs, vc, rhs, diag, P.lc[0], P.lc[1], P.lc[2]};
        for (int64_t r = tid; r < nd; r += BB) rowcount_one(ra, r);
        __syncthreads();
        for (int64_t r = tid; r < nd; r += BB) rowptr[r] = rowlen[r];
        __syncthreads();
        nnz = wg_scan_inplace<BB>(rowptr, nd);
        if (tid == 0) rowptr[nd] = (int32_t)nnz;
        if (nnz > P.nz_cap) st.status = DHIP_E_CAPACITY;
        __syncthreads();
    }
    st.ndof = (int32_t)nd;
    st.nnz = (int32_t)nnz;
    if (st.status == DHIP_OK) {
        RowArgs ra{m, cut, ncut, cellinfo, flags, dofs, rowowner, nd, loc, rowlen, rowptr, col,
                   vm, vl, vs, vc, rhs, diag, P.lc[0], P.lc[1], P.lc[2]};
        for (int64_t r = tid; r < nd; r += BB) {
            rowfill_one(ra, r);
            x[r] = 0.0;
        }
        __syncthreads();
        // SELL-64 copy of the FP32 matrix: entry k of row r at soff[r / 64] + 64 k + r % 64, so the
        // thread-per-row loads of a wave (64 consecutive rows) coalesce
        int32_t *scol = a.scol + P.sell_off, *soff = a.soff + P.slice_off;
        float *sval = a.sval + P.sell_off;
        const int64_t nsl = (nd + 63) / 64;
        for (int64_t sl = tid; sl < nsl; sl += BB) {
            int32_t w = 0;
            for (int64_t r = 64 * sl; r < 64 * sl + 64 && r < nd; ++r) {
                const int32_t len = rowptr[r + 1] - rowptr[r];
                w = len > w ? len : w;
            }
            soff[sl] = 64 * w;
        }
        __syncthreads();
        wg_scan_inplace<BB>(soff, nsl);
        for (int64_t r = tid; r < nd; r += BB) {
            const int64_t base = (int64_t)soff[r >> 6] + (r & 63);
            for (int32_t e = rowptr[r], k = 0; e < rowptr[r + 1]; ++e, ++k) {
                scol[base + 64 * (int64_t)k] = col[e];
                sval[base + 64 * (int64_t)k] = (float)vc[e];
            }
        }
        __syncthreads();
        st.t_asm = wall_clock64();
        if (solve) batch_pcg<BB>(a, p, lds_p, st);
    }
    st.t1 = wall_clock64();
    if (tid == 0) a.stat[p] = st;
    __syncthreads();
}

// Jacobi PCG of an assembled problem: FP64 residual and refinement, FP32 inner iterations on the
// SELL-64 copy; st receives iterations, residual and status.
template <int BB>
__device__ void batch_pcg(const BatchArgs &a, int p, float *lds_p, BatchStat &st) {
    const BatchProb P = a.prob[p];
    const int tid = threadIdx.x;
    const int64_t nd = st.ndof;
    const int32_t *rowptr = a.rowptr + P.row_off + p;
    const int32_t *col = a.col + P.nz_off;
    const double *vc = a.vc + P.nz_off;
    const double *rhs = a.rhs + P.row_off, *diag = a.diag + P.row_off;
    double *x = a.x + P.row_off, *r64 = a.r64 + P.row_off;
    float *d32 = a.d32 + P.row_off, *r32 = a.r32 + P.row_off, *z32 = a.z32 + P.row_off;
    float *p32 = a.p32 + P.row_off, *q32 = a.q32 + P.row_off;
    const int32_t *scol = a.scol + P.sell_off, *soff = a.soff + P.slice_off;
    const float *sval = a.sval + P.sell_off;
    {

        // ---- Jacobi PCG: FP64 residual and refinement, FP32 inner iterations ----
        int total_it = 0, outer = 0, status = 1;
        double normb = 1.0, resid = 0.0;
        bool firstpass = true;
        for (;;) {
            double p_rr = 0.0, p_bb = 0.0;
            for (int64_t r = tid; r < nd; r += BB) {
                const double s =
                    row_product<double, (BB >= 1024 ? 8 : MAXR)>(rowptr, col, vc, 0, r, [&](int32_t c) { return x[c]; });
                const double bv = rhs[r];
                const double rv = bv - s;
                r64[r] = rv;
                p_rr += rv * rv;
                p_bb += bv * bv;
            }
            p_rr = block_sum<BB>(p_rr);
            if (firstpass) {
                p_bb = block_sum<BB>(p_bb);
                normb = sqrt(p_bb);
                if (normb == 0.0 || !a.rel) normb = 1.0;
                firstpass = false;
            }
            const double rnorm = sqrt(p_rr);
            resid = rnorm / normb;
            if (resid <= a.tol) { status = 0; break; }
            if (total_it >= a.max_iter) { status = 1; break; }
            ++outer;
            const double scale = 1.0 / rnorm;
            // the search direction lives in LDS when the problem fits (the SpMV gathers read it)
            float *pv = nd <= a.lds_floats ? lds_p : p32;
            double p_zr = 0.0;
            for (int64_t r = tid; r < nd; r += BB) {
                const float rin = (float)(r64[r] * scale);
                const float zv = rin / (float)diag[r];
                r32[r] = rin;
                z32[r] = zv;
                d32[r] = 0.0f;
                pv[r] = 0.0f;
                p_zr += (double)zv * (double)rin;
            }
            double rho = block_sum<BB>(p_zr);
            float beta = 0.0f;
            for (int it = 1; it <= a.inner_max && total_it < a.max_iter; ++it) {
                ++total_it;
                constexpr int GU = BB >= 1024 ? 4 : 8;   // rows per group of the vector loops
                for (int64_t r0 = tid; r0 < nd; r0 += (int64_t)GU * BB) {
                    float zl[GU], pl[GU];
#pragma unroll
                    for (int u = 0; u < GU; ++u) {
                        const int64_t r = r0 + (int64_t)u * BB;
                        if (r < nd) {
                            zl[u] = z32[r];
                            pl[u] = pv[r];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < GU; ++u) {
                        const int64_t r = r0 + (int64_t)u * BB;
                        if (r < nd) pv[r] = zl[u] + beta * pl[u];
                    }
                }
                __syncthreads();
                const double lam =
                    block_sum<BB>(sell_spmv_group<BB, (BB >= 1024 ? 1 : 4)>(soff, scol, sval, rowptr, pv, q32, nd));
                const float alpha = (float)(rho / lam);
                double p_rr2 = 0.0, p_zr2 = 0.0;
                for (int64_t r0 = tid; r0 < nd; r0 += (int64_t)GU * BB) {
                    float dl[GU], rl[GU], ql[GU], pl[GU];
                    double gl[GU];
#pragma unroll
                    for (int u = 0; u < GU; ++u) {
                        const int64_t r = r0 + (int64_t)u * BB;
                        if (r < nd) {
                            dl[u] = d32[r];
                            rl[u] = r32[r];
                            ql[u] = q32[r];
                            gl[u] = diag[r];
                            pl[u] = pv[r];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < GU; ++u) {
                        const int64_t r = r0 + (int64_t)u * BB;
                        if (r < nd) {
                            d32[r] = dl[u] + alpha * pl[u];
                            const float rin = rl[u] - alpha * ql[u];
                            r32[r] = rin;
                            const float zv = rin / (float)gl[u];
                            z32[r] = zv;
                            p_rr2 += (double)rin * (double)rin;
                            p_zr2 += (double)zv * (double)rin;
                        }
                    }
                }
                block_sum2<BB>(p_rr2, p_zr2);
                if (sqrt(p_rr2) <= a.inner_tol) break;
                const double rho1 = rho;
                rho = p_zr2;
                beta = (float)(rho / rho1);
            }
            for (int64_t r = tid; r < nd; r += BB) x[r] += rnorm * (double)d32[r];
            __syncthreads();
        }
        st.iters = total_it;
        st.outer = outer;
        st.resid = resid;
        st.status = status == 0 ? DHIP_OK : DHIP_E_NOT_CONVERGED;
    }
}

template <int BB>
__global__ void __launch_bounds__(BB) k_batch(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int sh_p;
    for (;;) {
        if (threadIdx.x == 0) {
            const int q = atomicAdd(a.queue, 1);
            sh_p = q < a.nprob ? a.order[q] : -1;
        }
        __syncthreads();
        const int p = sh_p;
        __syncthreads();
        if (p < 0) break;
        batch_problem<BB>(a, p, reinterpret_cast<float *>(lds), !a.split);
    }
}

// The PCG of every assembled problem (split mode): problems from the second queue counter in the
// same largest-first order, BB threads each.
template <int BB>
__global__ void __launch_bounds__(BB) k_batch_solve(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int sh_p;
    for (;;) {
        if (threadIdx.x == 0) {
            const int q = atomicAdd(a.queue + 1, 1);
            sh_p = q < a.nprob ? a.order[q] : -1;
        }
        __syncthreads();
        const int p = sh_p;
        __syncthreads();
        if (p < 0) break;
        BatchStat st = a.stat[p];
        if (st.status == DHIP_OK) {
            batch_pcg<BB>(a, p, reinterpret_cast<float *>(lds), st);
            st.t1 = wall_clock64();
            if (threadIdx.x == 0) a.stat[p] = st;
        }
        __syncthreads();
    }
}

}  // namespace dhip
