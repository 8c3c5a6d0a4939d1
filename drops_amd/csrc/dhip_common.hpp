// dhip_common.hpp -- context, device tables, launch/profiling helpers and
// device-wide scans shared by the assembly and Krylov parts of libdhip.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/dhip.h"

namespace dhip {

// Blocks of a launch are dispatched to the 8 XCDs round robin (block b to XCD b % 8).  Position of
// block b when each XCD's blocks are to own consecutive row ranges: the blocks of the XCDs before
// its own, then b / 8.  A bijection on [0, nb) whatever the dispatch does.
__device__ inline unsigned xcd_position(unsigned b, unsigned nb) {
    const unsigned x = b % 8, q = nb / 8, rem = nb % 8;
    return x * q + (x < rem ? x : rem) + b / 8;
}

// ---------------------------------------------------------------------------
// Device tables (filled once per process from host-side derivations)
// ---------------------------------------------------------------------------
struct Tables {
    int8_t kuhn[6][4][3];      // corner offsets of the Kuhn tets (d1,d2,d3)
    int8_t lat_b2[10][4];      // lattice-2 vertices, barycentric x 2
    int8_t lat_t[8][4];        // lattice-2 sub-tets
    int8_t sub_edge[8][6];     // lattice edge id of sub-tet local edge
    int8_t edge_ab[25][2];     // lattice edge endpoints in first-appearance orientation
    int8_t rp_size[81];        // triangles per sign pattern
    int8_t rp_tri[81][2][3];   // triangle roots: 0-3 zero vertex, 4-9 cut edge + 4
    int8_t rp_bnd[81];         // facet lies in a face of the sub-tet
    int8_t lat_of_p2[10];      // lattice vertex of P2 dof (vertices, edges)
    int8_t inc[8][6];          // local index of corner delta in Kuhn tet k (-1 if absent)
    int8_t slot[6][4][4];      // neighbour slot 0..14 of (k, row vertex, col vertex)
    int8_t tet_node27[6][10];  // 3x3x3 cell-node index of lattice vertex l of tet k
    uint32_t tetmask27[6];     // the 10 cell nodes of tet k as a 27-bit mask
    int8_t p2_node27[6][10];   // 3x3x3 cell-node index of P2 dof d of tet k
    int8_t p2_of27[6][27];     // P2 dof of tet k at cell node (-1 if the tet has no node there)
    int8_t slot_delta[15][3];  // neighbour offset of each slot
    double lat_b[10][4];
    double q2n[4][3], q2w[4];  // Quad2_2DDataCL
    double q5n[7][3], q5w[7];  // Quad5_2DDataCL
    double v5n[15][4];         // Quad5DataCL nodes
    double v5c1, v5c2;         // Quad5CL::quad weights of node groups 1 and 2
};

extern __constant__ Tables c_tab;
void init_tables_once(int device);  // derives the tables on the host and uploads them
const Tables &host_tables();

// ---------------------------------------------------------------------------
// Mesh parameters passed by value to kernels
// ---------------------------------------------------------------------------
struct MeshP {
    double orig[3];
    double off[3][3];  // off[k][c] = 1/n_k * e_k[c]
    int32_t n[3];
    int64_t ncell;
};

// ---------------------------------------------------------------------------
// Device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t b) {
        if (b <= bytes && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            bytes = 0;
            if (e != hipSuccess) return e;
        }
        size_t nb = b < 256 ? 256 : b;
        hipError_t e = hipMalloc(&p, nb);
        if (e == hipSuccess) bytes = nb;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace dhip

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct dhip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    char err[512] = {0};
    int cu_count = 0;
    // resident blocks per CU of a persistent kernel at a dynamic LDS size (queried once)
    std::map<std::pair<const void *, size_t>, int> occupancy;
    int32_t *scan_total_dev = nullptr;   // device total of the last scan_exclusive_i32

    // profiling: per-kernel accumulated milliseconds / launches
    int prof = 0;                     // event timing: 0 off, 1 every launch, 2 persistent solvers only
    struct Pending { std::string name; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int64_t>> prof_acc;

    // mesh
    bool has_mesh = false;
    dhip::MeshP mesh{};
    double e[3][3] = {{0}};
    int64_t nvert = 0;

    // level set
    const double *phi = nullptr;
    int64_t phi_count = 0;
    dhip::DevBuf phi_own;

    uint64_t geo_gen = 1;   // bumped by every new mesh or level set
    // right-hand side values at the patch quadrature points (DHIP_RHS_VALUES)
    dhip::DevBuf qp_off, qp_vals;
    dhip::DevBuf qq_first, qq_gh, qq_list;   // mapped P2 pass: first cut slot of each slot's node, its recovered
                                             // gradient, the list of first slots (+ its count)
    int64_t qp_n = -1, qp_ncut = -1;
    uint64_t qp_gen = 0;
    double qp_band = -2.0;
    bool qp_have_vals = false;
    // the mapped P2 surface's base points handed out for a host load (dhip_quaqua_points / _values)
    dhip::DevBuf qqp_off, qqp_vals;
    int64_t qqp_n = -1, qqp_ncut = -1;
    uint64_t qqp_gen = 0;
    bool qqp_have_vals = false;
    int qqp_maxiter = -1, qqp_maxdamp = -1;
    double qqp_tol = 0., qqp_armijo = 0.;

    // assembly state: 0 none, 1 classified, 2 local, 3 numbered, 4 built
    int stage = 0;
    int space = 1;   // 1: P1 interface space, 2: P2 interface space (set by the local step)
    double band = -1.0;  // >= 0: the listed tets are the narrow band of this width (P1 only)
    int zero_policy = DHIP_ZERO_THROW;
    int64_t ncut = 0, ndof = 0, nnz = 0;
    double coef[3] = {1, 1, 1};
    dhip::DevBuf cellmask, blockcnt, cellinfo, cut, errflag;
    dhip::DevBuf loc, flags, qual;
    dhip::DevBuf first, isown, ownrank, dofs, rowowner;
    dhip::DevBuf rowlen, rowptr, col, val[4], rhs, diag;
    dhip::DevBuf scan_tmp, red_tmp, bar;
    dhip::DevBuf host_stage;  // pinned staging for scalars
    // Krylov work vectors (grow-only: no device allocation once warmed up)
    dhip::DevBuf kv[17];
    dhip::DevBuf gm_sval;   // right-preconditioned GMRES: a_rc / d_c
    dhip::DevBuf gm_sell_off, gm_sell_col, gm_sell_val;  // SELL-64 copy for large systems
    dhip::DevBuf gm_sell_val32;                          // its values rounded to float (FP32 Arnoldi)
    int gmres_precision = 0;                             // DHIP_FP64 or DHIP_FP32_IR for dhip_gmres*
    dhip::DevBuf gm_sell_pos;                            // SELL slot of every CSR entry
    uint64_t gm_sell_key[4] = {0, 0, 0, 0};              // pattern id, rows, entries, blocks of the layout
    // optional per-block wall-clock probe of the persistent solver
    bool probe_on = false;
    int64_t pcg_rows_per_block = 0;   // 0: one row per thread of a 512-thread block
    int batch_threads = 256;          // workgroup size of the batched solve (128, 256, 512)
    int batch_blocks_per_cu = 2;      // resident workgroups per CU of the batched solve
    int xcd_map = 1;                  // persistent solvers: each XCD's blocks own one contiguous run of rows
    int coherent_exchange = 2;        // persistent solvers' exchange: 0 fenced, 1 coherent atomics, 2 tagged records (FP64 PCG)
    int batch_split = 0;              // batched solve: assembly launch, then a separate PCG launch
    int batch_solve_threads = 1024;   // threads per problem of the separate PCG launch
    int batch_prepass = 0;            // batched solve: classification and local matrices over the whole grid first
    bool coop_refused = false;        // cooperative launch unavailable: same grid, regular launch
    int probe_blocks = 0;
    dhip::DevBuf probe;
    unsigned char *noalloc_scratch = nullptr;   // this device's copy of the allocation-free solver's scratch
    size_t tx_lds = 0;                          // dynamic LDS of the tagged PCG kernels
};

struct dhip_csr {
    dhip_ctx *ctx = nullptr;
    int64_t n = 0, nnz = 0;
    const int32_t *rowptr = nullptr, *col = nullptr;
    const double *val = nullptr;
    dhip::DevBuf own_rowptr, own_col, own_val;
    uint64_t pattern_id = 0;   // nonzero: rowptr/col never change for this handle (solver layouts may be reused)
};

namespace dhip {

int set_error(dhip_ctx *ctx, int code, const char *fmt, ...);

#define DHIP_HIPCHK(ctx, expr)                                                                     \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return ::dhip::set_error((ctx), DHIP_E_HIP, "%s failed: %s (%s:%d)", #expr,            \
                                     hipGetErrorString(_e), __FILE__, __LINE__);                   \
    } while (0)

// Launch with optional HIP-event timing on the context stream.
struct ProfScope {
    dhip_ctx *ctx;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(dhip_ctx *c, const char *n, bool solver = false);
    ~ProfScope();
};
hipError_t take_event(dhip_ctx *ctx, hipEvent_t *ev);
int flush_profile(dhip_ctx *ctx);

#define DHIP_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                                    \
    do {                                                                                           \
        ::dhip::ProfScope _ps((ctx), (name));                                                      \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), (shmem), (ctx)->stream, __VA_ARGS__);  \
        DHIP_HIPCHK((ctx), hipGetLastError());                                                     \
    } while (0)

inline unsigned blocks_for(int64_t n, int per_block) {
    return (unsigned)((n + per_block - 1) / per_block);
}

// ---------------------------------------------------------------------------
// Wave / block primitives (64-wide wavefronts)
// ---------------------------------------------------------------------------
__device__ inline int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

template <int BLOCK>
__device__ inline int block_excl_scan(int v, int *total) {
    __shared__ int wsum[BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < BLOCK / 64; ++i) {
        if (i < w) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    if (total) *total = tot;
    return off + inc - v;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int BLOCK>
__device__ inline double block_sum(double v) {
    __shared__ double ws[BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) ws[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < BLOCK / 64; ++i) s += ws[i];
    __syncthreads();
    return s;
}

// three block sums with one exchange through LDS (each rounded as block_sum rounds it)
template <int BLOCK>
__device__ inline void block_sum3(double &a, double &b, double &c) {
    __shared__ double ws[BLOCK / 64][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    if (lane == 0) {
        ws[w][0] = a;
        ws[w][1] = b;
        ws[w][2] = c;
    }
    __syncthreads();
    double sa = 0.0, sb = 0.0, sc = 0.0;
#pragma unroll
    for (int i = 0; i < BLOCK / 64; ++i) {
        sa += ws[i][0];
        sb += ws[i][1];
        sc += ws[i][2];
    }
    __syncthreads();
    a = sa;
    b = sb;
    c = sc;
}

// Exclusive scan of n int32 values (out may alias in); *total_host receives the sum.
int scan_exclusive_i32(dhip_ctx *ctx, const int32_t *in, int32_t *out, int64_t n, int64_t *total_host);

}  // namespace dhip
