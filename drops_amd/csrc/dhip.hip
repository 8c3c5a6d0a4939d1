// dhip.hip -- libdhip: MI355X (gfx950) trace-FEM assembly + Krylov core behind
// the C ABI of include/dhip.h.  Single translation unit; built with
//   hipcc --offload-arch=gfx950 -O3 -ffp-contract=off -fPIC -shared
#include "dhip_common.hpp"
#include "dhip_assembly.hpp"
#include "dhip_p2.hpp"
#include "dhip_quaqua.hpp"
#include "dhip_krylov.hpp"
#include "dhip_ch.hpp"
#include "dhip_ch_p2.hpp"
#include "dhip_mlp.hpp"
#include "dhip_batch.hpp"
#include "dhip_io.hpp"

#include <algorithm>
#include <cstddef>
#include <atomic>
#include <mutex>

namespace dhip {

__constant__ Tables c_tab;

// ---------------------------------------------------------------------------
// host-side derivation of the device tables
// ---------------------------------------------------------------------------
static Tables g_host_tab;
static std::once_flag g_tab_once;
static std::mutex g_dev_mutex;
static std::vector<int> g_dev_uploaded;

static const int EDGE_V[6][2] = {{0, 1}, {0, 2}, {1, 2}, {0, 3}, {1, 3}, {2, 3}};  // geom/topo.cpp:45-46

static void derive_tables(Tables &T) {
    std::memset(&T, 0, sizeof T);
    // Kuhn tets of a cell, vertex order of BrickBuilderCL::build_ser_impl (geom/builder.cpp:155-175)
    const int kuhn[6][4][3] = {{{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}},
                               {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {1, 1, 1}},
                               {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}},
                               {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}},
                               {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}},
                               {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}}};
    for (int k = 0; k < 6; ++k)
        for (int q = 0; q < 4; ++q)
            for (int c = 0; c < 3; ++c) T.kuhn[k][q][c] = (int8_t)kuhn[k][q][c];
    // order-2 principal lattice: points (x,y,z), 2 >= x >= y >= z >= 0, in rlex order,
    // barycentric (2-x, x-y, y-z, z)/2 (geom/principallattice.cpp:37-43)
    int pts[10][3], np = 0;
    for (int x = 0; x <= 2; ++x)
        for (int y = 0; y <= x; ++y)
            for (int z = 0; z <= y; ++z) {
                pts[np][0] = x; pts[np][1] = y; pts[np][2] = z;
                T.lat_b2[np][0] = (int8_t)(2 - x);
                T.lat_b2[np][1] = (int8_t)(x - y);
                T.lat_b2[np][2] = (int8_t)(y - z);
                T.lat_b2[np][3] = (int8_t)z;
                for (int c = 0; c < 4; ++c) T.lat_b[np][c] = (double)T.lat_b2[np][c] / 2.0;
                ++np;
            }
    auto pindex = [&](int x, int y, int z) {
        for (int i = 0; i < np; ++i)
            if (pts[i][0] == x && pts[i][1] == y && pts[i][2] == z) return i;
        return -1;
    };
    // sub-tets: unit steps along a permutation of the axes from every start point
    // that stays inside the lattice (geom/principallattice.cpp:46-80)
    const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int nt = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k <= j; ++k)
                for (int l = 0; l < 6; ++l) {
                    int p[3] = {i, j, k}, v[4];
                    v[0] = pindex(p[0], p[1], p[2]);
                    bool inside = true;
                    for (int s = 0; s < 3 && inside; ++s) {
                        ++p[perm[l][s]];
                        inside = 2 >= p[0] && p[0] >= p[1] && p[1] >= p[2];
                        if (inside) v[s + 1] = pindex(p[0], p[1], p[2]);
                    }
                    if (!inside) continue;
                    for (int q = 0; q < 4; ++q) T.lat_t[nt][q] = (int8_t)v[q];
                    ++nt;
                }
    // lattice edges numbered at first appearance (sub-tet order, local edge order);
    // the first sub-tet containing a cut edge is the one that creates its cut vertex
    int ne = 0;
    for (int t = 0; t < 8; ++t)
        for (int e = 0; e < 6; ++e) {
            const int a = T.lat_t[t][EDGE_V[e][0]], b = T.lat_t[t][EDGE_V[e][1]];
            int id = -1;
            for (int f = 0; f < ne; ++f)
                if ((T.edge_ab[f][0] == a && T.edge_ab[f][1] == b) || (T.edge_ab[f][0] == b && T.edge_ab[f][1] == a))
                    id = f;
            if (id < 0) {
                id = ne++;
                T.edge_ab[id][0] = (int8_t)a;
                T.edge_ab[id][1] = (int8_t)b;
            }
            T.sub_edge[t][e] = (int8_t)id;
        }
    // sign pattern -> triangles: roots = zero vertices, then sign-changing edges, in
    // increasing number; triangles are consecutive root triples
    // (SignTraitsCL<3>::compute_cuts, RefPatchBuilderCL<3>::assign)
    for (int pid = 0; pid < 81; ++pid) {
        int s[4], rem = pid;
        for (int q = 0; q < 4; ++q) {
            s[q] = rem % 3 - 1;
            rem /= 3;
        }
        // pid = (s0+1) + 3(s1+1) + 9(s2+1) + 27(s3+1) = s0 + 3 s1 + 9 s2 + 27 s3 + 40
        int roots[4], nz = 0, nr = 0;
        if (std::abs(s[0] + s[1] + s[2] + s[3]) != 4) {
            for (int q = 0; q < 4; ++q)
                if (s[q] == 0) roots[nr++] = q;
            nz = nr;
            for (int e = 0; e < 6; ++e)
                if (s[EDGE_V[e][0]] * s[EDGE_V[e][1]] == -1) roots[nr++] = 4 + e;
        }
        const int ntri = nr >= 3 ? nr - 2 : 0;
        T.rp_size[pid] = (int8_t)ntri;
        for (int tr = 0; tr < 2; ++tr)
            for (int j = 0; j < 3; ++j) T.rp_tri[pid][tr][j] = (int8_t)(tr < ntri ? roots[tr + j] : 0);
        T.rp_bnd[pid] = (int8_t)(nz == 3);
    }
    // P2 dof (vertices 0-3, edges 0-5) -> lattice vertex
    for (int l = 0; l < 10; ++l) {
        int a = -1, b = -1;
        for (int c = 0; c < 4; ++c)
            if (T.lat_b2[l][c]) (a < 0 ? a : b) = c;
        if (b < 0) T.lat_of_p2[a] = (int8_t)l;
        else
            for (int e = 0; e < 6; ++e)
                if (EDGE_V[e][0] == a && EDGE_V[e][1] == b) T.lat_of_p2[4 + e] = (int8_t)l;
    }
    // cell-node index (3x3x3 half-grid) of each lattice vertex of each Kuhn tet
    for (int k = 0; k < 6; ++k) {
        uint32_t m = 0;
        for (int l = 0; l < 10; ++l) {
            int o[3] = {0, 0, 0};
            for (int q = 0; q < 4; ++q)
                for (int c = 0; c < 3; ++c) o[c] += T.lat_b2[l][q] * kuhn[k][q][c];
            T.tet_node27[k][l] = (int8_t)(o[2] * 9 + o[1] * 3 + o[0]);
            m |= 1u << T.tet_node27[k][l];
        }
        T.tetmask27[k] = m;
        for (int i = 0; i < 27; ++i) T.p2_of27[k][i] = -1;
        for (int d = 0; d < 10; ++d) {
            T.p2_node27[k][d] = T.tet_node27[k][T.lat_of_p2[d]];
            T.p2_of27[k][T.p2_node27[k][d]] = (int8_t)d;
        }
    }
    // incidence: vertex = cell + delta, delta = (d&1, d>>1&1, d>>2&1)
    for (int d = 0; d < 8; ++d)
        for (int k = 0; k < 6; ++k) {
            T.inc[d][k] = -1;
            for (int q = 0; q < 4; ++q)
                if (kuhn[k][q][0] == (d & 1) && kuhn[k][q][1] == ((d >> 1) & 1) && kuhn[k][q][2] == ((d >> 2) & 1))
                    T.inc[d][k] = (int8_t)q;
        }
    // neighbour slots: the 15 vertex offsets of a Kuhn-mesh row (self + 7 edge directions x 2)
    int ns = 0;
    auto slot_of = [&](int a, int b, int c) {
        for (int s = 0; s < ns; ++s)
            if (T.slot_delta[s][0] == a && T.slot_delta[s][1] == b && T.slot_delta[s][2] == c) return s;
        T.slot_delta[ns][0] = (int8_t)a;
        T.slot_delta[ns][1] = (int8_t)b;
        T.slot_delta[ns][2] = (int8_t)c;
        return ns++;
    };
    slot_of(0, 0, 0);
    for (int k = 0; k < 6; ++k)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                T.slot[k][i][j] = (int8_t)slot_of(kuhn[k][j][0] - kuhn[k][i][0], kuhn[k][j][1] - kuhn[k][i][1],
                                                  kuhn[k][j][2] - kuhn[k][i][2]);
    // quadrature rules (num/discretize.cpp:153-211, 234-238, 263-296; num/discretize.tpp:754-762)
    const double s15 = std::sqrt(15.0);
    const double q2n[4][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}, {1. / 3., 1. / 3., 1. / 3.}};
    std::memcpy(T.q2n, q2n, sizeof q2n);
    T.q2w[0] = T.q2w[1] = T.q2w[2] = 1. / 24.;
    T.q2w[3] = 0.375;
    {
        const double A1 = (6.0 - s15) / 21.0, B1 = (9.0 + 2.0 * s15) / 21.0;
        const double A2 = (6.0 + s15) / 21.0, B2 = (9.0 - 2.0 * s15) / 21.0;
        const double q5n[7][3] = {{1. / 3., 1. / 3., 1. / 3.}, {A1, A1, B1}, {A1, B1, A1}, {B1, A1, A1},
                                  {A2, A2, B2}, {A2, B2, A2}, {B2, A2, A2}};
        std::memcpy(T.q5n, q5n, sizeof q5n);
        T.q5w[0] = 9. / 80.;
        T.q5w[1] = T.q5w[2] = T.q5w[3] = (155. - s15) / 2400.;
        T.q5w[4] = T.q5w[5] = T.q5w[6] = (155. + s15) / 2400.;
    }
    {
        const double A1 = (7.0 - s15) / 34.0, B1 = (13.0 + 3.0 * s15) / 34.0;
        const double A2 = (7.0 + s15) / 34.0, B2 = (13.0 - 3.0 * s15) / 34.0;
        const double A3 = (10.0 - 2.0 * s15) / 40.0, B3 = (10.0 + 2.0 * s15) / 40.0;
        const double v5[15][4] = {{0.25, 0.25, 0.25, 0.25},
                                  {A1, A1, A1, B1}, {A1, A1, B1, A1}, {A1, B1, A1, A1}, {B1, A1, A1, A1},
                                  {A2, A2, A2, B2}, {A2, A2, B2, A2}, {A2, B2, A2, A2}, {B2, A2, A2, A2},
                                  {A3, A3, B3, B3}, {A3, B3, A3, B3}, {A3, B3, B3, A3},
                                  {B3, A3, A3, B3}, {B3, A3, B3, A3}, {B3, B3, A3, A3}};
        std::memcpy(T.v5n, v5, sizeof v5);
        T.v5c1 = (2665.0 + 14.0 * s15) / 226800.0;
        T.v5c2 = (2665.0 - 14.0 * s15) / 226800.0;
    }
}

const Tables &host_tables() {
    std::call_once(g_tab_once, [] { derive_tables(g_host_tab); });
    return g_host_tab;
}

static hipError_t upload_tables(int device) {
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    if (std::find(g_dev_uploaded.begin(), g_dev_uploaded.end(), device) != g_dev_uploaded.end())
        return hipSuccess;
    const Tables &T = host_tables();
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &T, sizeof T, 0, hipMemcpyHostToDevice);
    if (e == hipSuccess) g_dev_uploaded.push_back(device);
    return e;
}

// ---------------------------------------------------------------------------
// errors, profiling
// ---------------------------------------------------------------------------
// messages of calls made without a context (file formats on host memory)
static thread_local char g_err_noctx[256] = "null context";

int set_error(dhip_ctx *ctx, int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    if (ctx) std::vsnprintf(ctx->err, sizeof ctx->err, fmt, ap);
    else std::vsnprintf(g_err_noctx, sizeof g_err_noctx, fmt, ap);
    va_end(ap);
    return code;
}

hipError_t take_event(dhip_ctx *ctx, hipEvent_t *ev) {
    if (!ctx->event_pool.empty()) {
        *ev = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return hipSuccess;
    }
    return hipEventCreate(ev);
}

ProfScope::ProfScope(dhip_ctx *c, const char *n, bool solver) : ctx(c), name(n) {
    if (!(ctx->prof == 1 || (ctx->prof == 2 && solver))) return;
    if (take_event(ctx, &a) != hipSuccess || take_event(ctx, &b) != hipSuccess) {
        a = b = nullptr;
        return;
    }
    (void)hipEventRecord(a, ctx->stream);
}

ProfScope::~ProfScope() {
    if (!a) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({name, a, b});
}

int flush_profile(dhip_ctx *ctx) {
    for (auto &p : ctx->pending) {
        float ms = 0.f;
        DHIP_HIPCHK(ctx, hipEventSynchronize(p.b));
        DHIP_HIPCHK(ctx, hipEventElapsedTime(&ms, p.a, p.b));
        auto &acc = ctx->prof_acc[p.name];
        acc.first += ms;
        acc.second += 1;
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
    return DHIP_OK;
}

// ---------------------------------------------------------------------------
// device-wide exclusive scan of int32
// ---------------------------------------------------------------------------
constexpr int SCAN_ITEMS = 4, SCAN_TILE = 256 * SCAN_ITEMS;

__global__ void __launch_bounds__(256) k_scan_reduce(const int32_t *in, int64_t n, int32_t *bsum) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int64_t j = base + (int64_t)i * 256 + threadIdx.x;
        if (j < n) s += in[j];
    }
    int tot;
    block_excl_scan<256>(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) k_scan_top(int32_t *bsum, int64_t nb, int32_t *total) {
    int carry = 0;
    for (int64_t base = 0; base < nb; base += 256) {
        const int64_t j = base + threadIdx.x;
        const int v = j < nb ? bsum[j] : 0;
        int tot;
        const int ex = block_excl_scan<256>(v, &tot);
        if (j < nb) bsum[j] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256) k_scan_final(const int32_t *in, int32_t *out, int64_t n,
                                                     const int32_t *bsum) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = base + i < n ? in[base + i] : 0;
        s += v[i];
    }
    int off = block_excl_scan<256>(s, nullptr) + bsum[blockIdx.x];
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = off;
        off += v[i];
    }
}

int scan_exclusive_i32(dhip_ctx *ctx, const int32_t *in, int32_t *out, int64_t n, int64_t *total_host) {
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    DHIP_HIPCHK(ctx, ctx->scan_tmp.ensure(sizeof(int32_t) * (nb + 2)));
    int32_t *bsum = ctx->scan_tmp.as<int32_t>();
    int32_t *tot = bsum + nb + 1;
    if (nb > 0) {
        DHIP_LAUNCH(ctx, "scan_reduce", k_scan_reduce, (unsigned)nb, 256, 0, in, n, bsum);
    }
    DHIP_LAUNCH(ctx, "scan_top", k_scan_top, 1, 256, 0, bsum, nb, tot);
    ctx->scan_total_dev = tot;
    if (nb > 0) {
        DHIP_LAUNCH(ctx, "scan_final", k_scan_final, (unsigned)nb, 256, 0, in, out, n, bsum);
    }
    if (total_host) {
        int32_t t = 0;
        DHIP_HIPCHK(ctx, hipMemcpyAsync(&t, tot, sizeof t, hipMemcpyDeviceToHost, ctx->stream));
        DHIP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        *total_host = t;
    }
    return DHIP_OK;
}

// the total of the last scan_exclusive_i32 and the error flags, with one synchronisation
static int read_total_and_flags(dhip_ctx *ctx, int64_t *total) {
    int32_t t = 0, f = 0;
    DHIP_HIPCHK(ctx, hipMemcpyAsync(&t, ctx->scan_total_dev, sizeof t, hipMemcpyDeviceToHost, ctx->stream));
    DHIP_HIPCHK(ctx, hipMemcpyAsync(&f, ctx->errflag.p, sizeof f, hipMemcpyDeviceToHost, ctx->stream));
    DHIP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *total = t;
    if (f & 2)
        return set_error(ctx, DHIP_E_ZERO_PATTERN, "found a full body in the zero level set, grid is too coarse");
    if (f & 1)
        return set_error(ctx, DHIP_E_ZERO_NODE, "levelset is exactly zero at node(s) of a cut tetra");
    return DHIP_OK;
}

static int check_err_flag(dhip_ctx *ctx) {
    int32_t f = 0;
    DHIP_HIPCHK(ctx, hipMemcpyAsync(&f, ctx->errflag.p, sizeof f, hipMemcpyDeviceToHost, ctx->stream));
    DHIP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (f & 2)
        return set_error(ctx, DHIP_E_ZERO_PATTERN,
                         "found a full body in the zero level set, grid is too coarse");
    if (f & 1)
        return set_error(ctx, DHIP_E_ZERO_NODE, "levelset is exactly zero at node(s) of a cut tetra");
    if (f & 4)
        return set_error(ctx, DHIP_E_NOT_CONVERGED,
                         "QuaQua mapping: rank-deficient matrix (a degenerate facet or a singular Newton step)");
    return DHIP_OK;
}

}  // namespace dhip

using namespace dhip;

static int noalloc_setup(dhip_ctx *c);
static size_t persistent_lds(dhip_ctx *c, const void *k);

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int dhip_create(dhip_ctx_t **out, int32_t device) {
    if (!out) return DHIP_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return DHIP_E_HIP;
    if (device < 0 || device >= ndev) return DHIP_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return DHIP_E_HIP;
    if (upload_tables(device) != hipSuccess) return DHIP_E_HIP;
    dhip_ctx *c = new dhip_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return DHIP_E_HIP;
    }
    c->cu_count = prop.multiProcessorCount;
    if (c->errflag.ensure(256) != hipSuccess || c->bar.ensure(256) != hipSuccess) {
        c->errflag.release();
        (void)hipStreamDestroy(c->stream);
        delete c;
        return DHIP_E_NOMEM;
    }
    // the tagged PCG kernels' dynamic LDS (their attribute raised here, once, so that the
    // allocation-free entry point launches them as they are)
    {
        const void *kt[6] = {(const void *)k_pcg64_tx<1>, (const void *)k_pcg64_tx<2>, (const void *)k_pcg64_tx<4>,
                             (const void *)k_pcg64_pipe<1>, (const void *)k_pcg64_pipe<2>,
                             (const void *)k_pcg64_pipe<4>};
        size_t lim = persistent_lds(c, kt[2]);
        for (const void *k : kt) lim = std::min(lim, persistent_lds(c, k));
        c->tx_lds = lim;
        for (const void *k : kt)
            if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim) != hipSuccess) {
                (void)hipGetLastError();
                c->tx_lds = LDS_BYTES;
            }
    }
    // once per context (and so per device it runs on): the device address of the allocation-free
    // solver's scratch array
    if (noalloc_setup(c) != DHIP_OK) {
        c->errflag.release();
        c->bar.release();
        (void)hipStreamDestroy(c->stream);
        delete c;
        return DHIP_E_HIP;
    }
    *out = c;
    return DHIP_OK;
}

void dhip_destroy(dhip_ctx_t *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    DevBuf *bufs[] = {&c->phi_own, &c->cellmask, &c->blockcnt, &c->cellinfo, &c->cut, &c->errflag,
                      &c->loc, &c->flags, &c->first, &c->isown, &c->ownrank, &c->dofs, &c->rowowner,
                      &c->rowlen, &c->rowptr, &c->col, &c->val[0], &c->val[1], &c->val[2], &c->val[3],
                      &c->rhs, &c->diag, &c->scan_tmp, &c->red_tmp, &c->bar, &c->host_stage, &c->qual,
                      &c->gm_sval, &c->probe, &c->gm_sell_off, &c->gm_sell_col, &c->gm_sell_val, &c->gm_sell_pos,
                      &c->gm_sell_val32, &c->qp_off, &c->qp_vals, &c->qq_first, &c->qq_gh, &c->qq_list, &c->qqp_off, &c->qqp_vals};
    for (DevBuf *b : bufs) b->release();
    for (auto &b : c->kv) b.release();
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *dhip_last_error(const dhip_ctx_t *c) { return c ? c->err : g_err_noctx; }

int dhip_profile_enable(dhip_ctx_t *c, int32_t on) {
    if (!c) return DHIP_E_ARG;
    if (!on) {
        int rc = flush_profile(c);
        if (rc) return rc;
    }
    if (on < 0 || on > 2) return set_error(c, DHIP_E_ARG, "profile mode must be 0, 1 or 2");
    c->prof = on;
    return DHIP_OK;
}

int dhip_profile_get(dhip_ctx_t *c, const char *name, double *total_ms, int64_t *count) {
    if (!c || !name) return DHIP_E_ARG;
    int rc = flush_profile(c);
    if (rc) return rc;
    auto it = c->prof_acc.find(name);
    if (total_ms) *total_ms = it == c->prof_acc.end() ? 0.0 : it->second.first;
    if (count) *count = it == c->prof_acc.end() ? 0 : it->second.second;
    return DHIP_OK;
}

int dhip_profile_reset(dhip_ctx_t *c) {
    if (!c) return DHIP_E_ARG;
    int rc = flush_profile(c);
    c->prof_acc.clear();
    return rc;
}

int dhip_malloc(dhip_ctx_t *c, void **dptr, int64_t bytes) {
    if (!c || !dptr || bytes < 0) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    DHIP_HIPCHK(c, hipMalloc(dptr, bytes > 0 ? (size_t)bytes : 1));
    return DHIP_OK;
}

int dhip_free(dhip_ctx_t *c, void *dptr) {
    if (!c) return DHIP_E_ARG;
    if (dptr) DHIP_HIPCHK(c, hipFree(dptr));
    return DHIP_OK;
}

int dhip_memcpy(dhip_ctx_t *c, void *dst, const void *src, int64_t bytes, int32_t kind) {
    if (!c || bytes < 0 || kind < 0 || kind > 2) return DHIP_E_ARG;
    if (bytes == 0) return DHIP_OK;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                         : hipMemcpyDeviceToDevice;
    DHIP_HIPCHK(c, hipMemcpyAsync(dst, src, (size_t)bytes, k, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

int dhip_synchronize(dhip_ctx_t *c) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_mesh_kuhn(dhip_ctx_t *c, const double orig[3], const double e[9], const int32_t n[3]) {
    if (!c || !orig || !e || !n) return DHIP_E_ARG;
    for (int k = 0; k < 3; ++k)
        if (n[k] <= 0) return set_error(c, DHIP_E_ARG, "cell counts must be positive");
    const int64_t ncell = (int64_t)n[0] * n[1] * n[2];
    // cut indices are packed as (index << 6 | mask) in 32 bits; half-grid indices fit int64
    if (ncell * 6 > (int64_t)1 << 31)
        return set_error(c, DHIP_E_CAPACITY, "mesh too large: %lld tets", (long long)(ncell * 6));
    MeshP &m = c->mesh;
    for (int cc = 0; cc < 3; ++cc) {
        m.orig[cc] = orig[cc];
        for (int k = 0; k < 3; ++k) {
            c->e[k][cc] = e[k * 3 + cc];
            m.off[k][cc] = 1.0 / (double)n[k] * e[k * 3 + cc];  // geom/builder.cpp:108-110
        }
        m.n[cc] = n[cc];
    }
    m.ncell = ncell;
    c->nvert = (int64_t)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    c->has_mesh = true;
    c->stage = 0;
    c->phi = nullptr;
    c->phi_count = 0;
    ++c->geo_gen;
    return DHIP_OK;
}

int dhip_levelset_p2(dhip_ctx_t *c, const double *phi, int64_t count, int32_t where) {
    if (!c || !phi) return DHIP_E_ARG;
    if (!c->has_mesh) return set_error(c, DHIP_E_STATE, "dhip_mesh_kuhn must precede dhip_levelset_p2");
    const int64_t need = (int64_t)(2 * c->mesh.n[0] + 1) * (2 * c->mesh.n[1] + 1) * (2 * c->mesh.n[2] + 1);
    if (count != need)
        return set_error(c, DHIP_E_ARG, "level set has %lld values, the half-grid needs %lld",
                         (long long)count, (long long)need);
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    if (where == DHIP_DEVICE) {
        c->phi = phi;
    } else {
        DHIP_HIPCHK(c, c->phi_own.ensure(sizeof(double) * need));
        DHIP_HIPCHK(c, hipMemcpyAsync(c->phi_own.p, phi, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
        DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
        c->phi = c->phi_own.as<double>();
    }
    c->phi_count = need;
    c->stage = 0;
    ++c->geo_gen;
    return DHIP_OK;
}

static int classify_impl(dhip_ctx *c, int32_t zero_policy, int64_t *ncut, bool sync_count, double band = -1.0) {
    if (!c->has_mesh || !c->phi) return set_error(c, DHIP_E_STATE, "mesh and level set required");
    const int64_t ncell = c->mesh.ncell;
    const unsigned nb = blocks_for(ncell, CLS_BLOCK);
    DHIP_HIPCHK(c, c->cellmask.ensure(ncell));
    DHIP_HIPCHK(c, c->blockcnt.ensure(sizeof(int32_t) * (nb + 1)));
    DHIP_HIPCHK(c, c->cellinfo.ensure(sizeof(uint32_t) * ncell));
    DHIP_LAUNCH(c, "zero_flags", k_zero_u32, 1, 64, 0, c->errflag.as<uint32_t>(), 64);
    ClassArgs ca{c->mesh, c->phi, c->cellmask.as<uint8_t>(), c->blockcnt.as<int32_t>(), c->errflag.as<int32_t>(),
                 zero_policy == DHIP_ZERO_THROW, band};
    c->stage = 0;
    c->band = band;
    DHIP_LAUNCH(c, "classify", k_classify, nb, CLS_BLOCK, 0, ca);
    int rc = scan_exclusive_i32(c, c->blockcnt.as<int32_t>(), c->blockcnt.as<int32_t>(), nb, nullptr);
    if (rc) return rc;
    (void)sync_count;
    // one readback: the cut-tet count (the scan's total) and the error flags
    int64_t total = 0;
    rc = read_total_and_flags(c, &total);
    if (rc) return rc;
    if (total >= ((int64_t)1 << 26))
        return set_error(c, DHIP_E_CAPACITY, "%lld cut tets exceed the 2^26 supported", (long long)total);
    DHIP_HIPCHK(c, c->cut.ensure(sizeof(int64_t) * (total + 1)));
    CompactArgs pa{ncell, c->cellmask.as<uint8_t>(), c->blockcnt.as<int32_t>(), c->cellinfo.as<uint32_t>(),
                   c->cut.as<int64_t>()};
    DHIP_LAUNCH(c, "compact", k_compact, nb, CLS_BLOCK, 0, pa);
    c->ncut = total;
    c->zero_policy = zero_policy;
    c->stage = 1;
    if (ncut) *ncut = total;
    return DHIP_OK;
}

int dhip_classify(dhip_ctx_t *c, int32_t zero_policy, int64_t *ncut) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = classify_impl(c, zero_policy, ncut, true);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

// The given right-hand side values, if they belong to the current cut list (DHIP_RHS_VALUES).
static int rhs_values_for(dhip_ctx *c, int32_t rhs_kind, const double **vals, const int32_t **off) {
    *vals = nullptr;
    *off = nullptr;
    if (rhs_kind != DHIP_RHS_VALUES) return DHIP_OK;
    if (!c->qp_have_vals || c->qp_gen != c->geo_gen || c->qp_band != c->band || c->qp_ncut != c->ncut)
        return set_error(c, DHIP_E_STATE,
                         "DHIP_RHS_VALUES needs dhip_quad_points and dhip_rhs_values on the current cut tets");
    *vals = c->qp_vals.as<double>();
    *off = c->qp_off.as<int32_t>();
    return DHIP_OK;
}

// check: read the error flags now; the assembly pipeline leaves that to the numbering's readback
static int local_impl(dhip_ctx *c, const double coef[3], int32_t rhs_kind, const double *rhs_param,
                      bool check = true) {
    if (c->stage < 1) return set_error(c, DHIP_E_STATE, "dhip_classify must precede dhip_local_p1");
    const int64_t N = c->ncut;
    DHIP_HIPCHK(c, c->loc.ensure(sizeof(double) * LOC_N * (N + 1)));
    DHIP_HIPCHK(c, c->flags.ensure(N + 1));
    for (int k = 0; k < 3; ++k) c->coef[k] = coef ? coef[k] : 1.0;
    const double *rv = nullptr;
    const int32_t *ro = nullptr;
    int rc0 = rhs_values_for(c, rhs_kind, &rv, &ro);
    if (rc0) return rc0;
    LocalArgs la{c->mesh, c->phi, c->cut.as<int64_t>(), N, c->coef[0], c->coef[1], c->coef[2], rhs_kind,
                 rhs_param ? rhs_param[0] : 0.0, c->loc.as<double>(), c->flags.as<uint8_t>(),
                 c->errflag.as<int32_t>(), c->band >= 0. ? 1 : 0, rv, ro};
    if (N > 0) DHIP_LAUNCH(c, "local_p1", k_local_p1, blocks_for(N, LOCAL_BLOCK), LOCAL_BLOCK, 0, la);
    if (check) {
        const int rc = check_err_flag(c);
        if (rc) return rc;
    }
    c->stage = 2;
    c->space = 1;
    return DHIP_OK;
}

// Maximal points of one tet (16 facets x 7): the slack after the given values, so that a tet whose
// patch turns out to be invalid (a sub-tet in the zero level set, reported by the local pass) reads
// inside the buffer.
constexpr int64_t QP_SLACK = 112;

int dhip_quad_points(dhip_ctx_t *c, int64_t *npoints, double *xyz, int32_t where) {
    if (!c || !npoints) return DHIP_E_ARG;
    if (c->stage < 1) return set_error(c, DHIP_E_STATE, "dhip_classify must precede dhip_quad_points");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int64_t N = c->ncut;
    DevBuf cnt;
    DHIP_HIPCHK(c, cnt.ensure(sizeof(int32_t) * (N + 1)));
    DHIP_HIPCHK(c, c->qp_off.ensure(sizeof(int32_t) * (N + 1)));
    LocalArgs la{c->mesh, c->phi, c->cut.as<int64_t>(), N, 1., 1., 1., DHIP_RHS_NONE, 0., nullptr, nullptr,
                 nullptr, 0};
    int64_t total = 0;
    if (N > 0) {
        DHIP_LAUNCH(c, "quad_points_count", k_quad_points<true>, blocks_for(N, LOCAL_BLOCK), LOCAL_BLOCK, 0, la,
                    cnt.as<int32_t>(), (const int32_t *)nullptr, (double *)nullptr);
        int rc = scan_exclusive_i32(c, cnt.as<int32_t>(), c->qp_off.as<int32_t>(), N, &total);
        if (rc) return rc;
    }
    const int32_t t32 = (int32_t)total;
    DHIP_HIPCHK(c, hipMemcpyAsync(c->qp_off.as<int32_t>() + N, &t32, sizeof t32, hipMemcpyHostToDevice, c->stream));
    c->qp_n = total;
    c->qp_ncut = N;
    c->qp_gen = c->geo_gen;
    c->qp_band = c->band;
    c->qp_have_vals = false;
    *npoints = total;
    if (xyz && total > 0) {
        DevBuf tmp;
        double *dst = xyz;
        if (where != DHIP_DEVICE) {
            DHIP_HIPCHK(c, tmp.ensure(sizeof(double) * 3 * total));
            dst = tmp.as<double>();
        }
        DHIP_LAUNCH(c, "quad_points", k_quad_points<false>, blocks_for(N, LOCAL_BLOCK), LOCAL_BLOCK, 0, la,
                    (int32_t *)nullptr, (const int32_t *)c->qp_off.as<int32_t>(), dst);
        if (where != DHIP_DEVICE)
            DHIP_HIPCHK(c, hipMemcpyAsync(xyz, dst, sizeof(double) * 3 * total, hipMemcpyDeviceToHost, c->stream));
        DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_rhs_values(dhip_ctx_t *c, const double *values, int64_t npoints, int32_t where) {
    if (!c || (npoints > 0 && !values) || npoints < 0) return DHIP_E_ARG;
    if (c->qp_n < 0 || c->qp_gen != c->geo_gen)
        return set_error(c, DHIP_E_STATE, "dhip_quad_points must precede dhip_rhs_values on this level set");
    if (npoints != c->qp_n)
        return set_error(c, DHIP_E_ARG, "%lld values for %lld quadrature points", (long long)npoints,
                         (long long)c->qp_n);
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    DHIP_HIPCHK(c, c->qp_vals.ensure(sizeof(double) * (npoints + QP_SLACK)));
    if (npoints > 0)
        DHIP_HIPCHK(c, hipMemcpyAsync(c->qp_vals.p, values, sizeof(double) * npoints,
                                      where == DHIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                      c->stream));
    DHIP_LAUNCH(c, "rhs_values_slack", k_zero_u32, 1, 256, 0, (uint32_t *)(c->qp_vals.as<double>() + npoints),
                (int)(2 * QP_SLACK));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    c->qp_have_vals = true;
    return flush_profile(c);
}

int dhip_local_p1(dhip_ctx_t *c, const double coef[3], int32_t rhs_kind, const double *rhs_param) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = local_impl(c, coef, rhs_kind, rhs_param);
    if (rc) return rc;
    return flush_profile(c);
}

static int number_impl(dhip_ctx *c, double omit_bound, int64_t *ndof) {
    if (c->stage < 2 || c->space != 1)
        return set_error(c, DHIP_E_STATE, "dhip_local_p1 must precede dhip_number_p1if");
    const int64_t N = c->ncut, S = 4 * N;
    DHIP_HIPCHK(c, c->first.ensure(sizeof(int32_t) * (S + 1)));
    DHIP_HIPCHK(c, c->isown.ensure(sizeof(int32_t) * (S + 1)));
    DHIP_HIPCHK(c, c->ownrank.ensure(sizeof(int32_t) * (S + 1)));
    DHIP_HIPCHK(c, c->dofs.ensure(sizeof(int32_t) * (S + 1)));
    if (N > 0) {
        QualArgs qa{c->mesh, c->cut.as<int64_t>(), N, c->loc.as<double>(), c->flags.as<uint8_t>(), omit_bound,
                    c->band >= 0. ? 1 : 0};
        DHIP_LAUNCH(c, "qualify", k_qualify, blocks_for(N, 256), 256, 0, qa);
        FirstArgs fa{c->mesh, c->cut.as<int64_t>(), N, c->cellinfo.as<uint32_t>(), c->flags.as<uint8_t>(),
                     c->first.as<int32_t>(), c->isown.as<int32_t>()};
        DHIP_LAUNCH(c, "first_owner", k_first, blocks_for(S, 256), 256, 0, fa);
    }
    int64_t nd = 0;
    int rc = scan_exclusive_i32(c, c->isown.as<int32_t>(), c->ownrank.as<int32_t>(), S, nullptr);
    if (rc) return rc;
    // one readback: the DOF count and the local step's error flags
    rc = read_total_and_flags(c, &nd);
    if (rc) return rc;
    DHIP_HIPCHK(c, c->rowowner.ensure(sizeof(int32_t) * (nd + 1)));
    if (N > 0) {
        DofArgs da{S, c->first.as<int32_t>(), c->ownrank.as<int32_t>(), c->dofs.as<int32_t>(),
                   c->rowowner.as<int32_t>()};
        DHIP_LAUNCH(c, "dofs", k_dofs, blocks_for(S, 256), 256, 0, da);
    }
    c->ndof = nd;
    c->stage = 3;
    if (ndof) *ndof = nd;
    return DHIP_OK;
}

int dhip_number_p1if(dhip_ctx_t *c, double omit_bound, int64_t *ndof) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = number_impl(c, omit_bound, ndof);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

static int build_impl(dhip_ctx *c, const double lc[3], int64_t *nnz) {
    if (c->stage < 3 || c->space != 1)
        return set_error(c, DHIP_E_STATE, "dhip_number_p1if must precede dhip_build_p1");
    const int64_t R = c->ndof;
    DHIP_HIPCHK(c, c->rowlen.ensure(sizeof(int32_t) * (R + 1)));
    DHIP_HIPCHK(c, c->rowptr.ensure(sizeof(int32_t) * (R + 1)));
    DHIP_HIPCHK(c, c->rhs.ensure(sizeof(double) * (R + 1)));
    DHIP_HIPCHK(c, c->diag.ensure(sizeof(double) * (R + 1)));
    RowArgs ra{c->mesh, c->cut.as<int64_t>(), c->ncut, c->cellinfo.as<uint32_t>(), c->flags.as<uint8_t>(),
               c->dofs.as<int32_t>(), c->rowowner.as<int32_t>(), R, c->loc.as<double>(), c->rowlen.as<int32_t>(),
               c->rowptr.as<int32_t>(), nullptr, nullptr, nullptr, nullptr, nullptr, c->rhs.as<double>(),
               c->diag.as<double>(), lc ? lc[0] : 1.0, lc ? lc[1] : 1.0, lc ? lc[2] : 1.0};
    if (R > 0) DHIP_LAUNCH(c, "rowcount", k_rows_w<false>, blocks_for(R, ROWW_WAVES), 64 * ROWW_WAVES, 0, ra);
    int64_t total = 0;
    int rc = scan_exclusive_i32(c, c->rowlen.as<int32_t>(), c->rowptr.as<int32_t>(), R, &total);
    if (rc) return rc;
    // rowptr[R] = nnz, from the scan's device total (no host-to-device copy)
    DHIP_LAUNCH(c, "rowptr_last", k_copy_i32, 1, 64, 0, c->scan_total_dev, c->rowptr.as<int32_t>() + R);
    DHIP_HIPCHK(c, c->col.ensure(sizeof(int32_t) * (total + 1)));
    for (int k = 0; k < 4; ++k) DHIP_HIPCHK(c, c->val[k].ensure(sizeof(double) * (total + 1)));
    ra.col = c->col.as<int32_t>();
    ra.vm = c->val[0].as<double>();
    ra.vl = c->val[1].as<double>();
    ra.vs = c->val[2].as<double>();
    ra.vc = c->val[3].as<double>();
    if (R > 0) DHIP_LAUNCH(c, "rowfill", k_rows_w<true>, blocks_for(R, ROWW_WAVES), 64 * ROWW_WAVES, 0, ra);
    c->nnz = total;
    c->stage = 4;
    if (nnz) *nnz = total;
    return DHIP_OK;
}

int dhip_build_p1(dhip_ctx_t *c, const double lc[3], int64_t *nnz) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = build_impl(c, lc, nnz);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_assemble_p1(dhip_ctx_t *c, int32_t zero_policy, const double coef[3], int32_t rhs_kind,
                     const double *rhs_param, double omit_bound, const double lc[3], int64_t *ncut,
                     int64_t *ndof, int64_t *nnz) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = classify_impl(c, zero_policy, ncut, false);
    if (!rc) rc = local_impl(c, coef, rhs_kind, rhs_param, false);
    if (!rc) rc = number_impl(c, omit_bound, ndof);
    if (!rc) rc = build_impl(c, lc, nnz);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_assemble_p1_band(dhip_ctx_t *c, int32_t zero_policy, const double coef[3], int32_t rhs_kind,
                          const double *rhs_param, double width, const double lc[3], int64_t *nband, int64_t *ndof,
                          int64_t *nnz) {
    if (!c) return DHIP_E_ARG;
    if (!(width >= 0.)) return set_error(c, DHIP_E_ARG, "narrow-band width must be >= 0");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = classify_impl(c, zero_policy, nband, false, width);
    if (!rc) rc = local_impl(c, coef, rhs_kind, rhs_param, false);
    if (!rc) rc = number_impl(c, -1.0, ndof);
    if (!rc) rc = build_impl(c, lc, nnz);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

// ---------------------------------------------------------------------------
// P2 interface space
// ---------------------------------------------------------------------------
static int local_p2_impl(dhip_ctx *c, const double coef[3], int32_t rhs_kind, const double *rhs_param) {
    if (c->stage < 1) return set_error(c, DHIP_E_STATE, "dhip_classify must precede dhip_local_p2");
    if (c->band >= 0.) return set_error(c, DHIP_E_STATE, "the narrow band is a P1 assembly");
    const int64_t N = c->ncut;
    DHIP_HIPCHK(c, c->loc.ensure(sizeof(double) * LOC2_N * (N + 1)));
    DHIP_HIPCHK(c, c->flags.ensure(N + 1));
    for (int k = 0; k < 3; ++k) c->coef[k] = coef ? coef[k] : 1.0;
    const double *rv = nullptr;
    const int32_t *ro = nullptr;
    int rc0 = rhs_values_for(c, rhs_kind, &rv, &ro);
    if (rc0) return rc0;
    LocalP2Args la{c->mesh, c->phi, c->cut.as<int64_t>(), N, c->coef[0], c->coef[1], c->coef[2], rhs_kind,
                   rhs_param ? rhs_param[0] : 0.0, c->loc.as<double>(), c->flags.as<uint8_t>(),
                   c->errflag.as<int32_t>(), rv, ro};
    if (N > 0) DHIP_LAUNCH(c, "local_p2", k_local_p2, (unsigned)N, 64, 0, la);
    int rc = check_err_flag(c);
    if (rc) return rc;
    c->stage = 2;
    c->space = 2;
    return DHIP_OK;
}

// the numbering's integrals alone (k_hat_p2), for a pass whose forms come from the mapped surface
static int hat_p2_impl(dhip_ctx *c) {
    if (c->stage < 1) return set_error(c, DHIP_E_STATE, "dhip_classify must precede the P2 numbering");
    if (c->band >= 0.) return set_error(c, DHIP_E_STATE, "the narrow band is a P1 assembly");
    const int64_t N = c->ncut;
    DHIP_HIPCHK(c, c->loc.ensure(sizeof(double) * LOC2_N * (N + 1)));
    DHIP_HIPCHK(c, c->flags.ensure(N + 1));
    LocalP2Args la{c->mesh, c->phi, c->cut.as<int64_t>(), N, 1., 1., 1., DHIP_RHS_NONE, 0.0, c->loc.as<double>(),
                   c->flags.as<uint8_t>(), c->errflag.as<int32_t>(), nullptr, nullptr};
    if (N > 0) DHIP_LAUNCH(c, "hat_p2", k_hat_p2, (unsigned)((N + HAT_TPW - 1) / HAT_TPW), 64, 0, la);
    int rc = check_err_flag(c);
    if (rc) return rc;
    c->stage = 2;
    c->space = 2;
    return DHIP_OK;
}

static int number_p2_impl(dhip_ctx *c, double omit_bound, int64_t *ndof) {
    if (c->stage < 2 || c->space != 2)
        return set_error(c, DHIP_E_STATE, "dhip_local_p2 must precede dhip_number_p2if");
    const int64_t N = c->ncut, S = 10 * N;
    DHIP_HIPCHK(c, c->qual.ensure(sizeof(uint16_t) * (N + 1)));
    DHIP_HIPCHK(c, c->first.ensure(sizeof(int32_t) * (S + 1)));
    DHIP_HIPCHK(c, c->isown.ensure(sizeof(int32_t) * (S + 1)));
    DHIP_HIPCHK(c, c->ownrank.ensure(sizeof(int32_t) * (S + 1)));
    DHIP_HIPCHK(c, c->dofs.ensure(sizeof(int32_t) * (S + 1)));
    if (N > 0) {
        QualP2Args qa{c->mesh, c->cut.as<int64_t>(), N, c->loc.as<double>(), c->flags.as<uint8_t>(),
                      c->qual.as<uint16_t>(), omit_bound};
        DHIP_LAUNCH(c, "qualify_p2", k_qualify_p2, blocks_for(N, 256), 256, 0, qa);
        FirstP2Args fa{c->mesh, c->cut.as<int64_t>(), N, c->cellinfo.as<uint32_t>(), c->flags.as<uint8_t>(),
                       c->qual.as<uint16_t>(), c->first.as<int32_t>(), c->isown.as<int32_t>()};
        DHIP_LAUNCH(c, "first_owner_p2", k_first_p2, blocks_for(S, 256), 256, 0, fa);
    }
    int64_t nd = 0;
    int rc = scan_exclusive_i32(c, c->isown.as<int32_t>(), c->ownrank.as<int32_t>(), S, &nd);
    if (rc) return rc;
    DHIP_HIPCHK(c, c->rowowner.ensure(sizeof(int32_t) * (nd + 1)));
    if (N > 0) {
        DofArgs da{S, c->first.as<int32_t>(), c->ownrank.as<int32_t>(), c->dofs.as<int32_t>(),
                   c->rowowner.as<int32_t>()};
        DHIP_LAUNCH(c, "dofs_p2", k_dofs, blocks_for(S, 256), 256, 0, da);
    }
    c->ndof = nd;
    c->stage = 3;
    if (ndof) *ndof = nd;
    return DHIP_OK;
}

static int build_p2_impl(dhip_ctx *c, const double lc[3], int64_t *nnz) {
    if (c->stage < 3 || c->space != 2)
        return set_error(c, DHIP_E_STATE, "dhip_number_p2if must precede dhip_build_p2");
    const int64_t R = c->ndof;
    DHIP_HIPCHK(c, c->rowlen.ensure(sizeof(int32_t) * (R + 1)));
    DHIP_HIPCHK(c, c->rowptr.ensure(sizeof(int32_t) * (R + 1)));
    DHIP_HIPCHK(c, c->rhs.ensure(sizeof(double) * (R + 1)));
    DHIP_HIPCHK(c, c->diag.ensure(sizeof(double) * (R + 1)));
    RowP2Args ra{c->mesh, c->cut.as<int64_t>(), c->ncut, c->cellinfo.as<uint32_t>(), c->flags.as<uint8_t>(),
                 c->dofs.as<int32_t>(), c->rowowner.as<int32_t>(), R, c->loc.as<double>(), c->rowlen.as<int32_t>(),
                 c->rowptr.as<int32_t>(), nullptr, nullptr, nullptr, nullptr, nullptr, c->rhs.as<double>(),
                 c->diag.as<double>(), lc ? lc[0] : 1.0, lc ? lc[1] : 1.0, lc ? lc[2] : 1.0};
    const unsigned nb = blocks_for(R, ROWP2_WAVES);
    if (R > 0) DHIP_LAUNCH(c, "rowcount_p2", k_rows_p2<false>, nb, 64 * ROWP2_WAVES, 0, ra);
    int64_t total = 0;
    int rc = scan_exclusive_i32(c, c->rowlen.as<int32_t>(), c->rowptr.as<int32_t>(), R, &total);
    if (rc) return rc;
    if (total >= ((int64_t)1 << 31) - 1)
        return set_error(c, DHIP_E_CAPACITY, "%lld nonzeros exceed int32 row pointers", (long long)total);
    {
        int32_t t32 = (int32_t)total;
        DHIP_HIPCHK(c, hipMemcpyAsync(c->rowptr.as<int32_t>() + R, &t32, sizeof t32, hipMemcpyHostToDevice,
                                      c->stream));
    }
    DHIP_HIPCHK(c, c->col.ensure(sizeof(int32_t) * (total + 1)));
    for (int k = 0; k < 4; ++k) DHIP_HIPCHK(c, c->val[k].ensure(sizeof(double) * (total + 1)));
    ra.col = c->col.as<int32_t>();
    ra.vm = c->val[0].as<double>();
    ra.vl = c->val[1].as<double>();
    ra.vs = c->val[2].as<double>();
    ra.vc = c->val[3].as<double>();
    if (R > 0) DHIP_LAUNCH(c, "rowfill_p2", k_rows_p2<true>, nb, 64 * ROWP2_WAVES, 0, ra);
    c->nnz = total;
    c->stage = 4;
    if (nnz) *nnz = total;
    return DHIP_OK;
}

static int qq_params(dhip_ctx *c, const dhip_quaqua_params_t *p, QQParams &P) {
    P = QQParams{20, 1e-7, 1e-2, 8};
    if (p) {
        if (p->maxiter < 0 || !(p->tol >= 0.) || p->max_damping < 0)
            return set_error(c, DHIP_E_ARG, "QuaQua parameters: maxiter, max_damping >= 0, tol >= 0");
        P = QQParams{p->maxiter, p->tol, p->armijo_c, p->max_damping};
    }
    return DHIP_OK;
}

// the recovered gradient at the P2 nodes of the cut tets (k_qq_first, k_qq_node_gradient)
static int qq_node_gradients(dhip_ctx *c) {
    const int64_t N = c->ncut;
    DHIP_HIPCHK(c, c->qq_first.ensure(sizeof(int32_t) * (10 * N + 1)));
    DHIP_HIPCHK(c, c->qq_gh.ensure(sizeof(double) * (30 * N + 3)));
    DHIP_HIPCHK(c, c->qq_list.ensure(sizeof(int32_t) * (10 * N + 2)));
    if (N > 0) {
        int32_t *cnt = c->qq_list.as<int32_t>() + 10 * N;
        DHIP_LAUNCH(c, "zero_count", k_zero_u32, 1, 64, 0, reinterpret_cast<uint32_t *>(cnt), 1);
        QQGradArgs ga{c->mesh, c->phi, c->cut.as<int64_t>(), N, c->cellinfo.as<uint32_t>(), c->qq_first.as<int32_t>(),
                      c->qq_list.as<int32_t>(), cnt, c->qq_gh.as<double>()};
        DHIP_LAUNCH(c, "first_qq", k_qq_first, blocks_for(10 * N, QQ_FIRST_THREADS), QQ_FIRST_THREADS, 0, ga);
        DHIP_LAUNCH(c, "node_gradient_qq", k_qq_node_gradient, QQN_BLOCKS, 64, 0, ga);
    }
    return DHIP_OK;
}

// the mapped-surface forms of every cut tet (after the planar pass has numbered the P2 DOFs): they
// replace the mass, Laplace-Beltrami and load slots, and flags keep only tets with a lattice-1 patch
static int local_p2qq_impl(dhip_ctx *c, const dhip_quaqua_params_t *p, int32_t rhs_kind, const double *rhs_param) {
    if (c->stage < 3 || c->space != 2)
        return set_error(c, DHIP_E_STATE, "dhip_number_p2if must precede dhip_local_p2_quaqua");
    if (rhs_kind < DHIP_RHS_NONE || rhs_kind > DHIP_RHS_VALUES)
        return set_error(c, DHIP_E_ARG, "unknown right-hand-side kind %d", (int)rhs_kind);
    QQParams P;
    int rc0 = qq_params(c, p, P);
    if (rc0) return rc0;
    const double *vals = nullptr;
    const int32_t *off = nullptr;
    if (rhs_kind == DHIP_RHS_VALUES) {
        // the values of f at the base points dhip_quaqua_points handed out for these cut tets and parameters
        if (!c->qqp_have_vals || c->qqp_gen != c->geo_gen || c->qqp_ncut != c->ncut || c->qqp_maxiter != P.maxiter ||
            c->qqp_tol != P.tol || c->qqp_armijo != P.armijo_c || c->qqp_maxdamp != P.max_damping)
            return set_error(c, DHIP_E_STATE,
                             "DHIP_RHS_VALUES on the mapped surface needs dhip_quaqua_points and dhip_quaqua_values on "
                             "the current cut tets with the same parameters");
        vals = c->qqp_vals.as<double>();
        off = c->qqp_off.as<int32_t>();
    }
    const int64_t N = c->ncut;
    DHIP_LAUNCH(c, "zero_flags", k_zero_u32, 1, 64, 0, c->errflag.as<uint32_t>(), 64);
    int rc = qq_node_gradients(c);
    if (rc) return rc;
    LocalP2QQArgs la{c->mesh, c->phi, c->cut.as<int64_t>(), N, rhs_kind, rhs_param ? rhs_param[0] : 0.0, P,
                     c->loc.as<double>(), c->flags.as<uint8_t>(), c->errflag.as<int32_t>(),
                     c->qq_first.as<int32_t>(), c->qq_gh.as<double>(), nullptr, vals, off};
    if (N > 0) DHIP_LAUNCH(c, "local_p2qq", k_local_p2qq, (unsigned)((N + QQ_TPW - 1) / QQ_TPW), 64, 0, la);
    rc = check_err_flag(c);
    if (rc) return rc;
    c->stage = 3;
    return DHIP_OK;
}

int dhip_quaqua_points(dhip_ctx_t *c, const dhip_quaqua_params_t *params, int64_t *npoints, double *xyz,
                       int32_t where) {
    if (!c || !npoints) return DHIP_E_ARG;
    if (c->stage < 1) return set_error(c, DHIP_E_STATE, "dhip_classify must precede dhip_quaqua_points");
    if (c->band >= 0.) return set_error(c, DHIP_E_STATE, "the narrow band is a P1 assembly");
    QQParams P;
    int rc = qq_params(c, params, P);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int64_t N = c->ncut;
    DevBuf cnt;
    DHIP_HIPCHK(c, cnt.ensure(sizeof(int32_t) * (N + 1)));
    DHIP_HIPCHK(c, c->qqp_off.ensure(sizeof(int32_t) * (N + 1)));
    int64_t total = 0;
    if (N > 0) {
        DHIP_LAUNCH(c, "quaqua_count", k_qq_count, blocks_for(N, 128), 128, 0, c->mesh, c->phi, c->cut.as<int64_t>(), N,
                    cnt.as<int32_t>());
        rc = scan_exclusive_i32(c, cnt.as<int32_t>(), c->qqp_off.as<int32_t>(), N, &total);
        if (rc) return rc;
        if (total > INT32_MAX)
            return set_error(c, DHIP_E_ARG, "%lld mapped quadrature points exceed the 32-bit offsets",
                             (long long)total);
    }
    const int32_t t32 = (int32_t)total;
    DHIP_HIPCHK(c, hipMemcpyAsync(c->qqp_off.as<int32_t>() + N, &t32, sizeof t32, hipMemcpyHostToDevice, c->stream));
    c->qqp_n = total;
    c->qqp_ncut = N;
    c->qqp_gen = c->geo_gen;
    c->qqp_have_vals = false;
    c->qqp_maxiter = P.maxiter;
    c->qqp_tol = P.tol;
    c->qqp_armijo = P.armijo_c;
    c->qqp_maxdamp = P.max_damping;
    *npoints = total;
    if (xyz && total > 0) {
        DevBuf tmp;
        double *dst = xyz;
        if (where != DHIP_DEVICE) {
            DHIP_HIPCHK(c, tmp.ensure(sizeof(double) * 3 * total));
            dst = tmp.as<double>();
        }
        DHIP_LAUNCH(c, "zero_flags", k_zero_u32, 1, 64, 0, c->errflag.as<uint32_t>(), 64);
        rc = qq_node_gradients(c);
        if (rc) return rc;
        LocalP2QQArgs la{c->mesh, c->phi, c->cut.as<int64_t>(), N, DHIP_RHS_NONE, 0.0, P, nullptr, nullptr,
                         c->errflag.as<int32_t>(), c->qq_first.as<int32_t>(), c->qq_gh.as<double>(), dst, nullptr,
                         c->qqp_off.as<int32_t>()};
        DHIP_LAUNCH(c, "quaqua_points", k_local_p2qq, (unsigned)((N + QQ_TPW - 1) / QQ_TPW), 64, 0, la);
        rc = check_err_flag(c);
        if (rc) return rc;
        if (where != DHIP_DEVICE)
            DHIP_HIPCHK(c, hipMemcpyAsync(xyz, dst, sizeof(double) * 3 * total, hipMemcpyDeviceToHost, c->stream));
        DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_quaqua_values(dhip_ctx_t *c, const double *values, int64_t npoints, int32_t where) {
    if (!c || (npoints > 0 && !values) || npoints < 0) return DHIP_E_ARG;
    if (c->qqp_n < 0 || c->qqp_gen != c->geo_gen)
        return set_error(c, DHIP_E_STATE, "dhip_quaqua_points must precede dhip_quaqua_values on this level set");
    if (npoints != c->qqp_n)
        return set_error(c, DHIP_E_ARG, "%lld values for %lld mapped quadrature points", (long long)npoints,
                         (long long)c->qqp_n);
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    DHIP_HIPCHK(c, c->qqp_vals.ensure(sizeof(double) * (npoints + 1)));
    if (npoints > 0)
        DHIP_HIPCHK(c, hipMemcpyAsync(c->qqp_vals.p, values, sizeof(double) * npoints,
                                      where == DHIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                      c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    c->qqp_have_vals = true;
    return flush_profile(c);
}

int dhip_p2_recovered_gradient(dhip_ctx_t *c, double *grad, int32_t where) {
    if (!c || !grad) return DHIP_E_ARG;
    if (!c->has_mesh || !c->phi) return set_error(c, DHIP_E_STATE, "mesh and level set required");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int64_t nn = (2LL * c->mesh.n[0] + 1) * (2LL * c->mesh.n[1] + 1) * (2LL * c->mesh.n[2] + 1);
    DevBuf tmp;
    double *dst = grad;
    if (where != DHIP_DEVICE) {
        DHIP_HIPCHK(c, tmp.ensure(sizeof(double) * 3 * nn));
        dst = tmp.as<double>();
    }
    RecGradArgs ra{c->mesh, c->phi, nn, dst};
    DHIP_LAUNCH(c, "recovered_gradient", k_recovered_gradient, blocks_for(nn, 128), 128, 0, ra);
    if (where != DHIP_DEVICE)
        DHIP_HIPCHK(c, hipMemcpyAsync(grad, dst, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_local_p2_quaqua(dhip_ctx_t *c, const dhip_quaqua_params_t *params, int32_t rhs_kind,
                         const double *rhs_param) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = local_p2qq_impl(c, params, rhs_kind, rhs_param);
    if (rc) return rc;
    return flush_profile(c);
}

int dhip_assemble_p2_quaqua(dhip_ctx_t *c, int32_t zero_policy, const dhip_quaqua_params_t *params,
                            int32_t rhs_kind, const double *rhs_param, double omit_bound, const double lc[3],
                            int64_t *ncut, int64_t *ndof, int64_t *nnz) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = classify_impl(c, zero_policy, ncut, false);
    if (!rc) rc = hat_p2_impl(c);
    if (!rc) rc = number_p2_impl(c, omit_bound, ndof);
    if (!rc) rc = local_p2qq_impl(c, params, rhs_kind, rhs_param);
    if (!rc) rc = build_p2_impl(c, lc, nnz);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_local_p2(dhip_ctx_t *c, const double coef[3], int32_t rhs_kind, const double *rhs_param) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = local_p2_impl(c, coef, rhs_kind, rhs_param);
    if (rc) return rc;
    return flush_profile(c);
}

int dhip_number_p2if(dhip_ctx_t *c, double omit_bound, int64_t *ndof) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = number_p2_impl(c, omit_bound, ndof);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_build_p2(dhip_ctx_t *c, const double lc[3], int64_t *nnz) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = build_p2_impl(c, lc, nnz);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_assemble_p2(dhip_ctx_t *c, int32_t zero_policy, const double coef[3], int32_t rhs_kind,
                     const double *rhs_param, double omit_bound, const double lc[3], int64_t *ncut,
                     int64_t *ndof, int64_t *nnz) {
    if (!c) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = classify_impl(c, zero_policy, ncut, false);
    if (!rc) rc = local_p2_impl(c, coef, rhs_kind, rhs_param);
    if (!rc) rc = number_p2_impl(c, omit_bound, ndof);
    if (!rc) rc = build_p2_impl(c, lc, nnz);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

static int d2h(dhip_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!dst || bytes == 0) return DHIP_OK;
    DHIP_HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    return DHIP_OK;
}

int dhip_get_cut_tets(dhip_ctx_t *c, int64_t *tets) {
    if (!c) return DHIP_E_ARG;
    if (c->stage < 1) return set_error(c, DHIP_E_STATE, "no classification");
    int rc = d2h(c, tets, c->cut.p, sizeof(int64_t) * c->ncut);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

int dhip_get_local(dhip_ctx_t *c, double *local57, uint8_t *flags) {
    if (!c) return DHIP_E_ARG;
    if (c->stage < 2 || c->space != 1) return set_error(c, DHIP_E_STATE, "no P1 local matrices");
    const int64_t N = c->ncut;
    if (local57 && N > 0) {
        std::vector<double> soa((size_t)LOC_N * N);
        int rc = d2h(c, soa.data(), c->loc.p, sizeof(double) * LOC_N * N);
        if (rc) return rc;
        DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < N; ++i)
            for (int e = 0; e < LOC_N; ++e) local57[i * LOC_N + e] = soa[(size_t)e * N + i];
    }
    int rc = d2h(c, flags, c->flags.p, N);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

int dhip_get_local_p2(dhip_ctx_t *c, double *local320, uint8_t *flags) {
    if (!c) return DHIP_E_ARG;
    if (c->stage < 2 || c->space != 2) return set_error(c, DHIP_E_STATE, "no P2 local matrices");
    const int64_t N = c->ncut;
    if (local320 && N > 0) {   // the device array is tet-major already
        int rc = d2h(c, local320, c->loc.p, sizeof(double) * LOC2_N * N);
        if (rc) return rc;
    }
    int rc = d2h(c, flags, c->flags.p, N);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

int dhip_get_dof_nodes(dhip_ctx_t *c, int64_t *node_of_dof) {
    if (!c || !node_of_dof) return DHIP_E_ARG;
    if (c->stage < 3) return set_error(c, DHIP_E_STATE, "no numbering");
    const int per = c->space == 2 ? 10 : 4;
    std::vector<int32_t> owner(c->ndof);
    std::vector<int64_t> cut(c->ncut);
    int rc = d2h(c, owner.data(), c->rowowner.p, sizeof(int32_t) * c->ndof);
    if (!rc) rc = d2h(c, cut.data(), c->cut.p, sizeof(int64_t) * c->ncut);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    const Tables &T = host_tables();
    const MeshP &m = c->mesh;
    const int64_t h1 = 2 * (int64_t)m.n[0] + 1, h2 = 2 * (int64_t)m.n[1] + 1;
    for (int64_t r = 0; r < c->ndof; ++r) {
        const int64_t s = owner[r], t = cut[s / per];
        const int k = (int)(t % 6), d = (int)(s % per);
        const int n27 = per == 10 ? T.p2_node27[k][d] : T.tet_node27[k][T.lat_of_p2[d]];
        const int64_t cell = t / 6;
        const int64_t j1 = 2 * (cell % m.n[0]) + n27 % 3;
        const int64_t j2 = 2 * ((cell / m.n[0]) % m.n[1]) + (n27 / 3) % 3;
        const int64_t j3 = 2 * (cell / ((int64_t)m.n[0] * m.n[1])) + n27 / 9;
        node_of_dof[r] = (j3 * h2 + j2) * h1 + j1;
    }
    return DHIP_OK;
}

int dhip_get_tet_dofs(dhip_ctx_t *c, int32_t *dofs) {
    if (!c) return DHIP_E_ARG;
    if (c->stage < 3) return set_error(c, DHIP_E_STATE, "no numbering");
    int rc = d2h(c, dofs, c->dofs.p, sizeof(int32_t) * (c->space == 2 ? 10 : 4) * c->ncut);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

int dhip_get_dof_vertices(dhip_ctx_t *c, int64_t *vertex_of_dof) {
    if (!c || !vertex_of_dof) return DHIP_E_ARG;
    if (c->stage < 3 || c->space != 1) return set_error(c, DHIP_E_STATE, "no P1 numbering");
    std::vector<int32_t> owner(c->ndof);
    std::vector<int64_t> cut(c->ncut);
    int rc = d2h(c, owner.data(), c->rowowner.p, sizeof(int32_t) * c->ndof);
    if (!rc) rc = d2h(c, cut.data(), c->cut.p, sizeof(int64_t) * c->ncut);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    const Tables &T = host_tables();
    const MeshP &m = c->mesh;
    for (int64_t r = 0; r < c->ndof; ++r) {
        const int64_t s = owner[r], t = cut[s >> 2];
        const int k = (int)(t % 6), i = (int)(s & 3);
        const int64_t cell = t / 6;
        const int64_t i1 = cell % m.n[0] + T.kuhn[k][i][0];
        const int64_t i2 = (cell / m.n[0]) % m.n[1] + T.kuhn[k][i][1];
        const int64_t i3 = cell / ((int64_t)m.n[0] * m.n[1]) + T.kuhn[k][i][2];
        vertex_of_dof[r] = (i3 * (m.n[1] + 1) + i2) * (m.n[0] + 1) + i1;
    }
    return DHIP_OK;
}

int dhip_get_csr(dhip_ctx_t *c, int32_t *rowptr, int32_t *colind, double *mass, double *lb, double *stab,
                 double *comb, double *rhs) {
    if (!c) return DHIP_E_ARG;
    if (c->stage < 4) return set_error(c, DHIP_E_STATE, "no assembled matrix");
    const size_t nz = (size_t)c->nnz, R = (size_t)c->ndof;
    int rc = 0;
    rc = rc ? rc : d2h(c, rowptr, c->rowptr.p, sizeof(int32_t) * (R + 1));
    rc = rc ? rc : d2h(c, colind, c->col.p, sizeof(int32_t) * nz);
    rc = rc ? rc : d2h(c, mass, c->val[0].p, sizeof(double) * nz);
    rc = rc ? rc : d2h(c, lb, c->val[1].p, sizeof(double) * nz);
    rc = rc ? rc : d2h(c, stab, c->val[2].p, sizeof(double) * nz);
    rc = rc ? rc : d2h(c, comb, c->val[3].p, sizeof(double) * nz);
    rc = rc ? rc : d2h(c, rhs, c->rhs.p, sizeof(double) * R);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

int dhip_device_arrays(dhip_ctx_t *c, const int32_t **rowptr, const int32_t **colind, const double **comb,
                       const double **rhs) {
    if (!c) return DHIP_E_ARG;
    if (c->stage < 4) return set_error(c, DHIP_E_STATE, "no assembled matrix");
    if (rowptr) *rowptr = c->rowptr.as<int32_t>();
    if (colind) *colind = c->col.as<int32_t>();
    if (comb) *comb = c->val[3].as<double>();
    if (rhs) *rhs = c->rhs.as<double>();
    return DHIP_OK;
}

// ---------------------------------------------------------------------------
// CSR objects
// ---------------------------------------------------------------------------
static int csr_finish(dhip_ctx *c, dhip_csr *A) {
    (void)c;
    (void)A;
    return DHIP_OK;
}

int dhip_csr_create(dhip_ctx_t *c, int64_t nrows, int64_t nnz, const int32_t *rowptr, const int32_t *colind,
                    const double *val, int32_t where, dhip_csr_t **out) {
    if (!c || !out || nrows < 0 || nnz < 0 || !rowptr || (nnz > 0 && (!colind || !val))) return DHIP_E_ARG;
    if (nnz >= ((int64_t)1 << 31)) return set_error(c, DHIP_E_CAPACITY, "nnz beyond int32");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    dhip_csr *A = new dhip_csr();
    A->ctx = c;
    A->n = nrows;
    A->nnz = nnz;
    auto fail = [&](int rc) { delete A; return rc; };
    if (A->own_rowptr.ensure(sizeof(int32_t) * (nrows + 1)) || A->own_col.ensure(sizeof(int32_t) * (nnz + 1)) ||
        A->own_val.ensure(sizeof(double) * (nnz + 1)))
        return fail(set_error(c, DHIP_E_NOMEM, "csr allocation failed"));
    const hipMemcpyKind k = where == DHIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (hipMemcpyAsync(A->own_rowptr.p, rowptr, sizeof(int32_t) * (nrows + 1), k, c->stream) ||
        (nnz > 0 && (hipMemcpyAsync(A->own_col.p, colind, sizeof(int32_t) * nnz, k, c->stream) ||
                     hipMemcpyAsync(A->own_val.p, val, sizeof(double) * nnz, k, c->stream))))
        return fail(set_error(c, DHIP_E_HIP, "csr upload failed"));
    A->rowptr = A->own_rowptr.as<int32_t>();
    A->col = A->own_col.as<int32_t>();
    A->val = A->own_val.as<double>();
    int rc = csr_finish(c, A);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = set_error(c, DHIP_E_HIP, "sync failed");
    if (rc) return fail(rc);
    *out = A;
    return flush_profile(c);
}

int dhip_csr_from_assembly(dhip_ctx_t *c, int32_t op, dhip_csr_t **out) {
    if (!c || !out || op < 0 || op > 3) return DHIP_E_ARG;
    if (c->stage < 4) return set_error(c, DHIP_E_STATE, "no assembled matrix");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    dhip_csr *A = new dhip_csr();
    A->ctx = c;
    A->n = c->ndof;
    A->nnz = c->nnz;
    A->rowptr = c->rowptr.as<int32_t>();
    A->col = c->col.as<int32_t>();
    A->val = c->val[op].as<double>();
    int rc = csr_finish(c, A);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = set_error(c, DHIP_E_HIP, "sync failed");
    if (rc) {
        delete A;
        return rc;
    }
    *out = A;
    return flush_profile(c);
}

void dhip_csr_destroy(dhip_csr_t *A) {
    if (!A) return;
    (void)hipSetDevice(A->ctx->device);
    (void)hipStreamSynchronize(A->ctx->stream);
    A->own_rowptr.release();
    A->own_col.release();
    A->own_val.release();
    delete A;
}

int dhip_csr_info(const dhip_csr_t *A, int64_t *nrows, int64_t *nnz) {
    if (!A) return DHIP_E_ARG;
    if (nrows) *nrows = A->n;
    if (nnz) *nnz = A->nnz;
    return DHIP_OK;
}

// ---------------------------------------------------------------------------
// vector core
// ---------------------------------------------------------------------------
// Stage caller vectors on the device (host pointers) or use them directly.
struct VecStage {
    dhip_ctx *c;
    DevBuf buf;
    ~VecStage() { buf.release(); }
};

static int stage_in(dhip_ctx *c, DevBuf &buf, const double *src, int64_t n, int32_t where, const double **dst) {
    if (where == DHIP_DEVICE) {
        *dst = src;
        return DHIP_OK;
    }
    DHIP_HIPCHK(c, buf.ensure(sizeof(double) * (n + 1)));
    if (n > 0) DHIP_HIPCHK(c, hipMemcpyAsync(buf.p, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    *dst = buf.as<double>();
    return DHIP_OK;
}

int dhip_spmv(dhip_ctx_t *c, const dhip_csr_t *A, const double *x, double *y, int32_t where) {
    if (!c || !A || !x || !y) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    DevBuf bx, by;
    const double *dx;
    int rc = stage_in(c, bx, x, A->n, where, &dx);
    if (rc) return rc;
    double *dy = y;
    if (where != DHIP_DEVICE) {
        DHIP_HIPCHK(c, by.ensure(sizeof(double) * (A->n + 1)));
        dy = by.as<double>();
    }
    if (A->n > 0)
    {
        const int rb = spmv_tile_rows(A->n, A->nnz);
        DHIP_LAUNCH(c, "spmv", k_spmv<false>, blocks_for(A->n, rb), SPMV_T, 0, A->n, rb, A->rowptr, A->col, A->val, dx,
                    nullptr, dy);
    }
    if (where != DHIP_DEVICE && A->n > 0)
        DHIP_HIPCHK(c, hipMemcpyAsync(y, dy, sizeof(double) * A->n, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    bx.release();
    by.release();
    return flush_profile(c);
}

int dhip_spmv_t(dhip_ctx_t *c, const dhip_csr_t *A, int64_t ncols, const double *x, double *y, int32_t where) {
    if (!c || !A || !x || !y || ncols < 0) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = A->n, nnz = A->nnz, m = ncols;
    DevBuf bx, by, cnt, trp, tcol, tval;
    const double *dx;
    int rc = stage_in(c, bx, x, n, where, &dx);
    if (rc) return rc;
    double *dy = y;
    if (where != DHIP_DEVICE) {
        DHIP_HIPCHK(c, by.ensure(sizeof(double) * (m + 1)));
        dy = by.as<double>();
    }
    if (m > 0) {
        DHIP_HIPCHK(c, cnt.ensure(sizeof(int32_t) * (m + 1)));
        DHIP_HIPCHK(c, trp.ensure(sizeof(int32_t) * (m + 1)));
        DHIP_HIPCHK(c, tcol.ensure(sizeof(int32_t) * (nnz + 1)));
        DHIP_HIPCHK(c, tval.ensure(sizeof(double) * (nnz + 1)));
        DHIP_LAUNCH(c, "spmv_t_zero", k_zero_u32, blocks_for(m + 1, 256), 256, 0, cnt.as<uint32_t>(), (int)(m + 1));
        DHIP_LAUNCH(c, "zero_flags", k_zero_u32, 1, 64, 0, c->errflag.as<uint32_t>(), 64);
        if (nnz > 0)
            DHIP_LAUNCH(c, "spmv_t_count", k_tcount, blocks_for(nnz, 256), 256, 0, nnz, m, A->col, cnt.as<int32_t>(),
                        c->errflag.as<int32_t>());
        int64_t tot = 0;
        rc = scan_exclusive_i32(c, cnt.as<int32_t>(), trp.as<int32_t>(), m, &tot);
        if (rc) return rc;
        int32_t fl = 0;
        DHIP_HIPCHK(c, hipMemcpyAsync(&fl, c->errflag.p, sizeof fl, hipMemcpyDeviceToHost, c->stream));
        DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
        if (fl) return set_error(c, DHIP_E_ARG, "dhip_spmv_t: a column index is not below ncols = %lld", (long long)m);
        const int32_t t32 = (int32_t)tot;
        DHIP_HIPCHK(c, hipMemcpyAsync(trp.as<int32_t>() + m, &t32, sizeof t32, hipMemcpyHostToDevice, c->stream));
        DHIP_LAUNCH(c, "spmv_t_zero", k_zero_u32, blocks_for(m + 1, 256), 256, 0, cnt.as<uint32_t>(), (int)(m + 1));
        if (n > 0)
            DHIP_LAUNCH(c, "spmv_t_fill", k_tfill, blocks_for(n, 256), 256, 0, n, A->rowptr, A->col, A->val,
                        (const int32_t *)trp.as<int32_t>(), cnt.as<int32_t>(), tcol.as<int32_t>(), tval.as<double>());
        DHIP_LAUNCH(c, "spmv_t_sort", k_tsort, blocks_for(m, 256), 256, 0, m, (const int32_t *)trp.as<int32_t>(),
                    tcol.as<int32_t>(), tval.as<double>());
        const int rb = spmv_tile_rows(m, nnz);
        DHIP_LAUNCH(c, "spmv_t", k_spmv<false>, blocks_for(m, rb), SPMV_T, 0, m, rb, (const int32_t *)trp.as<int32_t>(),
                    (const int32_t *)tcol.as<int32_t>(), (const double *)tval.as<double>(), dx, nullptr, dy);
        if (where != DHIP_DEVICE)
            DHIP_HIPCHK(c, hipMemcpyAsync(y, dy, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_dot(dhip_ctx_t *c, int64_t n, const double *x, const double *y, double *result, int32_t where) {
    if (!c || !x || !y || !result || n < 0) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    DevBuf bx, by;
    const double *dx, *dy;
    int rc = stage_in(c, bx, x, n, where, &dx);
    if (!rc) rc = stage_in(c, by, y, n, where, &dy);
    if (rc) return rc;
    const unsigned nb = std::max(1u, std::min(1024u, blocks_for(n, KB)));
    DHIP_HIPCHK(c, c->red_tmp.ensure(sizeof(double) * (nb + 2)));
    double *part = c->red_tmp.as<double>();
    DHIP_LAUNCH(c, "dot_partial", k_dot_partial, nb, KB, 0, n, dx, dy, part);
    DHIP_LAUNCH(c, "dot_reduce", k_reduce, 1, KB, 0, part, (int)nb, part + nb);
    DHIP_HIPCHK(c, hipMemcpyAsync(result, part + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    bx.release();
    by.release();
    return flush_profile(c);
}

int dhip_axpy(dhip_ctx_t *c, int64_t n, double a, const double *x, double *y, int32_t where) {
    if (!c || !x || !y || n < 0) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    DevBuf bx, by;
    const double *dx;
    int rc = stage_in(c, bx, x, n, where, &dx);
    if (rc) return rc;
    double *dy = y;
    if (where != DHIP_DEVICE) {
        DHIP_HIPCHK(c, by.ensure(sizeof(double) * (n + 1)));
        if (n > 0) DHIP_HIPCHK(c, hipMemcpyAsync(by.p, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        dy = by.as<double>();
    }
    if (n > 0) DHIP_LAUNCH(c, "axpy", k_axpy, blocks_for(n, KB), KB, 0, n, a, dx, dy);
    if (where != DHIP_DEVICE && n > 0)
        DHIP_HIPCHK(c, hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    bx.release();
    by.release();
    return flush_profile(c);
}

// ---------------------------------------------------------------------------
// Krylov solvers
// ---------------------------------------------------------------------------
}  // extern "C"

// Launch geometry of the persistent solvers: one row per thread when the rows
// fit 512 x (CUs) threads, else RPT rows per thread; one block per CU at most.
// Results the persistent solvers leave in the context's staging buffer (int32 slots 0..3: iterations,
// status, outer iterations, spare; then the final residual at byte 16), read back with one copy.
struct SolverResult {
    int32_t i[4];
    double resid;
};
static_assert(sizeof(SolverResult) == 24 && offsetof(SolverResult, resid) == 16, "staging layout");

static int persistent_geometry(dhip_ctx *c, int64_t n, unsigned *nb, int *rpt) {
    const int64_t cap_blocks = c->cu_count;
    // default 256 rows per block: measured optimum of compute (gathers) vs synchronisation
    // (block count) for 30k-row systems; drops_amd/tune.py re-tunes per matrix
    int64_t rows = c->pcg_rows_per_block > 0 ? c->pcg_rows_per_block : 256;
    rows = std::max<int64_t>(rows, (n + cap_blocks - 1) / cap_blocks);  // at most one block per CU
    int r = 1;
    while ((int64_t)r * PB < rows && r < 16) r *= 2;
    if ((int64_t)r * PB < rows)
        return set_error(c, DHIP_E_CAPACITY, "%lld rows exceed the persistent solver (max %lld)", (long long)n,
                         (long long)(16LL * PB * cap_blocks));
    *rpt = r;
    *nb = (unsigned)std::max<int64_t>(1, (n + rows - 1) / rows);
    return DHIP_OK;
}

template <class ArgT>
static int launch_persistent(dhip_ctx *c, const char *name, const void *const kernels[5], int rpt, unsigned nb,
                             ArgT &args, size_t lds = LDS_BYTES) {
    const int idx = rpt == 1 ? 0 : rpt == 2 ? 1 : rpt == 4 ? 2 : rpt == 8 ? 3 : 4;
    const void *k = kernels[idx];
    if (lds > (size_t)LDS_BYTES)
        DHIP_HIPCHK(c, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    const auto key = std::make_pair(k, lds);
    const auto hit = c->occupancy.find(key);
    if (hit != c->occupancy.end()) {
        per_cu = hit->second;
    } else {
        DHIP_HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, PB, lds));
        c->occupancy[key] = per_cu;
    }
    if ((int64_t)per_cu * c->cu_count < (int64_t)nb)
        return set_error(c, DHIP_E_HIP, "persistent kernel cannot keep %u blocks resident", nb);
    void *argv[] = {&args};
    ProfScope ps(c, name, true);
    if (!c->coop_refused) {
        const hipError_t e = hipLaunchCooperativeKernel(k, dim3(nb), dim3(PB), argv, lds, c->stream);
        if (e == hipSuccess) return DHIP_OK;
        // The cooperative-launch resource is per device and can be held by another process.
        // The grid is already capped at one block per CU (and checked against occupancy
        // above) and every wait is bounded, so launch the same grid normally instead.
        (void)hipGetLastError();
        c->coop_refused = true;
    }
    DHIP_HIPCHK(c, hipLaunchKernel(k, dim3(nb), dim3(PB), argv, lds, c->stream));
    return DHIP_OK;
}

// Dynamic LDS of a persistent kernel: the device's per-block limit less the kernel's static LDS,
// at most 128 KiB and never below the CSR-slice capacity.
static size_t persistent_lds(dhip_ctx *c, const void *k) {
    int maxs = 0;
    hipFuncAttributes fa{};
    if (hipDeviceGetAttribute(&maxs, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess ||
        hipFuncGetAttributes(&fa, k) != hipSuccess)
        return LDS_BYTES;
    int64_t dyn = (int64_t)maxs - (int64_t)fa.sharedSizeBytes - 256;
    dyn = std::min<int64_t>(dyn, 128 * 1024) & ~(int64_t)15;
    return (size_t)std::max<int64_t>(dyn, LDS_BYTES);
}

// Work buffers of the FP64 single-reduction PCG: the Jacobi diagonal and the exchanged r, w, s
// (double buffered), the synchronisation slots (one per block) and the 24-byte result record.
struct PcgFp64Bufs {
    double *diag;
    double *vec[6];   // R0, R1, W0, W1, S0, S1
    SyncSlot *slots;  // at least one per block of the launch
    SolverResult *res;
    uint4 *tx = nullptr;   // tagged records of r, w, s ([3][2][n], 96 n bytes); nullptr: none
};

// Diagonal scan, slot clearing (a fill kernel: no memset node) and the persistent FP64 PCG launch.
// `alloc_ok` false: nothing on this path may allocate (no profiling events, no occupancy cache insert).
// `max_blocks` > 0: the slots hold at most that many blocks.  `tagged`: the tagged-exchange kernel
// k_pcg64_tx (rows per thread 1, 2 or 4), else k_pcg64_sr with the context's exchange.
static int pcg_fp64_launch(dhip_ctx *c, const dhip_csr *A, int32_t pc, const double *db, double *dx,
                           const PcgFp64Bufs &w, int32_t max_iter, double tol, int32_t rel, bool alloc_ok,
                           unsigned max_blocks, bool tagged) {
    const int64_t n = A->n;
    unsigned nb = 1;
    int rpt = 1;
    int rc = persistent_geometry(c, n, &nb, &rpt);
    if (rc) return rc;
    if (max_blocks && nb > max_blocks)
        return set_error(c, DHIP_E_CAPACITY, "%u blocks exceed the %u synchronisation slots of the scratch", nb,
                         max_blocks);
    const unsigned slot_words = (unsigned)(nb * sizeof(SyncSlot) / 4);
    if (alloc_ok) {
        DHIP_LAUNCH(c, "csr_diag", k_diag, blocks_for(n, KB), KB, 0, n, A->rowptr, A->col, A->val, w.diag,
                    (float *)nullptr);
        DHIP_LAUNCH(c, "zero_slots", k_zero_u32, blocks_for(slot_words, 256), 256, 0,
                    reinterpret_cast<uint32_t *>(w.slots), (int)slot_words);
    } else {
        hipLaunchKernelGGL(k_diag, dim3(blocks_for(n, KB)), dim3(KB), 0, c->stream, n, A->rowptr, A->col, A->val,
                           w.diag, (float *)nullptr);
        DHIP_HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(slot_words, 256)), dim3(256), 0, c->stream,
                           reinterpret_cast<uint32_t *>(w.slots), (int)slot_words);
        DHIP_HIPCHK(c, hipGetLastError());
    }
    int32_t *outs = w.res->i;
    double *outd = &w.res->resid;
    const double *diag = pc == DHIP_PC_JACOBI ? w.diag : nullptr;
    tagged = tagged && w.tx && rpt <= 4;
    if (tagged && n > 0) {
        // Every solve numbers its generations from 1, so records left in the buffer by an earlier solve
        // carry tags this solve will wait for: a reader that arrives before the writer would take the
        // old value.  Clearing them (tag 0, never used) leaves only this solve's records acceptable.
        const int64_t words = 24 * n;   // three vectors of [2][n] 16-byte records
        if (alloc_ok) {
            DHIP_LAUNCH(c, "zero_records", k_zero_u32, blocks_for(words, 256), 256, 0,
                        reinterpret_cast<uint32_t *>(w.tx), (int)words);
        } else {
            hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(words, 256)), dim3(256), 0, c->stream,
                               reinterpret_cast<uint32_t *>(w.tx), (int)words);
            DHIP_HIPCHK(c, hipGetLastError());
        }
    }
    const int kr = rpt;   // selects the instantiation below
    PcgSrArgs pa{n, A->rowptr, A->col, A->val, diag, db, dx, {w.vec[0], w.vec[1]}, {w.vec[2], w.vec[3]},
                 {w.vec[4], w.vec[5]}, w.slots, max_iter, tol, rel, outs, outd, outs + 1, nullptr};
    pa.xcd_map = c->xcd_map;
    PcgTxArgs ta{n, A->rowptr, A->col, A->val, diag, db, dx, w.tx, w.tx + 2 * n, w.tx + 4 * n, w.slots,
                 max_iter, tol, rel, outs, outd, outs + 1, nullptr};
    ta.xcd_map = c->xcd_map;
    const void *ks_f[5] = {(const void *)k_pcg64_sr<1, false>, (const void *)k_pcg64_sr<2, false>,
                           (const void *)k_pcg64_sr<4, false>, (const void *)k_pcg64_sr<8, false>,
                           (const void *)k_pcg64_sr<16, false>};
    const void *ks_c[5] = {(const void *)k_pcg64_sr<1, true>, (const void *)k_pcg64_sr<2, true>,
                           (const void *)k_pcg64_sr<4, true>, (const void *)k_pcg64_sr<8, true>,
                           (const void *)k_pcg64_sr<16, true>};
    const void *ks_t[5] = {(const void *)k_pcg64_tx<1>, (const void *)k_pcg64_tx<2>, (const void *)k_pcg64_tx<4>,
                           nullptr, nullptr};
    const void *ks_p[5] = {(const void *)k_pcg64_pipe<1>, (const void *)k_pcg64_pipe<2>,
                           (const void *)k_pcg64_pipe<4>, nullptr, nullptr};
    // exchange 3: the pipelined arrangement over the same tagged records
    const bool pipe = tagged && c->coherent_exchange == 3;
    const void *const *ks = pipe ? ks_p : tagged ? ks_t : c->coherent_exchange ? ks_c : ks_f;
    const int idx = kr == 1 ? 0 : kr == 2 ? 1 : kr == 4 ? 2 : kr == 8 ? 3 : 4;
    // the tagged kernel keeps its slice and halo in up to 128 KiB of LDS
    const size_t lds = tagged ? c->tx_lds : (size_t)LDS_BYTES;
    ta.lds_bytes = (int64_t)lds;
    if (alloc_ok) {
        if (c->probe_on) {
            DHIP_HIPCHK(c, c->probe.ensure(sizeof(uint64_t) * 4 * (nb + 1)));
            pa.probe = ta.probe = c->probe.as<uint64_t>();
            c->probe_blocks = nb;
        }
        return tagged ? launch_persistent(c, "pcg64", ks, kr, nb, ta, lds)
                      : launch_persistent(c, "pcg64", ks, kr, nb, pa);
    }
    // allocation-free launch: occupancy queried directly, no event timing
    int per_cu = 0;
    DHIP_HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ks[idx], PB, lds));
    if ((int64_t)per_cu * c->cu_count < (int64_t)nb)
        return set_error(c, DHIP_E_HIP, "persistent kernel cannot keep %u blocks resident", nb);
    void *argv_s[] = {&pa};
    void *argv_t[] = {&ta};
    void **argv = tagged ? argv_t : argv_s;
    if (!c->coop_refused) {
        const hipError_t e = hipLaunchCooperativeKernel(ks[idx], dim3(nb), dim3(PB), argv, lds, c->stream);
        if (e == hipSuccess) return DHIP_OK;
        (void)hipGetLastError();
        c->coop_refused = true;
    }
    DHIP_HIPCHK(c, hipLaunchKernel(ks[idx], dim3(nb), dim3(PB), argv, lds, c->stream));
    return DHIP_OK;
}

// The FP64 solve with the result record read back into *out: the tagged kernel when the context
// asks for it and the buffers allow, relaunched as k_pcg64_sr when a row is longer than it handles.
static int pcg_fp64_run(dhip_ctx *c, const dhip_csr *A, int32_t pc, const double *db, double *dx,
                        const PcgFp64Bufs &w, int32_t max_iter, double tol, int32_t rel, bool alloc_ok,
                        unsigned max_blocks, SolverResult *out) {
    const bool tagged = c->coherent_exchange >= 2 && w.tx;
    int rc = pcg_fp64_launch(c, A, pc, db, dx, w, max_iter, tol, rel, alloc_ok, max_blocks, tagged);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipMemcpyAsync(out, w.res, sizeof *out, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    if (tagged && out->i[1] == 3) {
        rc = pcg_fp64_launch(c, A, pc, db, dx, w, max_iter, tol, rel, alloc_ok, max_blocks, false);
        if (rc) return rc;
        DHIP_HIPCHK(c, hipMemcpyAsync(out, w.res, sizeof *out, hipMemcpyDeviceToHost, c->stream));
        DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return DHIP_OK;
}

// Per-device scratch of dhip_pcg_noalloc: the slots of up to 256 blocks, the result record, then the
// diagonal and the six exchanged vectors of at most DHIP_NOALLOC_MAX_ROWS rows.
constexpr unsigned NOALLOC_MAX_BLOCKS = 256;
constexpr size_t NOALLOC_SLOT_BYTES = NOALLOC_MAX_BLOCKS * sizeof(SyncSlot);
constexpr size_t NOALLOC_RES_BYTES = 256;
constexpr size_t NOALLOC_VEC_BYTES = sizeof(double) * (size_t)DHIP_NOALLOC_MAX_ROWS;
constexpr size_t NOALLOC_BYTES = NOALLOC_SLOT_BYTES + NOALLOC_RES_BYTES + 7 * NOALLOC_VEC_BYTES;
// the tagged records of r, w, s (96 bytes a row) share the six vectors' 48 bytes a row: they fit
// systems of up to half the maximum rows
constexpr int64_t NOALLOC_TX_ROWS = DHIP_NOALLOC_MAX_ROWS / 2;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// Device address of this device's copy of `scratch` (the host-side name is only a stand-in).
// Called once per context, at creation, with the context's device current.
static int noalloc_setup(dhip_ctx *c) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        c->noalloc_scratch = nullptr;
        return set_error(c, DHIP_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    }
    c->noalloc_scratch = static_cast<unsigned char *>(p);
    return DHIP_OK;
}

extern "C" {

int dhip_pcg_noalloc(dhip_ctx_t *c, const dhip_csr_t *A, int32_t pc, const double *b, double *x,
                     int32_t *max_iter, double *tol, int32_t rel) {
    if (!c || !A || !b || !x || !max_iter || !tol) return DHIP_E_ARG;
    if (pc != DHIP_PC_NONE && pc != DHIP_PC_JACOBI) return set_error(c, DHIP_E_ARG, "unknown preconditioner");
    if (A->n > DHIP_NOALLOC_MAX_ROWS)
        return set_error(c, DHIP_E_CAPACITY, "%lld rows exceed the allocation-free solver's scratch (max %d rows)",
                         (long long)A->n, DHIP_NOALLOC_MAX_ROWS);
    if (!c->noalloc_scratch) return set_error(c, DHIP_E_STATE, "scratch address of this device is unknown");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = A->n;
    if (n == 0) {
        *max_iter = 0;
        *tol = 0.0;
        return DHIP_OK;
    }
    unsigned char *s = c->noalloc_scratch;
    PcgFp64Bufs w;
    w.slots = reinterpret_cast<SyncSlot *>(s);
    w.res = reinterpret_cast<SolverResult *>(s + NOALLOC_SLOT_BYTES);
    double *v = reinterpret_cast<double *>(s + NOALLOC_SLOT_BYTES + NOALLOC_RES_BYTES);
    w.diag = v;
    for (int k = 0; k < 6; ++k) w.vec[k] = v + (size_t)(k + 1) * DHIP_NOALLOC_MAX_ROWS;
    if (n <= NOALLOC_TX_ROWS) w.tx = reinterpret_cast<uint4 *>(w.vec[0]);
    SolverResult res;
    int rc = pcg_fp64_run(c, A, pc, b, x, w, *max_iter, *tol, rel, false, NOALLOC_MAX_BLOCKS, &res);
    if (rc) return rc;
    *max_iter = res.i[0];
    *tol = res.resid;
    if (res.i[1] == 2) return set_error(c, DHIP_E_HIP, "grid barrier timed out in the persistent solver");
    if (res.i[1] != 0)
        return set_error(c, DHIP_E_NOT_CONVERGED, "PCG did not converge: resid %g after %d", res.resid, res.i[0]);
    return DHIP_OK;
}

int dhip_pcg(dhip_ctx_t *c, const dhip_csr_t *A, int32_t pc, const double *b, double *x, int32_t where,
             int32_t *max_iter, double *tol, int32_t rel, int32_t precision) {
    if (!c || !A || !b || !x || !max_iter || !tol) return DHIP_E_ARG;
    if (pc != DHIP_PC_NONE && pc != DHIP_PC_JACOBI) return set_error(c, DHIP_E_ARG, "unknown preconditioner");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = A->n;
    DevBuf bb, bx;
    const double *db;
    int rc = stage_in(c, bb, b, n, where, &db);
    if (rc) return rc;
    double *dx = x;
    if (where != DHIP_DEVICE) {
        DHIP_HIPCHK(c, bx.ensure(sizeof(double) * (n + 1)));
        if (n > 0) DHIP_HIPCHK(c, hipMemcpyAsync(bx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        dx = bx.as<double>();
    }
    for (int k = 0; k < 6; ++k) DHIP_HIPCHK(c, c->kv[k].ensure(sizeof(double) * (n + 1)));
    DHIP_HIPCHK(c, c->red_tmp.ensure(sizeof(double) * (4 * 1024 + 16)));
    DHIP_HIPCHK(c, c->host_stage.ensure(64));
    int32_t *outs = c->host_stage.as<int32_t>();
    double *outd = reinterpret_cast<double *>(outs + 4);
    // Jacobi diagonal by a row scan on every solve (JACPcCL finds a_ii per Apply,
    // num/precond.h:163-168); FP32 copies for the refinement path
    DHIP_HIPCHK(c, c->kv[12].ensure(sizeof(double) * (n + 1)));
    DHIP_HIPCHK(c, c->kv[13].ensure(sizeof(float) * (n + 1)));
    const bool mixed = precision == DHIP_FP32_IR;
    if (mixed) DHIP_HIPCHK(c, c->kv[14].ensure(sizeof(float) * (A->nnz + 1)));
    if (n == 0) {
        *max_iter = 0;
        *tol = 0.0;
        return DHIP_OK;
    }
    double *diag = c->kv[12].as<double>();
    float *diag32 = c->kv[13].as<float>();
    float *val32 = c->kv[14].as<float>();
    int h_iter = 0, h_status = 1, h_outer = 0;
    double h_resid = 0.0;
    unsigned nb = 1;
    int rpt = 1;
    rc = persistent_geometry(c, n, &nb, &rpt);
    if (rc) return rc;
    DHIP_HIPCHK(c, c->bar.ensure(sizeof(SyncSlot) * (nb + 1)));
    if (precision == DHIP_FP64) {
        PcgFp64Bufs w;
        w.diag = diag;
        for (int k = 0; k < 6; ++k) w.vec[k] = c->kv[k].as<double>();
        w.slots = c->bar.as<SyncSlot>();
        w.res = reinterpret_cast<SolverResult *>(outs);
        if (c->coherent_exchange >= 2) {
            DHIP_HIPCHK(c, c->kv[16].ensure((size_t)96 * (n + 1)));
            w.tx = c->kv[16].as<uint4>();
        }
        SolverResult fres;
        rc = pcg_fp64_run(c, A, pc, db, dx, w, *max_iter, *tol, rel, true, 0, &fres);
        if (rc) return rc;
    } else {
        if (n > 0)
            DHIP_LAUNCH(c, "csr_diag", k_diag, blocks_for(n, KB), KB, 0, n, A->rowptr, A->col, A->val, diag,
                        mixed ? diag32 : nullptr);
        if (mixed && A->nnz > 0)
            DHIP_LAUNCH(c, "csr_to_f32", k_to_f32, blocks_for(A->nnz, KB), KB, 0, A->nnz, A->val, val32);
        DHIP_LAUNCH(c, "zero_slots", k_zero_u32, blocks_for((int64_t)nb * sizeof(SyncSlot) / 4, 256), 256, 0,
                    c->bar.as<uint32_t>(), (int)(nb * sizeof(SyncSlot) / 4));
    }
    if (precision == DHIP_FP64) {
        // launched by pcg_fp64_launch above
    } else if (precision == DHIP_FP64_CLASSIC) {
        PcgArgs pa{n, A->rowptr, A->col, A->val, pc == DHIP_PC_JACOBI ? diag : nullptr, db, dx,
                   c->kv[1].as<double>(), c->kv[2].as<double>(), c->kv[3].as<double>(), c->bar.as<SyncSlot>(),
                   *max_iter, *tol, rel, outs, outd, outs + 1};
        pa.xcd_map = c->xcd_map;
        const void *ks[5] = {(const void *)k_pcg64<1>, (const void *)k_pcg64<2>, (const void *)k_pcg64<4>,
                             (const void *)k_pcg64<8>, (const void *)k_pcg64<16>};
        rc = launch_persistent(c, "pcg64_classic", ks, rpt, nb, pa);
        if (rc) return rc;
    } else if (precision == DHIP_FP32_IR) {
        if (pc != DHIP_PC_JACOBI) return set_error(c, DHIP_E_ARG, "FP32 refinement path uses the Jacobi preconditioner");
        for (int k = 6; k < 9; ++k) DHIP_HIPCHK(c, c->kv[k].ensure(sizeof(float) * (n + 1)));
        PcgMixArgs pa{n, A->rowptr, A->col, A->val, diag, val32, db, dx, c->kv[6].as<float>(), c->kv[7].as<float>(),
                      c->kv[8].as<float>(), c->bar.as<SyncSlot>(), *max_iter, *tol, rel, 1e-4, 400, outs, outd,
                      outs + 1, outs + 2};
        pa.xcd_map = c->xcd_map;
        const void *ks[5] = {(const void *)k_pcg_mixed<1>, (const void *)k_pcg_mixed<2>, (const void *)k_pcg_mixed<4>,
                             (const void *)k_pcg_mixed<8>, (const void *)k_pcg_mixed<16>};
        rc = launch_persistent(c, "pcg_mixed", ks, rpt, nb, pa);
        if (rc) return rc;
    } else {
        return set_error(c, DHIP_E_ARG, "unknown precision");
    }
    SolverResult res;
    DHIP_HIPCHK(c, hipMemcpyAsync(&res, outs, sizeof res, hipMemcpyDeviceToHost, c->stream));
    if (where != DHIP_DEVICE && n > 0)
        DHIP_HIPCHK(c, hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    h_iter = res.i[0];
    h_status = res.i[1];
    h_outer = res.i[2];
    h_resid = res.resid;
    (void)h_outer;
    bb.release();
    bx.release();
    rc = flush_profile(c);
    if (rc) return rc;
    *max_iter = h_iter;
    *tol = h_resid;
    if (h_status == 2) return set_error(c, DHIP_E_HIP, "grid barrier timed out in the persistent solver");
    if (h_status != 0) return set_error(c, DHIP_E_NOT_CONVERGED, "PCG did not converge: resid %g after %d", h_resid, h_iter);
    return DHIP_OK;
}

int dhip_tune(dhip_ctx_t *c, const char *key, int64_t value) {
    if (!c || !key) return DHIP_E_ARG;
    if (!std::strcmp(key, "coherent_exchange")) {
        if (value < 0 || value > 3) return set_error(c, DHIP_E_ARG, "coherent_exchange must be 0, 1, 2 or 3");
        c->coherent_exchange = (int)value;
        return DHIP_OK;
    }
    if (!std::strcmp(key, "gmres_precision")) {
        if (value != DHIP_FP64 && value != DHIP_FP32_IR)
            return set_error(c, DHIP_E_ARG, "gmres_precision must be DHIP_FP64 or DHIP_FP32_IR");
        c->gmres_precision = (int)value;
        return DHIP_OK;
    }
    if (!std::strcmp(key, "xcd_map")) {
        if (value != 0 && value != 1) return set_error(c, DHIP_E_ARG, "xcd_map must be 0 or 1");
        c->xcd_map = (int)value;
        return DHIP_OK;
    }
    if (!std::strcmp(key, "pcg_rows_per_block")) {
        if (value < 0 || value > 16 * PB) return set_error(c, DHIP_E_ARG, "pcg_rows_per_block out of range");
        c->pcg_rows_per_block = value;
        return DHIP_OK;
    }
    if (!std::strcmp(key, "batch_threads")) {
        if (value != 128 && value != 256 && value != 512)
            return set_error(c, DHIP_E_ARG, "batch_threads must be 128, 256 or 512");
        c->batch_threads = (int)value;
        return DHIP_OK;
    }
    if (!std::strcmp(key, "batch_split")) {
        if (value != 0 && value != 1) return set_error(c, DHIP_E_ARG, "batch_split must be 0 or 1");
        c->batch_split = (int)value;
        return DHIP_OK;
    }
    if (!std::strcmp(key, "batch_solve_threads")) {
        if (value != 256 && value != 512 && value != 1024)
            return set_error(c, DHIP_E_ARG, "batch_solve_threads must be 256, 512 or 1024");
        c->batch_solve_threads = (int)value;
        return DHIP_OK;
    }
    if (!std::strcmp(key, "batch_prepass")) {
        if (value != 0 && value != 1) return set_error(c, DHIP_E_ARG, "batch_prepass must be 0 or 1");
        c->batch_prepass = (int)value;
        return DHIP_OK;
    }
    if (!std::strcmp(key, "batch_blocks_per_cu")) {
        if (value < 1 || value > 8) return set_error(c, DHIP_E_ARG, "batch_blocks_per_cu must be 1..8");
        c->batch_blocks_per_cu = (int)value;
        return DHIP_OK;
    }
    return set_error(c, DHIP_E_ARG, "unknown tuning key %s", key);
}

int dhip_solver_probe(dhip_ctx_t *c, int32_t on, uint64_t *ticks, int32_t *nblocks) {
    if (!c) return DHIP_E_ARG;
    if (ticks && c->probe_blocks > 0) {
        DHIP_HIPCHK(c, hipMemcpyAsync(ticks, c->probe.p, sizeof(uint64_t) * 4 * c->probe_blocks,
                                      hipMemcpyDeviceToHost, c->stream));
        DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (nblocks) *nblocks = c->probe_blocks;
    c->probe_on = on != 0;
    return DHIP_OK;
}

}  // extern "C"

// Persistent GMRES(m) on device vectors: diag == nullptr means no preconditioner.
// f32: Arnoldi cycles on FP32 basis and matrix values inside the FP64 restart residual (k_gmres F32).
static int gmres_core(dhip_ctx *c, const dhip_csr *A, const double *diag, const double *db, double *dx, int m,
                      int32_t *max_iter, double *tol, int32_t rel, int32_t left_pc, bool f32 = false) {
    const int64_t n = A->n;
    DHIP_HIPCHK(c, c->kv[15].ensure(sizeof(double) * (size_t)m * n + 8));
    unsigned nb = 1;
    int rpt = 1;
    int rc = persistent_geometry(c, n, &nb, &rpt);
    if (rc) return rc;
    DHIP_HIPCHK(c, c->bar.ensure(sizeof(SyncSlot) * (nb + 1)));
    DHIP_HIPCHK(c, c->red_tmp.ensure(sizeof(double) * 2 * (size_t)nb * KMAX + 64));
    DHIP_HIPCHK(c, c->host_stage.ensure(64));
    DHIP_LAUNCH(c, "zero_slots", k_zero_u32, blocks_for((int64_t)nb * sizeof(SyncSlot) / 4, 256), 256, 0,
                c->bar.as<uint32_t>(), (int)(nb * sizeof(SyncSlot) / 4));
    int32_t *outs = c->host_stage.as<int32_t>();
    double *outd = reinterpret_cast<double *>(outs + 4);
    GmresArgs ga{n, A->rowptr, A->col, A->val, diag, db, dx, c->kv[15].as<double>(),
                 c->bar.as<SyncSlot>(), c->red_tmp.as<double>(), m, *max_iter, *tol, rel, left_pc ? 1 : 0,
                 outs, outd, outs + 1, nullptr, nullptr, LDS_BYTES, nullptr, nullptr, nullptr, 0};
    ga.xcd_map = c->xcd_map;
    if (diag && !left_pc && A->nnz > 0) {
        DHIP_HIPCHK(c, c->gm_sval.ensure(sizeof(double) * (A->nnz + 1)));
        DHIP_LAUNCH(c, "gmres_scale", k_scale_vals, blocks_for(A->nnz, 256), 256, 0, A->nnz, A->col, A->val, diag,
                    c->gm_sval.as<double>());
        ga.sval = c->gm_sval.as<double>();
    }
    // Blocks whose CSR slice exceeds the LDS slice capacity read a SELL-64 copy instead (built
    // when the average block would not fit).
    const int64_t per = (n + nb - 1) / nb;
    const uint64_t key[4] = {A->pattern_id, (uint64_t)n, (uint64_t)A->nnz, (uint64_t)nb};
    if (A->nnz > 0 && (per + 1) * 4 + (A->nnz / nb) * 12 + 16 > (int64_t)LDS_BYTES && A->pattern_id != 0 &&
        std::equal(key, key + 4, c->gm_sell_key)) {
        // the same pattern and block layout as the last solve: only the values change
        const int spb = (int)((per + 63) / 64);
        float *sv32 = nullptr;
        if (f32) {
            DHIP_HIPCHK(c, c->gm_sell_val32.ensure(c->gm_sell_val.bytes / 2 + 8));
            sv32 = c->gm_sell_val32.as<float>();
        }
        DHIP_LAUNCH(c, "gmres_sell_vals", k_sell_vals, blocks_for(A->nnz, 256), 256, 0, A->nnz,
                    c->gm_sell_pos.as<int32_t>(), A->col, A->val, (diag && !left_pc) ? diag : nullptr,
                    c->gm_sell_val.as<double>(), sv32);
        ga.sell_val32 = sv32;
        ga.sell_off = c->gm_sell_off.as<int32_t>();
        ga.sell_col = c->gm_sell_col.as<int32_t>();
        ga.sell_val = c->gm_sell_val.as<double>();
        ga.sell_spb = spb;
    } else if (A->nnz > 0 && (per + 1) * 4 + (A->nnz / nb) * 12 + 16 > (int64_t)LDS_BYTES) {
        const int spb = (int)((per + 63) / 64);
        const int64_t nsl = (int64_t)nb * spb;
        for (int k = 0; k < 4; ++k) c->gm_sell_key[k] = 0;
        DHIP_HIPCHK(c, c->gm_sell_off.ensure(sizeof(int32_t) * (nsl + 1)));
        DHIP_LAUNCH(c, "gmres_sell_width", k_sell_width, blocks_for(nsl, 256), 256, 0, n, per, spb, nsl, A->rowptr,
                    c->gm_sell_off.as<int32_t>());
        int64_t tot = 0;
        rc = scan_exclusive_i32(c, c->gm_sell_off.as<int32_t>(), c->gm_sell_off.as<int32_t>(), nsl, &tot);
        if (rc) return rc;
        DHIP_HIPCHK(c, c->gm_sell_col.ensure(sizeof(int32_t) * (tot + 1)));
        DHIP_HIPCHK(c, c->gm_sell_val.ensure(sizeof(double) * (tot + 1)));
        float *sv32 = nullptr;
        if (f32) {
            DHIP_HIPCHK(c, c->gm_sell_val32.ensure(sizeof(float) * (tot + 1)));
            sv32 = c->gm_sell_val32.as<float>();
        }
        ga.sell_val32 = sv32;
        int32_t *pos = nullptr;
        if (A->pattern_id != 0) {
            DHIP_HIPCHK(c, c->gm_sell_pos.ensure(sizeof(int32_t) * (A->nnz + 1)));
            pos = c->gm_sell_pos.as<int32_t>();
        }
        DHIP_LAUNCH(c, "gmres_sell_fill", k_sell_fill, blocks_for(n, 256), 256, 0, n, per, spb, A->rowptr, A->col,
                    A->val, (diag && !left_pc) ? diag : nullptr, c->gm_sell_off.as<int32_t>(),
                    c->gm_sell_col.as<int32_t>(), c->gm_sell_val.as<double>(), pos, sv32);
        if (A->pattern_id != 0) std::copy(key, key + 4, c->gm_sell_key);
        ga.sell_off = c->gm_sell_off.as<int32_t>();
        ga.sell_col = c->gm_sell_col.as<int32_t>();
        ga.sell_val = c->gm_sell_val.as<double>();
        ga.sell_spb = spb;
    }
    if (c->probe_on) {
        DHIP_HIPCHK(c, c->probe.ensure(sizeof(uint64_t) * 4 * (nb + 1)));
        ga.probe = c->probe.as<uint64_t>();
        c->probe_blocks = nb;
    }
    const bool regv = rpt == 1 && m <= KREG;
    const void *ks_f[5] = {regv ? (const void *)k_gmres<1, KREG, false> : (const void *)k_gmres<1, 0, false>,
                           (const void *)k_gmres<2, 0, false>, (const void *)k_gmres<4, 0, false>,
                           (const void *)k_gmres<8, 0, false>, (const void *)k_gmres<16, 0, false>};
    const void *ks_c[5] = {regv ? (const void *)k_gmres<1, KREG, true> : (const void *)k_gmres<1, 0, true>,
                           (const void *)k_gmres<2, 0, true>, (const void *)k_gmres<4, 0, true>,
                           (const void *)k_gmres<8, 0, true>, (const void *)k_gmres<16, 0, true>};
    const void *ks_32[5] = {regv ? (const void *)k_gmres<1, KREG, true, true> : (const void *)k_gmres<1, 0, true, true>,
                            (const void *)k_gmres<2, 0, true, true>, (const void *)k_gmres<4, 0, true, true>,
                            (const void *)k_gmres<8, 0, true, true>, (const void *)k_gmres<16, 0, true, true>};
    const void *const *ks = f32 ? ks_32 : c->coherent_exchange ? ks_c : ks_f;
    const int kidx = rpt == 1 ? 0 : rpt == 2 ? 1 : rpt == 4 ? 2 : rpt == 8 ? 3 : 4;
    const size_t lds = regv ? (size_t)LDS_BYTES : persistent_lds(c, ks[kidx]);
    ga.lds_bytes = (int64_t)lds;
    rc = launch_persistent(c, "gmres", ks, rpt, nb, ga, lds);
    if (rc) return rc;
    SolverResult res;
    DHIP_HIPCHK(c, hipMemcpyAsync(&res, outs, sizeof res, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    const int32_t hs[2] = {res.i[0], res.i[1]};
    const double h_resid = res.resid;
    *max_iter = hs[0];
    *tol = h_resid;
    if (hs[1] == 2) return set_error(c, DHIP_E_HIP, "grid synchronisation timed out in the persistent solver");
    if (hs[1] != 0) return set_error(c, DHIP_E_NOT_CONVERGED, "GMRES did not converge: resid %g", h_resid);
    return DHIP_OK;
}

static int gmres_restart(dhip_ctx *c, int32_t restart, int32_t max_iter, int *m) {
    // m = min(restart, max_iter) as the reference; at least one Arnoldi step per cycle
    *m = std::min<int>(restart, max_iter);
    if (*m < 2 || *m > KMAX) return set_error(c, DHIP_E_ARG, "GMRES restart must be in [2, %d]", KMAX);
    return DHIP_OK;
}

// Shared host wrapper: stage vectors, run, copy the solution back.
static int gmres_entry(dhip_ctx *c, const dhip_csr *A, int32_t pc, const double *pcdiag, const double *b, double *x,
                       int32_t where, int32_t restart, int32_t *max_iter, double *tol, int32_t rel,
                       int32_t left_pc) {
    int m = 0;
    int rc = gmres_restart(c, restart, *max_iter, &m);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = A->n;
    if (n == 0) {
        *max_iter = 0;
        *tol = 0.0;
        return DHIP_OK;
    }
    DevBuf bb, bx, bd;
    const double *db;
    rc = stage_in(c, bb, b, n, where, &db);
    if (rc) return rc;
    double *dx = x;
    if (where != DHIP_DEVICE) {
        DHIP_HIPCHK(c, bx.ensure(sizeof(double) * (n + 1)));
        DHIP_HIPCHK(c, hipMemcpyAsync(bx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        dx = bx.as<double>();
    }
    const double *diag = nullptr;
    if (pcdiag) {
        rc = stage_in(c, bd, pcdiag, n, where, &diag);
        if (rc) return rc;
    } else if (pc == DHIP_PC_JACOBI) {
        DHIP_HIPCHK(c, c->kv[12].ensure(sizeof(double) * (n + 1)));
        DHIP_LAUNCH(c, "csr_diag", k_diag, blocks_for(n, KB), KB, 0, n, A->rowptr, A->col, A->val,
                    c->kv[12].as<double>(), nullptr);
        diag = c->kv[12].as<double>();
    }
    rc = gmres_core(c, A, diag, db, dx, m, max_iter, tol, rel, left_pc, c->gmres_precision == DHIP_FP32_IR);
    if ((rc == DHIP_OK || rc == DHIP_E_NOT_CONVERGED) && where != DHIP_DEVICE) {
        if (hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = set_error(c, DHIP_E_HIP, "copy of the GMRES solution failed");
    }
    bb.release();
    bx.release();
    bd.release();
    const int rp = flush_profile(c);
    return rc ? rc : rp;
}

extern "C" {

int dhip_gmres(dhip_ctx_t *c, const dhip_csr_t *A, int32_t pc, const double *b, double *x, int32_t where,
               int32_t restart, int32_t *max_iter, double *tol, int32_t rel, int32_t left_pc) {
    if (!c || !A || !b || !x || !max_iter || !tol) return DHIP_E_ARG;
    if (pc != DHIP_PC_NONE && pc != DHIP_PC_JACOBI) return set_error(c, DHIP_E_ARG, "unknown preconditioner");
    return gmres_entry(c, A, pc, nullptr, b, x, where, restart, max_iter, tol, rel, left_pc);
}

int dhip_gmres_diag(dhip_ctx_t *c, const dhip_csr_t *A, const double *pcdiag, const double *b, double *x,
                    int32_t where, int32_t restart, int32_t *max_iter, double *tol, int32_t rel, int32_t left_pc) {
    if (!c || !A || !pcdiag || !b || !x || !max_iter || !tol) return DHIP_E_ARG;
    return gmres_entry(c, A, DHIP_PC_JACOBI, pcdiag, b, x, where, restart, max_iter, tol, rel, left_pc);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Cahn-Hilliard time stepping (CahnHilliardcGP1CL, theta = 1, resting interface)
// ---------------------------------------------------------------------------
struct dhip_ch {
    dhip_ctx *ctx = nullptr;
    dhip_ch_params_t p{};
    ChCoef k{};
    MeshP mesh{};
    int64_t ncut = 0, ndof = 0, nnz = 0;
    // snapshot of the assembly the system was built from
    DevBuf phi, cut, cellinfo, flags, dofs, rowowner, massloc, rowptr, col, mass, lb, stab;
    // step system and state [mu; c]
    DevBuf geom, tet, brp, bcol, bval, pcdiag, rhs, x;
    // previous time levels for the learned start: hist[k] = state k+1 levels back (k < nhist)
    DevBuf hist[3];
    int nhist = 0;
    DevBuf feat, pred, xstart;
    DevBuf r3;   // the residuals of the network's and the extrapolation's candidates
    // P2 space: contribution lists of the entries and rows, row of each entry
    int space = 1;
    DevBuf eptr, econt, rptr, rcont, erow;
    dhip_csr K;
    bool prepared = false;
    // BDF loop: the solver's vector [mu; c], the BDF1 concentration, c_{n-1}, c_e, c_r, error parts
    DevBuf bdf_xl, bdf_c1, bdf_cprev, bdf_ce, bdf_cr, bdf_part;
    int64_t bdf_step = 0;
    double bdf_t = 0.0, bdf_dt = 0.0, bdf_e = 0.0;
    int32_t precision = DHIP_FP64;   // GMRES cycles: DHIP_FP64 or DHIP_FP32_IR (FP32 Arnoldi)
};

static int ch_copy(dhip_ctx *c, DevBuf &dst, const void *src, size_t bytes) {
    DHIP_HIPCHK(c, dst.ensure(bytes + 8));
    if (bytes) DHIP_HIPCHK(c, hipMemcpyAsync(dst.p, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return DHIP_OK;
}

// the P2 local mass of every tet, [N][100], out of the tet-major local array [N][LOC2_N]
static int ch_copy_mass_p2(dhip_ctx *c, DevBuf &dst, const void *loc, int64_t N) {
    DHIP_HIPCHK(c, dst.ensure(sizeof(double) * 100 * N + 8));
    if (N)
        DHIP_HIPCHK(c, hipMemcpy2DAsync(dst.p, sizeof(double) * 100, static_cast<const double *>(loc) + LOC2_MASS,
                                        sizeof(double) * LOC2_N, sizeof(double) * 100, (size_t)N,
                                        hipMemcpyDeviceToDevice, c->stream));
    return DHIP_OK;
}

static void ch_release(dhip_ch *h) {
    DevBuf *bufs[] = {&h->phi, &h->cut, &h->cellinfo, &h->flags, &h->dofs, &h->rowowner, &h->massloc, &h->rowptr,
                      &h->col, &h->mass, &h->lb, &h->stab, &h->geom, &h->tet, &h->brp, &h->bcol, &h->bval, &h->pcdiag,
                      &h->rhs, &h->x, &h->hist[0], &h->hist[1], &h->hist[2], &h->feat, &h->pred, &h->xstart,
                      &h->eptr, &h->econt, &h->rptr, &h->rcont, &h->erow, &h->bdf_xl, &h->bdf_c1,
                      &h->bdf_cprev, &h->bdf_ce, &h->bdf_cr, &h->bdf_part};
    for (DevBuf *b : bufs) b->release();
}

// The step system from the concentration conc; conc_r != nullptr: a BDF2 step with rhs3 = M conc_r.
static int ch_prepare_with(dhip_ch *h, const double *conc, const double *conc_r, const ChCoef &k) {
    dhip_ctx *c = h->ctx;
    const int64_t N = h->ncut, R = h->ndof;
    const int32_t bdf = conc_r != nullptr;
    if (h->space == 2) {
        ChTetP2Args ta{N, h->flags.as<uint8_t>(), h->dofs.as<int32_t>(), h->massloc.as<double>(), h->geom.as<double>(),
                       conc, conc_r, h->tet.as<double>()};
        if (N > 0) DHIP_LAUNCH(c, "ch_tet_p2", k_ch_tet_p2, blocks_for(N, CHTET2_WAVES), 64 * CHTET2_WAVES, 0, ta);
        ChEntryP2Args ea{h->nnz, h->erow.as<int32_t>(), h->rowptr.as<int32_t>(), h->col.as<int32_t>(),
                         h->eptr.as<int32_t>(), h->econt.as<int32_t>(), h->tet.as<double>(), h->mass.as<double>(),
                         h->stab.as<double>(), h->bval.as<double>(), h->pcdiag.as<double>(), k};
        if (h->nnz > 0) DHIP_LAUNCH(c, "ch_entries_p2", k_ch_entries_p2, blocks_for(h->nnz, 256), 256, 0, ea);
        ChRhsP2Args ra{R, h->rptr.as<int32_t>(), h->rcont.as<int32_t>(), h->tet.as<double>(), h->rhs.as<double>(), k,
                       bdf};
        if (R > 0) DHIP_LAUNCH(c, "ch_rhs_p2", k_ch_rhs_p2, blocks_for(R, 256), 256, 0, ra);
        h->prepared = true;
        return DHIP_OK;
    }
    ChTetArgs ta{N, h->flags.as<uint8_t>(), h->dofs.as<int32_t>(), h->massloc.as<double>(), h->geom.as<double>(),
                 conc, conc_r, h->tet.as<double>()};
    if (N > 0) DHIP_LAUNCH(c, "ch_tet", k_ch_tet, blocks_for(N, 256), 256, 0, ta);
    ChRowArgs ra{h->mesh, h->cut.as<int64_t>(), h->cellinfo.as<uint32_t>(), h->flags.as<uint8_t>(),
                 h->dofs.as<int32_t>(), h->rowowner.as<int32_t>(), R, N, h->tet.as<double>(),
                 h->rowptr.as<int32_t>(), h->mass.as<double>(), h->stab.as<double>(), h->bval.as<double>(),
                 h->pcdiag.as<double>(), h->rhs.as<double>(), k, bdf};
    if (R > 0) DHIP_LAUNCH(c, "ch_rows", k_ch_rows, blocks_for(R, 128), 128, 0, ra);
    h->prepared = true;
    return DHIP_OK;
}

static int ch_prepare_impl(dhip_ch *h) { return ch_prepare_with(h, h->x.as<double>() + h->ndof, nullptr, h->k); }

// the per-step coefficients of a time step dt
static ChCoef ch_coef(const dhip_ch_params_t &p, double dt) {
    ChCoef k{};
    const double eps = p.epsilon, rho = p.rho;
    k.sdt = p.sigma * dt;
    k.rdt = rho * dt;
    k.dtr = dt * rho;
    k.dt = dt;
    k.S = p.S;
    k.e2 = eps * eps;
    k.re2 = rho * eps * eps;
    k.e2r = eps * eps * rho;
    k.meps = -eps;
    return k;
}

// The current state becomes the newest history level (the ring keeps three).
static int ch_push_history(dhip_ch *h) {
    dhip_ctx *c = h->ctx;
    const size_t bytes = sizeof(double) * 2 * h->ndof;
    DevBuf last = h->hist[2];
    h->hist[2] = h->hist[1];
    h->hist[1] = h->hist[0];
    h->hist[0] = last;
    DHIP_HIPCHK(c, h->hist[0].ensure(bytes + 8));
    DHIP_HIPCHK(c, hipMemcpyAsync(h->hist[0].p, h->x.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    h->nhist = std::min(h->nhist + 1, 3);
    return DHIP_OK;
}

static int ch_solve_impl(dhip_ch *h, int32_t *iters, double *resid) {
    dhip_ctx *c = h->ctx;
    int m = 0;
    int rc = gmres_restart(c, h->p.restart, h->p.max_iter, &m);
    if (rc) return rc;
    int32_t it = h->p.max_iter;
    double tol = h->p.tol;
    rc = gmres_core(c, &h->K, h->pcdiag.as<double>(), h->rhs.as<double>(), h->x.as<double>(), m, &it, &tol,
                    h->p.rel, 0, h->precision == DHIP_FP32_IR);
    h->prepared = false;
    h->bdf_step = 0;
    h->bdf_e = 0.0;
    if (iters) *iters = it;
    if (resid) *resid = tol;
    return rc;
}

// P2 space: contribution lists of the entries and rows (counts, scans, fills, per-list sorts)
static int ch_contrib_lists(dhip_ch *h) {
    dhip_ctx *c = h->ctx;
    const int64_t N = h->ncut, R = h->ndof, Z = h->nnz;
    DHIP_HIPCHK(c, h->eptr.ensure(sizeof(int32_t) * (Z + 1)));
    DHIP_HIPCHK(c, h->rptr.ensure(sizeof(int32_t) * (R + 1)));
    DHIP_HIPCHK(c, h->erow.ensure(sizeof(int32_t) * (Z + 1)));
    DevBuf ecnt, rcnt;
    DHIP_HIPCHK(c, ecnt.ensure(sizeof(int32_t) * (Z + 1)));
    DHIP_HIPCHK(c, rcnt.ensure(sizeof(int32_t) * (R + 1)));
    auto zero = [&](DevBuf &b, int64_t n) {
        DHIP_LAUNCH(c, "ch_zero", k_zero_u32, blocks_for(n, 256), 256, 0, b.as<uint32_t>(), (int)n);
        return DHIP_OK;
    };
    int rc = zero(ecnt, Z + 1);
    if (!rc) rc = zero(rcnt, R + 1);
    if (rc) return rc;
    DHIP_LAUNCH(c, "zero_flags", k_zero_u32, 1, 64, 0, c->errflag.as<uint32_t>(), 64);
    ContribArgs ca{N, h->flags.as<uint8_t>(), h->dofs.as<int32_t>(), h->rowptr.as<int32_t>(), h->col.as<int32_t>(),
                   ecnt.as<int32_t>(), rcnt.as<int32_t>(), h->eptr.as<int32_t>(), h->rptr.as<int32_t>(),
                   nullptr, nullptr, c->errflag.as<int32_t>()};
    if (N > 0) DHIP_LAUNCH(c, "ch_contrib_count", k_ch_contrib<false>, blocks_for(100 * N, 256), 256, 0, ca);
    int64_t etot = 0, rtot = 0;
    rc = scan_exclusive_i32(c, ecnt.as<int32_t>(), h->eptr.as<int32_t>(), Z, &etot);
    if (!rc) rc = scan_exclusive_i32(c, rcnt.as<int32_t>(), h->rptr.as<int32_t>(), R, &rtot);
    if (rc) return rc;
    int32_t fl = 0;
    DHIP_HIPCHK(c, hipMemcpyAsync(&fl, c->errflag.p, sizeof fl, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    if (fl & 8) return set_error(c, DHIP_E_STATE, "a local entry has no slot in the P2 pattern");
    {
        const int32_t e32 = (int32_t)etot, r32 = (int32_t)rtot;
        DHIP_HIPCHK(c, hipMemcpyAsync(h->eptr.as<int32_t>() + Z, &e32, sizeof e32, hipMemcpyHostToDevice, c->stream));
        DHIP_HIPCHK(c, hipMemcpyAsync(h->rptr.as<int32_t>() + R, &r32, sizeof r32, hipMemcpyHostToDevice, c->stream));
    }
    DHIP_HIPCHK(c, h->econt.ensure(sizeof(int32_t) * (etot + 1)));
    DHIP_HIPCHK(c, h->rcont.ensure(sizeof(int32_t) * (rtot + 1)));
    rc = zero(ecnt, Z + 1);
    if (!rc) rc = zero(rcnt, R + 1);
    if (rc) return rc;
    ca.econt = h->econt.as<int32_t>();
    ca.rcont = h->rcont.as<int32_t>();
    if (N > 0) DHIP_LAUNCH(c, "ch_contrib_fill", k_ch_contrib<true>, blocks_for(100 * N, 256), 256, 0, ca);
    if (Z > 0) DHIP_LAUNCH(c, "ch_sort_lists", k_sort_lists, blocks_for(Z, 256), 256, 0, Z, h->eptr.as<int32_t>(),
                           h->econt.as<int32_t>());
    if (R > 0) DHIP_LAUNCH(c, "ch_sort_lists", k_sort_lists, blocks_for(R, 256), 256, 0, R, h->rptr.as<int32_t>(),
                           h->rcont.as<int32_t>());
    if (R > 0) DHIP_LAUNCH(c, "ch_entry_rows", k_entry_rows, blocks_for(R, 256), 256, 0, R, h->rowptr.as<int32_t>(),
                           h->erow.as<int32_t>());
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

static int ch_create(dhip_ctx *c, int32_t zero_policy, const dhip_ch_params_t *p, dhip_ch_t **out, int64_t *ndof,
                     int space) {
    if (!c || !p || !out) return DHIP_E_ARG;
    if (!(p->dt > 0.0) || p->restart < 2 || p->restart > KMAX || p->max_iter < 1 || !(p->tol > 0.0))
        return set_error(c, DHIP_E_ARG, "invalid Cahn-Hilliard parameters");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const double one3[3] = {1.0, 1.0, 1.0};
    int rc = classify_impl(c, zero_policy, nullptr, false);
    if (space == 2) {
        if (!rc && c->ncut >= (int64_t)(INT32_MAX / 100))
            return set_error(c, DHIP_E_CAPACITY, "%lld cut tets exceed the P2 stepper's key range",
                             (long long)c->ncut);
        if (!rc) rc = local_p2_impl(c, one3, DHIP_RHS_NONE, nullptr);
        if (!rc) rc = number_p2_impl(c, -1.0, nullptr);
        if (!rc) rc = build_p2_impl(c, one3, nullptr);
    } else {
        if (!rc) rc = local_impl(c, one3, DHIP_RHS_NONE, nullptr);
        if (!rc) rc = number_impl(c, -1.0, nullptr);
        if (!rc) rc = build_impl(c, one3, nullptr);
    }
    if (rc) return rc;
    const int nloc = space == 2 ? 10 : 4;
    dhip_ch *h = new dhip_ch();
    h->space = space;
    h->ctx = c;
    h->p = *p;
    h->mesh = c->mesh;
    h->ncut = c->ncut;
    h->ndof = c->ndof;
    h->nnz = c->nnz;
    const int64_t N = h->ncut, R = h->ndof, Z = h->nnz;
    h->k = ch_coef(*p, p->dt);
    auto fail = [&](int code) {
        ch_release(h);
        delete h;
        return code;
    };
    if ((rc = ch_copy(c, h->phi, c->phi, sizeof(double) * c->phi_count)) ||
        (rc = ch_copy(c, h->cut, c->cut.p, sizeof(int64_t) * N)) ||
        (rc = ch_copy(c, h->cellinfo, c->cellinfo.p, sizeof(uint32_t) * c->mesh.ncell)) ||
        (rc = ch_copy(c, h->flags, c->flags.p, N)) ||
        (rc = ch_copy(c, h->dofs, c->dofs.p, sizeof(int32_t) * nloc * N)) ||
        (rc = ch_copy(c, h->rowowner, c->rowowner.p, sizeof(int32_t) * R)) ||
        (rc = space == 2 ? ch_copy_mass_p2(c, h->massloc, c->loc.p, N)
                         : ch_copy(c, h->massloc, c->loc.p, sizeof(double) * nloc * nloc * N)) ||
        (rc = ch_copy(c, h->rowptr, c->rowptr.p, sizeof(int32_t) * (R + 1))) ||
        (rc = ch_copy(c, h->col, c->col.p, sizeof(int32_t) * Z)) ||
        (rc = ch_copy(c, h->mass, c->val[0].p, sizeof(double) * Z)) ||
        (rc = ch_copy(c, h->lb, c->val[1].p, sizeof(double) * Z)) ||
        (rc = ch_copy(c, h->stab, c->val[2].p, sizeof(double) * Z)))
        return fail(rc);
    const int geom_n = space == 2 ? CHG2_N : CHG_N, tet_n = space == 2 ? CHT2_N : CHT_N;
    if (h->geom.ensure(sizeof(double) * geom_n * (N + 1)) || h->tet.ensure(sizeof(double) * tet_n * (N + 1)) ||
        h->brp.ensure(sizeof(int32_t) * (2 * R + 1)) ||
        h->bcol.ensure(sizeof(int32_t) * (4 * Z + 1)) || h->bval.ensure(sizeof(double) * (4 * Z + 1)) ||
        h->pcdiag.ensure(sizeof(double) * (2 * R + 1)) || h->rhs.ensure(sizeof(double) * (2 * R + 1)) ||
        h->x.ensure(sizeof(double) * (2 * R + 1)))
        return fail(set_error(c, DHIP_E_NOMEM, "Cahn-Hilliard buffers"));
    if (hipMemsetAsync(h->x.p, 0, sizeof(double) * 2 * R, c->stream) != hipSuccess)
        return fail(set_error(c, DHIP_E_HIP, "state clear failed"));
    ChBlockArgs ba{R, Z, h->rowptr.as<int32_t>(), h->col.as<int32_t>(), h->mass.as<double>(), h->lb.as<double>(),
                   h->stab.as<double>(), h->brp.as<int32_t>(), h->bcol.as<int32_t>(), h->bval.as<double>(),
                   h->pcdiag.as<double>(), h->k};
    {
        ProfScope ps(c, "ch_block");
        hipLaunchKernelGGL(k_ch_block, dim3(blocks_for(R + 1, 128)), dim3(128), 0, c->stream, ba);
        if (hipGetLastError() != hipSuccess) return fail(set_error(c, DHIP_E_HIP, "k_ch_block launch failed"));
    }
    if (N > 0 && space == 2) {
        ChGeomP2Args ga{h->mesh, h->phi.as<double>(), h->cut.as<int64_t>(), N, h->flags.as<uint8_t>(),
                        h->geom.as<double>(), c->errflag.as<int32_t>()};
        ProfScope ps(c, "ch_geom_p2");
        hipLaunchKernelGGL(k_ch_geom_p2, dim3((unsigned)N), dim3(64), 0, c->stream, ga);
        if (hipGetLastError() != hipSuccess) return fail(set_error(c, DHIP_E_HIP, "k_ch_geom_p2 launch failed"));
    } else if (N > 0) {
        ChGeomArgs ga{h->mesh, h->phi.as<double>(), h->cut.as<int64_t>(), N, h->flags.as<uint8_t>(),
                      h->geom.as<double>()};
        ProfScope ps(c, "ch_geom");
        hipLaunchKernelGGL(k_ch_geom, dim3(blocks_for(N, LOCAL_BLOCK)), dim3(LOCAL_BLOCK), 0, c->stream, ga);
        if (hipGetLastError() != hipSuccess) return fail(set_error(c, DHIP_E_HIP, "k_ch_geom launch failed"));
    }
    if (space == 2) {
        rc = ch_contrib_lists(h);
        if (rc) return fail(rc);
    }
    static std::atomic<uint64_t> next_pattern{1};
    h->K.ctx = c;
    h->K.pattern_id = next_pattern.fetch_add(1);
    h->K.n = 2 * R;
    h->K.nnz = 4 * Z;
    h->K.rowptr = h->brp.as<int32_t>();
    h->K.col = h->bcol.as<int32_t>();
    h->K.val = h->bval.as<double>();
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(set_error(c, DHIP_E_HIP, "sync failed"));
    *out = h;
    if (ndof) *ndof = R;
    return flush_profile(c);
}

extern "C" {

int dhip_ch_create(dhip_ctx_t *c, int32_t zero_policy, const dhip_ch_params_t *p, dhip_ch_t **out, int64_t *ndof) {
    return ch_create(c, zero_policy, p, out, ndof, 1);
}

int dhip_ch_create_p2(dhip_ctx_t *c, int32_t zero_policy, const dhip_ch_params_t *p, dhip_ch_t **out,
                      int64_t *ndof) {
    return ch_create(c, zero_policy, p, out, ndof, 2);
}

void dhip_ch_destroy(dhip_ch_t *h) {
    if (!h) return;
    (void)hipSetDevice(h->ctx->device);
    (void)hipStreamSynchronize(h->ctx->stream);
    ch_release(h);
    delete h;
}

int dhip_ch_set_state(dhip_ch_t *h, const double *c_in, const double *mu_in, int32_t where) {
    if (!h || !c_in || !mu_in) return DHIP_E_ARG;
    dhip_ctx *c = h->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const hipMemcpyKind kd = where == DHIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const int64_t R = h->ndof;
    if (R > 0) {
        DHIP_HIPCHK(c, hipMemcpyAsync(h->x.as<double>(), mu_in, sizeof(double) * R, kd, c->stream));
        DHIP_HIPCHK(c, hipMemcpyAsync(h->x.as<double>() + R, c_in, sizeof(double) * R, kd, c->stream));
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    h->prepared = false;
    h->nhist = 0;
    h->bdf_step = 0;
    h->bdf_t = h->bdf_dt = h->bdf_e = 0.0;
    return DHIP_OK;
}

int dhip_ch_get_state(dhip_ch_t *h, double *c_out, double *mu_out, int32_t where) {
    if (!h) return DHIP_E_ARG;
    dhip_ctx *c = h->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const hipMemcpyKind kd = where == DHIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    const int64_t R = h->ndof;
    if (R > 0 && mu_out) DHIP_HIPCHK(c, hipMemcpyAsync(mu_out, h->x.as<double>(), sizeof(double) * R, kd, c->stream));
    if (R > 0 && c_out) DHIP_HIPCHK(c, hipMemcpyAsync(c_out, h->x.as<double>() + R, sizeof(double) * R, kd, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

int dhip_ch_prepare(dhip_ch_t *h) {
    if (!h) return DHIP_E_ARG;
    dhip_ctx *c = h->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = ch_prepare_impl(h);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_ch_system(dhip_ch_t *h, const int32_t **rowptr, const int32_t **col, const double **val, const double **rhs,
                   const double **pcdiag, double **state, int64_t *n2, int64_t *nnz2) {
    if (!h) return DHIP_E_ARG;
    if (rowptr) *rowptr = h->brp.as<int32_t>();
    if (col) *col = h->bcol.as<int32_t>();
    if (val) *val = h->bval.as<double>();
    if (rhs) *rhs = h->rhs.as<double>();
    if (pcdiag) *pcdiag = h->pcdiag.as<double>();
    if (state) *state = h->x.as<double>();
    if (n2) *n2 = 2 * h->ndof;
    if (nnz2) *nnz2 = 4 * h->nnz;
    return DHIP_OK;
}

int dhip_ch_solve(dhip_ch_t *h, const double *x0, int32_t where, int32_t *iters, double *resid) {
    if (!h) return DHIP_E_ARG;
    dhip_ctx *c = h->ctx;
    if (!h->prepared) return set_error(c, DHIP_E_STATE, "dhip_ch_prepare must precede dhip_ch_solve");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = h->ndof > 0 ? ch_push_history(h) : DHIP_OK;
    if (rc) return rc;
    if (x0 && h->ndof > 0)
        DHIP_HIPCHK(c, hipMemcpyAsync(h->x.p, x0, sizeof(double) * 2 * h->ndof,
                                      where == DHIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                      c->stream));
    rc = h->ndof > 0 ? ch_solve_impl(h, iters, resid) : DHIP_OK;
    h->prepared = false;
    const int rp = flush_profile(c);
    return rc ? rc : rp;
}

int dhip_ch_step(dhip_ch_t *h, int32_t nsteps, int32_t *iters, double *resid) {
    if (!h || nsteps < 0) return DHIP_E_ARG;
    dhip_ctx *c = h->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    for (int32_t s = 0; s < nsteps; ++s) {
        int rc = ch_prepare_impl(h);
        int32_t it = 0;
        double res = 0.0;
        if (!rc && h->ndof > 0) rc = ch_push_history(h);
        if (!rc && h->ndof > 0) rc = ch_solve_impl(h, &it, &res);
        if (iters) iters[s] = it;
        if (resid) resid[s] = res;
        if (rc) {
            flush_profile(c);
            return rc;
        }
    }
    return flush_profile(c);
}

}  // extern "C"

// F(e, dt), surfphasesep/surfphasesep.cpp:117-123
static double bdf_step_size(const dhip_bdf_params_t &p, double e, double dt) {
    if (!e) return p.step_size;
    const double f = p.rho * std::sqrt(p.tol / e) * dt;
    if (f > p.max_step) return p.max_step;
    if (f < p.min_step) return p.min_step;
    return f;
}

// One step of the loop at surfphasesep/surfphasesep.cpp:397-491 (see dhip_ch_bdf_step).
static int ch_bdf_one(dhip_ch *h, const dhip_bdf_params_t &bp, dhip_bdf_stat_t *stat) {
    dhip_ctx *c = h->ctx;
    const int64_t R = h->ndof;
    const size_t vb = sizeof(double) * R;
    int m = 0;
    int rc = gmres_restart(c, h->p.restart, h->p.max_iter, &m);
    if (rc) return rc;
    const unsigned nb = std::max(1u, std::min(1024u, blocks_for(R, KB)));
    DHIP_HIPCHK(c, h->bdf_xl.ensure(2 * vb + 8));
    DHIP_HIPCHK(c, h->bdf_c1.ensure(vb + 8));
    DHIP_HIPCHK(c, h->bdf_cprev.ensure(vb + 8));
    DHIP_HIPCHK(c, h->bdf_ce.ensure(vb + 8));
    DHIP_HIPCHK(c, h->bdf_cr.ensure(vb + 8));
    DHIP_HIPCHK(c, h->bdf_part.ensure(sizeof(double) * (2 * nb + 2)));
    double *xl = h->bdf_xl.as<double>(), *c1 = h->bdf_c1.as<double>(), *cn = h->x.as<double>() + R;
    double *part = h->bdf_part.as<double>();
    // the solver's vector: the reference keeps one lhs across solves, equal to the state between steps
    DHIP_HIPCHK(c, hipMemcpyAsync(xl, h->x.p, 2 * vb, hipMemcpyDeviceToDevice, c->stream));
    const int64_t i = h->bdf_step + 1;
    const double t_prev = h->bdf_t, dt_prev = h->bdf_dt;
    double dt = h->bdf_dt, e = h->bdf_e, resid = 0.0;
    int32_t tries = 0, it1 = 0, it2 = 0;
    auto solve = [&](int32_t *acc) {
        int32_t it = h->p.max_iter;
        double tl = h->p.tol;
        const int r = gmres_core(c, &h->K, h->pcdiag.as<double>(), h->rhs.as<double>(), xl, m, &it, &tl, h->p.rel, 0,
                                 h->precision == DHIP_FP32_IR);
        *acc += it;
        resid = tl;
        return r;
    };
    for (;;) {
        if (bp.max_tries > 0 && tries == bp.max_tries) {
            rc = set_error(c, DHIP_E_NOT_CONVERGED, "BDF step %lld: error %g above %g after %d attempts",
                           (long long)i, e, bp.tol, tries);
            break;
        }
        ++tries;
        if (i == 1 || bp.adaptive) dt = bdf_step_size(bp, e, dt);
        // BDF1: the implicit-Euler system of step dt at c_n
        rc = ch_prepare_with(h, cn, nullptr, ch_coef(h->p, dt));
        if (!rc) rc = solve(&it1);
        if (rc) break;
        DHIP_HIPCHK(c, hipMemcpyAsync(c1, xl + R, vb, hipMemcpyDeviceToDevice, c->stream));
        if (i == 1) break;
        const double r = dt / dt_prev;
        const double dte = (1. + r) * dt / (1. + 2. * r);
        const double w1 = 1. + r, w2 = -r, w3 = (1. + r) * (1. + r) / (1. + 2. * r), w4 = -(r * r) / (1. + 2. * r);
        DHIP_LAUNCH(c, "ch_bdf_combine", k_ch_bdf_combine, blocks_for(R, KB), KB, 0, R, (const double *)cn,
                    (const double *)h->bdf_cprev.as<double>(), w1, w2, w3, w4, h->bdf_ce.as<double>(),
                    h->bdf_cr.as<double>());
        rc = ch_prepare_with(h, h->bdf_ce.as<double>(), h->bdf_cr.as<double>(), ch_coef(h->p, dte));
        if (!rc) rc = solve(&it2);
        if (rc) break;
        DHIP_LAUNCH(c, "ch_mnorm", k_ch_mnorm, nb, KB, 0, R, (const int32_t *)h->rowptr.as<int32_t>(),
                    (const int32_t *)h->col.as<int32_t>(), (const double *)h->mass.as<double>(), (const double *)c1,
                    (const double *)(xl + R), part, part + nb);
        DHIP_LAUNCH(c, "ch_mnorm_reduce", k_reduce, 1, KB, 0, (const double *)part, (int)nb, part + 2 * nb);
        DHIP_LAUNCH(c, "ch_mnorm_reduce", k_reduce, 1, KB, 0, (const double *)(part + nb), (int)nb,
                    part + 2 * nb + 1);
        double dc[2];
        DHIP_HIPCHK(c, hipMemcpyAsync(dc, part + 2 * nb, sizeof dc, hipMemcpyDeviceToHost, c->stream));
        DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
        e = std::sqrt(dc[0] / dc[1]);
        if (!(bp.adaptive && e > bp.tol && bp.min_step < dt && dt < bp.max_step)) break;
    }
    if (stat) {
        stat->step = rc ? h->bdf_step : i;
        stat->t = rc ? t_prev : t_prev + dt;
        stat->dt = dt;
        stat->e = e;
        stat->tries = tries;
        stat->iters_bdf1 = it1;
        stat->iters_bdf2 = it2;
        stat->reserved = 0;
        stat->resid = resid;
    }
    h->prepared = false;
    if (rc) return rc;
    DHIP_HIPCHK(c, hipMemcpyAsync(h->bdf_cprev.p, cn, vb, hipMemcpyDeviceToDevice, c->stream));
    DHIP_HIPCHK(c, hipMemcpyAsync(h->x.p, xl, 2 * vb, hipMemcpyDeviceToDevice, c->stream));
    h->bdf_step = i;
    h->bdf_t = t_prev + dt;
    h->bdf_dt = dt;
    h->bdf_e = e;
    h->nhist = 0;
    return DHIP_OK;
}

extern "C" {

int dhip_ch_set_precision(dhip_ch_t *h, int32_t precision) {
    if (!h) return DHIP_E_ARG;
    if (precision != DHIP_FP64 && precision != DHIP_FP32_IR)
        return set_error(h->ctx, DHIP_E_ARG, "precision must be DHIP_FP64 or DHIP_FP32_IR");
    h->precision = precision;
    return DHIP_OK;
}

int dhip_ch_bdf_step(dhip_ch_t *h, const dhip_bdf_params_t *p, int32_t nsteps, dhip_bdf_stat_t *stats) {
    if (!h || !p || nsteps < 0) return DHIP_E_ARG;
    dhip_ctx *c = h->ctx;
    if (!(p->step_size > 0.0) || !(p->tol > 0.0) || !(p->rho > 0.0) || !(p->min_step >= 0.0) ||
        !(p->max_step > 0.0) || p->max_tries < 0)
        return set_error(c, DHIP_E_ARG, "invalid BDF parameters");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    for (int32_t s = 0; s < nsteps; ++s) {
        if (h->ndof == 0) {
            // no interface: the loop still advances its time
            const int64_t i = h->bdf_step + 1;
            if (i == 1 || p->adaptive) h->bdf_dt = bdf_step_size(*p, h->bdf_e, h->bdf_dt);
            h->bdf_step = i;
            h->bdf_t += h->bdf_dt;
            if (stats) stats[s] = dhip_bdf_stat_t{i, h->bdf_t, h->bdf_dt, h->bdf_e, 1, 0, 0, 0, 0.0};
            continue;
        }
        int rc = ch_bdf_one(h, *p, stats ? stats + s : nullptr);
        if (rc) {
            flush_profile(c);
            return rc;
        }
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned warm start
// ---------------------------------------------------------------------------
struct dhip_mlp {
    dhip_ctx *ctx = nullptr;
    MlpNorm norm{};
    int adam_step = 0;
    DevBuf theta, m, v, packed, grad, part;
    DevBuf scale, spart;  // step scale (s_mu, s_c) and its partial sums
    bool packed_ok = false;
};

static constexpr int SCALE_BLOCKS = 128;

// s_mu, s_c from x_n and x_{n-1} (both 2n, device); 1 where there is no history
static int mlp_scale_impl(dhip_mlp *q, const double *x0, const double *x1, int64_t n) {
    dhip_ctx *c = q->ctx;
    DHIP_HIPCHK(c, q->scale.ensure(sizeof(double) * 2));
    DHIP_HIPCHK(c, q->spart.ensure(sizeof(double) * 2 * SCALE_BLOCKS));
    DHIP_LAUNCH(c, "mlp_scale_part", k_mlp_scale_part, SCALE_BLOCKS, SCALE_B, 0, x0, x1, n, q->spart.as<double>());
    DHIP_LAUNCH(c, "mlp_scale_final", k_mlp_scale_final, 1, 64, 0, q->spart.as<double>(), SCALE_BLOCKS, n,
                q->scale.as<double>());
    return DHIP_OK;
}

static int mlp_pack(dhip_mlp *q) {
    if (q->packed_ok) return DHIP_OK;
    dhip_ctx *c = q->ctx;
    DHIP_LAUNCH(c, "mlp_pack", k_mlp_pack, blocks_for(MLP_NPACK, 256), 256, 0, q->theta.as<float>(),
                q->packed.as<uint16_t>());
    q->packed_ok = true;
    return DHIP_OK;
}

static int mlp_predict_impl(dhip_mlp *q, const float *X, int64_t S, float *Y) {
    dhip_ctx *c = q->ctx;
    int rc = mlp_pack(q);
    if (rc) return rc;
    if (S > 0)
        DHIP_LAUNCH(c, "mlp_infer", k_mlp_infer, std::min<unsigned>(blocks_for(S, 16 * INFER_WAVES), 2 * c->cu_count),
                    64 * INFER_WAVES, 0, X, S, q->packed.as<uint16_t>(), q->theta.as<float>(), Y);
    return DHIP_OK;
}

static int mlp_features_impl(dhip_mlp *q, const double *const hist[4], int32_t nhist, const double *res,
                             const double *pcdiag, int64_t n, float *X) {
    if (nhist < 1 || nhist > 4) return set_error(q->ctx, DHIP_E_ARG, "1 to 4 history levels");
    const double *h[4];
    for (int k = 0; k < 4; ++k) h[k] = hist[k < nhist ? k : nhist - 1];
    if (n == 0) return DHIP_OK;
    int rc = mlp_scale_impl(q, h[0], h[1], n);
    if (rc) return rc;
    FeatArgs fa{h[0], h[1], h[2], h[3], res, pcdiag, q->scale.as<double>(), n, q->norm, X};
    DHIP_LAUNCH(q->ctx, "mlp_features", k_mlp_features, blocks_for(n, 256), 256, 0, fa);
    return DHIP_OK;
}

// gradient (double, q->grad) and loss of a minibatch; returns the tile count
static int mlp_grad_impl(dhip_mlp *q, const float *X, const float *T, const int32_t *idx, int64_t B) {
    dhip_ctx *c = q->ctx;
    const int64_t tiles = (B + TS - 1) / TS;
    DHIP_HIPCHK(c, q->part.ensure(sizeof(float) * (size_t)(tiles + 1) * (MLP_NPARAM + 1)));
    TrainArgs ta{X, T, idx, B, q->theta.as<float>(), q->part.as<float>()};
    DHIP_LAUNCH(c, "mlp_train", k_mlp_train, (unsigned)tiles, TRAIN_T, 0, ta);
    DHIP_LAUNCH(c, "mlp_reduce", k_mlp_reduce, blocks_for(MLP_NPARAM + 1, 256), 256, 0, q->part.as<float>(),
                (int)tiles, q->grad.as<double>());
    return DHIP_OK;
}

extern "C" {

int dhip_mlp_create(dhip_ctx_t *c, dhip_mlp_t **out) {
    if (!c || !out) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    dhip_mlp *q = new dhip_mlp();
    q->ctx = c;
    for (int k = 0; k < MLP_NF; ++k) q->norm.in_scale[k] = 1.0;
    for (int k = 0; k < MLP_NO; ++k) q->norm.out_scale[k] = 1.0;
    if (q->theta.ensure(sizeof(float) * MLP_NPARAM) || q->m.ensure(sizeof(float) * MLP_NPARAM) ||
        q->v.ensure(sizeof(float) * MLP_NPARAM) || q->packed.ensure(sizeof(uint16_t) * MLP_NPACK) ||
        q->grad.ensure(sizeof(double) * (MLP_NPARAM + 1)) ||
        hipMemsetAsync(q->theta.p, 0, sizeof(float) * MLP_NPARAM, c->stream) ||
        hipMemsetAsync(q->m.p, 0, sizeof(float) * MLP_NPARAM, c->stream) ||
        hipMemsetAsync(q->v.p, 0, sizeof(float) * MLP_NPARAM, c->stream) || hipStreamSynchronize(c->stream)) {
        DevBuf *b[] = {&q->theta, &q->m, &q->v, &q->packed, &q->grad};
        for (DevBuf *x : b) x->release();
        delete q;
        return set_error(c, DHIP_E_NOMEM, "network buffers");
    }
    *out = q;
    return DHIP_OK;
}

void dhip_mlp_destroy(dhip_mlp_t *q) {
    if (!q) return;
    (void)hipSetDevice(q->ctx->device);
    (void)hipStreamSynchronize(q->ctx->stream);
    DevBuf *b[] = {&q->theta, &q->m, &q->v, &q->packed, &q->grad, &q->part, &q->scale, &q->spart};
    for (DevBuf *x : b) x->release();
    delete q;
}

int dhip_mlp_params(dhip_mlp_t *q, const float *theta_in, float *theta_out) {
    if (!q) return DHIP_E_ARG;
    dhip_ctx *c = q->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    if (theta_in) {
        DHIP_HIPCHK(c, hipMemcpyAsync(q->theta.p, theta_in, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice,
                                      c->stream));
        DHIP_HIPCHK(c, hipMemsetAsync(q->m.p, 0, sizeof(float) * MLP_NPARAM, c->stream));
        DHIP_HIPCHK(c, hipMemsetAsync(q->v.p, 0, sizeof(float) * MLP_NPARAM, c->stream));
        q->adam_step = 0;
        q->packed_ok = false;
    }
    if (theta_out)
        DHIP_HIPCHK(c, hipMemcpyAsync(theta_out, q->theta.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost,
                                      c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

int dhip_mlp_normalization(dhip_mlp_t *q, const double in_mean[12], const double in_scale[12],
                           const double out_mean[2], const double out_scale[2]) {
    if (!q || !in_mean || !in_scale || !out_mean || !out_scale) return DHIP_E_ARG;
    for (int k = 0; k < MLP_NF; ++k) {
        q->norm.in_mean[k] = in_mean[k];
        q->norm.in_scale[k] = in_scale[k];
    }
    for (int k = 0; k < MLP_NO; ++k) {
        if (!(out_scale[k] != 0.0)) return set_error(q->ctx, DHIP_E_ARG, "output scale must be nonzero");
        q->norm.out_mean[k] = out_mean[k];
        q->norm.out_scale[k] = out_scale[k];
    }
    return DHIP_OK;
}

int dhip_mlp_features(dhip_mlp_t *q, const double *const hist[4], int32_t nhist, const double *res,
                      const double *pcdiag, int64_t n, float *X) {
    if (!q || !hist || !res || !pcdiag || !X || n < 0) return DHIP_E_ARG;
    dhip_ctx *c = q->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = mlp_features_impl(q, hist, nhist, res, pcdiag, n, X);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_mlp_targets(dhip_mlp_t *q, const double *x_next, const double *const hist[3], int32_t nhist, int64_t n,
                     float *T) {
    if (!q || !x_next || !hist || !T || n < 0 || nhist < 1 || nhist > 3) return DHIP_E_ARG;
    dhip_ctx *c = q->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const double *h[3];
    for (int k = 0; k < 3; ++k) h[k] = hist[k < nhist ? k : nhist - 1];
    if (n > 0) {
        int rc = mlp_scale_impl(q, h[0], h[1], n);
        if (rc) return rc;
        TargetArgs ta{x_next, h[0], h[1], h[2], q->scale.as<double>(), n, q->norm, T};
        DHIP_LAUNCH(c, "mlp_targets", k_mlp_targets, blocks_for(n, 256), 256, 0, ta);
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_mlp_predict(dhip_mlp_t *q, const float *X, int64_t S, float *Y) {
    if (!q || !X || !Y || S < 0) return DHIP_E_ARG;
    dhip_ctx *c = q->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = mlp_predict_impl(q, X, S, Y);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_mlp_grad(dhip_mlp_t *q, const float *X, const float *T, const int32_t *idx, int64_t B, double *grad,
                  double *loss) {
    if (!q || !X || !T || !idx || B <= 0) return DHIP_E_ARG;
    dhip_ctx *c = q->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = mlp_grad_impl(q, X, T, idx, B);
    if (rc) return rc;
    if (grad)
        DHIP_HIPCHK(c, hipMemcpyAsync(grad, q->grad.p, sizeof(double) * MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
    double l = 0.0;
    DHIP_HIPCHK(c, hipMemcpyAsync(&l, q->grad.as<double>() + MLP_NPARAM, sizeof l, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = l / (double)B;
    return flush_profile(c);
}

int dhip_mlp_train_step(dhip_mlp_t *q, const float *X, const float *T, const int32_t *idx, int64_t B, float lr,
                        double *loss) {
    if (!q || !X || !T || !idx || B <= 0 || !(lr > 0.f)) return DHIP_E_ARG;
    dhip_ctx *c = q->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    int rc = mlp_grad_impl(q, X, T, idx, B);
    if (rc) return rc;
    AdamArgs aa{q->theta.as<float>(), q->m.as<float>(), q->v.as<float>(), q->grad.as<double>(), lr, 0.9f, 0.999f,
                1e-8f, ++q->adam_step};
    DHIP_LAUNCH(c, "mlp_adam", k_mlp_adam, blocks_for(MLP_NPARAM, 256), 256, 0, aa);
    q->packed_ok = false;
    if (loss) {
        double l = 0.0;
        DHIP_HIPCHK(c, hipMemcpyAsync(&l, q->grad.as<double>() + MLP_NPARAM, sizeof l, hipMemcpyDeviceToHost,
                                      c->stream));
        DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
        *loss = l / (double)B;
    }
    return flush_profile(c);
}

int dhip_ch_history(dhip_ch_t *h, int32_t level, const double **state) {
    if (!h || !state || level < 0 || level > 3) return DHIP_E_ARG;
    // missing levels repeat the oldest one held
    const int k = std::min<int>(level, h->nhist);
    *state = k == 0 ? h->x.as<double>() : h->hist[k - 1].as<double>();
    return DHIP_OK;
}

int dhip_ch_residual(dhip_ch_t *h, double *r) {
    if (!h || !r) return DHIP_E_ARG;
    dhip_ctx *c = h->ctx;
    if (!h->prepared) return set_error(c, DHIP_E_STATE, "dhip_ch_prepare must precede dhip_ch_residual");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int64_t n2 = 2 * h->ndof;
    if (n2 > 0) {
        const int rb = spmv_tile_rows(n2, h->K.nnz);
        DHIP_LAUNCH(c, "ch_residual", k_spmv<true>, blocks_for(n2, rb), SPMV_T, 0, n2, rb, h->K.rowptr, h->K.col,
                    h->K.val, h->x.as<double>(), h->rhs.as<double>(), r);
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

// Candidate starts for the prepared step from the current state x_n and the history (nothing is
// modified): the network's, the plain quadratic extrapolation and x_n; the one with the smallest
// residual is written to x0 (device) and its index to flag (device).
static int ch_learned_start(dhip_ch *h, dhip_mlp *q, double *x0, int32_t **flag_out) {
    dhip_ctx *c = h->ctx;
    const int64_t n = h->ndof, n2 = 2 * n;
    DHIP_HIPCHK(c, h->feat.ensure(sizeof(float) * MLP_NF * (n + 1)));
    DHIP_HIPCHK(c, h->pred.ensure(sizeof(float) * MLP_NO * (n + 1)));
    DHIP_HIPCHK(c, c->kv[11].ensure(sizeof(double) * (n2 + 1)));
    DHIP_HIPCHK(c, c->kv[10].ensure(sizeof(double) * 4 * SCALE_BLOCKS + sizeof(int32_t) * 4));
    DHIP_HIPCHK(c, c->kv[9].ensure(sizeof(double) * (n2 + 1)));
    double *res = c->kv[11].as<double>();
    double *part = c->kv[10].as<double>();
    double *xq = c->kv[9].as<double>();
    int32_t *flag = reinterpret_cast<int32_t *>(part + 4 * SCALE_BLOCKS);
    const double *xn = h->x.as<double>();
    // residual of the new system at the current state, features, prediction
    const int rb = spmv_tile_rows(n2, h->K.nnz);
    DHIP_LAUNCH(c, "ch_residual", k_spmv<true>, blocks_for(n2, rb), SPMV_T, 0, n2, rb, h->K.rowptr, h->K.col, h->K.val,
                xn, h->rhs.as<double>(), res);
    const double *hist[4] = {xn, h->hist[0].as<double>(), h->hist[1].as<double>(), h->hist[2].as<double>()};
    const int levels = 1 + h->nhist;
    int rc = mlp_features_impl(q, hist, levels, res, h->pcdiag.as<double>(), n, h->feat.as<float>());
    if (!rc) rc = mlp_predict_impl(q, h->feat.as<float>(), n, h->pred.as<float>());
    if (rc) return rc;
    const double *h1 = hist[std::min(1, levels - 1)], *h2 = hist[std::min(2, levels - 1)];
    CombineArgs ca{xn, h1, h2, h->pred.as<float>(), q->scale.as<double>(), n, q->norm, x0, xq};
    DHIP_LAUNCH(c, "mlp_combine", k_mlp_combine, blocks_for(n, 256), 256, 0, ca);
    // the candidates' residuals: x_n's is `res`; the network's and the extrapolation's through the same
    // tiled SpMV (row products summed in column order)
    DHIP_HIPCHK(c, h->r3.ensure(sizeof(double) * (2 * n2 + 1)));
    double *r1 = h->r3.as<double>(), *r2 = r1 + n2;
    DHIP_LAUNCH(c, "ch_residual", k_spmv<true>, blocks_for(n2, rb), SPMV_T, 0, n2, rb, h->K.rowptr, h->K.col, h->K.val,
                x0, h->rhs.as<double>(), r1);
    DHIP_LAUNCH(c, "ch_residual", k_spmv<true>, blocks_for(n2, rb), SPMV_T, 0, n2, rb, h->K.rowptr, h->K.col, h->K.val,
                xq, h->rhs.as<double>(), r2);
    DHIP_LAUNCH(c, "ch_three_norms", k_ch_three_norms, SCALE_BLOCKS, RES3_B, 0, n2, h->rhs.as<double>(), res, r1, r2,
                part);
    // the network's candidate is trusted while the extrapolation's residual exceeds 100 x tolerance
    const double gate = 100.0 * h->p.tol;
    DHIP_LAUNCH(c, "ch_pick", k_ch_pick, 1, 64, 0, part, SCALE_BLOCKS, gate, flag);
    DHIP_LAUNCH(c, "ch_keep", k_ch_keep, blocks_for(n2, 256), 256, 0, n2, flag, xn, xq, x0);
    *flag_out = flag;
    return DHIP_OK;
}

int dhip_ch_predict_start(dhip_ch_t *h, dhip_mlp_t *q, double *x0, int32_t where, int32_t *start) {
    if (!h || !q || !x0) return DHIP_E_ARG;
    dhip_ctx *c = h->ctx;
    if (q->ctx != c) return set_error(c, DHIP_E_ARG, "network and stepper belong to different contexts");
    if (!h->prepared) return set_error(c, DHIP_E_STATE, "dhip_ch_prepare must precede dhip_ch_predict_start");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int64_t n2 = 2 * h->ndof;
    int32_t st = 0;
    if (h->nhist == 0 || n2 == 0) {
        if (n2) DHIP_HIPCHK(c, hipMemcpyAsync(x0, h->x.p, sizeof(double) * n2,
                                              where == DHIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                              c->stream));
    } else {
        DHIP_HIPCHK(c, h->xstart.ensure(sizeof(double) * (n2 + 1)));
        int32_t *flag = nullptr;
        int rc = ch_learned_start(h, q, h->xstart.as<double>(), &flag);
        if (rc) return rc;
        DHIP_HIPCHK(c, hipMemcpyAsync(&st, flag, sizeof st, hipMemcpyDeviceToHost, c->stream));
        DHIP_HIPCHK(c, hipMemcpyAsync(x0, h->xstart.p, sizeof(double) * n2,
                                      where == DHIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                      c->stream));
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    if (start) *start = st;
    return flush_profile(c);
}

int dhip_ch_solve_learned(dhip_ch_t *h, dhip_mlp_t *q, int32_t *iters, double *resid, int32_t *start) {
    if (!h || !q) return DHIP_E_ARG;
    dhip_ctx *c = h->ctx;
    if (q->ctx != c) return set_error(c, DHIP_E_ARG, "network and stepper belong to different contexts");
    if (!h->prepared) return set_error(c, DHIP_E_STATE, "dhip_ch_prepare must precede dhip_ch_solve_learned");
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    if (start) *start = 0;
    if (h->nhist == 0 || h->ndof == 0) return dhip_ch_solve(h, nullptr, DHIP_DEVICE, iters, resid);
    DHIP_HIPCHK(c, h->xstart.ensure(sizeof(double) * (2 * h->ndof + 1)));
    int32_t *flag = nullptr;
    int rc = ch_learned_start(h, q, h->xstart.as<double>(), &flag);
    if (rc) return rc;
    if (start) DHIP_HIPCHK(c, hipMemcpyAsync(start, flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    // the verified stepper path: x_n into the history, the start into the state, GMRES
    return dhip_ch_solve(h, h->xstart.as<double>(), DHIP_DEVICE, iters, resid);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched problems: planning, one persistent launch per solve (optionally a replayed hipGraph)
// ---------------------------------------------------------------------------
struct dhip_batch {
    dhip_ctx *ctx = nullptr;
    int32_t nprob = 0;
    std::vector<BatchProb> hprob;
    DevBuf prob, order, queue, phi_own;
    DevBuf cellinfo, cut, flags, loc, first, slot, dofs, rowowner, rowlen, rowptr, col;
    DevBuf vm, vl, vs, vc, rhs, diag, x, r64, d32, r32, z32, p32, q32, scol, sval, soff, stat;
    DevBuf pre_mask, pre_blk, pre_flag;   // pre-pass: cell masks, block counts, zero / error flags
    BatchArgs args{};
    PrepassArgs pre{};
    int32_t max_cells = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int g_threads = 0, g_grid = 0, g_max_iter = 0, g_rel = 0, g_prepass = -1, g_split = -1, g_sthreads = 0;
    int64_t g_lds = 0;
    double g_tol = 0.0;
};

static void batch_release(dhip_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    DevBuf *bufs[] = {&b->prob, &b->order, &b->queue, &b->phi_own, &b->cellinfo, &b->cut, &b->flags, &b->loc,
                      &b->first, &b->slot, &b->dofs, &b->rowowner, &b->rowlen, &b->rowptr, &b->col, &b->vm,
                      &b->vl, &b->vs, &b->vc, &b->rhs, &b->diag, &b->x, &b->r64, &b->d32, &b->r32,
                      &b->z32, &b->p32, &b->q32, &b->scol, &b->sval, &b->soff, &b->stat, &b->pre_mask,
                      &b->pre_blk, &b->pre_flag};
    for (DevBuf *d : bufs) d->release();
}

static const void *batch_kernel(int threads) {
    return threads == 128 ? (const void *)k_batch<128>
                          : threads == 512 ? (const void *)k_batch<512> : (const void *)k_batch<256>;
}

static const void *batch_solve_kernel(int threads) {
    return threads == 1024 ? (const void *)k_batch_solve<1024>
                           : threads == 512 ? (const void *)k_batch_solve<512> : (const void *)k_batch_solve<256>;
}

// Dynamic LDS (the search direction) and grid of a per-problem batch kernel at a residency of
// `per_cu_want` workgroups per CU.
static int batch_launch_shape(dhip_ctx *c, const void *k, int threads, int per_cu_want, int32_t nprob, int64_t *lds,
                              int *grid) {
    int maxs = 0;
    hipFuncAttributes fa{};
    DHIP_HIPCHK(c, hipDeviceGetAttribute(&maxs, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device));
    DHIP_HIPCHK(c, hipFuncGetAttributes(&fa, k));
    int64_t l = ((int64_t)maxs / per_cu_want - (int64_t)fa.sharedSizeBytes - 1024) & ~(int64_t)255;
    l = std::max<int64_t>(0, std::min<int64_t>(l, 128 * 1024));
    if (l > 64 * 1024) DHIP_HIPCHK(c, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l));
    int per_cu = 0;
    DHIP_HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, (size_t)l));
    if (per_cu < 1) return set_error(c, DHIP_E_HIP, "batched kernel cannot be resident");
    *lds = l;
    *grid = (int)std::min<int64_t>(nprob, (int64_t)c->cu_count * std::min(per_cu, per_cu_want));
    return DHIP_OK;
}

extern "C" {

int dhip_batch_create(dhip_ctx_t *c, int32_t nprob, const dhip_batch_problem_t *probs, const double *phi,
                      int64_t phi_count, int32_t where, int32_t zero_policy, const double coef[3], int32_t rhs_kind,
                      const double *rhs_param, double omit_bound, dhip_batch_t **out) {
    if (!c || !probs || !phi || !out || nprob <= 0 || nprob > 65535) return DHIP_E_ARG;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    std::vector<BatchProb> hp(nprob);
    int64_t cells = 0, max_cells = 0, max_verts = 0;
    for (int32_t i = 0; i < nprob; ++i) {
        const dhip_batch_problem_t &q = probs[i];
        for (int k = 0; k < 3; ++k)
            if (q.n[k] <= 0) return set_error(c, DHIP_E_ARG, "problem %d: cell counts must be positive", i);
        const int64_t ncell = (int64_t)q.n[0] * q.n[1] * q.n[2];
        const int64_t hg = (2LL * q.n[0] + 1) * (2LL * q.n[1] + 1) * (2LL * q.n[2] + 1);
        if (ncell * 6 >= ((int64_t)1 << 26))
            return set_error(c, DHIP_E_CAPACITY, "problem %d: %lld tets exceed the 2^26 supported", i,
                             (long long)(ncell * 6));
        if (q.phi_offset < 0 || q.phi_offset + hg > phi_count)
            return set_error(c, DHIP_E_ARG, "problem %d: level set [%lld, %lld) outside the %lld values", i,
                             (long long)q.phi_offset, (long long)(q.phi_offset + hg), (long long)phi_count);
        MeshP &m = hp[i].m;
        for (int cc = 0; cc < 3; ++cc) {
            m.orig[cc] = q.orig[cc];
            for (int k = 0; k < 3; ++k) m.off[k][cc] = 1.0 / (double)q.n[k] * q.e[k * 3 + cc];  // geom/builder.cpp:108-110
            m.n[cc] = q.n[cc];
        }
        m.ncell = ncell;
        for (int k = 0; k < 3; ++k) hp[i].lc[k] = q.lc[k];
        hp[i].phi_off = q.phi_offset;
        hp[i].cell_off = cells;
        cells += ncell;
        max_cells = std::max(max_cells, ncell);
        max_verts = std::max<int64_t>(max_verts, (int64_t)(q.n[0] + 1) * (q.n[1] + 1) * (q.n[2] + 1));
    }
    dhip_batch *b = new dhip_batch();
    b->ctx = c;
    b->nprob = nprob;
    auto fail = [&](int code) {
        batch_release(b);
        delete b;
        return code;
    };
    const double *dphi = phi;
    if (where != DHIP_DEVICE) {
        if (b->phi_own.ensure(sizeof(double) * phi_count) ||
            hipMemcpyAsync(b->phi_own.p, phi, sizeof(double) * phi_count, hipMemcpyHostToDevice, c->stream))
            return fail(set_error(c, DHIP_E_HIP, "level-set upload failed"));
        dphi = b->phi_own.as<double>();
    }
    // ---- planning: cut tets and cut-tet vertices per problem ----
    DevBuf mask, cnt;
    if (mask.ensure(cells + 8) || cnt.ensure(sizeof(int32_t) * 3 * nprob) ||
        b->prob.ensure(sizeof(BatchProb) * nprob))
        return fail(set_error(c, DHIP_E_NOMEM, "batch planning buffers"));
    auto cleanup_plan = [&]() {
        mask.release();
        cnt.release();
    };
    if (hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * 3 * nprob, c->stream) ||
        hipMemcpyAsync(b->prob.p, hp.data(), sizeof(BatchProb) * nprob, hipMemcpyHostToDevice, c->stream)) {
        cleanup_plan();
        return fail(set_error(c, DHIP_E_HIP, "batch planning setup failed"));
    }
    int32_t *ncut_d = cnt.as<int32_t>(), *zero_d = ncut_d + nprob, *nvert_d = zero_d + nprob;
    hipLaunchKernelGGL(k_batch_cells, dim3(blocks_for(max_cells, BPLAN), nprob), dim3(BPLAN), 0, c->stream,
                       b->prob.as<BatchProb>(), dphi, mask.as<uint8_t>(), ncut_d, zero_d);
    hipLaunchKernelGGL(k_batch_verts, dim3(blocks_for(max_verts, BPLAN), nprob), dim3(BPLAN), 0, c->stream,
                       b->prob.as<BatchProb>(), (const uint8_t *)mask.p, nvert_d);
    std::vector<int32_t> hc(3 * nprob);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(hc.data(), cnt.p, sizeof(int32_t) * 3 * nprob, hipMemcpyDeviceToHost, c->stream) ||
        hipStreamSynchronize(c->stream)) {
        cleanup_plan();
        return fail(set_error(c, DHIP_E_HIP, "batch planning failed"));
    }
    cleanup_plan();
    int64_t cut_tot = 0, row_tot = 0, nz_tot = 0, sell_tot = 0, slice_tot = 0;
    for (int32_t i = 0; i < nprob; ++i) {
        BatchProb &P = hp[i];
        P.ncut_cap = hc[i];
        P.row_cap = hc[2 * nprob + i];
        P.nz_cap = 15 * P.row_cap;
        P.cut_off = cut_tot;
        P.row_off = row_tot;
        P.nz_off = nz_tot;
        P.sell_off = sell_tot;
        P.slice_off = slice_tot;
        cut_tot += P.ncut_cap;
        row_tot += P.row_cap;
        nz_tot += P.nz_cap;
        sell_tot += 15 * 64 * ((P.row_cap + 63) / 64);
        slice_tot += (P.row_cap + 63) / 64 + 1;
        if (P.nz_cap >= ((int64_t)1 << 31)) return fail(set_error(c, DHIP_E_CAPACITY, "problem %d too large", i));
    }
    b->hprob = hp;
    // largest first: the longest-processing-time order balances uneven problems over the workgroups
    std::vector<int32_t> ord(nprob);
    for (int32_t i = 0; i < nprob; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) { return hc[x] > hc[y]; });
    bool ok = !b->order.ensure(sizeof(int32_t) * nprob) && !b->queue.ensure(64) &&
              !b->cellinfo.ensure(sizeof(uint32_t) * (cells + 1)) && !b->cut.ensure(sizeof(int64_t) * (cut_tot + 1)) &&
              !b->flags.ensure(cut_tot + 1) && !b->loc.ensure(sizeof(double) * LOC_N * (cut_tot + 1)) &&
              !b->first.ensure(sizeof(int32_t) * 4 * (cut_tot + 1)) && !b->slot.ensure(sizeof(int32_t) * 4 * (cut_tot + 1)) &&
              !b->dofs.ensure(sizeof(int32_t) * 4 * (cut_tot + 1)) && !b->rowowner.ensure(sizeof(int32_t) * (row_tot + 1)) &&
              !b->rowlen.ensure(sizeof(int32_t) * (row_tot + 1)) &&
              !b->rowptr.ensure(sizeof(int32_t) * (row_tot + nprob + 1)) && !b->col.ensure(sizeof(int32_t) * (nz_tot + 1)) &&
              !b->vm.ensure(sizeof(double) * (nz_tot + 1)) && !b->vl.ensure(sizeof(double) * (nz_tot + 1)) &&
              !b->vs.ensure(sizeof(double) * (nz_tot + 1)) && !b->vc.ensure(sizeof(double) * (nz_tot + 1)) &&
              !b->scol.ensure(sizeof(int32_t) * (sell_tot + 1)) && !b->sval.ensure(sizeof(float) * (sell_tot + 1)) &&
              !b->soff.ensure(sizeof(int32_t) * (slice_tot + 1)) && !b->rhs.ensure(sizeof(double) * (row_tot + 1)) &&
              !b->diag.ensure(sizeof(double) * (row_tot + 1)) && !b->x.ensure(sizeof(double) * (row_tot + 1)) &&
              !b->r64.ensure(sizeof(double) * (row_tot + 1)) && !b->d32.ensure(sizeof(float) * (row_tot + 1)) &&
              !b->r32.ensure(sizeof(float) * (row_tot + 1)) && !b->z32.ensure(sizeof(float) * (row_tot + 1)) &&
              !b->p32.ensure(sizeof(float) * (row_tot + 1)) && !b->q32.ensure(sizeof(float) * (row_tot + 1)) &&
              !b->stat.ensure(sizeof(BatchStat) * nprob) && !b->pre_mask.ensure(cells + 8) &&
              !b->pre_blk.ensure(sizeof(int32_t) * ((int64_t)nprob * blocks_for(max_cells, BPLAN) + 1)) &&
              !b->pre_flag.ensure(sizeof(int32_t) * (2 * (int64_t)nprob + 1));
    if (!ok) return fail(set_error(c, DHIP_E_NOMEM, "batch buffers (%lld cut tets, %lld rows, %lld entries)",
                                   (long long)cut_tot, (long long)row_tot, (long long)nz_tot));
    if (hipMemcpyAsync(b->prob.p, hp.data(), sizeof(BatchProb) * nprob, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(b->order.p, ord.data(), sizeof(int32_t) * nprob, hipMemcpyHostToDevice, c->stream) ||
        hipMemsetAsync(b->stat.p, 0, sizeof(BatchStat) * nprob, c->stream) || hipStreamSynchronize(c->stream))
        return fail(set_error(c, DHIP_E_HIP, "batch table upload failed"));
    BatchArgs &a = b->args;
    a.prob = b->prob.as<BatchProb>();
    a.order = b->order.as<int32_t>();
    a.nprob = nprob;
    a.queue = b->queue.as<int32_t>();
    a.phi = dphi;
    a.cellinfo = b->cellinfo.as<uint32_t>();
    a.cut = b->cut.as<int64_t>();
    a.flags = b->flags.as<uint8_t>();
    a.loc = b->loc.as<double>();
    a.first = b->first.as<int32_t>();
    a.slot = b->slot.as<int32_t>();
    a.dofs = b->dofs.as<int32_t>();
    a.rowowner = b->rowowner.as<int32_t>();
    a.rowlen = b->rowlen.as<int32_t>();
    a.rowptr = b->rowptr.as<int32_t>();
    a.col = b->col.as<int32_t>();
    a.vm = b->vm.as<double>();
    a.vl = b->vl.as<double>();
    a.vs = b->vs.as<double>();
    a.vc = b->vc.as<double>();
    a.rhs = b->rhs.as<double>();
    a.diag = b->diag.as<double>();
    a.x = b->x.as<double>();
    a.r64 = b->r64.as<double>();
    a.d32 = b->d32.as<float>();
    a.r32 = b->r32.as<float>();
    a.z32 = b->z32.as<float>();
    a.p32 = b->p32.as<float>();
    a.q32 = b->q32.as<float>();
    a.scol = b->scol.as<int32_t>();
    a.sval = b->sval.as<float>();
    a.soff = b->soff.as<int32_t>();
    a.stat = b->stat.as<BatchStat>();
    a.coef0 = coef ? coef[0] : 1.0;
    a.coef1 = coef ? coef[1] : 1.0;
    a.coef2 = coef ? coef[2] : 1.0;
    a.rhs_kind = rhs_kind;
    a.rhs_p0 = rhs_param ? rhs_param[0] : 0.0;
    a.omit_bound = omit_bound;
    a.zero_throw = zero_policy == DHIP_ZERO_THROW;
    a.inner_tol = 1e-4;
    a.inner_max = 400;
    a.pre_zero = b->pre_flag.as<int32_t>();
    a.pre_err = a.pre_zero + nprob;
    PrepassArgs &q = b->pre;
    q.prob = a.prob;
    q.nprob = nprob;
    q.phi = dphi;
    q.cellmask = b->pre_mask.as<uint8_t>();
    q.blkcnt = b->pre_blk.as<int32_t>();
    q.max_blocks = (int32_t)blocks_for(max_cells, BPLAN);
    q.zero = b->pre_flag.as<int32_t>();
    q.perr = q.zero + nprob;
    q.cellinfo = a.cellinfo;
    q.cut = a.cut;
    q.cut_total = cut_tot;
    q.flags = a.flags;
    q.loc = a.loc;
    q.coef0 = a.coef0;
    q.coef1 = a.coef1;
    q.coef2 = a.coef2;
    q.rhs_kind = a.rhs_kind;
    q.rhs_p0 = a.rhs_p0;
    b->max_cells = (int32_t)max_cells;
    *out = b;
    return DHIP_OK;
}

void dhip_batch_destroy(dhip_batch_t *b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_release(b);
    delete b;
}

int dhip_batch_solve(dhip_batch_t *b, int32_t max_iter, double tol, int32_t rel, int32_t use_graph) {
    if (!b || max_iter < 1 || !(tol > 0.0)) return DHIP_E_ARG;
    dhip_ctx *c = b->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    const int threads = c->batch_threads;
    const void *k = batch_kernel(threads);
    // dynamic LDS for the search direction: what the requested residency leaves of the CU's LDS
    int64_t lds = 0;
    int grid = 0;
    int rc = batch_launch_shape(c, k, threads, c->batch_blocks_per_cu, b->nprob, &lds, &grid);
    if (rc) return rc;
    const int split = c->batch_split;
    const int sthreads = c->batch_solve_threads;
    const void *ks = batch_solve_kernel(sthreads);
    int64_t slds = 0;
    int sgrid = 0;
    if (split) {
        rc = batch_launch_shape(c, ks, sthreads, 1, b->nprob, &slds, &sgrid);
        if (rc) return rc;
    }
    b->args.split = split;
    b->args.lds_floats = (split ? slds : lds) / 4;
    b->args.max_iter = max_iter;
    b->args.tol = tol;
    b->args.rel = rel;
    b->args.prepass = c->batch_prepass;
    BatchArgs a = b->args;
    void *argv[] = {&a};
    const PrepassArgs q = b->pre;
    auto enqueue = [&]() -> hipError_t {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->stream, b->queue.as<uint32_t>(), 16);
        if (a.prepass) {
            const int nflag = 2 * b->nprob;
            hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(nflag, 256)), dim3(256), 0, c->stream,
                               b->pre_flag.as<uint32_t>(), nflag);
            const dim3 g2(q.max_blocks, b->nprob);
            hipLaunchKernelGGL(k_batch_cut_count, g2, dim3(BPLAN), 0, c->stream, q);
            hipLaunchKernelGGL(k_batch_cut_scan, dim3(b->nprob), dim3(BPLAN), 0, c->stream, q);
            hipLaunchKernelGGL(k_batch_cut_fill, g2, dim3(BPLAN), 0, c->stream, q);
            if (q.cut_total > 0)
                hipLaunchKernelGGL(k_batch_local, dim3(blocks_for(q.cut_total, LOCAL_BLOCK)), dim3(LOCAL_BLOCK), 0,
                                   c->stream, q);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = hipLaunchKernel(k, dim3(grid), dim3(threads), argv, split ? 0 : (size_t)lds, c->stream);
        if (e != hipSuccess || !split) return e;
        return hipLaunchKernel(ks, dim3(sgrid), dim3(sthreads), argv, (size_t)slds, c->stream);
    };
    if (!use_graph) {
        ProfScope ps(c, "batch");
        DHIP_HIPCHK(c, enqueue());
    } else {
        const bool same = b->exec && b->g_threads == threads && b->g_grid == grid && b->g_max_iter == max_iter &&
                          b->g_tol == tol && b->g_rel == rel && b->g_lds == lds && b->g_prepass == a.prepass &&
                          b->g_split == split && b->g_sthreads == sthreads;
        if (!same) {
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
            DHIP_HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t e = enqueue();
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            DHIP_HIPCHK(c, e);
            DHIP_HIPCHK(c, e2);
            b->graph = g;
            DHIP_HIPCHK(c, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
            b->g_threads = threads;
            b->g_grid = grid;
            b->g_max_iter = max_iter;
            b->g_tol = tol;
            b->g_rel = rel;
            b->g_lds = lds;
            b->g_prepass = a.prepass;
            b->g_split = split;
            b->g_sthreads = sthreads;
        }
        ProfScope ps(c, "batch_graph");
        DHIP_HIPCHK(c, hipGraphLaunch(b->exec, c->stream));
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return flush_profile(c);
}

int dhip_batch_stats(dhip_batch_t *b, dhip_batch_stat_t *out) {
    if (!b || !out) return DHIP_E_ARG;
    dhip_ctx *c = b->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    std::vector<BatchStat> st(b->nprob);
    DHIP_HIPCHK(c, hipMemcpyAsync(st.data(), b->stat.p, sizeof(BatchStat) * b->nprob, hipMemcpyDeviceToHost,
                                  c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int32_t i = 0; i < b->nprob; ++i) {
        out[i].status = st[i].status;
        out[i].ncut = st[i].ncut;
        out[i].ndof = st[i].ndof;
        out[i].nnz = st[i].nnz;
        out[i].iters = st[i].iters;
        out[i].outer = st[i].outer;
        out[i].resid = st[i].resid;
        out[i].start_s = (double)st[i].t0 * 1e-8;
        out[i].end_s = (double)st[i].t1 * 1e-8;
        out[i].assembled_s = (double)st[i].t_asm * 1e-8;
    }
    return DHIP_OK;
}

int dhip_batch_problem(dhip_batch_t *b, int32_t p, int32_t *rowptr, int32_t *colind, double *comb, double *rhs,
                       double *x) {
    if (!b || p < 0 || p >= b->nprob) return DHIP_E_ARG;
    dhip_ctx *c = b->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    BatchStat st{};
    DHIP_HIPCHK(c, hipMemcpyAsync(&st, b->stat.as<BatchStat>() + p, sizeof st, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st.status != DHIP_OK && st.status != DHIP_E_NOT_CONVERGED)
        return set_error(c, DHIP_E_STATE, "problem %d has no assembled system (status %d)", p, st.status);
    const BatchProb &P = b->hprob[p];
    const size_t nd = (size_t)st.ndof, nz = (size_t)st.nnz;
    int rc = 0;
    rc = rc ? rc : d2h(c, rowptr, b->rowptr.as<int32_t>() + P.row_off + p, sizeof(int32_t) * (nd + 1));
    rc = rc ? rc : d2h(c, colind, b->col.as<int32_t>() + P.nz_off, sizeof(int32_t) * nz);
    rc = rc ? rc : d2h(c, comb, b->vc.as<double>() + P.nz_off, sizeof(double) * nz);
    rc = rc ? rc : d2h(c, rhs, b->rhs.as<double>() + P.row_off, sizeof(double) * nd);
    rc = rc ? rc : d2h(c, x, b->x.as<double>() + P.row_off, sizeof(double) * nd);
    if (rc) return rc;
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}


// ---------------------------------------------------------------------------
// on-disk formats (MatrixCL::exportMTX / out / in, VecDescBaseCL::Write / Read)
// ---------------------------------------------------------------------------
static int io_error(dhip_ctx *c, int rc, const char *what) {
    const int code = rc == dhipio::IO_E_COUNT   ? DHIP_E_ARG
                     : rc == dhipio::IO_E_SPACE ? DHIP_E_CAPACITY
                     : rc == dhipio::IO_E_NOMEM ? DHIP_E_NOMEM
                                                : DHIP_E_IO;
    return set_error(c, code, "%s: %s", what, dhipio::io_message(rc));
}

static bool bad_format(int32_t f) { return f != DHIP_MTX_MARKET && f != DHIP_MTX_PLAIN; }

// runs f, turning allocation failures into DHIP_E_NOMEM (no exception crosses the C ABI)
extern "C++" {
template <class F>
static int guarded(dhip_ctx *c, F &&f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return set_error(c, DHIP_E_NOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return set_error(c, DHIP_E_IO, "%s", e.what());
    }
}
}

static int read_text(dhip_ctx *c, const char *path, std::string &s) {
    const int rc = dhipio::slurp(path, s);
    return rc ? io_error(c, rc, path) : DHIP_OK;
}

static int matrix_to_host(const dhip_csr *A, std::vector<int32_t> &rp, std::vector<int32_t> &col,
                          std::vector<double> &val) {
    dhip_ctx *c = A->ctx;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    rp.resize((size_t)A->n + 1);
    col.resize((size_t)A->nnz + 1);
    val.resize((size_t)A->nnz + 1);
    DHIP_HIPCHK(c, hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int32_t) * (A->n + 1), hipMemcpyDeviceToHost, c->stream));
    if (A->nnz > 0) {
        DHIP_HIPCHK(c, hipMemcpyAsync(col.data(), A->col, sizeof(int32_t) * A->nnz, hipMemcpyDeviceToHost, c->stream));
        DHIP_HIPCHK(c, hipMemcpyAsync(val.data(), A->val, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, c->stream));
    }
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

int dhip_mtx_format(int32_t format, int64_t nrows, int64_t ncols, const int32_t *rowptr, const int32_t *colind,
                    const double *val, char *buf, int64_t cap, int64_t *len) {
    if (bad_format(format) || nrows < 0 || ncols < 0 || !rowptr || !len)
        return set_error(nullptr, DHIP_E_ARG, "dhip_mtx_format: bad arguments");
    if (!buf) {
        *len = dhipio::matrix_bound(format, nrows, rowptr[nrows]);
        return DHIP_OK;
    }
    if (rowptr[nrows] > 0 && (!colind || !val)) return set_error(nullptr, DHIP_E_ARG, "dhip_mtx_format: no entries");
    return guarded(nullptr, [&] {
        dhipio::Sink out;
        out.buf = buf;
        out.cap = cap;
        const int rc = dhipio::format_matrix(out, format, nrows, ncols, rowptr, colind, val);
        *len = out.len;
        return rc ? io_error(nullptr, rc, "dhip_mtx_format") : DHIP_OK;
    });
}

int dhip_mtx_write(const char *path, int32_t format, int64_t nrows, int64_t ncols, const int32_t *rowptr,
                   const int32_t *colind, const double *val) {
    if (!path || bad_format(format) || nrows < 0 || ncols < 0 || !rowptr)
        return set_error(nullptr, DHIP_E_ARG, "dhip_mtx_write: bad arguments");
    if (rowptr[nrows] > 0 && (!colind || !val)) return set_error(nullptr, DHIP_E_ARG, "dhip_mtx_write: no entries");
    return guarded(nullptr, [&] {
        const int rc = dhipio::to_file(path, [&](dhipio::Sink &out) {
            return dhipio::format_matrix(out, format, nrows, ncols, rowptr, colind, val);
        });
        return rc ? io_error(nullptr, rc, path) : DHIP_OK;
    });
}

int dhip_mtx_parse(const char *buf, int64_t len, int32_t format, int64_t *nrows, int64_t *ncols, int64_t *nnz,
                   int32_t *rowptr, int32_t *colind, double *val) {
    if ((!buf && len > 0) || len < 0 || !nrows || !ncols || !nnz || bad_format(format) ||
        (rowptr && (!colind || !val)))
        return set_error(nullptr, DHIP_E_ARG, "dhip_mtx_parse: bad arguments");
    return guarded(nullptr, [&] {
        const std::string s(buf ? buf : "", (size_t)len);
        const int rc = dhipio::parse_matrix(s, format, nrows, ncols, nnz, rowptr, colind, val);
        return rc ? io_error(nullptr, rc, "dhip_mtx_parse") : DHIP_OK;
    });
}

int dhip_mtx_read(const char *path, int32_t format, int64_t *nrows, int64_t *ncols, int64_t *nnz, int32_t *rowptr,
                  int32_t *colind, double *val) {
    if (!path || !nrows || !ncols || !nnz || bad_format(format) || (rowptr && (!colind || !val)))
        return set_error(nullptr, DHIP_E_ARG, "dhip_mtx_read: bad arguments");
    return guarded(nullptr, [&] {
        std::string s;
        int rc = read_text(nullptr, path, s);
        if (rc) return rc;
        rc = dhipio::parse_matrix(s, format, nrows, ncols, nnz, rowptr, colind, val);
        return rc ? io_error(nullptr, rc, path) : DHIP_OK;
    });
}

int dhip_csr_export(const dhip_csr_t *A, const char *path, int32_t format) {
    if (!A || !path) return DHIP_E_ARG;
    dhip_ctx *c = A->ctx;
    if (bad_format(format)) return set_error(c, DHIP_E_ARG, "unknown format");
    return guarded(c, [&] {
        std::vector<int32_t> rp, col;
        std::vector<double> val;
        int rc = matrix_to_host(A, rp, col, val);
        if (rc) return rc;
        rc = dhipio::to_file(path, [&](dhipio::Sink &out) {
            return dhipio::format_matrix(out, format, A->n, A->n, rp.data(), col.data(), val.data());
        });
        return rc ? io_error(c, rc, path) : DHIP_OK;
    });
}

int dhip_csr_import(dhip_ctx_t *c, const char *path, int32_t format, dhip_csr_t **A) {
    if (!c || !path || !A) return DHIP_E_ARG;
    if (bad_format(format)) return set_error(c, DHIP_E_ARG, "unknown format");
    return guarded(c, [&] {
        std::string s;
        int rc = read_text(c, path, s);
        if (rc) return rc;
        int64_t nr = 0, nc = 0, nz = 0;
        if ((rc = dhipio::parse_matrix(s, format, &nr, &nc, &nz, nullptr, nullptr, nullptr)))
            return io_error(c, rc, path);
        if (nr != nc)
            return set_error(c, DHIP_E_ARG, "%s: %lld x %lld is not square", path, (long long)nr, (long long)nc);
        std::vector<int32_t> rp((size_t)nr + 1), col((size_t)nz + 1);
        std::vector<double> val((size_t)nz + 1);
        if ((rc = dhipio::parse_matrix(s, format, &nr, &nc, &nz, rp.data(), col.data(), val.data())))
            return io_error(c, rc, path);
        return dhip_csr_create(c, nr, nz, rp.data(), col.data(), val.data(), DHIP_HOST, A);
    });
}

// host copy of a caller vector (device: copied down; host: used in place)
static int vector_on_host(dhip_ctx *c, const double *x, int64_t n, int32_t where, std::vector<double> &h,
                          const double **src) {
    *src = x;
    if (where != DHIP_DEVICE) return DHIP_OK;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    h.resize((size_t)n + 1);
    if (n > 0) DHIP_HIPCHK(c, hipMemcpyAsync(h.data(), x, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    *src = h.data();
    return DHIP_OK;
}

static int vector_to_caller(dhip_ctx *c, double *x, const std::vector<double> &h, int64_t n, int32_t where) {
    if (where != DHIP_DEVICE || n == 0) return DHIP_OK;
    DHIP_HIPCHK(c, hipSetDevice(c->device));
    DHIP_HIPCHK(c, hipMemcpyAsync(x, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    DHIP_HIPCHK(c, hipStreamSynchronize(c->stream));
    return DHIP_OK;
}

static bool bad_vec_args(dhip_ctx *c, const double *x, int64_t n, int32_t where) {
    return n < 0 || (n > 0 && !x) || (where != DHIP_HOST && where != DHIP_DEVICE) || (where == DHIP_DEVICE && !c);
}

int dhip_vec_format(dhip_ctx_t *c, const double *x, int64_t n, int32_t where, int32_t binary, char *buf, int64_t cap,
                    int64_t *len) {
    if (bad_vec_args(c, x, n, where) || !len) return set_error(c, DHIP_E_ARG, "dhip_vec_format: bad arguments");
    if (!buf) {
        *len = dhipio::vec_bound(n, binary != 0);
        return DHIP_OK;
    }
    return guarded(c, [&] {
        std::vector<double> h;
        const double *src;
        int rc = vector_on_host(c, x, n, where, h, &src);
        if (rc) return rc;
        dhipio::Sink out;
        out.buf = buf;
        out.cap = cap;
        rc = dhipio::format_vec(out, src, n, binary != 0);
        *len = out.len;
        return rc ? io_error(c, rc, "dhip_vec_format") : DHIP_OK;
    });
}

int dhip_vec_write(dhip_ctx_t *c, const char *path, const double *x, int64_t n, int32_t where, int32_t binary) {
    if (!path || bad_vec_args(c, x, n, where)) return set_error(c, DHIP_E_ARG, "dhip_vec_write: bad arguments");
    return guarded(c, [&] {
        std::vector<double> h;
        const double *src;
        int rc = vector_on_host(c, x, n, where, h, &src);
        if (rc) return rc;
        rc = dhipio::to_file(path, [&](dhipio::Sink &out) { return dhipio::format_vec(out, src, n, binary != 0); });
        return rc ? io_error(c, rc, path) : DHIP_OK;
    });
}

int dhip_vec_parse(dhip_ctx_t *c, const char *buf, int64_t len, double *x, int64_t n, int32_t where, int32_t binary) {
    if ((!buf && len > 0) || len < 0 || bad_vec_args(c, x, n, where))
        return set_error(c, DHIP_E_ARG, "dhip_vec_parse: bad arguments");
    return guarded(c, [&] {
        const std::string s(buf ? buf : "", (size_t)len);
        std::vector<double> h(where == DHIP_DEVICE ? (size_t)n + 1 : 0);
        int rc = dhipio::parse_vec(s, where == DHIP_DEVICE ? h.data() : x, n, binary != 0);
        if (rc) return io_error(c, rc, "dhip_vec_parse");
        return vector_to_caller(c, x, h, n, where);
    });
}

int dhip_vec_read(dhip_ctx_t *c, const char *path, double *x, int64_t n, int32_t where, int32_t binary) {
    if (!path || bad_vec_args(c, x, n, where)) return set_error(c, DHIP_E_ARG, "dhip_vec_read: bad arguments");
    return guarded(c, [&] {
        std::string s;
        int rc = read_text(c, path, s);
        if (rc) return rc;
        std::vector<double> h(where == DHIP_DEVICE ? (size_t)n + 1 : 0);
        rc = dhipio::parse_vec(s, where == DHIP_DEVICE ? h.data() : x, n, binary != 0);
        if (rc) return io_error(c, rc, path);
        return vector_to_caller(c, x, h, n, where);
    });
}

int dhip_vec_count(const char *path, int32_t binary, int64_t *n) {
    if (!path || !n) return set_error(nullptr, DHIP_E_ARG, "dhip_vec_count: bad arguments");
    // the count is at the start of the file
    FILE *f = std::fopen(path, "rb");
    if (!f) return io_error(nullptr, dhipio::IO_E_OPEN, path);
    char head[64];
    const size_t k = std::fread(head, 1, sizeof head, f);
    std::fclose(f);
    const int rc = dhipio::parse_vec_count(std::string(head, k), binary != 0, n);
    return rc ? io_error(nullptr, rc, path) : DHIP_OK;
}

}  // extern "C"
