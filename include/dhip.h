/*
 * dhip.h -- C ABI of the MI355X (gfx950) trace-FEM assembly + Krylov library.
 *
 * The library replaces, for one GPU, the DROPS (56th/drops) hot path
 *   accumulate( {InterfaceCommonDataP1CL, InterfaceMatrixAccuCL<Local*P1CL>,
 *                InterfaceVectorAccuCL<LocalVectorP1CL>} )      (assembly)
 *   PCGSolverCL<JACPcCL>::Solve / GMResSolverCL::Solve           (Krylov core)
 * behind plain C entry points: pointers, sizes and integer error codes, no
 * C++ or torch types.  The reference is header-template C++ with no FFI of
 * its own; each entry point names the reference interface it stands in for.
 *
 * Conventions
 *   - Every call returns DHIP_OK (0) or a DHIP_E_* code; dhip_last_error()
 *     gives the message of the last failure on that context.
 *   - A context owns one HIP stream and all device buffers it allocates; calls
 *     on one context are stream ordered and synchronous on return.  A context
 *     must not be used by two host threads at once.
 *   - `where` arguments say whether caller pointers are host (DHIP_HOST) or
 *     device (DHIP_DEVICE) memory.
 *   - Level sets are P2 nodal vectors on a Kuhn brick stored as a dense
 *     float64 array over the half-spaced grid, shape (2n3+1, 2n2+1, 2n1+1),
 *     x fastest: vertices at even indices, Kuhn edge midpoints at odd ones.
 *   - Tets are numbered t = 6*((i3*n2 + i2)*n1 + i1) + k (geom/builder.cpp:150-176),
 *     vertices v = (i3*(n2+1) + i2)*(n1+1) + i1.
 *   - CSR matrices use int32 row pointers and column indices, columns sorted
 *     ascending within each row, explicit zeros kept (SparseMatBuilderCL).
 */
#ifndef DHIP_H
#define DHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define DHIP_OK 0
#define DHIP_E_ZERO_NODE 1       /* std::invalid_argument "levelset is exactly zero at node(s)",
                                    geom/subtriangulation.tpp:44 */
#define DHIP_E_ZERO_PATTERN 2    /* DROPSErrCL "found a full body in the zero level set",
                                    geom/reftetracut.h:152-153 */
#define DHIP_E_ARG 3
#define DHIP_E_NOT_CONVERGED 4   /* solver returned false, num/krylovsolver.h:176-178 */
#define DHIP_E_NOMEM 5
#define DHIP_E_HIP 6
#define DHIP_E_STATE 7           /* call out of order (e.g. number before classify) */
#define DHIP_E_CAPACITY 8        /* size beyond what the library supports */
#define DHIP_E_IO 9              /* file cannot be opened, written or parsed */

#define DHIP_HOST 0
#define DHIP_DEVICE 1

/* zero-node policy: DHIP_ZERO_THROW reproduces the reference (a cut tet with an
 * exact-zero P2 node is an error); DHIP_ZERO_ALLOW treats zeros as roots
 * (zero vertices and boundary facets of the patch, geom/reftetracut.cpp:53-60). */
#define DHIP_ZERO_THROW 0
#define DHIP_ZERO_ALLOW 1

/* operators of the P1 interface assembly */
#define DHIP_OP_MASS 0   /* LocalInterfaceMassP1CL, surfactant/ifacetransp.h:881-906 */
#define DHIP_OP_LB 1     /* LocalLaplaceBeltramiP1CL, surfactant/ifacetransp.h:1657-1691 */
#define DHIP_OP_STAB 2   /* LocalNormalLaplaceBulkP1CL, surfactant/ifacetransp.h:2032-2051 */
#define DHIP_OP_COMB 3   /* LinComb of the three, num/spmat.h:819-825 + 1458-1560 */

/* right-hand sides (LocalVectorP1CL, surfactant/ifacetransp.h:768-798) */
#define DHIP_RHS_NONE 0
#define DHIP_RHS_3XY_MINUS_X3 1  /* f(x) = 3 x y - x^3 */
#define DHIP_RHS_CONST 2         /* f(x) = param[0] */
#define DHIP_RHS_VALUES 3        /* f given at the patches' quadrature points (dhip_quad_points, dhip_rhs_values) */

/* preconditioners */
#define DHIP_PC_NONE 0           /* DummyPcCL, num/precond.h:1032-1065 */
#define DHIP_PC_JACOBI 1         /* JACPcCL, num/precond.h:874 */

/* Krylov variants (the `precision` argument of dhip_pcg) */
#define DHIP_FP64 0              /* FP64, one fused reduction per iteration (Chronopoulos-Gear
                                    arrangement of the PCG recurrences) */
#define DHIP_FP32_IR 1           /* FP32 inner iterations, FP64 residual/refinement */
#define DHIP_FP64_CLASSIC 2      /* FP64, the reference's operation order: two reductions per
                                    iteration (num/krylovsolver.h:150-191) */

typedef struct dhip_ctx dhip_ctx_t;
typedef struct dhip_csr dhip_csr_t;

/* ---- context ---- */
int dhip_create(dhip_ctx_t **ctx, int32_t device);
void dhip_destroy(dhip_ctx_t *ctx);
const char *dhip_last_error(const dhip_ctx_t *ctx);
/* Per-kernel HIP-event timing on the context's stream: 0 = off, 1 = every library kernel
 * launch is bracketed by events, 2 = only the persistent solver launches ("pcg64",
 * "pcg64_classic", "pcg_mixed", "gmres"), so that timing them adds nothing to the other
 * launches.  dhip_profile_get returns the accumulated milliseconds and launch count of
 * kernel `name`. */
int dhip_profile_enable(dhip_ctx_t *ctx, int32_t on);
int dhip_profile_get(dhip_ctx_t *ctx, const char *name, double *total_ms, int64_t *count);
int dhip_profile_reset(dhip_ctx_t *ctx);
/* Wall-clock probe of the persistent FP64 PCG and of GMRES (100 MHz ticks per block:
 * setup (GMRES: the SpMV part of the compute time), compute phases, synchronisations, total).  dhip_solver_probe(ctx, 1, ...)
 * arms it for the following solves; `ticks` (4 per block, may be NULL) receives
 * the last solve's values and `nblocks` its block count. */
int dhip_solver_probe(dhip_ctx_t *ctx, int32_t on, uint64_t *ticks, int32_t *nblocks);
/* Tuning parameters (chosen by drops_amd/tune.py): "pcg_rows_per_block" -- rows owned
 * by one block of the persistent solvers (0 = library default, 256); "xcd_map" -- 1 (default):
 * the blocks of each XCD own one contiguous run of row ranges, so most gathered neighbour rows
 * come from the XCD's own L2; 0: range i belongs to block i (same results up to the order in
 * which the blocks' partial sums are added); "coherent_exchange" -- how the persistent solvers'
 * blocks exchange vectors: 0 release/acquire fences around every grid reduction, 1 coherent
 * (L2-scope) loads and stores of the exchanged values without fences, 2 (default) as 1 and the FP64
 * PCG with generation-tagged 16-byte records and its block's gathered values staged in LDS (the
 * same iterates bit for bit while a block's slice fits its LDS; the single-reduction kernel
 * otherwise), 3 as 2 with the FP64 PCG in the pipelined arrangement of CG (one vector of records
 * per iteration, the wait for the block partials overlapped with the next product; iterates equal
 * CG's up to rounding, and b - A x is checked, restarting from it if needed, before it stops);
 * "gmres_precision" -- DHIP_FP64 (default) or DHIP_FP32_IR for dhip_gmres and
 * dhip_gmres_diag, as dhip_ch_set_precision; "batch_threads",
 * "batch_blocks_per_cu", "batch_prepass", "batch_split", "batch_solve_threads" -- see
 * dhip_batch_solve. */
int dhip_tune(dhip_ctx_t *ctx, const char *key, int64_t value);
/* Device buffers for callers without their own allocator (bench, tests). */
int dhip_malloc(dhip_ctx_t *ctx, void **dptr, int64_t bytes);
int dhip_free(dhip_ctx_t *ctx, void *dptr);
int dhip_memcpy(dhip_ctx_t *ctx, void *dst, const void *src, int64_t bytes, int32_t kind);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int dhip_synchronize(dhip_ctx_t *ctx);

/* ---- mesh and level set ----
 * BrickBuilderCL(orig, e1, e2, e3, n1, n2, n3) (geom/builder.cpp:99-178);
 * e is row-major {e1, e2, e3}. */
int dhip_mesh_kuhn(dhip_ctx_t *ctx, const double orig[3], const double e[9], const int32_t n[3]);
/* The P2 level set (LevelsetP2CL::Phi, levelset/levelset.h:69) on the half-grid.
 * DHIP_DEVICE: the library keeps the pointer (no copy); it must stay valid. */
int dhip_levelset_p2(dhip_ctx_t *ctx, const double *phi, int64_t count, int32_t where);

/* ---- assembly (replaces accumulate() over InterfaceCommonDataP1CL and the
 *      P1 interface matrix/vector accumulators, num/accumulator.h:215-268) ---- */
/* Sign classification of every tet on its 10 P2 values (equalSigns,
 * geom/subtriangulation.tpp:32-36) and stream compaction of the cut tets in
 * mesh order.  Returns DHIP_E_ZERO_NODE under DHIP_ZERO_THROW when a cut tet
 * has an exact-zero node (distance(), geom/subtriangulation.tpp:39-46). */
int dhip_classify(dhip_ctx_t *ctx, int32_t zero_policy, int64_t *ncut);
/* Interface patch (SPatchCL::make_patch<MergeCutPolicyCL>) and all local
 * quantities of every cut tet: 4x4 mass (coef[0] = alpha), Laplace-Beltrami
 * (coef[1] = D), normal-gradient stabilisation (coef[2] = D), load vector,
 * the interface-numbering integrals of phi_i^2 and the patch area. */
int dhip_local_p1(dhip_ctx_t *ctx, const double coef[3], int32_t rhs_kind, const double *rhs_param);
/* Right-hand sides from any function (LocalVectorP1CL / LocalVectorP2CL with an InstatScalarFunction f,
 * surfactant/ifacetransp.h:768-798, which evaluates f at the vertices of the composite Quad5 domain of
 * each tet's patch): dhip_quad_points lists those points in world coordinates -- cut tets in cut-list
 * order, per tet the patch facets in make_patch order, 7 Quad5_2D nodes per facet -- after dhip_classify
 * (or an assembly) on the current mesh and level set; xyz == NULL gives the count only, else xyz holds
 * 3 * npoints doubles (host or device by where).  The caller evaluates f there and hands the values to
 * dhip_rhs_values; the next local pass with rhs_kind DHIP_RHS_VALUES on the same cut list (P1, P2 or
 * narrow band) sums them exactly as it sums a built-in kind. */
int dhip_quad_points(dhip_ctx_t *ctx, int64_t *npoints, double *xyz, int32_t where);
int dhip_rhs_values(dhip_ctx_t *ctx, const double *values, int64_t npoints, int32_t where);
/* First-encounter P1 interface numbering, IdxDescCL::CreateNumbering ->
 * CreateNumbOnInterfaceVertex (misc/problem.cpp:443-503, 725-745).
 * omit_bound < 0 numbers every vertex with a positive phi_i^2 integral. */
int dhip_number_p1if(dhip_ctx_t *ctx, double omit_bound, int64_t *ndof);
/* CSR pattern + values of the three operators, their linear combination
 * comb = 1*(lc[0]*MASS + lc[1]*LB) + lc[2]*STAB (LinComb(lc0,M,lc1,A,lc2,C)),
 * and the load vector (update_global_matrix, surfactant/ifacetransp.tpp:31-43;
 * SparseMatBuilderCL::Build, num/spmat.h:621-677). */
int dhip_build_p1(dhip_ctx_t *ctx, const double lc[3], int64_t *nnz);
/* classify + local + number + build in one call. */
int dhip_assemble_p1(dhip_ctx_t *ctx, int32_t zero_policy, const double coef[3], int32_t rhs_kind,
                     const double *rhs_param, double omit_bound, const double lc[3],
                     int64_t *ncut, int64_t *ndof, int64_t *nnz);
/* The narrow-band assembly of SurfactantNarrowBandStblP1CL::Update (surfactant/ifacetransp.cpp:1110-1172):
 * the listed tets are those with distance() <= width (NarrowBandCommonDataP1CL,
 * surfactant/ifacetransp.h:647-700; cut tets have distance 0), numbered by
 * CreateNumbNearInterfaceVertex (misc/problem.cpp:509-555: every vertex of every band tet, first
 * encounter in mesh order); mass, Laplace-Beltrami and load come from the cut tets
 * (InterfaceMatrixAccuCL), the normal stabilisation from every band tet (NarrowBandMatrixAccuP1CL<
 * LocalNormalLaplaceBulkP1CL>, surfactant/ifacetransp.h:703-760).  The CSR pattern is the
 * stabilisation's (it contains the others'); mass and LB hold 0.0 on the entries only band tets
 * touch, where the reference's own mass and LB matrices have no entry.  *nband = listed tets.
 * The getters below (CSR, cut tets = band tets, local data; flags bit 5 = band tet without patch)
 * apply as after dhip_assemble_p1. */
int dhip_assemble_p1_band(dhip_ctx_t *ctx, int32_t zero_policy, const double coef[3], int32_t rhs_kind,
                          const double *rhs_param, double width, const double lc[3], int64_t *nband,
                          int64_t *ndof, int64_t *nnz);

/* ---- P2 interface space on the same cut tets and planar patch ----
 * Local forms of LocalAssembler (surfactant/ifacetransp.h:914-1619) with the P2 basis
 * FE_P2CL (num/fe.h:287-296): mass M_P2P2 (:1579-1588, times coef[0]), Laplace-Beltrami
 * with the projected P2 gradients of buildSurfGradP2 (:1049-1053, times coef[1]), normal
 * = normalised P2 level-set gradient (build3DNormal, :1062-1076), the normal-gradient
 * stabilisation as the true volume integral of C_n_P1P1 (:1545-1555) on the 15-point
 * Quad5 rule (times coef[2]), load vector; composite Quad5 surface rule throughout.
 * The P2 dof order is vertices 0-3, then edges (0,1),(0,2),(1,2),(0,3),(1,3),(2,3). */
int dhip_local_p2(dhip_ctx_t *ctx, const double coef[3], int32_t rhs_kind, const double *rhs_param);
/* First-encounter P2 interface numbering, CreateNumbOnInterfaceP2 (misc/problem.cpp:571-636):
 * a half-grid node gets the next number on its first (tet, local dof) in mesh order whose
 * composite-Quad5 integral of the squared basis function exceeds the limit. */
int dhip_number_p2if(dhip_ctx_t *ctx, double omit_bound, int64_t *ndof);
/* CSR of the P2 operators and load vector, as dhip_build_p1 (same result accessors). */
int dhip_build_p2(dhip_ctx_t *ctx, const double lc[3], int64_t *nnz);
/* classify + local_p2 + number_p2if + build_p2 in one call. */
int dhip_assemble_p2(dhip_ctx_t *ctx, int32_t zero_policy, const double coef[3], int32_t rhs_kind,
                     const double *rhs_param, double omit_bound, const double lc[3],
                     int64_t *ncut, int64_t *ndof, int64_t *nnz);

/* ---- P2 forms on the QuaQua-mapped surface (higher-order geometry) ----
 * InterfaceCommonDataP2CL with QuaQuaMapperCL (surfactant/ifacetransp.h:200-276,
 * levelset/levelsetmapper.h:52-140, levelset/levelsetmapper.cpp:551-638) and
 * LocalMassP2CL / LocalLaplaceBeltramiP2CL / LocalVectorP2CL (ifacetransp.h:2302-2403), as
 * surfactant/surfactant.cpp:1596-1640 assembles them: each cut tet's lattice-1 patch (the linear
 * interpolant of the level set at its vertices), composite Quad5 points mapped by damped Newton to
 * base points on the zero level of the P2 level set along the recovered gradient
 * (averaging_P2_gradient_recovery, num/gradient_recovery.cpp:82-89), |det| of the map's Jacobian
 * and its gradient transformation W^-1 (gradient_trafo, ifacetransp.cpp:894-918).  The load value is
 * f at the base point.  The Laplace-Beltrami form carries no coefficient (the reference's
 * LocalLaplaceBeltramiP2CL ignores its D); the stabilisation slot is zero.  As in the reference the
 * base-point search stays in the starting tetra (locate_new_point, levelsetmapper.cpp:463-490, never
 * searches the neighbourhood it is given), so the tet's polynomials are extrapolated. */
typedef struct {
    int32_t maxiter;     /* LevelsetMapper.Iter: 20 in the reference's parameter files */
    double tol;          /* LevelsetMapper.Tol: 1e-7 */
    double armijo_c;     /* LevelsetMapper.ArmijoConstant: 1e-2 */
    int32_t max_damping; /* damping steps per Newton step: 8 */
} dhip_quaqua_params_t;
/* averaging_P2_gradient_recovery (num/gradient_recovery.cpp:82-89): the gradient of the P2 level set
 * at every half-grid node, averaged over the tets that hold the node as OswaldProjectionP2AccuCL does
 * (num/oswald_projection.h:102-125, a running mean in tet order); grad: 3 doubles per node, nodes in
 * the level set's order. */
int dhip_p2_recovered_gradient(dhip_ctx_t *ctx, double *grad, int32_t where);
/* A load function evaluated by the caller (LocalVectorP2CL with an InstatScalarFunction,
 * ifacetransp.h:2302-2311: f at the base points): dhip_quaqua_points (after dhip_classify) lists the
 * base points of every mapped quadrature point, the cut tets in order and each tet's points in the
 * composite order (7 per facet of its lattice-1 patch); *npoints = their number, xyz (may be NULL)
 * receives 3 coordinates each.  dhip_quaqua_values hands f at those points back (npoints values);
 * an assembly with DHIP_RHS_VALUES on the same level set and parameters then uses them. */
int dhip_quaqua_points(dhip_ctx_t *ctx, const dhip_quaqua_params_t *params, int64_t *npoints, double *xyz,
                       int32_t where);
int dhip_quaqua_values(dhip_ctx_t *ctx, const double *values, int64_t npoints, int32_t where);
/* The mapped mass, Laplace-Beltrami and load of every cut tet into the local slots the P2 CSR build
 * reads (after dhip_number_p2if; dhip_build_p2 follows).  params NULL: the values above.  rhs_kind
 * DHIP_RHS_NONE, _3XY_MINUS_X3, _CONST or _VALUES (given by dhip_quaqua_values). */
int dhip_local_p2_quaqua(dhip_ctx_t *ctx, const dhip_quaqua_params_t *params, int32_t rhs_kind,
                         const double *rhs_param);
/* classify + local_p2 (for the numbering) + number_p2if + local_p2_quaqua + build_p2 in one call:
 * the CSR holds mass (slot 0), Laplace-Beltrami (slot 1), zero stabilisation (slot 2) and their
 * combination with lc; a row's entries come from the tets with a lattice-1 patch only. */
int dhip_assemble_p2_quaqua(dhip_ctx_t *ctx, int32_t zero_policy, const dhip_quaqua_params_t *params,
                            int32_t rhs_kind, const double *rhs_param, double omit_bound, const double lc[3],
                            int64_t *ncut, int64_t *ndof, int64_t *nnz);

/* ---- results of the assembly (host copies; any pointer may be NULL) ---- */
int dhip_get_cut_tets(dhip_ctx_t *ctx, int64_t *tets);
/* per cut tet: 16 mass, 16 LB, 16 stab, 4 rhs, 4 phi_i^2 integrals, area (57 doubles),
 * and flags (bit0 non-empty patch, bits1-4 vertex qualifies for numbering) */
int dhip_get_local(dhip_ctx_t *ctx, double *local57, uint8_t *flags);
/* P2 space: per cut tet 100 mass, 100 LB, 100 stab, 10 rhs, 10 squared-basis integrals
 * (320 doubles, row-major 10x10 in P2 dof order) and flags (bit0 non-empty patch) */
int dhip_get_local_p2(dhip_ctx_t *ctx, double *local320, uint8_t *flags);
/* DOF of each (cut tet, local dof) pair (4 per tet for P1, 10 for P2), -1 where not numbered */
int dhip_get_tet_dofs(dhip_ctx_t *ctx, int32_t *dofs);
/* vertex id of each DOF (P1 space) */
int dhip_get_dof_vertices(dhip_ctx_t *ctx, int64_t *vertex_of_dof);
/* half-grid node index (j3 (2 n2 + 1) + j2) (2 n1 + 1) + j1 of each DOF (either space) */
int dhip_get_dof_nodes(dhip_ctx_t *ctx, int64_t *node_of_dof);
int dhip_get_csr(dhip_ctx_t *ctx, int32_t *rowptr, int32_t *colind, double *mass, double *lb,
                 double *stab, double *comb, double *rhs);
/* Device pointers of the assembled arrays (valid until the next assembly). */
int dhip_device_arrays(dhip_ctx_t *ctx, const int32_t **rowptr, const int32_t **colind,
                       const double **comb, const double **rhs);

/* ---- sparse matrices (MatrixCL = SparseMatBaseCL<double>, num/spmat.h:683-878) ---- */
int dhip_csr_create(dhip_ctx_t *ctx, int64_t nrows, int64_t nnz, const int32_t *rowptr,
                    const int32_t *colind, const double *val, int32_t where, dhip_csr_t **A);
/* wraps the assembled operator `op` (DHIP_OP_*) without copying; valid until the
 * next assembly on the context */
int dhip_csr_from_assembly(dhip_ctx_t *ctx, int32_t op, dhip_csr_t **A);
/* Handles must be destroyed before the context that created them. */
void dhip_csr_destroy(dhip_csr_t *A);
int dhip_csr_info(const dhip_csr_t *A, int64_t *nrows, int64_t *nnz);

/* ---- vector core (num/spmat.h:164-235, 1683-1722) ---- */
int dhip_spmv(dhip_ctx_t *ctx, const dhip_csr_t *A, const double *x, double *y, int32_t where);
/* y = A^T x (transp_mul / y_ATx, num/spmat.h:1729-1757): y has ncols entries, x has A's row count.
 * The transposed pattern is built for the call (counts, scan, fill, each row sorted by original row)
 * so that every y_j is summed in the row order of y_ATx, bit for bit; a column index >= ncols is
 * DHIP_E_ARG. */
int dhip_spmv_t(dhip_ctx_t *ctx, const dhip_csr_t *A, int64_t ncols, const double *x, double *y, int32_t where);
int dhip_dot(dhip_ctx_t *ctx, int64_t n, const double *x, const double *y, double *result,
             int32_t where);
int dhip_axpy(dhip_ctx_t *ctx, int64_t n, double a, const double *x, double *y, int32_t where);

/* ---- Krylov solvers ----
 * PCG(A, x, b, ex, M, max_iter, tol, rel), num/krylovsolver.h:107-196 and
 * PCGSolverCL<PC>::Solve, :1805-1835.  x holds the start vector on entry and
 * the solution on return; *max_iter / *tol are IN limits and OUT used
 * iterations / final residual (relative if rel != 0).  Returns DHIP_OK on
 * convergence, DHIP_E_NOT_CONVERGED otherwise (the reference's `false`).
 * precision DHIP_FP32_IR: FP32 inner PCG sweeps inside an FP64 iterative
 * refinement loop; *max_iter then counts inner iterations. */
int dhip_pcg(dhip_ctx_t *ctx, const dhip_csr_t *A, int32_t pc, const double *b, double *x,
             int32_t where, int32_t *max_iter, double *tol, int32_t rel, int32_t precision);

/* Allocation-free FP64 PCG.  No counterpart in the reference: it serves callers that cannot
 * allocate device memory (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same solver, arguments and results as dhip_pcg with precision DHIP_FP64 and
 * where = DHIP_DEVICE (b and x are device pointers), but the Jacobi diagonal, the exchanged work
 * vectors, the synchronisation slots and the result record live in one fixed scratch array of the
 * library (a static device array, one copy per GPU, held until the process ends; about 56 MiB).
 * Systems above DHIP_NOALLOC_MAX_ROWS rows are rejected with DHIP_E_CAPACITY.  The call allocates
 * nothing, on the device or the host.  Every call on one GPU uses that GPU's array (the slots are
 * cleared by a fill kernel on each call): calls that overlap in time on one GPU, from several
 * contexts or host threads, race and must be serialised by the caller. */
#define DHIP_NOALLOC_MAX_ROWS 1048576
int dhip_pcg_noalloc(dhip_ctx_t *ctx, const dhip_csr_t *A, int32_t pc, const double *b, double *x,
                     int32_t *max_iter, double *tol, int32_t rel);
/* GMRES(A, x, b, ex, M, m, max_iter, tol, rel, calc2norm=false, method),
 * num/krylovsolver.h:320-463 (serial variant: m-1 Arnoldi steps per cycle). */
int dhip_gmres(dhip_ctx_t *ctx, const dhip_csr_t *A, int32_t pc, const double *b, double *x,
               int32_t where, int32_t restart, int32_t *max_iter, double *tol, int32_t rel,
               int32_t left_pc);

/* GMRES as dhip_gmres, preconditioned by z_i = r_i / pcdiag_i with a caller-given diagonal
 * (JACPcCL applied blockwise: DiagBlockPreCL, num/oseenprecond.h:1051-1073, with a Jacobi
 * sweep of each block's preconditioning matrix in place of its inner solve). */
int dhip_gmres_diag(dhip_ctx_t *ctx, const dhip_csr_t *A, const double *pcdiag, const double *b, double *x,
                    int32_t where, int32_t restart, int32_t *max_iter, double *tol, int32_t rel,
                    int32_t left_pc);

/* ---- Cahn-Hilliard time stepping on the P1 interface space ----
 * CahnHilliardcGP1CL (surfactant/ifacetransp.h:2947-2990, surfactant/ifacetransp.cpp:1526-1869)
 * with theta = 1 (implicit Euler) on a resting interface: no wind (the convection and
 * mass-divergence matrices vanish) and no source terms.  Unknowns [mu; c].  Per step
 *   [ sigma dt LM(c_old) + rho dt C    M                    ] [mu]   [ M c_old                    ]
 *   [ -eps M                           eps^2 L + S M + rho eps^2 C ] [c ] = [ -M f'(c_old) + S M c_old ]
 * with LM the mobility matrix (LocalLaplaceMobilityP1CL, mobility (1-c)c clipped at 0),
 * f'(x) = (x-1) x (x-1/2) (Potential_prime_function, :3360-3364), solved by right-preconditioned
 * GMRES(restart) with the block Jacobi preconditioner built from the diagonals of the reference's
 * Precond3 = M + sigma dt LM + dt rho C and Precond4 = M + eps^2 L + eps^2 rho C (:1567-1570).
 * The normal of the mobility and stabilisation terms is the normalised P2 level-set gradient.
 * The block matrix has the layout of SparseMatBaseCL(A, B, C, D) (num/spmat.h:719-751). */
typedef struct dhip_ch dhip_ch_t;
typedef struct {
    double dt;        /* time step */
    double epsilon;   /* interface width eps */
    double rho;       /* volume stabilisation factor (Scaling * h^Power) */
    double sigma;     /* mobility scaling ("Mobility") */
    double S;         /* time-derivative stabilisation ("S = beta_s") */
    int32_t restart;  /* GMRES restart m (m - 1 Arnoldi steps per cycle) */
    int32_t max_iter;
    double tol;
    int32_t rel;      /* residual relative to |b| */
} dhip_ch_params_t;
/* Assembles the P1 interface space of the context's mesh + level set (as dhip_assemble_p1 with
 * coefficients 1, no load vector) and keeps its own copy of everything a step needs.  The state
 * starts at zero. */
int dhip_ch_create(dhip_ctx_t *ctx, int32_t zero_policy, const dhip_ch_params_t *params, dhip_ch_t **ch,
                   int64_t *ndof);
/* The same stepper on the P2 interface space (dhip_assemble_p2 with coefficients 1): P2 mass,
 * Laplace-Beltrami and stabilisation blocks, and the mobility matrix with the P2 interpolant of
 * the nodal mobilities, LM_ij = sum_q w_q (sum_d Mob(c_d) h_d)(P grad h_i . P grad h_j) on the
 * composite Quad5 rule (the P2 form of LocalLaplaceMobilityP1CL, surfactant/ifacetransp.h:1731-1766,
 * with LocalAssembler's P2 surface gradients, :1049-1053).  Every other dhip_ch_* call applies. */
int dhip_ch_create_p2(dhip_ctx_t *ctx, int32_t zero_policy, const dhip_ch_params_t *params, dhip_ch_t **ch,
                      int64_t *ndof);
void dhip_ch_destroy(dhip_ch_t *ch);
int dhip_ch_set_state(dhip_ch_t *ch, const double *c, const double *mu, int32_t where);
int dhip_ch_get_state(dhip_ch_t *ch, double *c, double *mu, int32_t where);
/* InitStep + Update: the mobility matrix, block A_, Precond3 diagonal and right-hand sides from
 * the current concentration. */
int dhip_ch_prepare(dhip_ch_t *ch);
/* Device pointers of the step system: block CSR (2n rows, 4 nnz), rhs, preconditioner diagonal,
 * and the state vector [mu; c] (valid until dhip_ch_destroy). */
int dhip_ch_system(dhip_ch_t *ch, const int32_t **rowptr, const int32_t **col, const double **val,
                   const double **rhs, const double **pcdiag, double **state, int64_t *n2, int64_t *nnz2);
/* DoStep's solve (BlockMatrixSolverCL::Solve, num/oseensolver.h:223-239): start from x0 = [mu; c]
 * (NULL: the current state, i.e. the previous step) and commit the solution as the new state. */
int dhip_ch_solve(dhip_ch_t *ch, const double *x0, int32_t where, int32_t *iters, double *resid);
/* nsteps x (prepare + solve from the previous step); iters/resid (may be NULL) get one entry per step. */
int dhip_ch_step(dhip_ch_t *ch, int32_t nsteps, int32_t *iters, double *resid);
/* GMRES precision of every solve of this stepper (default DHIP_FP64).  DHIP_FP32_IR: the
 * Arnoldi cycles of GMRES(m) store the Krylov basis and read the block matrix (scaled by the
 * preconditioner) in FP32, accumulating in FP64; each restart recomputes b - K x and updates x in
 * FP64 -- iterative refinement whose inner solve is one FP32 cycle -- and the tolerance is
 * accepted only on the recomputed FP64 residual.  The reference's GMRES (num/krylovsolver.h:
 * 320-463) is FP64 throughout; the solution meets the same FP64 residual bound. */
int dhip_ch_set_precision(dhip_ch_t *ch, int32_t precision);

/* ---- BDF1/BDF2 with the adaptive step size ----
 * The time loop of surfphasesep/surfphasesep.cpp:396-491 on this stepper's system, replacing the
 * reference's per-attempt reassembly + Belos solves.  Each attempt of step i (i = 1, 2, ...):
 *   dt = F(e, dt) on the first step or when adaptive, F from :117-123 (step_size for e == 0, else
 *        rho sqrt(tol/e) dt clamped to [min_step, max_step]);
 *   BDF1: the implicit-Euler system of step dt from c_n, solved from the solver's previous vector;
 *   i == 1: the BDF1 solution is taken (:446-450);
 *   BDF2 (r = dt/dt_prev): mobility, f' and S M c at c_e = (1+r) c_n - r c_{n-1}, effective step
 *        dt_eff = (1+r) dt / (1+2r) = 1/alpha and rhs3 = M c_r, c_r = ((1+r)^2 c_n - r^2 c_{n-1})/(1+2r)
 *        (F_chi of :455 times dt_eff), solved from the BDF1 solution;
 *   e = sqrt((d, M d) / (c2, M c2)), d = c2 - c1 with the assembled mass matrix (:477-479);
 * repeated while adaptive && e > tol && min_step < dt < max_step (:483).  The last solve's [mu; c]
 * becomes the state, c_n becomes c_{n-1}.  The system's unknowns are scaled as in dhip_ch_step
 * (the time-derivative row times dt_eff).  The reference retries without bound: max_tries > 0
 * ends a step after that many rejected attempts with DHIP_E_NOT_CONVERGED, state unchanged; a
 * GMRES solve that does not converge ends the step the same way.  dhip_ch_set_state restarts the
 * loop (step 0, t = 0, e = 0); dhip_ch_solve and dhip_ch_step, which keep the creation dt, make the
 * next BDF step a first step again (step 0, e = 0, t kept). */
typedef struct {
    double step_size;   /* Time.StepSize */
    double tol;         /* Time.Adaptive.Tol */
    double rho;         /* Time.Adaptive.rho */
    double min_step;    /* Time.Adaptive.MinStepSize (0: none) */
    double max_step;    /* Time.Adaptive.MaxStepSize (+inf: none) */
    int32_t adaptive;   /* Time.Adaptive.Use */
    int32_t max_tries;  /* 0: no bound */
} dhip_bdf_params_t;
typedef struct {
    int64_t step;       /* steps taken since the state was set (the reference's i) */
    double t, dt, e;    /* time reached, accepted step, its error (0 until the first BDF2 solve) */
    int32_t tries;      /* attempts of this step */
    int32_t iters_bdf1, iters_bdf2;  /* GMRES iterations summed over the attempts */
    int32_t reserved;
    double resid;       /* last solve's residual */
} dhip_bdf_stat_t;
/* nsteps time steps; stats (may be NULL) get one entry per step taken. */
int dhip_ch_bdf_step(dhip_ch_t *ch, const dhip_bdf_params_t *params, int32_t nsteps, dhip_bdf_stat_t *stats);

/* ---- learned warm start (no reference counterpart: an addition of this library) ----
 * A per-DOF network 12 -> 64 -> 64 -> 64 -> 2 (tanh hidden layers, bf16 weights and layer inputs,
 * float32 accumulation on the matrix cores) predicts, for the step n -> n+1, a correction of the
 * linear extrapolation of [mu; c] in units of the step scale s = (RMS of mu_n - mu_{n-1},
 * RMS of c_n - c_{n-1}):
 *     x_{n+1} ~ x_n + s * (dx_n/s + y * out_scale + out_mean),  dx_k = x_k - x_{k-1}.
 * GMRES still runs to the same tolerance.  The start is x_n or the plain extrapolation, whichever
 * has the smaller residual, replaced by the network's candidate where that has a smaller residual
 * still and the extrapolation's residual exceeds 100 x tol x |b| (the transient).  Parameter layout (float32, 9282):
 * W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[2][64] b4[2] (weights [out][in]).
 * Features per DOF i (float64, then (f - in_mean) * in_scale, stored float32 [n][12]):
 *   c_n, dc_n/s_c, dc_{n-1}/s_c, dc_{n-2}/s_c, mu_n, dmu_n/s_mu, dmu_{n-1}/s_mu, dmu_{n-2}/s_mu,
 *   (r_i/d_i)/s_mu, (r_{n+i}/d_{n+i})/s_c, d_i, d_{n+i}
 * with r = b - K x_n the residual of the new step's system at the current state and d its
 * preconditioner diagonal; missing history levels repeat the oldest one. */
#define DHIP_MLP_NPARAM 9282
#define DHIP_MLP_NF 12
typedef struct dhip_mlp dhip_mlp_t;
int dhip_mlp_create(dhip_ctx_t *ctx, dhip_mlp_t **mlp);
void dhip_mlp_destroy(dhip_mlp_t *mlp);
/* set (theta_in, host, resets the optimiser) and/or read (theta_out, host) the parameters */
int dhip_mlp_params(dhip_mlp_t *mlp, const float *theta_in, float *theta_out);
int dhip_mlp_normalization(dhip_mlp_t *mlp, const double in_mean[12], const double in_scale[12],
                           const double out_mean[2], const double out_scale[2]);
/* device pointers: hist[0..nhist) states newest first (2n), residual and pcdiag (2n), X [n][12] */
int dhip_mlp_features(dhip_mlp_t *mlp, const double *const hist[4], int32_t nhist, const double *res,
                      const double *pcdiag, int64_t n, float *X);
/* device pointers: training targets T[n][2] = ((x_next - x_n)/s - dx_n/s - out_mean) / out_scale
 * from hist[0..nhist) = x_n, x_{n-1}(, x_{n-2}) */
int dhip_mlp_targets(dhip_mlp_t *mlp, const double *x_next, const double *const hist[3], int32_t nhist,
                     int64_t n, float *T);
/* device pointers: Y[S][2] (normalised corrections) */
int dhip_mlp_predict(dhip_mlp_t *mlp, const float *X, int64_t S, float *Y);
/* minibatch of B samples idx (device) from X, T (device): gradient (host, may be NULL) and loss
 * 1/2 mean |y - t|^2; dhip_mlp_train_step applies one Adam step (beta 0.9, 0.999). */
int dhip_mlp_grad(dhip_mlp_t *mlp, const float *X, const float *T, const int32_t *idx, int64_t B,
                  double *grad, double *loss);
int dhip_mlp_train_step(dhip_mlp_t *mlp, const float *X, const float *T, const int32_t *idx, int64_t B,
                        float lr, double *loss);
/* device pointer of the state `level` steps back (0 = current) as the stepper holds them */
int dhip_ch_history(dhip_ch_t *ch, int32_t level, const double **state);
/* r = b - K x for the prepared step at the current state (device pointer, 2n) */
int dhip_ch_residual(dhip_ch_t *ch, double *r);
/* The start dhip_ch_solve_learned would use for the prepared step, into x0 (2n, host or device per
 * `where`); state and history are not modified.  *start as below. */
int dhip_ch_predict_start(dhip_ch_t *ch, dhip_mlp_t *mlp, double *x0, int32_t where, int32_t *start);
/* dhip_ch_solve from the start chosen as above (x_n on the first step after dhip_ch_set_state);
 * *start (may be NULL) = 0 x_n, 1 network, 2 linear extrapolation */
int dhip_ch_solve_learned(dhip_ch_t *ch, dhip_mlp_t *mlp, int32_t *iters, double *resid, int32_t *start);

/* ---- batched problems (configuration 4: many uneven P1 problems in one launch) ----
 * No reference counterpart: each problem is what a DROPS caller does once per level set --
 * IdxDescCL::CreateNumbering (misc/problem.cpp:725-745), accumulate() of the P1 interface mass,
 * Laplace-Beltrami and stabilisation accumulators (num/accumulator.h:265-268), LinComb, and
 * PCGSolverCL<JACPcCL>::Solve (num/krylovsolver.h:1805-1835) -- for a set of level sets.  One
 * workgroup owns one problem at a time (workgroup barriers only), problems are taken largest first
 * from a queue, and the solve is Jacobi PCG with FP32 inner iterations and FP64 residuals and
 * refinement from x = 0.  The assembled arrays of each problem equal those of dhip_assemble_p1. */
typedef struct dhip_batch dhip_batch_t;
typedef struct {
    double orig[3];
    double e[9];          /* row-major {e1, e2, e3} as dhip_mesh_kuhn */
    int32_t n[3];
    int64_t phi_offset;   /* first value of this problem's half-grid level set in the shared array */
    double lc[3];         /* the solved operator 1*(lc[0]*MASS + lc[1]*LB) + lc[2]*STAB of this problem */
} dhip_batch_problem_t;
typedef struct {
    int32_t status;       /* DHIP_OK, DHIP_E_NOT_CONVERGED, DHIP_E_ZERO_NODE, DHIP_E_ZERO_PATTERN, DHIP_E_CAPACITY */
    int32_t ncut, ndof, nnz;
    int32_t iters, outer; /* FP32 inner iterations in total, FP64 refinement steps */
    double resid;         /* final FP64 relative residual */
    double start_s, end_s;  /* device wall clock at the start and end of the problem in the last solve */
    double assembled_s;     /* device wall clock when its system was assembled */
} dhip_batch_stat_t;
/* Plans the batch: counts each problem's cut tets and cut-tet vertices on the device and sizes
 * the per-problem scratch.  phi holds all level sets (phi_count values, DHIP_HOST: copied;
 * DHIP_DEVICE: kept, must stay valid).  coef/rhs_kind/omit_bound as dhip_assemble_p1. */
int dhip_batch_create(dhip_ctx_t *ctx, int32_t nprob, const dhip_batch_problem_t *probs, const double *phi,
                      int64_t phi_count, int32_t where, int32_t zero_policy, const double coef[3],
                      int32_t rhs_kind, const double *rhs_param, double omit_bound, dhip_batch_t **batch);
/* Assembles and solves every problem in one persistent launch (relative residual tol, at most
 * max_iter inner iterations per problem).  use_graph: capture the launch into a hipGraph on the
 * first call and replay it while the arguments stay the same.  Workgroup size and residency come
 * from dhip_tune("batch_threads" 128/256/512, "batch_blocks_per_cu" 1..8); "batch_prepass" 1 runs
 * classification, compaction and local matrices of all problems over the whole grid before the
 * per-problem launch (same results; 0, the default, does them inside each problem's workgroup).
 * "batch_split" 1 ends the per-problem launch after assembly and runs the PCGs in a second
 * launch with "batch_solve_threads" (256/512/1024, default 1024) threads per problem, largest
 * first (bit-identical to the single launch when both use the same number of threads; with
 * another number the reductions add their partial sums in another order). */
int dhip_batch_solve(dhip_batch_t *batch, int32_t max_iter, double tol, int32_t rel, int32_t use_graph);
/* per-problem results of the last solve (nprob entries) */
int dhip_batch_stats(dhip_batch_t *batch, dhip_batch_stat_t *stats);
/* host copies of problem p's CSR (comb values), load vector and solution (sizes from the stats) */
int dhip_batch_problem(dhip_batch_t *batch, int32_t p, int32_t *rowptr, int32_t *colind, double *comb,
                       double *rhs, double *x);
void dhip_batch_destroy(dhip_batch_t *batch);

/* ---- on-disk formats ----
 * Matrices: DHIP_MTX_MARKET is MatrixCL::exportMTX, i.e. operator<< (num/spmat.h:1084-1089,
 * 1213-1226): "%%MatrixMarket matrix coordinate real general", "rows cols nnz", one line
 * "row+1 col+1 value" per entry in CSR order, value as printf %.15e.  DHIP_MTX_PLAIN is
 * out()/in() (num/spmat.h:1228-1255): "rows cols nnz", then row pointers, column indices and
 * values (%g, the stream's default 6 digits), each item followed by a blank, one line per array.
 * MatrixCL::exportMAT needs the MATLAB engine and has no counterpart here.
 * Vectors: VecDescBaseCL::Write / Read (misc/problem.h:928-975): binary = the count as a
 * 64-bit size_t and the raw doubles; text = the count, six blanks, every value as %.15g
 * followed by a blank, a newline.  Reading checks the stored count against n (the reference's
 * "Number of Unknowns does not match": DHIP_E_ARG).  Calls on host memory need no context
 * (ctx may be NULL; dhip_last_error(NULL) then holds the message of this thread's last failure). */
#define DHIP_MTX_MARKET 0
#define DHIP_MTX_PLAIN 1
/* host arrays; rowptr[nrows] entries.  _format writes to a caller buffer of cap bytes
 * (buf == NULL: *len = an upper bound of the size; a short buffer gives DHIP_E_CAPACITY) and
 * sets *len to the bytes written: what operator<< / out() put on a stream. */
int dhip_mtx_write(const char *path, int32_t format, int64_t nrows, int64_t ncols, const int32_t *rowptr,
                   const int32_t *colind, const double *val);
int dhip_mtx_format(int32_t format, int64_t nrows, int64_t ncols, const int32_t *rowptr, const int32_t *colind,
                    const double *val, char *buf, int64_t cap, int64_t *len);
/* host arrays.  rowptr == NULL: sizes only (*nnz = the stored count, an upper bound).  Otherwise
 * rowptr (nrows+1), colind and val (*nnz from the first call) are filled and *nnz is the final
 * count: MatrixMarket entries may come in any order and are sorted by row, then column, with
 * repeated (row, col) entries summed in file order.  _parse reads len bytes of a buffer. */
int dhip_mtx_read(const char *path, int32_t format, int64_t *nrows, int64_t *ncols, int64_t *nnz, int32_t *rowptr,
                  int32_t *colind, double *val);
int dhip_mtx_parse(const char *buf, int64_t len, int32_t format, int64_t *nrows, int64_t *ncols, int64_t *nnz,
                   int32_t *rowptr, int32_t *colind, double *val);
/* a device matrix to a file / a square matrix from a file into a new handle */
int dhip_csr_export(const dhip_csr_t *A, const char *path, int32_t format);
int dhip_csr_import(dhip_ctx_t *ctx, const char *path, int32_t format, dhip_csr_t **A);
/* vectors from / to files or caller buffers (as dhip_mtx_format / dhip_mtx_parse) */
int dhip_vec_write(dhip_ctx_t *ctx, const char *path, const double *x, int64_t n, int32_t where, int32_t binary);
int dhip_vec_format(dhip_ctx_t *ctx, const double *x, int64_t n, int32_t where, int32_t binary, char *buf,
                    int64_t cap, int64_t *len);
int dhip_vec_read(dhip_ctx_t *ctx, const char *path, double *x, int64_t n, int32_t where, int32_t binary);
int dhip_vec_parse(dhip_ctx_t *ctx, const char *buf, int64_t len, double *x, int64_t n, int32_t where,
                   int32_t binary);
/* the count stored in a vector file */
int dhip_vec_count(const char *path, int32_t binary, int64_t *n);

#ifdef __cplusplus
}
#endif
#endif /* DHIP_H */
