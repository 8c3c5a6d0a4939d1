/*
 * tracefem_oracle.h -- CPU oracle for the trace-FEM assembly + Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the DROPS
 * (56th/drops) algorithms, used by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py to check results.  No product code path
 * links, loads or calls it.
 *
 * Every function names the reference file:line it restates (paths relative
 * to the reference's src/ directory).  Parity pins (see tests/test_oracle_*):
 *   - lattice / sign / patch tables: brute force from their definitions
 *   - P1 interface numbering: tests/specifications/serial/TestinterfaceP1FE.ref
 *     (0, 4, 8, 4, 1156)
 *   - surface quadrature + patch: Testprincipallattice.ref (area 3.0909)
 *   - Quad2/Quad5 2D and Quad5 volume rules: TestQuad5_2D.ref, TestQuad5.ref
 *     (exactness on degree <= 5, P2 mass, brick integral; kept in tests/golden)
 *   - mass matrix: two independent routes (per-facet Quad5_2D vs composite
 *     Quad2), as tests/interfaceP1FE.cpp:140-230 does
 *   - Laplace-Beltrami: closed form on planar patches + the O(h^2)
 *     convergence to the continuous sphere solution
 *   - normal stabilisation: closed form for planar level sets
 *   - CSR builder / SpMV / Kahan dot: dense brute force
 *   - PCG / GMRES: dense direct solve on small systems
 *   - narrow band: band and near-interface numbering by brute force, width 0 against the
 *     interface assembly, surface forms under relabelling, planar closed form of the band
 *     stabilisation (tests/test_oracle_band_pins.py)
 * The file formats are restated separately with C++ iostreams (format_oracle.cpp).
 */
#ifndef TRACEFEM_ORACLE_H
#define TRACEFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_E_ZERO_NODE = 1,     /* std::invalid_argument, geom/subtriangulation.tpp:44 */
    ORC_E_ZERO_PATTERN = 2,  /* DROPSErrCL, geom/reftetracut.h:152-153 */
    ORC_E_ARG = 3,
    ORC_E_NOT_CONVERGED = 4,
    ORC_E_NOMEM = 5,
    ORC_E_CAPACITY = 6
};

/* zero_policy: 0 = reference behaviour (a cut tet with an exact-zero P2 node
 * raises ORC_E_ZERO_NODE, misc/problem.cpp:472 via distance()); 1 = allow
 * (historical behaviour, used to pin zero-vertex / boundary-facet handling). */
enum { ORC_ZERO_THROW = 0, ORC_ZERO_ALLOW = 1 };

/* Kuhn brick mesh: BrickBuilderCL(orig, e1, e2, e3, n1, n2, n3),
 * geom/builder.cpp:99-178.  e[k] is the k-th edge vector of the brick. */
typedef struct {
    double orig[3];
    double e[3][3];
    int32_t n[3];
} orc_mesh_t;

/* Operator selection bits for the local matrices. */
enum { ORC_OP_MASS = 1, ORC_OP_LB = 2, ORC_OP_STAB = 4 };

/* right-hand-side functions evaluated at surface quadrature points */
enum { ORC_RHS_NONE = 0, ORC_RHS_3XY_MINUS_X3 = 1, ORC_RHS_CONST = 2 };

/* ---- tables (exposed for the table tests) ---- */
/* quadrature tables: Quad5_2D nodes (barycentric on a triangle) and weights (sum 1/2), Quad5
 * volume nodes (barycentric) and the weights of node groups 1 (nodes 1-4) and 2 (nodes 5-8);
 * node 0 has 8/405, nodes 9-14 have 5/567 (num/discretize.cpp:153-296, num/discretize.tpp:754-762) */
void orc_quad_tables(double q5n[7][3], double q5w[7], double v5n[15][4], double v5c[2]);
void orc_lattice_tables(int32_t bary2[10][4], int32_t tets[8][4]);
/* 81 patterns indexed by pattern_idx+40: size, 2 triangles x 3, boundary flag */
void orc_refpatch_table(int32_t size[81], int32_t tri[81][2][3], int32_t bnd[81]);

/* ---- mesh / level set ---- */
void orc_vertex_coord(const orc_mesh_t *m, int64_t i1, int64_t i2, int64_t i3, double x[3]);
/* world coordinates of half-grid point (j1,j2,j3): vertex or edge midpoint */
void orc_halfgrid_coord(const orc_mesh_t *m, int64_t j1, int64_t j2, int64_t j3, double x[3]);

/* ---- classification: all tets whose 10 P2 values do not share one strict sign ---- */
int orc_classify(const orc_mesh_t *m, const double *phi, int zero_policy,
                 int64_t cap, int64_t *ncut, int64_t *cut_tets);

/* ---- P1 interface numbering (misc/problem.cpp:443-503) ---- */
int orc_number_p1if(const orc_mesh_t *m, const double *phi, double omit_bound, int zero_policy,
                    int32_t *dof_of_vertex, int64_t *ndof);

/* ---- per-tet local data (debug / fine-grained parity) ---- */
typedef struct {
    int32_t nfacet;
    int32_t nvert;
    double area;          /* sum over facets of absdet/2 * (0.5 on boundary facets) */
    double mass[16];
    double lb[16];
    double stab[16];
    double rhs[4];
    double hat_sq[4];     /* Quad2 composite integral of phi_i^2 (numbering test) */
} orc_local_t;

int orc_tet_local(const orc_mesh_t *m, const double *phi, int64_t tet, const double coef[3],
                  int rhs_kind, const double *rhs_param, int zero_policy, orc_local_t *out);

/* patch of one tet: facets as world/barycentric triangles (debug / geometry pins) */
typedef struct {
    int32_t nfacet;
    int32_t bnd[16];
    double world[16][3][3];
    double bary[16][3][4];
    double absdet[16];
    double normal[16][3];
    double ls[10];        /* level set on the lattice-2 vertices */
} orc_patch_t;

int orc_tet_patch(const orc_mesh_t *m, const double *phi, int64_t tet, int zero_policy, orc_patch_t *out);

/* ---- assembly (accumulate() with InterfaceCommonDataP1CL + matrix/vector
 *      accumulators, num/accumulator.h:215-228, surfactant/ifacetransp.h:533-641) ---- */
int orc_assemble_p1(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                    const double coef[3], int rhs_kind, const double *rhs_param, int zero_policy,
                    int64_t cap_nnz, int64_t *nnz, int64_t *rowptr, int32_t *colind,
                    double *val_mass, double *val_lb, double *val_stab, double *rhs);

/* ---- colour classes (ColorClassesCL::compute_color_classes, geom/multigrid.cpp:1238-1319) and the
 *      colour-parallel accumulate() of the OpenMP build (num/accumulator.h:161-188): class_start has
 *      room for 129 entries, class_tets for 6 n1 n2 n3 ---- */
int orc_color_classes(const orc_mesh_t *m, int64_t *ncolors, int64_t *class_start, int64_t *class_tets);
int orc_assemble_p1_colored(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                            const double coef[3], int rhs_kind, const double *rhs_param, int zero_policy,
                            int64_t ncolors, const int64_t *class_start, const int64_t *class_tets,
                            int64_t cap_nnz, int64_t *nnz, int64_t *rowptr, int32_t *colind,
                            double *val_mass, double *val_lb, double *val_stab, double *rhs);

/* ---- narrow band (SurfactantNarrowBandStblP1CL::Update, surfactant/ifacetransp.cpp:1110-1172):
 *      band = tets with distance() <= width (cut tets have distance 0); numbering of every vertex of
 *      every band tet (CreateNumbNearInterfaceVertex, misc/problem.cpp:509-555); mass / LB / load over
 *      the cut tets on their own pattern (_m), the normal stabilisation over the band on its own
 *      pattern (_s) ---- */
int orc_classify_band(const orc_mesh_t *m, const double *phi, double width, int zero_policy, int64_t cap,
                      int64_t *nband, int64_t *band_tets, uint8_t *is_cut);
int orc_number_p1band(const orc_mesh_t *m, const double *phi, double width, int zero_policy,
                      int32_t *dof_of_vertex, int64_t *ndof);
int orc_assemble_p1band(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                        double width, const double coef[3], int rhs_kind, const double *rhs_param, int zero_policy,
                        int64_t cap_nnz, int64_t *nnz_s, int64_t *rowptr_s, int32_t *colind_s, double *val_stab,
                        int64_t *nnz_m, int64_t *rowptr_m, int32_t *colind_m, double *val_mass, double *val_lb,
                        double *rhs);

/* surface integral of 1 over the patch (tests/principallattice.cpp:296-330) */
int orc_surface_area(const orc_mesh_t *m, const double *phi, int zero_policy, double *area);

/* ---- sparse linear algebra (num/spmat.h, num/krylovsolver.h) ---- */
void orc_spmv(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
              const double *x, double *y);
double orc_kahan_dot(int64_t n, const double *x, const double *y);
/* PCG with Jacobi (num/krylovsolver.h:107-196 + num/precond.h:155-170) */
int orc_pcg_jacobi(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
                   const double *b, double *x, int32_t *max_iter, double *tol, int32_t rel);
/* serial restarted GMRES, right/left Jacobi preconditioning (num/krylovsolver.h:320-463) */
int orc_gmres_jacobi(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
                     const double *b, double *x, int32_t restart, int32_t *max_iter, double *tol,
                     int32_t rel, int32_t left_pc, int32_t use_pc);

/* ---- P2 interface space on the planar patch ---- */
typedef struct {
    int32_t nfacet;
    double mass[100], lb[100], stab[100];   /* 10 x 10, P2 dof order: vertices, edges (0,1),(0,2),(1,2),(0,3),(1,3),(2,3) */
    double rhs[10], hat_sq[10];
} orc_local_p2_t;
/* CreateNumbOnInterfaceP2 (misc/problem.cpp:571-636); dof_of_node over the half-grid nodes */
int orc_number_p2if(const orc_mesh_t *m, const double *phi, double omit_bound, int zero_policy,
                    int32_t *dof_of_node, int64_t *ndof);
int orc_tet_local_p2(const orc_mesh_t *m, const double *phi, int64_t tet, const double coef[3], int rhs_kind,
                     const double *rhs_param, int zero_policy, orc_local_p2_t *out);
int orc_assemble_p2(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_node, int64_t ndof,
                    const double coef[3], int rhs_kind, const double *rhs_param, int zero_policy,
                    int64_t cap_nnz, int64_t *nnz, int64_t *rowptr, int32_t *colind,
                    double *val_mass, double *val_lb, double *val_stab, double *rhs);

/* ---- P2 interface space on the QuaQua-mapped surface (InterfaceCommonDataP2CL + QuaQuaMapperCL,
 *      surfactant/ifacetransp.h:200-276, levelset/levelsetmapper.h:52-140) ---- */
typedef struct {
    int32_t maxiter;     /* LevelsetMapper.Iter (20 in the reference's parameter files) */
    double tol;          /* LevelsetMapper.Tol (1e-7) */
    double armijo_c;     /* LevelsetMapper.ArmijoConstant (1e-2) */
    int32_t max_damping; /* the mapper's max_damping_steps (8) */
} orc_quaqua_t;
/* averaging_P2_gradient_recovery (num/gradient_recovery.cpp:82-89): 3 doubles per half-grid node */
int orc_p2_recovered_gradient(const orc_mesh_t *m, const double *phi, double *grad);
/* base point (barycentric in the tet, and world), quasi-distance dh and Jacobian dph (row-major) of
 * barycentric point xb of the tet (QuaQuaMapperCL::base_point_newton and ::jacobian) */
int orc_quaqua_point(const orc_mesh_t *m, const double *phi, const double *grad, int64_t tet, const double xb[4],
                     const orc_quaqua_t *P, double bxb[4], double y[3], double *dh, double dph[9], int32_t *iters);
/* W^-1 of gradient_trafo (surfactant/ifacetransp.cpp:894-918) at xb for the plane normal nl (row-major) */
int orc_quaqua_winv(const orc_mesh_t *m, const double *phi, const double *grad, int64_t tet, const double xb[4],
                    const double nl[3], const orc_quaqua_t *P, double winv[9]);
/* LocalMassP2CL, LocalLaplaceBeltramiP2CL, LocalVectorP2CL of one tet (nfacet 0: empty lattice-1 patch) */
int orc_tet_local_p2qq(const orc_mesh_t *m, const double *phi, const double *grad, int64_t tet,
                       const orc_quaqua_t *P, int rhs_kind, const double *rhs_param, int zero_policy,
                       orc_local_p2_t *out);
int orc_assemble_p2qq(const orc_mesh_t *m, const double *phi, const double *grad, const int32_t *dof_of_node,
                      int64_t ndof, const orc_quaqua_t *P, int rhs_kind, const double *rhs_param, int zero_policy,
                      int64_t cap_nnz, int64_t *nnz, int64_t *rowptr, int32_t *colind, double *val_mass,
                      double *val_lb, double *rhs);

/* GMRES as above, preconditioned by z = b / d with a given diagonal d (block Jacobi) */
int orc_gmres_diag(int64_t n, const int64_t *rowptr, const int32_t *colind, const double *val,
                   const double *pcdiag, const double *b, double *x, int32_t restart, int32_t *max_iter,
                   double *tol, int32_t rel, int32_t left_pc);

/* ---- Cahn-Hilliard step pieces (CahnHilliardcGP1CL, surfactant/ifacetransp.cpp:1526-1689) ---- */
/* LocalLaplaceMobilityP1CL on one tet, cv = concentration at its 4 vertices */
int orc_tet_mobility(const orc_mesh_t *m, const double *phi, int64_t tet, const double cv[4], int zero_policy,
                     double out[16]);
/* mobility matrix on the given pattern, M c and M f'(c) as LocalMatVecP1CL accumulations */
int orc_ch_assemble(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                    const int64_t *rowptr, const int32_t *colind, const double *c, int zero_policy,
                    double *lm, double *timeder, double *well);

/* the block system of one step (block CSR, rhs, block-Jacobi diagonal) and one full step */
int orc_ch_system(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                  const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                  const double *stab, const double *c, double dt, double eps, double rho, double sigma, double S,
                  int64_t *brp, int32_t *bcol, double *bval, double *rhs, double *pcdiag);
int orc_ch_step(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_vertex, int64_t ndof,
                const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                const double *stab, double *x, double dt, double eps, double rho, double sigma, double S,
                int32_t restart, int32_t *max_iter, double *tol, int32_t rel);

/* P2 interface space: the same pieces with dof_of_node over the half-grid, the P2 local mass and
 * the P2 mobility matrix (mobility = P2 interpolant of the nodal mobilities) */
int orc_tet_mobility_p2(const orc_mesh_t *m, const double *phi, int64_t tet, const double cv[10], int zero_policy,
                        double out[100]);
int orc_ch_assemble_p2(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_node, int64_t ndof,
                       const int64_t *rowptr, const int32_t *colind, const double *c, int zero_policy,
                       double *lm, double *timeder, double *well);
int orc_ch_system_p2(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_node, int64_t ndof,
                     const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                     const double *stab, const double *c, double dt, double eps, double rho, double sigma,
                     double S, int64_t *brp, int32_t *bcol, double *bval, double *rhs, double *pcdiag);
int orc_ch_step_p2(const orc_mesh_t *m, const double *phi, const int32_t *dof_of_node, int64_t ndof,
                   const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                   const double *stab, double *x, double dt, double eps, double rho, double sigma, double S,
                   int32_t restart, int32_t *max_iter, double *tol, int32_t rel);

/* ---- BDF1/BDF2 with the adaptive step size (surfphasesep/surfphasesep.cpp:112-123, 396-491) ---- */
typedef struct {
    double step_size;  /* Time.StepSize: F(0, dt) */
    double tol;        /* Time.Adaptive.Tol */
    double rho;        /* Time.Adaptive.rho */
    double min_step;   /* Time.Adaptive.MinStepSize (0: none) */
    double max_step;   /* Time.Adaptive.MaxStepSize (inf: none) */
    int32_t adaptive;  /* Time.Adaptive.Use */
    int32_t max_tries; /* attempts per step before ORC_E_NOT_CONVERGED (0: no bound, as the reference) */
} orc_bdf_params_t;
typedef struct {
    int64_t step;      /* steps taken (the reference's i) */
    double t, dt, e;   /* time reached, last accepted step, last error (0 until a BDF2 step) */
} orc_bdf_state_t;
typedef struct {
    int32_t tries, iters_bdf1, iters_bdf2;  /* attempts, GMRES iterations summed over them */
    double resid, dt, e;                      /* last solve's residual, last attempt's dt and e */
} orc_bdf_stat_t;

void orc_bdf2_weights(double dt, double dt_prev, double w[5]);
double orc_bdf_step_size(const orc_bdf_params_t *p, double e, double dt);
int orc_ch_system_bdf(int space, const orc_mesh_t *m, const double *phi, const int32_t *dof_of, int64_t ndof,
                      const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                      const double *stab, const double *c_e, const double *c_r, double dt_eff, double eps,
                      double rho, double sigma, double S, int64_t *brp, int32_t *bcol, double *bval, double *rhs,
                      double *pcdiag);
int orc_ch_bdf_step(int space, const orc_mesh_t *m, const double *phi, const int32_t *dof_of, int64_t ndof,
                    const int64_t *rowptr, const int32_t *colind, const double *mass, const double *lb,
                    const double *stab, double *x, double *cprev, orc_bdf_state_t *st, const orc_bdf_params_t *bp,
                    double eps, double rho, double sigma, double S, int32_t restart, int32_t max_iter, double tol,
                    int32_t rel, orc_bdf_stat_t *stat);

#ifdef __cplusplus
}
#endif
#endif
