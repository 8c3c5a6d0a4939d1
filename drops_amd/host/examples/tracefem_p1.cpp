// tracefem_p1.cpp -- a DROPS-style program for the P1 trace-FEM path written against
// the host mirror (dropship.hpp), in the shape of SetupLBP1 / the interfaceP1FE test:
// Kuhn brick, P2 level set, interface numbering, accumulator tuple, LinComb, PCG.
//
//   tracefem_p1 sphere20                    -> the 1156-DOF known answer (TestinterfaceP1FE.ref)
//   tracefem_p1 c1 <n> <outdir>             -> configuration-1 problem, writes CSR/rhs/solution, and the
//                                              operator / load vector in the reference's file formats
//   tracefem_p1 band <n> <width> <outdir>   -> the narrow-band tuple of SurfactantNarrowBandStblP1CL::Update
//   tracefem_p1 p2 <n> <outdir>             -> the same sphere on the P2 interface space (IdxDescCL(P2IF_FE),
//                                              P2 accumulators), writes CSR/rhs/solution
//   tracefem_p1 accus <n> <p1|p2>           -> user-written accumulators (cut-tet count, patch area, visit order)
//                                              in one tuple with the device-assembled interface accumulators
//   tracefem_p1 colors <n>                  -> ColorClassesCL of the n^3 brick, checked (host only)
//   tracefem_p1 ch <n> <steps> <outdir>     -> CahnHilliardcGP1CL on the torus: SetInitialValue, DoStep x steps;
//                                              writes the initial and final c, mu and the iterations
//   tracefem_p1 block <n> <outdir>          -> BlockMatrixSolverCL<GMResSolverCL<DiagBlockJacobiPcCL>> on the
//                                              CH blocks from the accumulators, transp_mul and transpose
//   tracefem_p1 distribute <n> <dist> <outdir> -> IdxDescCL::DistributeDOFs on the torus: marked tets and the
//                                              grid vertex of every DOF
//   tracefem_p1 userrhs <n> <p1|p2>         -> the load vector of a host function f(x, t) (LocalVectorP1CL /
//                                              LocalVectorP2CL with a std::function) against the device kind
#include <cstdio>
#include <fstream>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../dropship.hpp"

using namespace DROPS;

static void write_bin(const std::string &path, const void *p, size_t bytes) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw DROPSErrCL("cannot write " + path);
    std::fwrite(p, 1, bytes, f);
    std::fclose(f);
}

static int sphere20() {
    const double o[3] = {0, 0, 0}, e1[3] = {1, 0, 0}, e2[3] = {0, 1, 0}, e3[3] = {0, 0, 1};
    MultiGridCL mg(BrickBuilderCL(o, e1, e2, e3, 20, 20, 20));
    LevelsetP2CL lset;
    lset.Init(mg, [](const double *p, double) {
        const double x = p[0] - 0.5, y = p[1] - 0.5, z = p[2] - 0.5;
        return std::sqrt(x * x + y * y + z * z) - 0.31;
    });
    IdxDescCL ifaceidx(P1IF_FE);
    ifaceidx.CreateNumbering(0, mg, &lset.Phi);
    std::printf("(sphere) NumUnknowns: %zu\n", ifaceidx.NumUnknowns());
    return 0;
}

static int c1(int n, const std::string &out) {
    // [-2,2]^3 moved by -s, s = 1e-3 h (1,2,3)/sqrt(14) (mesh transform p -> p - s)
    const double h = 4.0 / n, sn = 1e-3 * h / std::sqrt(14.0);
    const double o[3] = {-2.0 - sn, -2.0 - 2 * sn, -2.0 - 3 * sn};
    const double e1[3] = {4, 0, 0}, e2[3] = {0, 4, 0}, e3[3] = {0, 0, 4};
    MultiGridCL mg(BrickBuilderCL(o, e1, e2, e3, n, n, n));
    LevelsetP2CL lset;
    lset.Init(mg, [](const double *p, double) { return (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 1.0; });
    BndDataCL lsbnd;

    IdxDescCL ifaceidx(P1IF_FE);
    ifaceidx.CreateNumbering(0, mg, &lset.Phi, &lsbnd);
    std::printf("NumUnknowns: %zu\n", ifaceidx.NumUnknowns());

    MatDescCL M(&ifaceidx, &ifaceidx), A(&ifaceidx, &ifaceidx), C(&ifaceidx, &ifaceidx);
    VecDescCL b(&ifaceidx);
    TetraAccumulatorTupleCL accus;
    InterfaceCommonDataP1CL cdata(lset.Phi, lsbnd);
    accus.push_back(&cdata);
    InterfaceMatrixAccuCL<LocalInterfaceMassP1CL, InterfaceCommonDataP1CL> mass(&M, LocalInterfaceMassP1CL(), cdata, "mass");
    InterfaceMatrixAccuCL<LocalLaplaceBeltramiP1CL, InterfaceCommonDataP1CL> lb(&A, LocalLaplaceBeltramiP1CL(1.), cdata, "LB");
    InterfaceMatrixAccuCL<LocalNormalLaplaceBulkP1CL, InterfaceCommonDataP1CL> stab(&C, LocalNormalLaplaceBulkP1CL(1.), cdata, "stab");
    DeviceFunctionCL f;
    f.kind = DHIP_RHS_3XY_MINUS_X3;
    InterfaceVectorAccuCL<LocalVectorP1CL, InterfaceCommonDataP1CL> rhs(&b, LocalVectorP1CL(f), cdata, "rhs");
    accus.push_back(&mass);
    accus.push_back(&lb);
    accus.push_back(&stab);
    accus.push_back(&rhs);
    accumulate(accus, mg, 0, &lsbnd);
    std::printf("nnz: %zu\n", M.Data.num_nonzeros());

    MatrixCL L;
    L.LinComb(1.0, M.Data, 1.0, A.Data, h, C.Data);
    VectorCL x(L.num_rows());
    JACPcCL pc;
    PCGSolverCL<JACPcCL> solver(pc, 10000, 1e-10, true);
    solver.Solve(L, x, b.Data, DummyExchangeCL());
    std::printf("PCG iterations: %d\nPCG residual: %.6e\n", solver.GetIter(), solver.GetResid());
    // true residual on the device
    const VectorCL Lx = L * x;
    VectorCL r(b.Data);
    r -= Lx;
    std::printf("true relative residual: %.6e\n", norm(r) / norm(b.Data));

    if (!out.empty()) {
        std::vector<int64_t> rp(L.num_rows() + 1), ci(L.num_nonzeros());
        for (size_t i = 0; i <= L.num_rows(); ++i) rp[i] = (int64_t)L.raw_row()[i];
        for (size_t k = 0; k < L.num_nonzeros(); ++k) ci[k] = (int64_t)L.raw_col()[k];
        write_bin(out + "/rowptr.bin", rp.data(), rp.size() * 8);
        write_bin(out + "/colind.bin", ci.data(), ci.size() * 8);
        write_bin(out + "/mass.bin", M.Data.raw_val(), M.Data.num_nonzeros() * 8);
        write_bin(out + "/lb.bin", A.Data.raw_val(), A.Data.num_nonzeros() * 8);
        write_bin(out + "/stab.bin", C.Data.raw_val(), C.Data.num_nonzeros() * 8);
        write_bin(out + "/comb.bin", L.raw_val(), L.num_nonzeros() * 8);
        write_bin(out + "/rhs.bin", b.Data.raw(), b.Data.size() * 8);
        write_bin(out + "/x.bin", x.raw(), x.size() * 8);
        write_bin(out + "/vertex_of_dof.bin", ifaceidx.vertex_of_dof().data(), ifaceidx.vertex_of_dof().size() * 8);
        // the reference's file formats: exportMTX, out(), WriteFEToFile (binary and text)
        L.exportMTX(out + "/L.mtx");
        {
            std::ofstream os(out + "/L.txt");
            DROPS::out(os, L);
        }
        WriteFEToFile(b, mg, out + "/rhs.dat", true);
        WriteFEToFile(b, mg, out + "/rhs.txt", false);
        VecDescCL back(&ifaceidx);
        ReadFEFromFile(back, mg, out + "/rhs.dat", true);
        bool same = back.Data.size() == b.Data.size();
        for (size_t i = 0; same && i < b.Data.size(); ++i) same = back.Data[i] == b.Data[i];
        MatrixCL L2;
        {
            std::ifstream is(out + "/L.txt");
            DROPS::in(is, L2);
        }
        same = same && L2.num_rows() == L.num_rows() && L2.num_nonzeros() == L.num_nonzeros();
        for (size_t k = 0; same && k < L.num_nonzeros(); ++k) same = L2.col_ind(k) == L.col_ind(k);
        std::printf("formats round trip: %s\n", same ? "ok" : "MISMATCH");
    }
    return 0;
}

// SurfactantNarrowBandStblP1CL::Update's tuple (surfactant/ifacetransp.cpp:1128-1166): interface
// mass / LB / load with InterfaceCommonDataP1CL, the stabilisation with the narrow band, on the
// near-interface numbering
static int band(int n, double width, const std::string &out) {
    const double h = 4.0 / n, sn = 1e-3 * h / std::sqrt(14.0);
    const double o[3] = {-2.0 - sn, -2.0 - 2 * sn, -2.0 - 3 * sn};
    const double e1[3] = {4, 0, 0}, e2[3] = {0, 4, 0}, e3[3] = {0, 0, 4};
    MultiGridCL mg(BrickBuilderCL(o, e1, e2, e3, n, n, n));
    LevelsetP2CL lset;
    lset.Init(mg, [](const double *p, double) { return (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 1.0; });
    BndDataCL lsbnd;
    IdxDescCL idx(P1IF_FE);
    idx.CreateNumbering(0, mg, &lset.Phi, &lsbnd, width);
    std::printf("NumUnknowns: %zu\n", idx.NumUnknowns());
    MatDescCL M(&idx, &idx), A(&idx, &idx), C(&idx, &idx);
    VecDescCL b(&idx);
    TetraAccumulatorTupleCL accus;
    InterfaceCommonDataP1CL cdata(lset.Phi, lsbnd);
    accus.push_back(&cdata);
    InterfaceMatrixAccuCL<LocalInterfaceMassP1CL, InterfaceCommonDataP1CL> mass(&M, LocalInterfaceMassP1CL(), cdata, "mass");
    accus.push_back(&mass);
    InterfaceMatrixAccuCL<LocalLaplaceBeltramiP1CL, InterfaceCommonDataP1CL> lb(&A, LocalLaplaceBeltramiP1CL(1.), cdata, "LB");
    accus.push_back(&lb);
    NarrowBandCommonDataP1CL bdata(lset.Phi, lsbnd, width);
    accus.push_back(&bdata);
    NarrowBandMatrixAccuP1CL<LocalNormalLaplaceBulkP1CL> sb(&C, LocalNormalLaplaceBulkP1CL(1.), bdata, "NormalLaplaceBulk");
    accus.push_back(&sb);
    DeviceFunctionCL f;
    f.kind = DHIP_RHS_3XY_MINUS_X3;
    InterfaceVectorAccuCL<LocalVectorP1CL, InterfaceCommonDataP1CL> rhs(&b, LocalVectorP1CL(f), cdata, "load");
    accus.push_back(&rhs);
    accumulate(accus, mg, 0, &lsbnd);
    std::printf("nnz: %zu\n", C.Data.num_nonzeros());
    MatrixCL L;
    L.LinComb(1.0, M.Data, 1.0, A.Data, h, C.Data);
    if (!out.empty()) {
        std::vector<int64_t> rp(L.num_rows() + 1), ci(L.num_nonzeros());
        for (size_t i = 0; i <= L.num_rows(); ++i) rp[i] = (int64_t)L.raw_row()[i];
        for (size_t k = 0; k < L.num_nonzeros(); ++k) ci[k] = (int64_t)L.raw_col()[k];
        write_bin(out + "/rowptr.bin", rp.data(), rp.size() * 8);
        write_bin(out + "/colind.bin", ci.data(), ci.size() * 8);
        write_bin(out + "/mass.bin", M.Data.raw_val(), M.Data.num_nonzeros() * 8);
        write_bin(out + "/lb.bin", A.Data.raw_val(), A.Data.num_nonzeros() * 8);
        write_bin(out + "/stab.bin", C.Data.raw_val(), C.Data.num_nonzeros() * 8);
        write_bin(out + "/comb.bin", L.raw_val(), L.num_nonzeros() * 8);
        write_bin(out + "/rhs.bin", b.Data.raw(), b.Data.size() * 8);
        write_bin(out + "/vertex_of_dof.bin", idx.vertex_of_dof().data(), idx.vertex_of_dof().size() * 8);
    }
    return 0;
}

// surfactant/surfactant.cpp:1590-1640: P2 mass, Laplace-Beltrami and load on the QuaQua-mapped sphere
static int quaqua(int n, const std::string &out) {
    const double h = 4.0 / n, sn = 1e-3 * h / std::sqrt(14.0);
    const double o[3] = {-2.0 - sn, -2.0 - 2 * sn, -2.0 - 3 * sn};
    const double e1[3] = {4, 0, 0}, e2[3] = {0, 4, 0}, e3[3] = {0, 0, 4};
    MultiGridCL mg(BrickBuilderCL(o, e1, e2, e3, n, n, n));
    LevelsetP2CL lset;
    lset.Init(mg, [](const double *p, double) { return (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 1.0; });
    BndDataCL lsbnd;
    IdxDescCL ifacep2idx(P2IF_FE);
    ifacep2idx.CreateNumbering(0, mg, &lset.Phi, &lsbnd);
    std::printf("P2-NumUnknowns: %zu\n", ifacep2idx.NumUnknowns());
    // Recover the gradient of the level set function
    IdxDescCL vecp2idx(vecP2_FE);
    vecp2idx.CreateNumbering(0, mg);
    VecDescCL lsgradrec(&vecp2idx);
    averaging_P2_gradient_recovery(mg, lset.Phi, lsbnd, lsgradrec);
    // Compute neighborhoods of the tetras at the interface
    const PrincipalLatticeCL &lat = PrincipalLatticeCL::instance(1);
    TetraToTetrasT tetra_neighborhoods;
    compute_tetra_neighborhoods(mg, lset.Phi, lsbnd, lat, tetra_neighborhoods);
    QuaQuaMapperCL quaqua(mg, lset.Phi, lsgradrec, &tetra_neighborhoods, 20, 1e-7, false, 1e-2);
    TetraAccumulatorTupleCL accus;
    InterfaceCommonDataP2CL cdatap2(lset.Phi, lsbnd, quaqua, lat);
    accus.push_back(&cdatap2);
    MatDescCL Mp2(&ifacep2idx, &ifacep2idx), Ap2(&ifacep2idx, &ifacep2idx);
    InterfaceMatrixAccuCL<LocalMassP2CL, InterfaceCommonDataP2CL> accuMp2(&Mp2, LocalMassP2CL(), cdatap2, "Mp2");
    accus.push_back(&accuMp2);
    InterfaceMatrixAccuCL<LocalLaplaceBeltramiP2CL, InterfaceCommonDataP2CL> accuAp2(&Ap2, LocalLaplaceBeltramiP2CL(1.),
                                                                                    cdatap2, "Ap2");
    accus.push_back(&accuAp2);
    VecDescCL bp2(&ifacep2idx);
    DeviceFunctionCL f;
    f.kind = DHIP_RHS_3XY_MINUS_X3;
    InterfaceVectorAccuCL<LocalVectorP2CL, InterfaceCommonDataP2CL> acculoadp2(&bp2, LocalVectorP2CL(f), cdatap2);
    accus.push_back(&acculoadp2);
    accumulate(accus, mg, 0, &lsbnd);
    double area = 0.;
    for (size_t k = 0; k < Mp2.Data.num_nonzeros(); ++k) area += Mp2.Data.raw_val()[k];
    std::printf("nnz: %zu\nmapped area - 4 pi: %.6e\n", Mp2.Data.num_nonzeros(), area - 4.0 * std::acos(-1.0));
    // the same load from a host function, evaluated at the base points the library lists
    VecDescCL bu(&ifacep2idx);
    TetraAccumulatorTupleCL accu;
    accu.push_back(&cdatap2);
    InstatScalarFunction fu = [](const double *x, double) { return 3.0 * x[0] * x[1] - x[0] * x[0] * x[0]; };
    InterfaceVectorAccuCL<LocalVectorP2CL, InterfaceCommonDataP2CL> accuu(&bu, LocalVectorP2CL(fu, 0.), cdatap2);
    accu.push_back(&accuu);
    accumulate(accu, mg, 0, &lsbnd);
    double du = 0.;
    for (size_t i = 0; i < bu.Data.size(); ++i) du = std::max(du, std::fabs(bu.Data[i] - bp2.Data[i]));
    std::printf("user rhs vs device rhs: %.3e\n", du);
    if (!out.empty()) {
        std::vector<int64_t> rp(Mp2.Data.num_rows() + 1), ci(Mp2.Data.num_nonzeros());
        for (size_t i = 0; i <= Mp2.Data.num_rows(); ++i) rp[i] = (int64_t)Mp2.Data.raw_row()[i];
        for (size_t k = 0; k < Mp2.Data.num_nonzeros(); ++k) ci[k] = (int64_t)Mp2.Data.raw_col()[k];
        write_bin(out + "/rowptr.bin", rp.data(), rp.size() * 8);
        write_bin(out + "/colind.bin", ci.data(), ci.size() * 8);
        write_bin(out + "/mass.bin", Mp2.Data.raw_val(), Mp2.Data.num_nonzeros() * 8);
        write_bin(out + "/lb.bin", Ap2.Data.raw_val(), Ap2.Data.num_nonzeros() * 8);
        write_bin(out + "/rhs.bin", bp2.Data.raw(), bp2.Data.size() * 8);
        write_bin(out + "/grad.bin", lsgradrec.Data.raw(), lsgradrec.Data.size() * 8);
    }
    return 0;
}

static int p2(int n, const std::string &out) {
    const double h = 4.0 / n, sn = 1e-3 * h / std::sqrt(14.0);
    const double o[3] = {-2.0 - sn, -2.0 - 2 * sn, -2.0 - 3 * sn};
    const double e1[3] = {4, 0, 0}, e2[3] = {0, 4, 0}, e3[3] = {0, 0, 4};
    MultiGridCL mg(BrickBuilderCL(o, e1, e2, e3, n, n, n));
    LevelsetP2CL lset;
    lset.Init(mg, [](const double *p, double) { return (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 1.0; });
    BndDataCL lsbnd;
    IdxDescCL ifaceidx(P2IF_FE);
    ifaceidx.CreateNumbering(0, mg, &lset.Phi, &lsbnd);
    std::printf("NumUnknowns: %zu\n", ifaceidx.NumUnknowns());
    MatDescCL M(&ifaceidx, &ifaceidx), A(&ifaceidx, &ifaceidx), C(&ifaceidx, &ifaceidx);
    VecDescCL b(&ifaceidx);
    TetraAccumulatorTupleCL accus;
    InterfaceCommonDataP2CL cdata(lset.Phi, lsbnd);
    accus.push_back(&cdata);
    InterfaceMatrixAccuCL<LocalMassP2CL, InterfaceCommonDataP2CL> mass(&M, LocalMassP2CL(), cdata, "mass");
    InterfaceMatrixAccuCL<LocalLaplaceBeltramiP2CL, InterfaceCommonDataP2CL> lb(&A, LocalLaplaceBeltramiP2CL(1.), cdata, "LB");
    InterfaceMatrixAccuCL<LocalNormalGradientStabP2CL, InterfaceCommonDataP2CL> stab(&C, LocalNormalGradientStabP2CL(1.), cdata, "stab");
    DeviceFunctionCL f;
    f.kind = DHIP_RHS_3XY_MINUS_X3;
    InterfaceVectorAccuCL<LocalVectorP2CL, InterfaceCommonDataP2CL> rhs(&b, LocalVectorP2CL(f), cdata, "rhs");
    accus.push_back(&mass);
    accus.push_back(&lb);
    accus.push_back(&stab);
    accus.push_back(&rhs);
    accumulate(accus, mg, 0, &lsbnd);
    std::printf("nnz: %zu\n", M.Data.num_nonzeros());
    MatrixCL L;
    L.LinComb(1.0, M.Data, 1.0, A.Data, h, C.Data);
    VectorCL x(L.num_rows());
    JACPcCL pc;
    PCGSolverCL<JACPcCL> solver(pc, 20000, 1e-10, true);
    solver.Solve(L, x, b.Data, DummyExchangeCL());
    std::printf("PCG iterations: %d\nPCG residual: %.6e\n", solver.GetIter(), solver.GetResid());
    if (!out.empty()) {
        std::vector<int64_t> rp(L.num_rows() + 1), ci(L.num_nonzeros());
        for (size_t i = 0; i <= L.num_rows(); ++i) rp[i] = (int64_t)L.raw_row()[i];
        for (size_t k = 0; k < L.num_nonzeros(); ++k) ci[k] = (int64_t)L.raw_col()[k];
        write_bin(out + "/rowptr.bin", rp.data(), rp.size() * 8);
        write_bin(out + "/colind.bin", ci.data(), ci.size() * 8);
        write_bin(out + "/mass.bin", M.Data.raw_val(), M.Data.num_nonzeros() * 8);
        write_bin(out + "/lb.bin", A.Data.raw_val(), A.Data.num_nonzeros() * 8);
        write_bin(out + "/stab.bin", C.Data.raw_val(), C.Data.num_nonzeros() * 8);
        write_bin(out + "/rhs.bin", b.Data.raw(), b.Data.size() * 8);
        write_bin(out + "/x.bin", x.raw(), x.size() * 8);
        write_bin(out + "/node_of_dof.bin", ifaceidx.node_of_dof().data(), ifaceidx.node_of_dof().size() * 8);
    }
    return 0;
}

// torus R = 1, r = 0.3 on [-2,2]^3 moved by -s, s = 1e-3 h (1,2,3)/sqrt(14), as the c1 problem
static void torus_setup(int n, MultiGridCL *&mg, LevelsetP2CL &lset) {
    const double h = 4.0 / n, sn = 1e-3 * h / std::sqrt(14.0);
    const double o[3] = {-2.0 - sn, -2.0 - 2 * sn, -2.0 - 3 * sn};
    const double e1[3] = {4, 0, 0}, e2[3] = {0, 4, 0}, e3[3] = {0, 0, 4};
    mg = new MultiGridCL(BrickBuilderCL(o, e1, e2, e3, n, n, n));
    lset.Init(*mg, [](const double *p, double) {
        const double q = std::sqrt(p[0] * p[0] + p[1] * p[1]) - 1.0;
        return (q * q + p[2] * p[2]) - 0.3 * 0.3;
    });
}

// the smooth initial concentration of the CH examples
static double ch_c0(const double *p, double) { return 0.5 + 0.05 * std::sin(3.0 * p[0]) * std::cos(2.0 * p[1] + p[2]); }

static int ch(int n, int steps, const std::string &out) {
    MultiGridCL *mg = nullptr;
    LevelsetP2CL lset;
    torus_setup(n, mg, lset);
    BndDataCL bnd;
    const double h = 4.0 / n, dt = 1e-4, width = 0.;
    CahnHilliardcGP1CL ch(*mg, 1.0, 1.0, 0.05, nullptr, bnd, lset.Phi, bnd, nullptr, width, h, 1.0, 20000, 1e-13);
    ch.SetRestart(50);
    ch.CreateNumbering(&bnd);
    ch.SetInitialValue([](const double *, double) { return 0.0; }, ch_c0);
    std::printf("NumUnknowns: %zu\n", ch.idx.NumUnknowns());
    write_bin(out + "/c0.bin", ch.ic.Data.raw(), ch.ic.Data.size() * 8);
    std::vector<int32_t> its;
    double t = 0.;
    for (int s = 0; s < steps; ++s) {
        ch.InitTimeStep();
        t += dt;
        ch.DoStep(t);
        its.push_back(ch.GetSolver().GetIter());
        std::printf("step %d: iter %d resid %.3e\n", s + 1, ch.GetSolver().GetIter(), ch.GetSolver().GetResid());
    }
    write_bin(out + "/c.bin", ch.ic.Data.raw(), ch.ic.Data.size() * 8);
    write_bin(out + "/mu.bin", ch.imu.Data.raw(), ch.imu.Data.size() * 8);
    write_bin(out + "/iters.bin", its.data(), its.size() * 4);
    write_bin(out + "/vertex_of_dof.bin", ch.idx.vertex_of_dof().data(), ch.idx.vertex_of_dof().size() * 8);
    delete mg;
    return 0;
}

// The blocks of CahnHilliardcGP1CL::DoStep with the mobility matrix replaced by the Laplace-Beltrami
// matrix (constant mobility 1), assembled by the accumulators, solved by BlockMatrixSolverCL.
static int block(int n, const std::string &out) {
    MultiGridCL *mg = nullptr;
    LevelsetP2CL lset;
    torus_setup(n, mg, lset);
    BndDataCL bnd;
    IdxDescCL idx(P1IF_FE);
    idx.CreateNumbering(0, *mg, &lset.Phi, &bnd);
    MatDescCL M(&idx, &idx), A(&idx, &idx), C(&idx, &idx);
    TetraAccumulatorTupleCL accus;
    InterfaceCommonDataP1CL cdata(lset.Phi, bnd);
    accus.push_back(&cdata);
    InterfaceMatrixAccuCL<LocalInterfaceMassP1CL, InterfaceCommonDataP1CL> mass(&M, LocalInterfaceMassP1CL(), cdata, "mass");
    InterfaceMatrixAccuCL<LocalLaplaceBeltramiP1CL, InterfaceCommonDataP1CL> lb(&A, LocalLaplaceBeltramiP1CL(1.), cdata, "LB");
    InterfaceMatrixAccuCL<LocalNormalLaplaceBulkP1CL, InterfaceCommonDataP1CL> st(&C, LocalNormalLaplaceBulkP1CL(1.), cdata, "stab");
    accus.push_back(&mass);
    accus.push_back(&lb);
    accus.push_back(&st);
    accumulate(accus, *mg, 0, &bnd);
    const size_t nd = idx.NumUnknowns();
    const double dt = 1e-4, eps = 0.05, rho = 4.0 / n, sigma = 1.0, S = 1.0;
    MatrixCL A_, B_, C_, D_;
    A_.LinComb(sigma * dt, A.Data, rho * dt, C.Data);
    B_.LinComb(1.0, M.Data, 0.0, M.Data);
    C_.LinComb(0., A.Data, -eps, M.Data);
    D_.LinComb(eps * eps, A.Data, S, M.Data, rho * eps * eps, C.Data);
    VectorCL c0(nd), mu(nd), c(nd);
    for (size_t i = 0; i < nd; ++i) c0[i] = 0.5 + 0.05 * std::sin(0.37 * (double)i);
    c = c0;
    const VectorCL rhs3 = M.Data * c0;
    VectorCL rhs4(nd);
    for (size_t i = 0; i < nd; ++i) rhs4[i] = S * rhs3[i];
    // the block-Jacobi diagonal: Precond3 = M + sigma dt A + dt rho C, Precond4 = M + eps^2 A + eps^2 rho C
    MatrixCL P3, P4;
    P3.LinComb(1.0, M.Data, sigma * dt, A.Data, dt * rho, C.Data);
    P4.LinComb(1.0, M.Data, eps * eps, A.Data, eps * eps * rho, C.Data);
    VectorCL d(2 * nd);
    for (size_t i = 0; i < nd; ++i)
        for (size_t k = P3.row_beg(i); k < P3.row_beg(i + 1); ++k)
            if (P3.col_ind(k) == i) {
                d[i] = P3.raw_val()[k];
                d[nd + i] = P4.raw_val()[k];
            }
    DiagBlockJacobiPcCL pc(d);
    GMResSolverCL<DiagBlockJacobiPcCL> gm(pc, 50, 20000, 1e-10, true, false, RightPreconditioning);
    BlockMatrixSolverCL<GMResSolverCL<DiagBlockJacobiPcCL>> solver(gm);
    solver.Solve(A_, B_, C_, D_, mu, c, rhs3, rhs4, DummyExchangeCL(), DummyExchangeCL());
    std::printf("block GMRES iterations: %d\nblock GMRES residual: %.6e\n", solver.GetIter(), solver.GetResid());
    // y_ATx on the device and transpose on the host: (C_^T x) against (transpose(C_)) x
    const VectorCL yt = transp_mul(C_, c);
    MatrixCL CT;
    transpose(C_, CT);
    const VectorCL yt2 = CT * c;
    double dmax = 0.;
    for (size_t i = 0; i < nd; ++i) dmax = std::max(dmax, std::fabs(yt[i] - yt2[i]));
    std::printf("transp_mul vs transpose: %.3e\n", dmax);
    auto dump = [&](const std::string &name, const MatrixCL &X) {
        std::vector<int64_t> rp(X.num_rows() + 1), ci(X.num_nonzeros());
        for (size_t i = 0; i <= X.num_rows(); ++i) rp[i] = (int64_t)X.raw_row()[i];
        for (size_t k = 0; k < X.num_nonzeros(); ++k) ci[k] = (int64_t)X.raw_col()[k];
        write_bin(out + "/" + name + "_rowptr.bin", rp.data(), rp.size() * 8);
        write_bin(out + "/" + name + "_colind.bin", ci.data(), ci.size() * 8);
        write_bin(out + "/" + name + "_val.bin", X.raw_val(), X.num_nonzeros() * 8);
    };
    dump("A", A_);
    dump("B", B_);
    dump("C", C_);
    dump("D", D_);
    write_bin(out + "/rhs3.bin", rhs3.raw(), nd * 8);
    write_bin(out + "/rhs4.bin", rhs4.raw(), nd * 8);
    write_bin(out + "/mu.bin", mu.raw(), nd * 8);
    write_bin(out + "/c.bin", c.raw(), nd * 8);
    write_bin(out + "/yt.bin", yt.raw(), nd * 8);
    delete mg;
    return 0;
}

static int distribute(int n, double dist, const std::string &out) {
    MultiGridCL *mg = nullptr;
    LevelsetP2CL lset;
    torus_setup(n, mg, lset);
    IdxDescCL idx(P1IF_FE);
    const size_t ntet = idx.DistributeDOFs(0, *mg, &lset.Phi, dist);
    std::printf("marked tets: %zu\nNumUnknowns: %zu\n", ntet, idx.NumUnknowns());
    write_bin(out + "/vertex_of_dof.bin", idx.vertex_of_dof().data(), idx.vertex_of_dof().size() * 8);
    delete mg;
    return 0;
}

// The configuration-1 sphere's load vector of f = 3xy - x^3 given as a host function and as the
// device-evaluated kind; prints the largest difference (0 when the points and the sums agree).
template <class CommonT, class LocalT>
static double load_vector(MultiGridCL &mg, LevelsetP2CL &lset, IdxDescCL &idx, const LocalT &local) {
    BndDataCL bnd;
    VecDescCL b(&idx);
    TetraAccumulatorTupleCL accus;
    CommonT cdata(lset.Phi, bnd);
    accus.push_back(&cdata);
    InterfaceVectorAccuCL<LocalT, CommonT> rhs(&b, local, cdata, "load");
    accus.push_back(&rhs);
    accumulate(accus, mg, 0, &bnd);
    double s = 0.;
    for (size_t i = 0; i < b.Data.size(); ++i) s += std::fabs(b.Data[i]);
    static VectorCL last;
    double d = 0.;
    if (last.size() == b.Data.size())
        for (size_t i = 0; i < b.Data.size(); ++i) d = std::max(d, std::fabs(last[i] - b.Data[i]));
    last.resize(b.Data.size());
    last = b.Data;
    std::printf("sum |b| = %.17g\n", s);
    return d;
}

static int userrhs(int n, const std::string &space) {
    const double h = 4.0 / n, sn = 1e-3 * h / std::sqrt(14.0);
    const double o[3] = {-2.0 - sn, -2.0 - 2 * sn, -2.0 - 3 * sn};
    const double e1[3] = {4, 0, 0}, e2[3] = {0, 4, 0}, e3[3] = {0, 0, 4};
    MultiGridCL mg(BrickBuilderCL(o, e1, e2, e3, n, n, n));
    LevelsetP2CL lset;
    lset.Init(mg, [](const double *p, double) { return (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 1.0; });
    BndDataCL bnd;
    const bool p2 = space == "p2";
    IdxDescCL idx(p2 ? P2IF_FE : P1IF_FE);
    idx.CreateNumbering(0, mg, &lset.Phi, &bnd);
    DeviceFunctionCL kind;
    kind.kind = DHIP_RHS_3XY_MINUS_X3;
    InstatScalarFunction f = [](const double *x, double) { return 3.0 * x[0] * x[1] - x[0] * x[0] * x[0]; };
    double d;
    if (p2) {
        load_vector<InterfaceCommonDataP2CL>(mg, lset, idx, LocalVectorP2CL(kind));
        d = load_vector<InterfaceCommonDataP2CL>(mg, lset, idx, LocalVectorP2CL(f, 0.));
    } else {
        load_vector<InterfaceCommonDataP1CL>(mg, lset, idx, LocalVectorP1CL(kind));
        d = load_vector<InterfaceCommonDataP1CL>(mg, lset, idx, LocalVectorP1CL(f, 0.));
    }
    std::printf("user rhs vs device rhs: %.3e\n", d);
    return 0;
}

// A user accumulator as a DROPS application writes one (num/accumulator.h:42-66): it counts the
// visited tets, the tets with an interface patch and sums the patch areas through the tuple's
// interface common data, and records the visit order and the begin/finalize calls.
template <class CommonT>
class SurfaceStatsAccuCL : public TetraAccumulatorCL {
    const CommonT *cdata_;
    SurfaceStatsAccuCL *root_;
    int tid_;

  public:
    size_t visited = 0, cut = 0, begins = 0, finals = 0, clones = 0;
    bool mesh_order = true;
    size_t last = 0;
    double area = 0.;
    std::vector<size_t> per_thread_visits;
    std::vector<double> per_thread_area;
    std::vector<size_t> per_thread_cut;
    std::vector<uint8_t> seen;

    explicit SurfaceStatsAccuCL(const CommonT &c) : cdata_(&c), root_(this), tid_(0) {}
    void begin_accumulation() override {
        ++begins;
        const int nth = detail::max_threads();
        per_thread_visits.assign((size_t)nth, 0);
        per_thread_area.assign((size_t)nth, 0.);
        per_thread_cut.assign((size_t)nth, 0);
    }
    void visit(const TetraCL &t) override {
        SurfaceStatsAccuCL &R = *root_;
        if (R.seen.empty()) return;
        const auto &cd = cdata_->get_clone();
        ++R.per_thread_visits[(size_t)tid_];
        ++R.seen[t.GetId()];
        if (tid_ == 0 && R.clones == 0) {
            if (R.per_thread_visits[0] > 1 && t.GetId() <= R.last) R.mesh_order = false;
            R.last = t.GetId();
        }
        if (!cd.empty()) {
            ++R.per_thread_cut[(size_t)tid_];
            R.per_thread_area[(size_t)tid_] += cd.surface_area();
        }
    }
    TetraAccumulatorCL *clone(int clone_id) override {
        auto *c = new SurfaceStatsAccuCL(*cdata_);
        c->root_ = this;
        c->tid_ = clone_id;
        ++clones;
        return c;
    }
    void finalize_accumulation() override {
        ++finals;
        for (size_t i = 0; i < per_thread_visits.size(); ++i) {
            visited += per_thread_visits[i];
            cut += per_thread_cut[i];
            area += per_thread_area[i];
        }
    }
    void size_for(const MultiGridCL &mg) { seen.assign(mg.GetNumTriangTetra(), 0); }
};

template <class CommonT, class LocalMassT, class LocalLBT>
static int accus_case(int n, FiniteElementT fe) {
    const double h = 4.0 / n, sn = 1e-3 * h / std::sqrt(14.0);
    const double o[3] = {-2.0 - sn, -2.0 - 2 * sn, -2.0 - 3 * sn};
    const double e1[3] = {4, 0, 0}, e2[3] = {0, 4, 0}, e3[3] = {0, 0, 4};
    MultiGridCL mg(BrickBuilderCL(o, e1, e2, e3, n, n, n));
    LevelsetP2CL lset;
    lset.Init(mg, [](const double *p, double) { return (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 1.0; });
    BndDataCL lsbnd;
    IdxDescCL ifaceidx(fe);
    ifaceidx.CreateNumbering(0, mg, &lset.Phi, &lsbnd);
    MatDescCL M(&ifaceidx, &ifaceidx), A(&ifaceidx, &ifaceidx);
    TetraAccumulatorTupleCL accus;
    CommonT cdata(lset.Phi, lsbnd);
    accus.push_back(&cdata);
    InterfaceMatrixAccuCL<LocalMassT, CommonT> mass(&M, LocalMassT(1.), cdata, "mass");
    accus.push_back(&mass);
    auto *stats = new SurfaceStatsAccuCL<CommonT>(cdata);
    stats->size_for(mg);
    accus.push_back_acquire(stats);
    InterfaceMatrixAccuCL<LocalLBT, CommonT> lb(&A, LocalLBT(1.), cdata, "LB");
    accus.push_back(&lb);
    accumulate(accus, mg, 0, &lsbnd);
    size_t once = 0;
    for (uint8_t v : stats->seen) once += v == 1;
    double msum = 0.;
    for (size_t k = 0; k < M.Data.num_nonzeros(); ++k) msum += M.Data.raw_val()[k];
    std::printf("threads: %d\ncolors: %zu\ntets: %zu\nvisited: %zu\nvisited once: %zu\ncut: %zu\n"
                "begin: %zu\nfinalize: %zu\nclones: %zu\nmesh order: %d\narea: %.17g\nmass sum: %.17g\n"
                "unknowns: %zu\nnnz: %zu\n",
                detail::max_threads(), detail::max_threads() > 1 ? mg.GetColorClasses().num_colors() : (size_t)1,
                mg.GetNumTriangTetra(), stats->visited, once, stats->cut, stats->begins, stats->finals,
                stats->clones, stats->mesh_order ? 1 : 0, stats->area, msum, ifaceidx.NumUnknowns(),
                M.Data.num_nonzeros());
    return 0;
}

// ColorClassesCL on the n^3 brick: every tet in exactly one class, no two tets of a class share a
// vertex, classes sorted, the greedy choice (smallest colour free at the tet's vertices when it is
// reached in mesh order) reproduced by a second pass over the classes.
static int colors(int n) {
    const double o[3] = {0, 0, 0}, e1[3] = {1, 0, 0}, e2[3] = {0, 1, 0}, e3[3] = {0, 0, 1};
    MultiGridCL mg(BrickBuilderCL(o, e1, e2, e3, n, n, n));
    const ColorClassesCL &cc = mg.GetColorClasses();
    const size_t nt = mg.GetNumTriangTetra();
    std::vector<int> color_of(nt, -1);
    bool ok = true;
    for (size_t c = 0; c < cc.num_colors(); ++c) {
        const auto &cls = cc[c];
        ok &= std::is_sorted(cls.begin(), cls.end()) && !cls.empty();
        std::vector<uint8_t> vtx(mg.GetNumVertices(), 0);
        for (size_t t : cls) {
            ok &= color_of[t] == -1;
            color_of[t] = (int)c;
            const TetraCL T = mg.GetTetra(t);
            for (int i = 0; i < 4; ++i) ok &= vtx[T.GetVertexNum(i)]++ == 0;
        }
    }
    std::vector<std::vector<int>> at_vertex(mg.GetNumVertices());
    for (size_t t = 0; t < nt && ok; ++t) {
        ok &= color_of[t] >= 0;
        const TetraCL T = mg.GetTetra(t);
        std::vector<int> used;
        for (int i = 0; i < 4; ++i)
            for (int c : at_vertex[T.GetVertexNum(i)]) used.push_back(c);
        int c = 0;
        while (std::find(used.begin(), used.end(), c) != used.end()) ++c;
        ok &= c == color_of[t];
        for (int i = 0; i < 4; ++i) at_vertex[T.GetVertexNum(i)].push_back(color_of[t]);
    }
    std::printf("colors: %zu\ntets: %zu\nvalid: %d\nsizes:", cc.num_colors(), nt, ok ? 1 : 0);
    for (size_t c = 0; c < cc.num_colors(); ++c) std::printf(" %zu", cc[c].size());
    std::printf("\n");
    return ok ? 0 : 1;
}

int main(int argc, char **argv) {
    try {
        if (argc >= 2 && !std::strcmp(argv[1], "sphere20")) return sphere20();
        if (argc >= 3 && !std::strcmp(argv[1], "c1")) return c1(std::atoi(argv[2]), argc >= 4 ? argv[3] : "");
        if (argc >= 3 && !std::strcmp(argv[1], "p2")) return p2(std::atoi(argv[2]), argc >= 4 ? argv[3] : "");
        if (argc >= 3 && !std::strcmp(argv[1], "quaqua")) return quaqua(std::atoi(argv[2]), argc >= 4 ? argv[3] : "");
        if (argc >= 4 && !std::strcmp(argv[1], "accus")) {
            if (!std::strcmp(argv[3], "p2"))
                return accus_case<InterfaceCommonDataP2CL, LocalMassP2CL, LocalLaplaceBeltramiP2CL>(std::atoi(argv[2]),
                                                                                                    P2IF_FE);
            return accus_case<InterfaceCommonDataP1CL, LocalInterfaceMassP1CL, LocalLaplaceBeltramiP1CL>(
                std::atoi(argv[2]), P1IF_FE);
        }
        if (argc >= 3 && !std::strcmp(argv[1], "colors")) return colors(std::atoi(argv[2]));
        if (argc >= 5 && !std::strcmp(argv[1], "ch")) return ch(std::atoi(argv[2]), std::atoi(argv[3]), argv[4]);
        if (argc >= 4 && !std::strcmp(argv[1], "block")) return block(std::atoi(argv[2]), argv[3]);
        if (argc >= 4 && !std::strcmp(argv[1], "userrhs")) return userrhs(std::atoi(argv[2]), argv[3]);
        if (argc >= 5 && !std::strcmp(argv[1], "distribute"))
            return distribute(std::atoi(argv[2]), std::atof(argv[3]), argv[4]);
        if (argc >= 4 && !std::strcmp(argv[1], "band"))
            return band(std::atoi(argv[2]), std::atof(argv[3]), argc >= 5 ? argv[4] : "");
        std::fprintf(stderr,
                     "usage: %s sphere20 | c1 <n> [outdir] | p2 <n> [outdir] | band <n> <width> [outdir] | "
                     "accus <n> <p1|p2> | colors <n>\n",
                     argv[0]);
        return 2;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
