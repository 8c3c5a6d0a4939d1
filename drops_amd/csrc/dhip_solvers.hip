// dhip_solvers.hip -- the persistent Krylov solver kernels (FP64 PCG in the reference's order and in
// the single-reduction arrangement, FP32-inner / FP64-refined PCG, GMRES), compiled apart from the
// rest of libdhip.  The file is compiled once per DHIP_SOLVER_PART (0 .. 14), each compilation
// instantiating one share of the kernels, so that the shares build in parallel; the host code in
// dhip.hip launches them through their declarations in dhip_krylov.hpp.
#define DHIP_SOLVER_KERNELS
#include "dhip_krylov.hpp"

#ifndef DHIP_SOLVER_PART
#error "compile with -DDHIP_SOLVER_PART=<0..14>"
#endif

namespace dhip {

#if DHIP_SOLVER_PART == 0
template __global__ void k_pcg64<1>(PcgArgs);
template __global__ void k_pcg64<2>(PcgArgs);
template __global__ void k_pcg64<4>(PcgArgs);
template __global__ void k_pcg64<8>(PcgArgs);
template __global__ void k_pcg64<16>(PcgArgs);
template __global__ void k_pcg_mixed<1>(PcgMixArgs);
template __global__ void k_pcg_mixed<2>(PcgMixArgs);
template __global__ void k_pcg_mixed<4>(PcgMixArgs);
template __global__ void k_pcg_mixed<8>(PcgMixArgs);
template __global__ void k_pcg_mixed<16>(PcgMixArgs);
#elif DHIP_SOLVER_PART == 1
template __global__ void k_pcg64_sr<1, false>(PcgSrArgs);
template __global__ void k_pcg64_sr<2, false>(PcgSrArgs);
template __global__ void k_pcg64_sr<4, false>(PcgSrArgs);
template __global__ void k_pcg64_sr<8, false>(PcgSrArgs);
template __global__ void k_pcg64_sr<16, false>(PcgSrArgs);
#elif DHIP_SOLVER_PART == 2
template __global__ void k_pcg64_sr<1, true>(PcgSrArgs);
template __global__ void k_pcg64_sr<2, true>(PcgSrArgs);
template __global__ void k_pcg64_sr<4, true>(PcgSrArgs);
template __global__ void k_pcg64_sr<8, true>(PcgSrArgs);
template __global__ void k_pcg64_sr<16, true>(PcgSrArgs);
#elif DHIP_SOLVER_PART == 3
template __global__ void k_gmres<1, KREG, false>(GmresArgs);
template __global__ void k_gmres<1, KREG, true>(GmresArgs);
#elif DHIP_SOLVER_PART == 4
template __global__ void k_gmres<1, 0, false>(GmresArgs);
template __global__ void k_gmres<2, 0, false>(GmresArgs);
template __global__ void k_gmres<1, 0, true>(GmresArgs);
template __global__ void k_gmres<2, 0, true>(GmresArgs);
#elif DHIP_SOLVER_PART == 5
template __global__ void k_gmres<4, 0, false>(GmresArgs);
template __global__ void k_gmres<4, 0, true>(GmresArgs);
#elif DHIP_SOLVER_PART == 6
template __global__ void k_gmres<8, 0, false>(GmresArgs);
#elif DHIP_SOLVER_PART == 7
template __global__ void k_gmres<16, 0, false>(GmresArgs);
#elif DHIP_SOLVER_PART == 9
template __global__ void k_gmres<8, 0, true>(GmresArgs);
#elif DHIP_SOLVER_PART == 10
template __global__ void k_gmres<16, 0, true>(GmresArgs);
#elif DHIP_SOLVER_PART == 8
template __global__ void k_pcg64_tx<1>(PcgTxArgs);
template __global__ void k_pcg64_tx<2>(PcgTxArgs);
template __global__ void k_pcg64_tx<4>(PcgTxArgs);
#elif DHIP_SOLVER_PART == 11
// FP32 Arnoldi cycles (coherent exchange)
template __global__ void k_gmres<1, KREG, true, true>(GmresArgs);
template __global__ void k_gmres<1, 0, true, true>(GmresArgs);
template __global__ void k_gmres<2, 0, true, true>(GmresArgs);
template __global__ void k_gmres<4, 0, true, true>(GmresArgs);
#elif DHIP_SOLVER_PART == 12
template __global__ void k_gmres<8, 0, true, true>(GmresArgs);
#elif DHIP_SOLVER_PART == 13
template __global__ void k_gmres<16, 0, true, true>(GmresArgs);
#elif DHIP_SOLVER_PART == 14
template __global__ void k_pcg64_pipe<1>(PcgTxArgs);
template __global__ void k_pcg64_pipe<2>(PcgTxArgs);
template __global__ void k_pcg64_pipe<4>(PcgTxArgs);
#endif

}  // namespace dhip
