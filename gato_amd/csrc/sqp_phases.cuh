// SQP phase routines on CDNA4: one 16-lane "knot group" per knot (four groups per wave64),
// shared by the one-kernel-per-phase path and the persistent fused kernel.
//
// Reference semantics (Mo0dy/GATO):
//   KKT setup      gato/bsqp/kernels/setup_kkt.cuh:14-108, gato/dynamics/integrator.cuh:21-257,
//                  gato/dynamics/iiwa14/iiwa14_plant.cuh:104-155, 339-450 (indy7_plant.cuh:141-450)
//   Schur + P      gato/bsqp/kernels/schur_linsys.cuh:13-260
//   PCG            gato/bsqp/kernels/pcg.cuh:13-148
//   dz             gato/bsqp/kernels/schur_linsys.cuh:315-431
//   merit          gato/bsqp/kernels/merit.cuh:16-92, iiwa14_plant.cuh:276-327
//   line search    gato/bsqp/kernels/line_search.cuh:12-98
//
// Layouts (fp32): A_k (nx x nx), B_k (nx x nu), S/P blocks (nx x nx) are row-major;
// the position block of Q_k (nq x nq) is row-major, the velocity block and R_k are diagonal.
// Q~_k = Q_k + rho diag(1_nq, 0) (linalg.cuh:85-96): its inverse is blkdiag(Qip, 1/Qv).
#pragma once
#include <type_traits>
#include <utility>

#include "dynamics.cuh"

namespace gato {

constexpr int KG = 16;       // lanes per knot group
constexpr int NUM_ALPHAS = 8;  // settings.h:15
// exact in fp64; every use converts to the scalar type (fp32: the float constants)
constexpr double RHO_INIT = 1e-3, RHO_FACTOR = 1.2, RHO_MIN = 1e-8, RHO_MAX = 10.0;  // settings.h:18-22
constexpr double PCG_ABS_TOL = 1e-6;  // pcg.cuh:27

struct CostW {
    float q_cost, N_cost, q_lim, vel_lim, ctrl_lim;
    int terminal_mode;  // 0: reference (terminal block at x_{N-2} with q_cost); 1: at x_{N-1} with N_cost
    int hess_mode;      // 0: reference rank-1 EE Hessian; 1: Gauss-Newton J^T J
};

// Order LDS traffic between lanes of one wavefront (knot groups never straddle a wave).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// s + sum_{c < NX} x_c * m[c] in c order, where x_c is x in lane c of this lane's 16-lane row: one
// v_fmac_f32_dpp with row_newbcast:c per term, so a block row times a vector that is spread one entry per lane over
// the row reads the vector from the row's lanes instead of from LDS. Bitwise equal to the chain s = fma(m[c], x_c, s).
// Lanes 0 .. NX-1 of the row must be active. The s_nop before covers the hazard of a DPP read of a VGPR written by a
// VALU instruction in the previous two wait states, the one after a DPP read of s right after the block.
#define GATO_FMAC_BCAST(J, C) "v_fmac_f32_dpp %0, %1, %" #J " row_newbcast:" #C " row_mask:0xf bank_mask:0xf\n"
template <int NX>
__device__ __forceinline__ float row_bcast_dot(float s, float x, const float* m);
template <>
__device__ __forceinline__ float row_bcast_dot<14>(float s, float x, const float* m) {
    asm volatile("s_nop 1\n" GATO_FMAC_BCAST(2, 0) GATO_FMAC_BCAST(3, 1) GATO_FMAC_BCAST(4, 2) GATO_FMAC_BCAST(5, 3)
                 GATO_FMAC_BCAST(6, 4) GATO_FMAC_BCAST(7, 5) GATO_FMAC_BCAST(8, 6) GATO_FMAC_BCAST(9, 7)
                 GATO_FMAC_BCAST(10, 8) GATO_FMAC_BCAST(11, 9) GATO_FMAC_BCAST(12, 10) GATO_FMAC_BCAST(13, 11)
                 GATO_FMAC_BCAST(14, 12) GATO_FMAC_BCAST(15, 13) "s_nop 1\n"
                 : "+v"(s)
                 : "v"(x), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]),
                   "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]), "v"(m[12]), "v"(m[13]));
    return s;
}
template <>
__device__ __forceinline__ float row_bcast_dot<12>(float s, float x, const float* m) {
    asm volatile("s_nop 1\n" GATO_FMAC_BCAST(2, 0) GATO_FMAC_BCAST(3, 1) GATO_FMAC_BCAST(4, 2) GATO_FMAC_BCAST(5, 3)
                 GATO_FMAC_BCAST(6, 4) GATO_FMAC_BCAST(7, 5) GATO_FMAC_BCAST(8, 6) GATO_FMAC_BCAST(9, 7)
                 GATO_FMAC_BCAST(10, 8) GATO_FMAC_BCAST(11, 9) GATO_FMAC_BCAST(12, 10) GATO_FMAC_BCAST(13, 11)
                 "s_nop 1\n"
                 : "+v"(s)
                 : "v"(x), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]),
                   "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]));
    return s;
}
// NC independent chains side by side in one asm block, chain i: s[i] + sum_{j < NJ} x_i(lane j) * m[j] in j order.
// Interleaving hides the dependent-FMA latency that a single chain exposes when few waves share a SIMD; each chain
// is summed exactly as row_bcast_dot sums it.
template <int NJ, int NC>
__device__ __forceinline__ void row_bcast_dots(float* s, const float* x, const float* m);
template <>
__device__ __forceinline__ void row_bcast_dots<6, 2>(float* s, const float* x, const float* m) {
    asm volatile("s_nop 1\n"
                 "v_fmac_f32_dpp %0, %2, %4 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %4 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %5 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %5 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %6 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %6 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %7 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %7 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %8 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %8 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %9 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %9 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "s_nop 1\n"
                 : "+v"(s[0]), "+v"(s[1])
                 : "v"(x[0]), "v"(x[1]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]));
}
template <>
__device__ __forceinline__ void row_bcast_dots<6, 4>(float* s, const float* x, const float* m) {
    asm volatile("s_nop 1\n"
                 "v_fmac_f32_dpp %0, %4, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "s_nop 1\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3])
                 : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]));
}
template <>
__device__ __forceinline__ void row_bcast_dots<7, 2>(float* s, const float* x, const float* m) {
    asm volatile("s_nop 1\n"
                 "v_fmac_f32_dpp %0, %2, %4 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %4 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %5 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %5 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %6 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %6 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %7 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %7 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %8 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %8 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %9 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %9 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %10 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %10 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "s_nop 1\n"
                 : "+v"(s[0]), "+v"(s[1])
                 : "v"(x[0]), "v"(x[1]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]));
}
template <>
__device__ __forceinline__ void row_bcast_dots<7, 3>(float* s, const float* x, const float* m) {
    asm volatile("s_nop 1\n"
                 "v_fmac_f32_dpp %0, %3, %6 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %4, %6 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %5, %6 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %3, %7 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %4, %7 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %5, %7 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %3, %8 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %4, %8 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %5, %8 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %3, %9 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %4, %9 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %5, %9 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %3, %10 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %4, %10 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %5, %10 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %3, %11 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %4, %11 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %5, %11 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %3, %12 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %4, %12 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %5, %12 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "s_nop 1\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2])
                 : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]));
}
template <>
__device__ __forceinline__ void row_bcast_dots<7, 4>(float* s, const float* x, const float* m) {
    asm volatile("s_nop 1\n"
                 "v_fmac_f32_dpp %0, %4, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "s_nop 1\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3])
                 : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]));
}
template <>
__device__ __forceinline__ void row_bcast_dots<12, 4>(float* s, const float* x, const float* m) {
    asm volatile("s_nop 1\n"
                 "v_fmac_f32_dpp %0, %4, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %15 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %15 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %15 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %15 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %18 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %18 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %18 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %18 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %19 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %19 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %19 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %19 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "s_nop 1\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3])
                 : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]), "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]));
}
template <>
__device__ __forceinline__ void row_bcast_dots<14, 2>(float* s, const float* x, const float* m) {
    asm volatile("s_nop 1\n"
                 "v_fmac_f32_dpp %0, %2, %4 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %4 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %5 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %5 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %6 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %6 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %7 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %7 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %8 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %8 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %9 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %9 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %10 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %10 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %11 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %11 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %12 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %12 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %13 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %13 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %14 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %14 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %15 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %15 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %16 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %16 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %2, %17 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %3, %17 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
                 "s_nop 1\n"
                 : "+v"(s[0]), "+v"(s[1])
                 : "v"(x[0]), "v"(x[1]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]), "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]), "v"(m[12]), "v"(m[13]));
}
template <>
__device__ __forceinline__ void row_bcast_dots<14, 4>(float* s, const float* x, const float* m) {
    asm volatile("s_nop 1\n"
                 "v_fmac_f32_dpp %0, %4, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %8 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %9 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %10 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %11 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %12 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %13 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %14 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %15 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %15 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %15 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %15 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %16 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %18 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %18 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %18 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %18 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %19 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %19 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %19 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %19 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %20 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %20 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %20 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %20 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %0, %4, %21 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %1, %5, %21 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %2, %6, %21 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
                 "v_fmac_f32_dpp %3, %7, %21 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
                 "s_nop 1\n"
                 : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3])
                 : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]), "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]), "v"(m[12]), "v"(m[13]));
}
// all NR chains of a block product, four at a time
template <int NJ, int NR>
__device__ __forceinline__ void row_bcast_dots_all(float* s, const float* x, const float* m) {
    constexpr int Q = NR / 4, R = NR % 4;
#pragma unroll
    for (int i = 0; i < Q; i++) row_bcast_dots<NJ, 4>(s + 4 * i, x + 4 * i, m);
    if constexpr (R > 0) row_bcast_dots<NJ, R>(s + 4 * Q, x + 4 * Q, m);
}
// x in lane P of this lane's 16-lane row (a DPP move the compiler sees, so it inserts the hazard waits itself)
template <int P>
__device__ __forceinline__ float row_lane(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x150 + P, 0xF, 0xF,
                                                                 false));
}
// s[k] += sum_j x[j] of lane C_k of this lane's 16-lane row times m[j], j = 0 .. NX - 1 in order: the broadcast lane is
// fixed per chain and the broadcast register runs along it (the row of M times the row of another lane: M X^T with
// both matrices held a row per lane), four chains per asm block (two for the last two chains of NX = 14).
template <int NX, int C0, int C1, int C2, int C3>
__device__ __forceinline__ std::enable_if_t<NX == 14> lane_dots4(float* s, const float* x, const float* m) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, %4, %18 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %4, %18 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %4, %18 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %4, %18 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %5, %19 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %5, %19 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %5, %19 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %5, %19 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %6, %20 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %6, %20 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %6, %20 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %6, %20 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %7, %21 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %7, %21 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %7, %21 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %7, %21 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %8, %22 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %8, %22 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %8, %22 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %8, %22 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %9, %23 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %9, %23 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %9, %23 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %9, %23 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %10, %24 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %10, %24 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %10, %24 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %10, %24 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %11, %25 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %11, %25 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %11, %25 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %11, %25 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %12, %26 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %12, %26 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %12, %26 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %12, %26 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %13, %27 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %13, %27 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %13, %27 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %13, %27 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %14, %28 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %14, %28 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %14, %28 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %14, %28 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %15, %29 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %15, %29 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %15, %29 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %15, %29 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %16, %30 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %16, %30 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %16, %30 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %16, %30 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %17, %31 row_newbcast:%32 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %17, %31 row_newbcast:%33 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %17, %31 row_newbcast:%34 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %17, %31 row_newbcast:%35 row_mask:0xf bank_mask:0xf\n"
        "s_nop 1\n"
        : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3])
        : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]), "v"(x[8]), "v"(x[9]), "v"(x[10]), "v"(x[11]), "v"(x[12]), "v"(x[13]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]), "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]), "v"(m[12]), "v"(m[13]), "i"(C0), "i"(C1), "i"(C2), "i"(C3));
}
template <int NX, int C0, int C1>
__device__ __forceinline__ std::enable_if_t<NX == 14> lane_dots2(float* s, const float* x, const float* m) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, %2, %16 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %2, %16 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %3, %17 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %3, %17 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %4, %18 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %4, %18 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %5, %19 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %5, %19 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %6, %20 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %6, %20 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %7, %21 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %7, %21 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %8, %22 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %8, %22 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %9, %23 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %9, %23 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %10, %24 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %10, %24 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %11, %25 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %11, %25 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %12, %26 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %12, %26 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %13, %27 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %13, %27 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %14, %28 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %14, %28 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %15, %29 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %15, %29 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "s_nop 1\n"
        : "+v"(s[0]), "+v"(s[1])
        : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]), "v"(x[8]), "v"(x[9]), "v"(x[10]), "v"(x[11]), "v"(x[12]), "v"(x[13]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]), "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]), "v"(m[12]), "v"(m[13]), "i"(C0), "i"(C1));
}
template <int NX, int C0, int C1, int C2, int C3>
__device__ __forceinline__ std::enable_if_t<NX == 12> lane_dots4(float* s, const float* x, const float* m) {
    asm volatile(
        "s_nop 1\n"
        "v_fmac_f32_dpp %0, %4, %16 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %4, %16 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %4, %16 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %4, %16 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %5, %17 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %5, %17 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %5, %17 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %5, %17 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %6, %18 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %6, %18 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %6, %18 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %6, %18 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %7, %19 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %7, %19 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %7, %19 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %7, %19 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %8, %20 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %8, %20 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %8, %20 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %8, %20 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %9, %21 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %9, %21 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %9, %21 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %9, %21 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %10, %22 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %10, %22 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %10, %22 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %10, %22 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %11, %23 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %11, %23 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %11, %23 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %11, %23 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %12, %24 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %12, %24 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %12, %24 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %12, %24 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %13, %25 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %13, %25 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %13, %25 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %13, %25 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %14, %26 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %14, %26 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %14, %26 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %14, %26 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %0, %15, %27 row_newbcast:%28 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %1, %15, %27 row_newbcast:%29 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %2, %15, %27 row_newbcast:%30 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f32_dpp %3, %15, %27 row_newbcast:%31 row_mask:0xf bank_mask:0xf\n"
        "s_nop 1\n"
        : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3])
        : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]), "v"(x[8]), "v"(x[9]), "v"(x[10]), "v"(x[11]), "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]), "v"(m[8]), "v"(m[9]), "v"(m[10]), "v"(m[11]), "i"(C0), "i"(C1), "i"(C2), "i"(C3));
}
template <int NX, int... C>
__device__ __forceinline__ void lane_dots_seq(float* s, const float* x, const float* m, std::integer_sequence<int, C...>) {
    constexpr int c[] = {C...};
    constexpr int n = sizeof...(C);
    if constexpr (n == 4) lane_dots4<NX, c[0], c[1], c[2], c[3]>(s, x, m);
    else lane_dots2<NX, c[0], c[1]>(s, x, m);
}
// all NX chains: s[c] += sum_j x[j] of lane c times m[j] (c = 0 .. NX - 1)
template <int NX>
__device__ __forceinline__ void lane_dots_all(float* s, const float* x, const float* m) {
    lane_dots_seq<NX>(s, x, m, std::integer_sequence<int, 0, 1, 2, 3>{});
    lane_dots_seq<NX>(s + 4, x, m, std::integer_sequence<int, 4, 5, 6, 7>{});
    lane_dots_seq<NX>(s + 8, x, m, std::integer_sequence<int, 8, 9, 10, 11>{});
    if constexpr (NX == 14) lane_dots_seq<NX>(s + 12, x, m, std::integer_sequence<int, 12, 13>{});
}
#undef GATO_FMAC_BCAST

// max and |x| in the scalar type (fmaxf / fabsf for fp32)
__device__ __forceinline__ float smax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double smax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float smin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double smin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float sabs(float a) { return fabsf(a); }
__device__ __forceinline__ double sabs(double a) { return fabs(a); }

// ---------------------------------------------------------------- barrier terms
// -log(max(v - lo, 1e-10)) - log(max(hi - v, 1e-10))  (iiwa14_plant.cuh:104-111)
template <typename S = float>
__device__ __forceinline__ S barrier(S v, S lo, S hi) {
    S a = v - lo, b = hi - v;
    a = a <= S(1e-10f) ? S(1e-10f) : a;
    b = b <= S(1e-10f) ? S(1e-10f) : b;
    if constexpr (std::is_same_v<S, float>) return -logf(a) - logf(b);
    else return -log(a) - log(b);
}
// gradient: iiwa14 sign-preserving clamp (iiwa14_plant.cuh:114-136), indy7 plain clamp (indy7_plant.cuh:141-148)
template <int PL, typename S = float>
__device__ __forceinline__ S barrier_g(S v, S lo, S hi) {
    S a = v - lo, b = hi - v;
    const S eps = S(1e-6f);
    if (PL == 1) {
        a = a <= eps ? eps : a;
        b = b <= eps ? eps : b;
    } else {
        a = a >= S(0) ? (a < eps ? eps : a) : (a > -eps ? -eps : a);
        b = b >= S(0) ? (b < eps ? eps : b) : (b > -eps ? -eps : b);
    }
    return -rcp(a) + rcp(b);  // fp32: fast-math reciprocals (-use_fast_math)
}
// iiwa14 curvature 1/d_lo^2 + 1/d_hi^2 with |d| >= 1e-6 (iiwa14_plant.cuh:141-155)
template <typename S = float>
__device__ __forceinline__ S barrier_h(S v, S lo, S hi) {
    S a = fabs(v - lo), b = fabs(hi - v);
    a = a < S(1e-6f) ? S(1e-6f) : a;
    b = b < S(1e-6f) ? S(1e-6f) : b;
    return rcp(a * a) + rcp(b * b);
}

template <typename Tarr, typename S = float>
__device__ __forceinline__ S pick(const Tarr& arr, int n, int idx) {
    S r = S(0);
#pragma unroll
    for (int i = 0; i < n; i++) r = (i == idx) ? arr[i] : r;
    return r;
}

template <int PL, typename S = float>
__device__ __forceinline__ S plim(int which, int i) {
    using P = Plant<PL>;
    S r = S(0);
#pragma unroll
    for (int k = 0; k < P::NQ; k++) {
        const S v = which == 0 ? P::template qlo<S>(k) : which == 1 ? P::template qhi<S>(k)
                  : which == 2 ? P::template vlo<S>(k) : which == 3 ? P::template vhi<S>(k)
                  : which == 4 ? P::template ulo<S>(k) : P::template uhi<S>(k);
        r = (k == i) ? v : r;
    }
    return r;
}

// LDS floats per knot used by lin_knot: RNEA table (24 nq), Cholesky factor (nq(nq+1)/2), the transported-axes
// table (8 nq(nq+1)/2), which also hosts the RNEA force buffer before it is built and the gradient / Jacobian
// scratch after the tangents are done, and the knot's joint accelerations (8).
template <int PL>
__host__ __device__ constexpr int lin_lds_floats() {
    constexpr int NQ = Plant<PL>::NQ;
    return 24 * NQ + 9 * (NQ * (NQ + 1) / 2) + 8;
}

// Knots per chunk of the linearization in the fused kernels (k_fused, k_fused_mixed): up to 32 whatever the knot-group
// count G, so that a 256-lane work-group (G = 16) runs the shared dynamics of all knots of an N = 32 problem in one
// chunk (one single-wave latency instead of two); at most 64 (the chunk's lanes of lin_dyn), at most N.
__host__ __device__ constexpr int fused_lin_slots(int N, int G) {
    const int t = G > 32 ? G : 32;
    return N < t ? N : t;
}

// The dynamics of knot k <= N-2 that every column of its linearization shares, by ONE lane: qdd = M^-1 (u - c) with
// the Cholesky factor of M, the RNEA table at (q, qd, qdd) and the transported axes, written to the knot's LDS table
// shg (lin_lds_floats). xk points at [x_k u_k].
template <int PL, typename S = float>
__device__ __forceinline__ void lin_dyn(const S* xk, const S* fext, S* shg) {
    constexpr int NQ = Plant<PL>::NQ, NX = Plant<PL>::NX;
    S q[NQ], qd[NQ], u[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) {
        q[i] = xk[i];
        qd[i] = xk[NQ + i];
        u[i] = xk[NX + i];
    }
    Kin<PL, S> K;
    K.set(q);
    S* tab = shg;
    S* shL = tab + RneaTab<PL>::SIZE;
    S* ax = shL + NQ * (NQ + 1) / 2;
    S* sqdd = ax + 8 * (NQ * (NQ + 1) / 2);
    S qdd[NQ];
    FLdsT<S> fs{ax, 1};
    forward_dynamics_fs<PL>(K, qd, u, fext, qdd, shL, fs);
    __builtin_amdgcn_sched_barrier(0);
    rnea_publish<PL>(K, qd, qdd, fext, tab);
    publish_axes<PL>(K, ax);
#pragma unroll
    for (int i = 0; i < NQ; i++) sqdd[i] = qdd[i];
}

// ---------------------------------------------------------------- KKT setup, one knot
// k <= N-2: A_k, B_k, c_{k+1} and the cost model at (x_k, u_k, ref_k);
// k == N-1: terminal cost model (quirk Q4: at x_{N-2} with ref_{N-1} and q_cost) and c_0 = x_0 - x_s.
// xk points at [x_k u_k x_{k+1}] (k <= N-2) or is unused; xt is the terminal-block state. DYN: the knot group computes
// the shared dynamics itself (every lane, identical values); otherwise lin_dyn has written them to shg beforehand.
template <int PL, typename S = float, bool DYN = true>
__device__ __forceinline__ void lin_knot(int k, int N, int lane, S dt, const CostW& cw, const S* xk,
                                         const S* xt, const S* refk, const S* xrefk, const S* Qd,
                                         const S* Rd, const S* fext, const S* x0, const S* xs,
                                         S* A, S* Bm, S* cn, S* c0, S* Qp, S* Qv, S* Rg,
                                         S* qv, S* rv, S* shg) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU;
    const bool term = (k == N - 1);
    const S* xsrc = term ? xt : xk;
    S q[NQ], qd[NQ], u[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) {
        q[i] = xsrc[i];
        qd[i] = xsrc[NQ + i];
        const S ui = xk[NX + i];  // xk is a valid knot for the terminal block too: loaded, then discarded there
        u[i] = term ? S(0) : ui;
    }
    // the lane's operands of the integrator error and of the cost model are loaded from clamped lane indices right
    // before each section (a load under a lane or knot condition compiles to a branch that waits for its own load)
    const int lq = lane < NQ ? lane : 0, lx = lane < NX ? lane : 0;
    Kin<PL, S> K;
    K.set(q);
    // knot-group LDS: RNEA table, Cholesky factor of M (identical in every lane of the group,
    // published by lane 0 so that each lane's tangent pass keeps only its own derivative buffer)
    S* tab = shg;
    S* shL = tab + RneaTab<PL>::SIZE;
    S* ax = shL + NQ * (NQ + 1) / 2;
    if (!term) {
        S qdd[NQ];
        if constexpr (DYN) {
            FLdsT<S> fs{ax, 1};
            forward_dynamics_fs<PL>(K, qd, u, fext, qdd, shL, fs);
            __builtin_amdgcn_sched_barrier(0);
            rnea_publish<PL>(K, qd, qdd, fext, tab);
            publish_axes<PL>(K, ax);
            wave_sync();
        } else {
            const S* sqdd = ax + 8 * (NQ * (NQ + 1) / 2);
#pragma unroll
            for (int i = 0; i < NQ; i++) qdd[i] = sqdd[i];
        }
        const int col = lane < NX ? lane : 0;
        const int j = col < NQ ? col : col - NQ;
        const int dir = col < NQ ? 0 : 1;
        S dq[NQ];
        rnea_tangent_ax<PL>(K, qd, qdd, tab, ax, j, dir, dq);
#pragma unroll
        for (int i = 0; i < NQ; i++) dq[i] = -dq[i];
        chol_solve<NQ>(shL, dq);  // d qdd / d x_col
        S mc[NQ];
#pragma unroll
        for (int i = 0; i < NQ; i++) mc[i] = (i == j) ? S(1) : S(0);
        chol_solve<NQ>(shL, mc);  // M^-1 e_j
        // materialise both columns here: otherwise the compiler sinks the whole tangent sweep into the
        // lane-predicated stores below and keeps every intermediate of it live across the branch
#pragma unroll
        for (int i = 0; i < NQ; i++) {
            dq[i] = opaque(dq[i]);
            mc[i] = opaque(mc[i]);
        }
        const S h2 = S(0.5) * dt * dt;
        if (lane < NX) {  // column `lane` of A (integrator.cuh:143-169)
#pragma unroll
            for (int r = 0; r < NQ; r++) {
                S top = (r == lane ? S(1) : S(0)) + ((lane >= NQ && r == lane - NQ) ? dt : S(0)) + h2 * dq[r];
                S bot = ((NQ + r) == lane ? S(1) : S(0)) + dt * dq[r];
                A[r * NX + lane] = top;
                A[(NQ + r) * NX + lane] = bot;
            }
        }
        if (lane < NU) {  // column `lane` of B (integrator.cuh:170-184)
#pragma unroll
            for (int r = 0; r < NQ; r++) {
                Bm[r * NU + lane] = h2 * mc[r];
                Bm[(NQ + r) * NU + lane] = dt * mc[r];
            }
        }
        const S ql_ = xsrc[lq], qdl_ = xsrc[NQ + lq], xnq = xk[NX + NU + lq], xnv = xk[NX + NU + NQ + lq];
        if (lane < NQ) {  // c_{k+1} = x_{k+1} - x+ (integrator.cuh:49-62)
            const S ql = ql_, qdl = qdl_, al = pick<S[NQ], S>(qdd, NQ, lane);
            cn[lane] = xnq - (ql + dt * qdl + S(0.5) * al * dt * dt);
            cn[NQ + lane] = xnv - (qdl + dt * al);
        }
    }
    // cost model (iiwa14_plant.cuh:339-424); joint-space diagonal weights generalise qd_cost/u_cost.
    // Lane scalars are re-read from memory and the kinematics restart from opaque copies, so nothing of
    // the dynamics block above stays live into this section.
    const S w = S(term ? (cw.terminal_mode ? cw.N_cost : cw.q_cost) : cw.q_cost);
    const S* xr = xrefk ? xrefk : Qd;
    const S ql_ = xsrc[lq], qdl_ = xsrc[NQ + lq];
    const S rf0 = refk[0], rf1 = refk[1], rf2 = refk[2];
    const S qdw_q = Qd[lq], qdw_v = Qd[NQ + lq], rdw = Rd[lq], xrq = xr[lq], xrv = xr[NQ + lq], ul_ = xk[NX + lq];
    const S x0l = x0[lx], xsl = xs[lx];
    S xsm[NQ];
#pragma unroll
    for (int m = 0; m < NQ; m++) xsm[m] = xsrc[m];
    S p[3], Jc[3];
    {
        Kin<PL, S> K2;
#pragma unroll
        for (int i = 0; i < NQ; i++) {
            K2.s[i] = opaque(K.s[i]);
            K2.c[i] = opaque(K.c[i]);
        }
        ee_position<PL>(K2, lane < NQ ? lane : -1, p, Jc);
    }
    const S e0 = p[0] - rf0, e1 = p[1] - rf1, e2 = p[2] - rf2;
    const S gl = Jc[0] * e0 + Jc[1] * e1 + Jc[2] * e2;
    if (lane < NQ) {
        shg[lane] = gl;
        if (cw.hess_mode) { shg[NQ + 3 * lane] = Jc[0]; shg[NQ + 3 * lane + 1] = Jc[1]; shg[NQ + 3 * lane + 2] = Jc[2]; }
    }
    wave_sync();
    if (lane < NQ) {
        const S ql = ql_, qdl = qdl_;
        const S bgl = barrier_g<PL, S>(ql, plim<PL, S>(0, lane), plim<PL, S>(1, lane));
#pragma unroll
        for (int m = 0; m < NQ; m++) {
            S v;
            if (cw.hess_mode)
                v = w * (Jc[0] * shg[NQ + 3 * m] + Jc[1] * shg[NQ + 3 * m + 1] + Jc[2] * shg[NQ + 3 * m + 2]);
            else
                v = w * gl * shg[m];  // rank-1 EE term (quirk Q5)
            if (PL == 1) v += cw.q_lim * bgl * barrier_g<PL, S>(xsm[m], P::template qlo<S>(m), P::template qhi<S>(m));
            if (m == lane) {
                v += qdw_q;
                if (PL != 1) v += cw.q_lim * barrier_h<S>(ql, plim<PL, S>(0, lane), plim<PL, S>(1, lane));
            }
            Qp[lane * NQ + m] = v;
        }
        const S xr_q = xrefk ? xrq : S(0), xr_v = xrefk ? xrv : S(0);
        qv[lane] = w * gl + cw.q_lim * bgl + qdw_q * (ql - xr_q);
        const S bv = barrier_g<PL, S>(qdl, plim<PL, S>(2, lane), plim<PL, S>(3, lane));
        qv[NQ + lane] = qdw_v * (qdl - xr_v) + cw.vel_lim * bv;
        Qv[lane] = qdw_v + cw.vel_lim * (PL == 1 ? bv * bv : barrier_h<S>(qdl, plim<PL, S>(2, lane), plim<PL, S>(3, lane)));
        if (!term) {
            const S ul = ul_;
            const S bu = barrier_g<PL, S>(ul, plim<PL, S>(4, lane), plim<PL, S>(5, lane));
            rv[lane] = rdw * ul + cw.ctrl_lim * bu;
            Rg[lane] = rdw + cw.ctrl_lim * (PL == 1 ? bu * bu : barrier_h<S>(ul, plim<PL, S>(4, lane), plim<PL, S>(5, lane)));
        } else {
            rv[lane] = S(0);
            Rg[lane] = S(1);
        }
    }
    if (term && lane < NX) c0[lane] = x0l - xsl;
    wave_sync();
}

// In-place Gauss-Jordan inverse without pivoting of an n x n matrix held one row per lane
// (lane r < n owns row r in registers); the inverse overwrites the rows. Equivalent in exact
// arithmetic to the reference's elimination on [A | I] (linalg.cuh:365-519). piv: n floats of LDS.
// fp32: the pivot row reaches the other lanes of the knot group by DPP row broadcasts (row_lane) instead of an LDS
// round trip and two wave syncs per pivot. Every lane forms the scaled pivot row entry piv_c = row_p[c] * (1/row_p[p])
// itself, the same product the pivot lane stores, and updates row[c] = fma(-f, piv_c, row[c]) as the LDS form does:
// the results are the same bit for bit. Lanes past NX update rows nobody reads (they are never broadcast sources).
template <int NX, int P>
__device__ __forceinline__ void gj_pivot_dpp(int lane, float* row) {
    const float inv = rcp(row_lane<P>(row[P]));  // fast-math reciprocal of the pivot
    const float f = row[P];
    const bool pv = lane == P;
#pragma unroll
    for (int c = 0; c < NX; c++) {
        if (c == P) {
            row[c] = pv ? inv : fmaf(-f, inv, 0.f);
        } else {
            const float pc = row_lane<P>(row[c]) * inv;
            row[c] = pv ? pc : fmaf(-f, pc, row[c]);
        }
    }
}
template <int NX, int... P>
__device__ __forceinline__ void gj_dpp(int lane, float* row, std::integer_sequence<int, P...>) {
    (gj_pivot_dpp<NX, P>(lane, row), ...);
}

template <int NX, typename S>
__device__ __forceinline__ void gj_inplace(int lane, S* row, S* piv) {
    if constexpr (std::is_same_v<S, float>) {
        gj_dpp<NX>(lane, row, std::make_integer_sequence<int, NX>{});
        return;
    }
#pragma unroll
    for (int p = 0; p < NX; p++) {
        if (lane == p) {
            const S inv = rcp(row[p]);  // fast-math reciprocal of the pivot
            row[p] = S(1);
#pragma unroll
            for (int c = 0; c < NX; c++) {
                row[c] *= inv;
                piv[c] = row[c];
            }
        }
        wave_sync();
        if (lane < NX && lane != p) {
            const S f = row[p];
            row[p] = S(0);
#pragma unroll
            for (int c = 0; c < NX; c++) row[c] -= f * piv[c];
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------- Schur step 1: Q~_k^-1, R_k^-1
// Gauss-Jordan without pivoting on the nq x nq position block (linalg.cuh:457-519), lane l owns row l.
template <int PL, typename S>
__device__ __forceinline__ void invert_q_knot(int lane, bool has_r, S rho, const S* Qp, const S* Qv,
                                              const S* Rg, S* Qip, S* Qiv, S* Ri, S* piv) {
    constexpr int NQ = Plant<PL>::NQ;
    S row[NQ], t[NQ];
    // the row is loaded from a clamped index with every load in flight at once, then selected: the value barrier
    // (opaque) keeps the compiler from sinking each load into a lane-conditional branch that waits for it
    const int lr = lane < NQ ? lane : 0;
#pragma unroll
    for (int c = 0; c < NQ; c++) t[c] = Qp[lr * NQ + c];
#pragma unroll
    for (int c = 0; c < NQ; c++) t[c] = opaque(t[c]);
#pragma unroll
    for (int c = 0; c < NQ; c++) row[c] = lane < NQ ? t[c] + (c == lane ? rho : S(0)) : S(0);
    gj_inplace<NQ>(lane, row, piv);
    if (lane < NQ) {
#pragma unroll
        for (int c = 0; c < NQ; c++) Qip[lane * NQ + c] = row[c];
        Qiv[lane] = rcp(Qv[lane]);
        if (has_r) Ri[lane] = rcp(Rg[lane]);
    }
}

// Row r of the block-diagonal Q~^-1 = blkdiag(Qip, diag(Qiv)), loaded without branches on r.
template <int PL, typename S>
__device__ __forceinline__ void qinv_row(const S* Qip, const S* Qiv, int r, S* out) {
    constexpr int NQ = Plant<PL>::NQ;
    const int rq = r < NQ ? r : 0;
#pragma unroll
    for (int m = 0; m < NQ; m++) {
        const S a = Qip[rq * NQ + m];
        const S d = Qiv[m];
        out[m] = r < NQ ? a : S(0);
        out[NQ + m] = (NQ + m == r) ? d : S(0);
    }
}

// ---------------------------------------------------------------- Schur step 2: block row k+1 (k <= N-2)
// theta_k = A Q~_k^-1 A^T + B R^-1 B^T + Q~_{k+1}^-1, phi_k = A Q~_k^-1,
// gamma_{k+1} = c_{k+1} - Q~_{k+1}^-1 q_{k+1} + phi_k q_k + B R^-1 r_k,
// S_{k+1,k+1} = -theta_k, S_{k+1,k} = phi_k, P_{k+1,k+1} = -(theta_k + rho diag(1_nq, 0))^-1 (quirk Q8).
// The knot's A, B and position block of Q~_k^-1 are staged in the knot group's LDS tile (SchurTile) so the
// products read them as broadcasts from LDS instead of from global memory.
template <int PL>
struct SchurTile {
    static constexpr int NQ = Plant<PL>::NQ, NX = Plant<PL>::NX;
    static constexpr int AS = 16, BS = 8, QS = 8;            // row strides (floats)
    static constexpr int A_OFF = 0, B_OFF = NX * AS, Q_OFF = B_OFF + NX * BS, PIV_OFF = Q_OFF + NQ * QS;
    static constexpr int SIZE = PIV_OFF + 2 * NX;            // floats per knot group
};

// Store an NX x NX block whose row r is held by lane r (r < NX) of the knot group, times sgn, to dst (16-byte aligned,
// row-major): the rows go through the knot group's LDS stage (NX * NX floats, 16-byte aligned) and leave as contiguous
// 128-bit stores, instead of NX scattered 32-bit stores per lane at a stride of NX floats.
template <int NX>
__device__ __forceinline__ void store_block(int lane, const float* row, float sgn, float* stage, float* dst) {
    static_assert((NX * NX) % 4 == 0, "whole float4s");
    if (lane < NX) {
#pragma unroll
        for (int c = 0; c < NX; c++) stage[lane * NX + c] = sgn * row[c];
    }
    wave_sync();
    const float4* s4 = reinterpret_cast<const float4*>(stage);
    float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
    for (int i = lane; i < NX * NX / 4; i += 16) d4[i] = s4[i];
    wave_sync();
}

// schur_row in fp32 with every block product on DPP row broadcasts (row_bcast_dot) instead of the knot group's LDS
// tile: lane l is the broadcast source of row l of Q~_k^-1 (for phi_k = A Q~_k^-1), of A[m][l] and B[m][l] (for the
// products with A^T and B^T) and of the Gauss-Jordan pivot rows; each sum runs in the LDS form's order, so the results
// are the same bit for bit, without the tile's bank conflicts and wave syncs.
template <int PL>
__device__ __forceinline__ void schur_row_dpp(int lane, float rho, const float* A, const float* Bm, const float* Qip0,
                                              const float* Qiv0, const float* Ri0, const float* Qip1,
                                              const float* Qiv1, const float* q0, const float* q1, const float* r0,
                                              const float* c1, float* Sd1, float* So1, float* Pd1, float* g1,
                                              float* phi, float* prow, float* stage) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU;
    const int r = lane < NX ? lane : 0, rq = lane < NQ ? lane : 0, ru = lane < NU ? lane : 0;
    float Ar[NX], Qr[NQ];
#pragma unroll
    for (int l = 0; l < NX; l++) Ar[l] = A[r * NX + l];
#pragma unroll
    for (int c = 0; c < NQ; c++) Qr[c] = Qip0[rq * NQ + c];
#pragma unroll
    for (int c = 0; c < NQ; c++) {
        phi[c] = 0.f;
        phi[NQ + c] = Ar[NQ + c] * Qiv0[c];
    }
    row_bcast_dots_all<NQ, NQ>(phi, Qr, Ar);
    float brr[NU];
#pragma unroll
    for (int c = 0; c < NU; c++) brr[c] = Bm[r * NU + c] * Ri0[c];
    float th[NX];
    qinv_row<PL, float>(Qip1, Qiv1, r, th);
    float g = c1[r];
#pragma unroll
    for (int m = 0; m < NX; m++) {
        g -= th[m] * q1[m];
        g += phi[m] * q0[m];
    }
    float xa[NX], xb[NX];  // this lane's entries of A[m][.] and B[m][.] as broadcast sources
#pragma unroll
    for (int m = 0; m < NX; m++) {
        xa[m] = A[m * NX + r];
        xb[m] = Bm[m * NU + ru];
    }
    row_bcast_dots_all<NX, NX>(th, xa, phi);
    row_bcast_dots_all<NU, NX>(th, xb, brr);
#pragma unroll
    for (int l = 0; l < NU; l++) g += brr[l] * r0[l];
    if (lane < NX) g1[r] = g;
    store_block<NX>(lane, th, -1.f, stage, Sd1);
    store_block<NX>(lane, phi, 1.f, stage, So1);
#pragma unroll
    for (int m = 0; m < NX; m++) prow[m] = th[m] + ((m == r && m < NQ) ? rho : 0.f);
    gj_inplace<NX>(lane, prow, static_cast<float*>(nullptr));
#pragma unroll
    for (int m = 0; m < NX; m++) prow[m] = -prow[m];
    store_block<NX>(lane, prow, 1.f, stage, Pd1);
}

template <int PL, typename S>
__device__ __forceinline__ void schur_row(int lane, S rho, const S* A, const S* Bm, const S* Qip0,
                                          const S* Qiv0, const S* Ri0, const S* Qip1, const S* Qiv1,
                                          const S* q0, const S* q1, const S* r0, const S* c1,
                                          S* Sd1, S* So1, S* Pd1, S* g1, S* tile, S* phi,
                                          S* prow) {
    using P = Plant<PL>;
    using T = SchurTile<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU;
    if constexpr (std::is_same_v<S, float>) {
        schur_row_dpp<PL>(lane, rho, A, Bm, Qip0, Qiv0, Ri0, Qip1, Qiv1, q0, q1, r0, c1, Sd1, So1, Pd1, g1, phi, prow,
                          tile);
        return;
    }
    S* As = tile + T::A_OFF;
    S* Bs = tile + T::B_OFF;
    S* Qs = tile + T::Q_OFF;
    S* piv = tile + T::PIV_OFF;
    const int r = lane < NX ? lane : 0;
    if (lane < NX) {
#pragma unroll
        for (int l = 0; l < NX; l++) As[r * T::AS + l] = A[r * NX + l];
#pragma unroll
        for (int l = 0; l < NU; l++) Bs[r * T::BS + l] = Bm[r * NU + l];
        if (lane < NQ) {
#pragma unroll
            for (int l = 0; l < NQ; l++) Qs[r * T::QS + l] = Qip0[r * NQ + l];
        }
    }
    wave_sync();
#pragma unroll
    for (int c = 0; c < NQ; c++) {
        S s = S(0);
#pragma unroll
        for (int l = 0; l < NQ; l++) s += As[r * T::AS + l] * Qs[l * T::QS + c];
        phi[c] = s;
        phi[NQ + c] = As[r * T::AS + NQ + c] * Qiv0[c];
    }
    S brr[NU];
#pragma unroll
    for (int c = 0; c < NU; c++) brr[c] = Bs[r * T::BS + c] * Ri0[c];
    S th[NX];
    qinv_row<PL, S>(Qip1, Qiv1, r, th);
    S g = c1[r];
#pragma unroll
    for (int m = 0; m < NX; m++) {
        asm volatile("" ::: "memory");  // one row of A and B at a time: their LDS loads are not hoisted
        S s = th[m];
        g -= s * q1[m];
#pragma unroll
        for (int l = 0; l < NX; l++) s += phi[l] * As[m * T::AS + l];
#pragma unroll
        for (int l = 0; l < NU; l++) s += brr[l] * Bs[m * T::BS + l];
        th[m] = s;
        g += phi[m] * q0[m];
    }
#pragma unroll
    for (int l = 0; l < NU; l++) g += brr[l] * r0[l];
    if (lane < NX) {
        g1[r] = g;
#pragma unroll
        for (int m = 0; m < NX; m++) {
            Sd1[r * NX + m] = -th[m];
            So1[r * NX + m] = phi[m];
        }
    }
#pragma unroll
    for (int m = 0; m < NX; m++) prow[m] = th[m] + ((m == r && m < NQ) ? rho : S(0));
    gj_inplace<NX>(lane, prow, piv);
#pragma unroll
    for (int m = 0; m < NX; m++) prow[m] = -prow[m];
    if (lane < NX) {
#pragma unroll
        for (int m = 0; m < NX; m++) Pd1[r * NX + m] = prow[m];
    }
    wave_sync();
}

// Block row 0: S_00 = -Q~_0^-1, P_00 = -Q~_0, gamma_0 = c_0 - Q~_0^-1 q_0 (schur_linsys.cuh:164-210)
template <int PL, typename S>
__device__ __forceinline__ void schur_row0(int lane, S rho, const S* Qp0, const S* Qv0, const S* Qip0,
                                           const S* Qiv0, const S* q0, const S* c0, S* Sd0,
                                           S* Pd0, S* g0) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX;
    if (lane >= NX) return;
    const int r = lane;
    S g = c0[r];
    S qir[NX];
    qinv_row<PL, S>(Qip0, Qiv0, r, qir);
    // the blocks' entries loaded unconditionally from clamped indices, then selected (a load under the branches
    // below compiles to a branch that waits for its own load)
    const int rq = r < NQ ? r : 0, rv = (r >= NQ && r < NX) ? r - NQ : 0;
    S qprow[NQ];
#pragma unroll
    for (int m = 0; m < NQ; m++) qprow[m] = Qp0[rq * NQ + m];
    const S qvr = opaque(Qv0[rv]);
#pragma unroll
    for (int m = 0; m < NQ; m++) qprow[m] = opaque(qprow[m]);  // not sunk into the branches below
#pragma unroll
    for (int m = 0; m < NX; m++) {
        const S qi = qir[m];
        g -= qi * q0[m];
        Sd0[r * NX + m] = -qi;
        S qt = S(0);
        if (r < NQ && m < NQ) qt = qprow[m < NQ ? m : 0] + (r == m ? rho : S(0));
        else if (r >= NQ && m == r) qt = qvr;
        Pd0[r * NX + m] = -qt;
    }
    g0[r] = g;
}

// ---------------------------------------------------------------- Schur step 3: P_{k+1,k} = -P_{k+1,k+1} phi_k P_{k,k}
// P_{k,k} is staged in the knot group's tile; the product phi_k P_{k,k} then replaces it there.
template <int PL, typename S>
__device__ __forceinline__ void precond_offdiag(int lane, const S* phi, const S* prow, const S* Pd0,
                                                S* tile, S* Po1) {
    constexpr int NX = Plant<PL>::NX, AS = SchurTile<PL>::AS;
    const int r = lane < NX ? lane : 0;
    if constexpr (std::is_same_v<S, float>) {
        // DPP form, as schur_row_dpp: lane m broadcasts row m of P_{k,k}, then row m of t = phi_k P_{k,k}
        float pk[NX], t[NX], o[NX];
#pragma unroll
        for (int c = 0; c < NX; c++) {
            pk[c] = Pd0[r * NX + c];
            t[c] = 0.f;
            o[c] = 0.f;
        }
        row_bcast_dots_all<NX, NX>(t, pk, phi);
        row_bcast_dots_all<NX, NX>(o, t, prow);
        store_block<NX>(lane, o, -1.f, tile, Po1);
        return;
    }
    if (lane < NX) {
#pragma unroll
        for (int c = 0; c < NX; c++) tile[r * AS + c] = Pd0[r * NX + c];
    }
    wave_sync();
    S t[NX];
#pragma unroll
    for (int c = 0; c < NX; c++) {
        asm volatile("" ::: "memory");
        S s = S(0);
#pragma unroll
        for (int m = 0; m < NX; m++) s += phi[m] * tile[m * AS + c];
        t[c] = opaque(s);  // computed here, not sunk into the lane-predicated store below
    }
    wave_sync();
    if (lane < NX) {
#pragma unroll
        for (int c = 0; c < NX; c++) tile[lane * AS + c] = t[c];
    }
    wave_sync();
    if (lane < NX) {
#pragma unroll
        for (int c = 0; c < NX; c++) {
            S s = S(0);
#pragma unroll
            for (int m = 0; m < NX; m++) s += prow[m] * tile[m * AS + c];
            Po1[lane * NX + c] = -s;
        }
    }
    wave_sync();
}

// ---------------------------------------------------------------- dz, one knot (schur_linsys.cuh:315-431)
// v = q_k - lambda_k + A_k^T lambda_{k+1}; dx_k = -Q~_k^-1 v;  w = r_k + B_k^T lambda_{k+1}; du_k = -R_k^-1 w.
// Returns the lane's max |stationarity| and |defect| entries (the KKT residual of the problem is
// the max over knots and lanes).
template <int PL, typename S>
__device__ __forceinline__ S dz_knot(int k, int N, int lane, const S* A, const S* Bm, const S* Qip,
                                         const S* Qiv, const S* Ri, const S* q, const S* rr,
                                         const S* c, const S* lam0, const S* lam1, S* dzk,
                                         S* shv) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU;
    const bool last = (k == N - 1);
    // every operand is loaded up front from clamped addresses inside knot k's blocks (which exist for every k < N;
    // lam1 points at a valid knot) and used under the same conditions as before: loads placed under those conditions
    // become branches that each wait for their own loads
    const int lx = lane < NX ? lane : 0, lq = lane < NQ ? lane : 0, lu = lane < NU ? lane : 0;
    const int lv = (lane >= NQ && lane < NX) ? lane - NQ : 0;
    S acol[NX], bcol[NX], l1[NX], qrow[NQ];
#pragma unroll
    for (int i = 0; i < NX; i++) {
        acol[i] = A[i * NX + lx];
        bcol[i] = Bm[i * NU + lu];
        l1[i] = lam1[i];
    }
#pragma unroll
    for (int l = 0; l < NQ; l++) qrow[l] = Qip[lq * NQ + l];
    const S qv = q[lx], l0 = lam0[lx], cv = c[lx], qiv = Qiv[lv], rrv = rr[lu], ri = Ri[lu];
    S res = S(0);
    if (lane < NX) {
        S v = qv - l0;
        if (!last) {
#pragma unroll
            for (int i = 0; i < NX; i++) v += acol[i] * l1[i];
        }
        shv[lane] = v;
        res = smax(sabs(v), sabs(cv));
    }
    wave_sync();
    if (lane < NX) {
        S s;
        if (lane < NQ) {
            s = S(0);
#pragma unroll
            for (int l = 0; l < NQ; l++) s += qrow[l] * shv[l];
        } else {
            s = qiv * shv[lane];
        }
        dzk[lane] = -s;
    }
    if (!last && lane < NU) {
        S w = rrv;
#pragma unroll
        for (int i = 0; i < NX; i++) w += bcol[i] * l1[i];
        dzk[NX + lane] = -ri * w;
        res = smax(res, sabs(w));
    }
    wave_sync();
    return res;
}

// ---------------------------------------------------------------- merit, one (knot, alpha) item
// l_k + mu * |c|_1 at z + alpha dz (merit.cuh:16-92, iiwa14_plant.cuh:276-327).
template <int PL, typename S>
__device__ __forceinline__ S merit_item(int k, int N, S alpha, S dt, const CostW& cw, S mu,
                                            const S* xuk, const S* dzk, const S* refk,
                                            const S* xrefk, const S* Qd, const S* Rd,
                                            const S* fext, const S* x0, const S* dz0, const S* xs) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU;
    const bool last = (k == N - 1);
    // Every operand is loaded up front from an address that is valid for every item (the last knot has no control
    // and no successor in the trajectory, so its control and successor reads are pointed at its own state and the
    // values discarded; an absent joint reference reads Qd instead and is discarded): loads under conditions
    // compile to branches that each wait for their own loads.
    // (each group of operands is loaded right before the code that uses it, so that none stays live across the
    // forward dynamics)
    const S* xu_u = xuk + (last ? 0 : NX);
    const S* dz_u = dzk + (last ? 0 : NX);
    const S* xr = xrefk ? xrefk : Qd;
    S q[NQ], qd[NQ], u[NQ], xrv[NX], qdw[NX], rdw[NQ], fe[6], rf[3];
#pragma unroll
    for (int i = 0; i < NX; i++) {
        xrv[i] = xr[i];
        qdw[i] = Qd[i];
    }
#pragma unroll
    for (int i = 0; i < NQ; i++) rdw[i] = Rd[i];
#pragma unroll
    for (int i = 0; i < 6; i++) fe[i] = fext[i];
#pragma unroll
    for (int i = 0; i < 3; i++) rf[i] = refk[i];
#pragma unroll
    for (int i = 0; i < NQ; i++) {
        q[i] = xuk[i] + alpha * dzk[i];
        qd[i] = xuk[NQ + i] + alpha * dzk[NQ + i];
        const S uu = xu_u[i] + alpha * dz_u[i];
        u[i] = last ? S(0) : uu;
    }
    S cost = S(0);
#pragma unroll
    for (int i = 0; i < NQ; i++) {
        const S e = qd[i] - (xrefk ? xrv[NQ + i] : S(0));
        const S eq = q[i] - (xrefk ? xrv[i] : S(0));
        cost += S(0.5) * qdw[NQ + i] * e * e + S(0.5) * qdw[i] * eq * eq;
        cost += S(cw.q_lim) * barrier(q[i], P::template qlo<S>(i), P::template qhi<S>(i));
        cost += S(cw.vel_lim) * barrier(qd[i], P::template vlo<S>(i), P::template vhi<S>(i));
        if (!last) {
            cost += S(0.5) * rdw[i] * u[i] * u[i];
            cost += S(cw.ctrl_lim) * barrier(u[i], P::template ulo<S>(i), P::template uhi<S>(i));
        }
    }
    Kin<PL, S> K;
    K.set(q);
    S p[3];
    ee_position<PL>(K, -1, p, static_cast<S*>(nullptr));
    const S w = S(last ? cw.N_cost : cw.q_cost);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const S e = p[i] - rf[i];
        cost += S(0.5) * w * e * e;
    }
    S cons = S(0);
    if (!last) {
        S qdd[NQ], L[NQ * (NQ + 1) / 2];
        forward_dynamics<PL>(K, qd, u, fe, qdd, L);
        const S* xn = xuk + NX + NU;
        const S* dn = dzk + NX + NU;
        S xnv[NX], dnv[NX];
#pragma unroll
        for (int i = 0; i < NX; i++) {
            xnv[i] = xn[i];
            dnv[i] = dn[i];
        }
#pragma unroll
        for (int i = 0; i < NQ; i++) {
            const S qn = xnv[i] + alpha * dnv[i], qdn = xnv[NQ + i] + alpha * dnv[NQ + i];
            cons += sabs(qn - (q[i] + dt * qd[i] + S(0.5) * qdd[i] * dt * dt));
            cons += sabs(qdn - (qd[i] + dt * qdd[i]));
        }
    } else {
#pragma unroll
        for (int i = 0; i < NX; i++) cons += sabs(x0[i] + alpha * dz0[i] - xs[i]);
    }
    return cost + mu * cons;
}

// ---------------------------------------------------------------- work-group reductions (deterministic)
// Wave64 reductions on the DPP network: butterflies within each row of 16 lanes (quad_perm,
// row_half_mirror, row_mirror), then the four row results combined through readlane. The result is
// wave-uniform. All 64 lanes must be active.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
// fp64: the two 32-bit halves moved by the same DPP pattern
template <int CTRL>
__device__ __forceinline__ double dpp_f(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float lane_f(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lane_f(double v, int l) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
constexpr int DPP_QUAD_1032 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_2301 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_MIRROR = 0x140;

template <typename S>
__device__ __forceinline__ S wave_sum(S v) {
    v += dpp_f<DPP_QUAD_1032>(v);
    v += dpp_f<DPP_QUAD_2301>(v);
    v += dpp_f<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_f<DPP_ROW_MIRROR>(v);
    return (lane_f(v, 0) + lane_f(v, 16)) + (lane_f(v, 32) + lane_f(v, 48));
}
// Sum over one 16-lane row (a knot group); every lane of the row receives the sum.
template <typename S>
__device__ __forceinline__ S row16_sum(S v) {
    v += dpp_f<DPP_QUAD_1032>(v);
    v += dpp_f<DPP_QUAD_2301>(v);
    v += dpp_f<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_f<DPP_ROW_MIRROR>(v);
    return v;
}
template <typename S>
__device__ __forceinline__ S wave_max(S v) {
    v = smax(v, dpp_f<DPP_QUAD_1032>(v));
    v = smax(v, dpp_f<DPP_QUAD_2301>(v));
    v = smax(v, dpp_f<DPP_ROW_HALF_MIRROR>(v));
    v = smax(v, dpp_f<DPP_ROW_MIRROR>(v));
    return smax(smax(lane_f(v, 0), lane_f(v, 16)), smax(lane_f(v, 32), lane_f(v, 48)));
}

// Sum over the work-group; red holds one value per wave. Every thread returns the total.
template <typename S>
__device__ __forceinline__ S block_sum(S v, S* red) {
    const int nw = (blockDim.x + 63) >> 6;
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    S s = S(0);
    for (int i = 0; i < nw; i++) s += red[i];
    return s;
}
template <typename S>
__device__ __forceinline__ S block_max(S v, S* red) {
    const int nw = (blockDim.x + 63) >> 6;
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    S s = S(0);
    for (int i = 0; i < nw; i++) s = smax(s, red[i]);
    return s;
}

}  // namespace gato
