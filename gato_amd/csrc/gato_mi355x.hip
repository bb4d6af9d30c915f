// Batched SQP solver for one MI355X: kernels, host loop and the C-ABI of libgato_mi355x.so.
//
// Two execution paths share every phase routine (sqp_phases.cuh):
//   GATO_MODE_PLAIN  one kernel per phase (KKT, Schur, PCG, dz, check, merit, line search) and a
//                    host loop with a device->host read of the converged count every iteration,
//                    the structure of the reference's BSQP::solve (gato/bsqp/bsqp.cuh:103-197);
//   GATO_MODE_FUSED  one persistent kernel, one work-group per problem, running the initial merit,
//                    every SQP iteration and the final merit on the device with no host round trip.
// A work-group owns one problem; its 16-lane knot groups each own one knot (looping when N > 32).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gato_mi355x.h"
#include "sqp_phases.cuh"

// Work-groups are up to 512 lanes (8 wavefronts, two per SIMD). Asking for at least four waves per SIMD
// caps every phase kernel at 128 VGPRs, so two problems share a CU; left free, the scheduler spends
// registers on latency hiding and drops to one work-group per CU.
#define GATO_TWO_WG_PER_CU __attribute__((amdgpu_waves_per_eu(4)))
// The phase kernels templated on the scalar type: fp32 as above; the fp64 instantiations run at most 16 knot
// groups (256 lanes) and may use 256 VGPRs (two waves per SIMD), which their doubled register footprint needs.
#define GATO_PHASE_BOUNDS(S) \
    __launch_bounds__(sizeof(S) == 8 ? 256 : 512) __attribute__((amdgpu_waves_per_eu(sizeof(S) == 8 ? 2 : 4)))

namespace gato {

template <typename S>
struct ArgsT {
    int N, B, traj, max_sqp, max_pcg, exit_rule, adapt_rho;
    int skip_po;  // the Schur phase leaves P's off-diagonal blocks out (a PCG that applies P in factored form)
    S dt;
    float kkt_tol, solve_ratio;
    S pcg_abs_tol;  // absolute part of the PCG exit test (1e-6 in the solve, pcg.cuh:27; tiny for the adjoint)
    CostW cw;
    S* xu;
    const S *xs, *ref, *fext, *Qd, *Rd, *xref;
    S *lam, *rho, *drho;
    const S *drho_init, *mu, *tol;
    S *A, *Bm, *c, *Qp, *Qv, *Rg, *q, *r, *Qip, *Qiv, *Ri, *Sd, *So, *Pd, *Po, *gam, *dz, *mpart, *phi;
    int *conv, *active, *solved;
    const int* stop;  // graph mode: set once the batch has converged; every later phase launch returns at once
    int* stop_w;
    int* st_pcg;
    S *st_kres, *st_lsm, *st_step;
    int* st_sqp;
    S *st_init, *st_final;
};
using Args = ArgsT<float>;

// the kernels' dynamic LDS as an array of the scalar type (one declaration for every kernel)
template <typename S>
__device__ __forceinline__ S* dyn_lds() {
    extern __shared__ __align__(16) float sm_dyn[];
    return reinterpret_cast<S*>(sm_dyn);
}

template <int PL, typename S = float>
struct View {  // per-problem pointers
    static constexpr int NQ = Plant<PL>::NQ, NX = Plant<PL>::NX, NU = Plant<PL>::NU;
    S* xu;
    const S *xs, *ref, *fext, *Qd, *Rd, *xref;
    S *lam, *A, *Bm, *c, *Qp, *Qv, *Rg, *q, *r, *Qip, *Qiv, *Ri, *Sd, *So, *Pd, *Po, *gam, *dz, *mpart;
    __device__ View(const ArgsT<S>& a, int b) {
        const int N = a.N;
        xu = a.xu + (size_t)b * a.traj;
        xs = a.xs + (size_t)b * NX;
        ref = a.ref + (size_t)b * 6 * N;
        fext = a.fext + (size_t)b * 6;
        Qd = a.Qd + (size_t)b * NX;
        Rd = a.Rd + (size_t)b * NU;
        xref = a.xref ? a.xref + (size_t)b * N * NX : nullptr;
        lam = a.lam + (size_t)b * N * NX;
        A = a.A + (size_t)b * N * NX * NX;
        Bm = a.Bm + (size_t)b * N * NX * NU;
        c = a.c + (size_t)b * N * NX;
        Qp = a.Qp + (size_t)b * N * NQ * NQ;
        Qv = a.Qv + (size_t)b * N * NQ;
        Rg = a.Rg + (size_t)b * N * NU;
        q = a.q + (size_t)b * N * NX;
        r = a.r + (size_t)b * N * NU;
        Qip = a.Qip + (size_t)b * N * NQ * NQ;
        Qiv = a.Qiv + (size_t)b * N * NQ;
        Ri = a.Ri + (size_t)b * N * NU;
        Sd = a.Sd + (size_t)b * N * NX * NX;
        So = a.So + (size_t)b * N * NX * NX;
        Pd = a.Pd + (size_t)b * N * NX * NX;
        Po = a.Po + (size_t)b * N * NX * NX;
        gam = a.gam + (size_t)b * N * NX;
        dz = a.dz + (size_t)b * (a.traj + NU);
        mpart = a.mpart + (size_t)b * NUM_ALPHAS * N;
    }
};

// graph mode: the iteration's kernels are all enqueued up front; once the batch-level exit fires they return
template <typename S>
__device__ __forceinline__ bool stopped(const ArgsT<S>& a) { return a.stop && *a.stop; }

// ------------------------------------------------------------------ phases on one problem
// The KKT set-up of every knot (setup_kkt.cuh:14-108) in chunks of ts knots (ts <= 64, ts LDS tables of
// lin_lds_floats): the dynamics that all columns of a knot share (forward dynamics, RNEA table, transported axes) once
// per knot by one lane of the first wave (lin_dyn), then the knot's derivative columns and cost model one lane per
// column in its knot group (lin_knot). Computing the shared part per knot group instead would repeat it in all 16
// lanes of the group, and it is about 60% of a lane's instructions.
template <int PL, typename S = float>
__device__ __forceinline__ void phase_lin(const ArgsT<S>& a, int b, S* sm, int ts) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU, TL = lin_lds_floats<PL>();
    View<PL, S> v(a, b);
    const int N = a.N, tid = threadIdx.x, g = tid / KG, lane = tid % KG, G = blockDim.x / KG;
    const S* xt = v.xu + (size_t)(a.cw.terminal_mode ? N - 1 : N - 2) * (NX + NU);
    for (int k0 = 0; k0 < N; k0 += ts) {
        const int kn = min(N, k0 + ts);
        if (k0 > 0) __syncthreads();  // the previous chunk's tables have been read
        if (tid < kn - k0 && k0 + tid < N - 1) lin_dyn<PL, S>(v.xu + (size_t)(k0 + tid) * (NX + NU), v.fext, sm + tid * TL);
        __syncthreads();
        for (int k = k0 + g; k < kn; k += G) {
            // the lane index is re-laundered per knot so that lane-only constants of the knot routine are
            // recomputed inside the loop rather than hoisted out of it and held in registers throughout
            const int ln = opaque(lane);
            const S* xk = v.xu + (size_t)(k < N - 1 ? k : 0) * (NX + NU);
            lin_knot<PL, S, false>(k, N, ln, a.dt, a.cw, xk, xt, v.ref + 6 * k, v.xref ? v.xref + k * NX : nullptr,
                                   v.Qd, v.Rd, v.fext, v.xu, v.xs, v.A + (size_t)k * NX * NX,
                                   v.Bm + (size_t)k * NX * NU, v.c + (k + 1) * NX, v.c, v.Qp + k * NQ * NQ,
                                   v.Qv + k * NQ, v.Rg + k * NU, v.q + k * NX, v.r + k * NU, sm + (k - k0) * TL);
        }
    }
}

template <int PL, typename S = float>
__device__ __forceinline__ void phase_schur(const ArgsT<S>& a, int b, S rho, S* sm) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU;
    View<PL, S> v(a, b);
    const int N = a.N, g = threadIdx.x / KG, lane = threadIdx.x % KG, G = blockDim.x / KG;
    S* tile = sm + g * SchurTile<PL>::SIZE;
    S* piv = tile + SchurTile<PL>::PIV_OFF;
    for (int k = g; k < N; k += G)
        invert_q_knot<PL>(lane, k < N - 1, rho, v.Qp + k * NQ * NQ, v.Qv + k * NQ, v.Rg + k * NU,
                          v.Qip + k * NQ * NQ, v.Qiv + k * NQ, v.Ri + k * NU, piv);
    __syncthreads();
    for (int k = g; k < N; k += G) {
        if (k < N - 1) {
            S phi[NX], prow[2 * NX];
            schur_row<PL>(lane, rho, v.A + (size_t)k * NX * NX, v.Bm + (size_t)k * NX * NU, v.Qip + k * NQ * NQ,
                          v.Qiv + k * NQ, v.Ri + k * NU, v.Qip + (k + 1) * NQ * NQ, v.Qiv + (k + 1) * NQ,
                          v.q + k * NX, v.q + (k + 1) * NX, v.r + k * NU, v.c + (k + 1) * NX,
                          v.Sd + (size_t)(k + 1) * NX * NX, v.So + (size_t)(k + 1) * NX * NX,
                          v.Pd + (size_t)(k + 1) * NX * NX, v.gam + (k + 1) * NX, tile, phi, prow);
        } else {
            schur_row0<PL>(lane, rho, v.Qp, v.Qv, v.Qip, v.Qiv, v.q, v.c, v.Sd, v.Pd, v.gam);
        }
    }
    __syncthreads();
    if (a.skip_po) return;
    for (int k = g; k < N - 1; k += G) {
        const int rr = lane < NX ? lane : 0;
        S phi[NX], prow[NX];
        const S* So1 = v.So + (size_t)(k + 1) * NX * NX;
        const S* Pd1 = v.Pd + (size_t)(k + 1) * NX * NX;
#pragma unroll
        for (int m = 0; m < NX; m++) {
            phi[m] = So1[rr * NX + m];
            prow[m] = Pd1[rr * NX + m];
        }
        precond_offdiag<PL>(lane, phi, prow, v.Pd + (size_t)k * NX * NX, tile, v.Po + (size_t)(k + 1) * NX * NX);
    }
}

// y = (D, L) block-tridiagonal times the padded LDS vector x at row (k, l)
template <int NX, typename S>
__device__ __forceinline__ S btd_row(const S* D, const S* L, const S* x, int N, int k, int l) {
    constexpr int XP = 16;
    const S* Dr = D + (size_t)k * NX * NX + l * NX;
    const S* xk = x + (k + 1) * XP;
    // the neighbour blocks' entries are loaded from addresses inside the batch's blocks for every k (knot 0's left
    // block and the last knot's own block stand in) with all loads in flight at once; value barriers keep the
    // compiler from sinking them into the conditional sums, where each would wait for its own load
    const S* Lr = L + (size_t)k * NX * NX + l * NX;
    const S* Lc = L + (size_t)(k + 1 < N ? k + 1 : k) * NX * NX + l;
    S lr[NX], lc[NX];
#pragma unroll
    for (int c = 0; c < NX; c++) {
        lr[c] = Lr[c];
        lc[c] = Lc[c * NX];
    }
    S s = S(0);
#pragma unroll
    for (int c = 0; c < NX; c++) s += Dr[c] * xk[c];
#pragma unroll
    for (int c = 0; c < NX; c++) {
        lr[c] = opaque(lr[c]);
        lc[c] = opaque(lc[c]);
    }
    if (k > 0) {
        const S* xm = x + k * XP;
#pragma unroll
        for (int c = 0; c < NX; c++) s += lr[c] * xm[c];
    }
    if (k < N - 1) {
        const S* xp = x + (k + 2) * XP;
#pragma unroll
        for (int c = 0; c < NX; c++) s += lc[c] * xp[c];
    }
    return s;
}

template <int PL>
__device__ void btd_apply(int N, const float* Gp, const float* Dp, const float* rv, float* zv, float* pv);

// Warm-started PCG on S lambda = gamma (pcg.cuh:13-148). Returns the iteration count. BTD: precondition with the
// twisted factorization of S (Gp, Dp: the problem's blocks from k_btd_factor, btd_apply) instead of P, for the
// backward pass; the LDS then also holds z and the sweeps' middle vectors (3 (N + 2) 16 + 48 floats in all).
template <int PL, int NPASS, bool BTD = false, typename S = float>
__device__ __forceinline__ int phase_pcg(const ArgsT<S>& a, int b, S eps, S* sm, const S* Gp = nullptr,
                                         const S* Dp = nullptr) {
    constexpr int NX = Plant<PL>::NX, XP = 16;
    View<PL, S> v(a, b);
    const int N = a.N, tid = threadIdx.x, g = tid / KG, lane = tid % KG, G = blockDim.x / KG;
    S* sp = sm;
    S* sr = sp + (N + 2) * XP;
    S* red = sr + (N + 2) * XP;
    S* sz = red + 16;
    S* pv = sz + (N + 2) * XP;
    for (int i = tid; i < 2 * (N + 2) * XP; i += blockDim.x) sm[i] = S(0);
    __syncthreads();
    S x[NPASS], r[NPASS], z[NPASS], p[NPASS], Ap[NPASS];
    bool val[NPASS];
    int kk[NPASS];
#pragma unroll
    for (int s = 0; s < NPASS; s++) {
        kk[s] = g + s * G;
        val[s] = (lane < NX) && (kk[s] < N);
        x[s] = val[s] ? v.lam[kk[s] * NX + lane] : S(0);
        if (val[s]) sp[(kk[s] + 1) * XP + lane] = x[s];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NPASS; s++) r[s] = val[s] ? v.gam[kk[s] * NX + lane] - btd_row<NX, S>(v.Sd, v.So, sp, N, kk[s], lane) : S(0);
#pragma unroll
    for (int s = 0; s < NPASS; s++)
        if (val[s]) sr[(kk[s] + 1) * XP + lane] = r[s];
    __syncthreads();
    if constexpr (BTD) btd_apply<PL>(N, Gp, Dp, sr, sz, pv);
    S loc = S(0);
#pragma unroll
    for (int s = 0; s < NPASS; s++) {
        if constexpr (BTD) z[s] = val[s] ? sz[(kk[s] + 1) * XP + lane] : S(0);
        else z[s] = val[s] ? btd_row<NX, S>(v.Pd, v.Po, sr, N, kk[s], lane) : S(0);
        p[s] = z[s];
        loc += r[s] * z[s];
    }
    S rho = block_sum(loc, red);
    if (sabs(rho) < a.pcg_abs_tol) return 0;
    const S rho0 = sabs(rho);
    int it = 0;
    for (int i = 0; i < a.max_pcg; i++) {
        it++;
#pragma unroll
        for (int s = 0; s < NPASS; s++)
            if (val[s]) sp[(kk[s] + 1) * XP + lane] = p[s];
        __syncthreads();
        loc = S(0);
#pragma unroll
        for (int s = 0; s < NPASS; s++) {
            Ap[s] = val[s] ? btd_row<NX, S>(v.Sd, v.So, sp, N, kk[s], lane) : S(0);
            loc += p[s] * Ap[s];
        }
        const S alpha = rho / block_sum(loc, red);
#pragma unroll
        for (int s = 0; s < NPASS; s++) {
            x[s] += alpha * p[s];
            r[s] -= alpha * Ap[s];
            if (val[s]) sr[(kk[s] + 1) * XP + lane] = r[s];
        }
        __syncthreads();
        if (it >= a.max_pcg) break;  // x is final at the cap: its last preconditioner application is not needed
        if constexpr (BTD) btd_apply<PL>(N, Gp, Dp, sr, sz, pv);
        loc = S(0);
#pragma unroll
        for (int s = 0; s < NPASS; s++) {
            if constexpr (BTD) z[s] = val[s] ? sz[(kk[s] + 1) * XP + lane] : S(0);
            else z[s] = val[s] ? btd_row<NX, S>(v.Pd, v.Po, sr, N, kk[s], lane) : S(0);
            loc += r[s] * z[s];
        }
        const S rn = block_sum(loc, red);
        if (sabs(rn) < a.pcg_abs_tol + eps * rho0) break;
        const S beta = rn / rho;
        rho = rn;
#pragma unroll
        for (int s = 0; s < NPASS; s++) p[s] = z[s] + beta * p[s];
    }
#pragma unroll
    for (int s = 0; s < NPASS; s++)
        if (val[s]) v.lam[kk[s] * NX + lane] = x[s];
    return it;
}

// Sum over the work-group with one barrier: each 16-lane row is summed on the DPP network and its lane 0
// writes one slot (at most 32 rows); after the barrier lanes 0..31 of every wave pick up the slots and two
// more row sums plus one add give the total, identical in every lane. red (32 floats) is double-buffered by
// the caller (alternate halves between consecutive calls), so no barrier is needed before the writes.
__device__ __forceinline__ float block_sum1(float v, float* red) {
    const int rows = blockDim.x >> 4;
    v = row16_sum(v);
    if ((threadIdx.x & 15) == 0) red[threadIdx.x >> 4] = v;
    __syncthreads();
    const int l = threadIdx.x & 63;
    float s = l < rows ? red[l] : 0.f;
    s = row16_sum(s);
    return lane_f(s, 0) + lane_f(s, 16);
}

// Same iteration as phase_pcg with the lane's rows of S and P held in registers for the whole
// solve (one knot per knot group: N <= 32) and two work-group barriers per iteration (the two dot products').
// The sums run in the same order as btd_row, so the results are identical to phase_pcg.
template <int PL>
__device__ __forceinline__ int phase_pcg_reg(const Args& a, int b, float eps, float* sm) {
    constexpr int NX = Plant<PL>::NX, XP = 16;
    View<PL> v(a, b);
    const int N = a.N, tid = threadIdx.x, k = tid / KG, lane = tid % KG;
    float* sp = sm;
    float* sr = sp + (N + 2) * XP;
    float* red = sr + (N + 2) * XP;  // 2 x 32 floats: two double-buffered reduction buffers
    for (int i = tid; i < 2 * (N + 2) * XP; i += blockDim.x) sm[i] = 0.f;
    const bool val = (lane < NX) && (k < N);
    const int l = lane < NX ? lane : 0;
    const int kk = k < N ? k : 0;
    // rows: main block row l, left block row l (k > 0), right block column l (k < N-1)
    float sd[NX], sl[NX], su[NX], pd[NX], pl[NX], pu[NX];
    {
        const float* D = v.Sd + (size_t)kk * NX * NX + l * NX;
        const float* L = v.So + (size_t)kk * NX * NX + l * NX;
        const float* U = v.So + (size_t)(kk + 1 < N ? kk + 1 : kk) * NX * NX + l;
        const float* PD = v.Pd + (size_t)kk * NX * NX + l * NX;
        const float* PL_ = v.Po + (size_t)kk * NX * NX + l * NX;
        const float* PU = v.Po + (size_t)(kk + 1 < N ? kk + 1 : kk) * NX * NX + l;
        const bool hasl = kk > 0, hasu = kk < N - 1;
        // every address is inside the batch's blocks (kk, l clamped), so all 84 loads are issued unconditionally and
        // the missing neighbours zeroed by selects: a load under the condition becomes a branch per column, each
        // waiting for its loads
#pragma unroll
        for (int c = 0; c < NX; c++) {
            const float l0 = L[c], l1 = PL_[c], u0 = U[c * NX], u1 = PU[c * NX];
            sd[c] = D[c];
            pd[c] = PD[c];
            sl[c] = hasl ? l0 : 0.f;
            pl[c] = hasl ? l1 : 0.f;
            su[c] = hasu ? u0 : 0.f;
            pu[c] = hasu ? u1 : 0.f;
        }
    }
    // block-tridiagonal row times a vector held one entry per lane (own: this knot's entry, 0 on padding lanes):
    // the knot's own block from the row's lanes by DPP broadcast, the neighbours' entries (one 32-bit LDS load
    // each) likewise; the terms are summed in the order diagonal, left, right block, column by column
    const int ln = lane;  // 0 .. 15; padding lanes read the zero padding of the LDS vectors
    float* sz = sr + (N + 2) * XP + 64;  // z, after the two reduction buffers
    for (int i = tid; i < (N + 2) * XP; i += blockDim.x) sz[i] = 0.f;
    // block row times a vector held one entry per lane: own entry, and the entries of knots k-1 and k+1 at this
    // lane's index (xm, xp), which the knot group's lanes broadcast to each other
    auto mv = [&](const float* M0, const float* M1, const float* M2, float own, float xm, float xp) {
        float s = row_bcast_dot<NX>(0.f, own, M0);
        s = row_bcast_dot<NX>(s, xm, M1);
        return row_bcast_dot<NX>(s, xp, M2);
    };
    const float lam0 = v.lam[kk * NX + l], gam0 = v.gam[kk * NX + l];  // issued with the rows, before the barriers
    float x = val ? lam0 : 0.f;
    __syncthreads();
    if (val) sp[(kk + 1) * XP + l] = x;
    __syncthreads();
    const float sx = mv(sd, sl, su, x, sp[kk * XP + ln], sp[(kk + 2) * XP + ln]);
    float r = val ? gam0 - sx : 0.f;
    if (val) sr[(kk + 1) * XP + l] = r;
    __syncthreads();
    // Each lane keeps its own copies of r and p at its index in the neighbour knots (rm, rp, pm, pp) and updates
    // them with the neighbours' Ap and z, which are published before the two reductions' barriers anyway, by the
    // same operations the owners apply: the copies equal the owners' values bit for bit, and no barrier is needed
    // to publish r and p (two barriers per iteration instead of four).
    float rm = sr[kk * XP + ln], rp = sr[(kk + 2) * XP + ln];
    const float pr0 = mv(pd, pl, pu, r, rm, rp);
    float z = val ? pr0 : 0.f;
    if (val) sz[(kk + 1) * XP + l] = z;
    int slot = 0;
    float rho = block_sum1(r * z, red + 32 * slot);
    slot ^= 1;
    float p = z, pm = sz[kk * XP + ln], pp = sz[(kk + 2) * XP + ln];
    if (fabsf(rho) < a.pcg_abs_tol) return 0;
    const float rho0 = fabsf(rho);
    int it = 0;
    for (int i = 0; i < a.max_pcg; i++) {
        it++;
        const float sp_ = mv(sd, sl, su, p, pm, pp);
        const float Ap = val ? sp_ : 0.f;
        if (val) sp[(kk + 1) * XP + l] = Ap;
        // quotients by reciprocal and product, as the reference's fast-math build divides (-use_fast_math)
        const float alpha = rho * __builtin_amdgcn_rcpf(block_sum1(p * Ap, red + 32 * slot));
        slot ^= 1;
        x += alpha * p;
        r -= alpha * Ap;
        rm -= alpha * sp[kk * XP + ln];
        rp -= alpha * sp[(kk + 2) * XP + ln];
        if (it >= a.max_pcg) break;  // x is final at the cap: its last preconditioner application is not needed
        const float pr = mv(pd, pl, pu, r, rm, rp);
        z = val ? pr : 0.f;
        if (val) sz[(kk + 1) * XP + l] = z;
        const float rn = block_sum1(r * z, red + 32 * slot);
        slot ^= 1;
        if (fabsf(rn) < a.pcg_abs_tol + eps * rho0) break;
        const float beta = rn * __builtin_amdgcn_rcpf(rho);
        rho = rn;
        p = z + beta * p;
        pm = sz[kk * XP + ln] + beta * pm;
        pp = sz[(kk + 2) * XP + ln] + beta * pp;
    }
    if (val) v.lam[kk * NX + l] = x;
    return it;
}

// Sum over a 1024-lane work-group (64 knot groups) with one barrier, as block_sum1 over 64 slots. red: 64 floats,
// double-buffered by the caller.
__device__ __forceinline__ float block_sum64(float v, float* red) {
    v = row16_sum(v);
    if ((threadIdx.x & 15) == 0) red[threadIdx.x >> 4] = v;
    __syncthreads();
    float s = row16_sum(red[threadIdx.x & 63]);
    return (lane_f(s, 0) + lane_f(s, 16)) + (lane_f(s, 32) + lane_f(s, 48));
}

// LDS floats of phase_pcg_res: the problem's off-diagonal blocks of S, three padded vectors, two reduction buffers
template <int PL>
__host__ __device__ constexpr int pcg_res_lds_floats(int N) {
    return N * Plant<PL>::NX * Plant<PL>::NX + 3 * (N + 2) * 16 + 128;
}

// The iteration of phase_pcg with the problem held on its CU for the whole solve: G knot groups (GC = 64: 1024 lanes,
// the sums over 64 groups of block_sum64; GC = 0: blockDim.x / 16 <= 32 groups, block_sum1's), knot group g owning
// knots g, g + G, ..., g + (NP - 1) G (N <= NP G), the diagonal blocks of S and P and each lane's columns of the
// upper off-diagonal blocks in registers (lane l: row l), the off-diagonal blocks of S (So, 100 KB for iiwa14 at
// N = 128) in LDS, nothing streamed from HBM per iteration. Every block row
// times a vector uses DPP broadcasts within the knot group (row_bcast_dot), the neighbour knots' entries read as one
// 32-bit LDS load per lane. S p sums the diagonal, left and right blocks in btd_row's order. P is applied in factored
// form: the preconditioner's off-diagonal blocks are P_{k+1,k} = -P_{k+1,k+1} So_{k+1} P_{k,k} by construction
// (schur_linsys.cuh:213-260, precond_offdiag), so with w = Pd r (block diagonal),
//   (P r)_k = Pd_k r_k + P_{k,k-1} r_{k-1} + P_{k+1,k}^T r_{k+1} = Pd_k (r_k - So_k w_{k-1} - So_{k+1}^T w_{k+1}),
// exactly in real arithmetic (Pd_k is symmetric); in fp32 it differs from the explicit product by rounding. The
// neighbours' w are carried like r (w -= alpha Pd Ap), so the iteration has two work-group barriers, its two dot
// products'.
template <int PL, int NP = 2, int GC = 64>
__device__ __forceinline__ int phase_pcg_res(const Args& a, int b, float eps, float* sm) {
    constexpr int NX = Plant<PL>::NX, NB = NX * NX, XP = 16;
    View<PL> v(a, b);
    const int N = a.N, tid = threadIdx.x, g = tid / KG, lane = tid % KG, G = GC ? GC : (int)blockDim.x / KG;
    auto bsum = [](float x, float* red) { return GC == 64 ? block_sum64(x, red) : block_sum1(x, red); };
    float* So = sm;
    float* sp = So + N * NB;        // x at the start
    float* sw = sp + (N + 2) * XP;  // w = Pd r
    float* sz = sw + (N + 2) * XP;  // z
    float* red = sz + (N + 2) * XP;  // 2 x 64 floats
    {
        // knot 0 has no left neighbour: its block is stored as zeros, so block row 0 needs no test of k > 0 (taken
        // in the 1024-lane form, GC = 64, where the tests measured faster; the same values either way)
        const float4* src = reinterpret_cast<const float4*>(v.So);
        float4* dst = reinterpret_cast<float4*>(So);
        for (int i = tid; i < N * NB / 4; i += blockDim.x) dst[i] = i < NB / 4 ? make_float4(0.f, 0.f, 0.f, 0.f) : src[i];
        for (int i = tid; i < 3 * (N + 2) * XP; i += blockDim.x) sp[i] = 0.f;
    }
    const int l = lane < NX ? lane : 0;
    bool val[NP];
    int kk[NP];
    float sd[NP][NX], pd[NP][NX];
#pragma unroll
    for (int s = 0; s < NP; s++) {
        kk[s] = g + s * G < N ? g + s * G : 0;
        val[s] = lane < NX && g + s * G < N;
        const float* D = v.Sd + (size_t)kk[s] * NB + l * NX;
        const float* PD = v.Pd + (size_t)kk[s] * NB + l * NX;
#pragma unroll
        for (int c = 0; c < NX; c++) {
            sd[s][c] = D[c];
            pd[s][c] = PD[c];
        }
    }
    // column l of So_{k+1} (k < N-1) for every knot of the group, in registers for the whole solve
    float uc[NP][NX];
#pragma unroll
    for (int s = 0; s < NP; s++) {
        const int k = kk[s];
        const float* Uc = v.So + (size_t)(k + 1 < N ? k + 1 : k) * NB + l;
#pragma unroll
        for (int c = 0; c < NX; c++) uc[s][c] = k < N - 1 ? Uc[c * NX] : 0.f;
    }
    // the off-diagonal blocks of block row kk[s] times a vector whose entries at this lane's index in knots k-1 and k+1
    // are ym and yp: row l of So_k (from LDS; zeros for k = 0), then column l of So_{k+1} (registers), from acc
    auto offdiag = [&](int s, float ym, float yp, float acc) {
        const int k = kk[s];
        float L[NX];
        const float2* Lr = reinterpret_cast<const float2*>(So + k * NB + l * NX);  // 8-byte aligned: NX is even
#pragma unroll
        for (int c = 0; c < NX / 2; c++) {
            const float2 t = Lr[c];
            L[2 * c] = GC == 64 && k == 0 ? 0.f : t.x;
            L[2 * c + 1] = GC == 64 && k == 0 ? 0.f : t.y;
        }
        acc = row_bcast_dot<NX>(acc, ym, L);
        return row_bcast_dot<NX>(acc, yp, uc[s]);
    };
    float x[NP], r[NP], p[NP], z[NP], pm[NP], pp[NP];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NP; s++) {
        x[s] = val[s] ? v.lam[kk[s] * NX + l] : 0.f;
        if (val[s]) sp[(kk[s] + 1) * XP + l] = x[s];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NP; s++) {
        const float sx = offdiag(s, sp[kk[s] * XP + lane], sp[(kk[s] + 2) * XP + lane],
                                 row_bcast_dot<NX>(0.f, x[s], sd[s]));
        r[s] = val[s] ? v.gam[kk[s] * NX + l] - sx : 0.f;
    }
    // w = Pd r at this lane's index in knots k-1 and k+1 (wm, wp): published once here (one barrier), then kept up to
    // date as r is, w -= alpha Pd Ap, with Pd Ap published by its owners before the barrier of p.Ap, so that
    // z = Pd (r - So_k wm - So_{k+1}^T wp) needs no exchange of its own; z is published for the neighbours' copies
    // of p (pm, pp), which are updated by the owners' operations, so no barrier publishes p either
    float wm[NP], wp[NP];
#pragma unroll
    for (int s = 0; s < NP; s++) {
        const float w = row_bcast_dot<NX>(0.f, r[s], pd[s]);
        if (val[s]) sw[(kk[s] + 1) * XP + l] = w;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NP; s++) {
        wm[s] = sw[kk[s] * XP + lane];
        wp[s] = sw[(kk[s] + 2) * XP + lane];
    }
    auto precond = [&](float* zz) {
        float loc = 0.f;
#pragma unroll
        for (int s = 0; s < NP; s++) {
            const float t = offdiag(s, wm[s], wp[s], 0.f);
            const float u = val[s] ? r[s] - t : 0.f;
            const float zs = row_bcast_dot<NX>(0.f, u, pd[s]);
            zz[s] = val[s] ? zs : 0.f;
            if (val[s]) sz[(kk[s] + 1) * XP + l] = zz[s];
            loc += r[s] * zz[s];
        }
        return loc;
    };
    int slot = 0;
    float rho = bsum(precond(z), red + 64 * slot);
    slot ^= 1;
#pragma unroll
    for (int s = 0; s < NP; s++) {
        p[s] = z[s];
        pm[s] = sz[kk[s] * XP + lane];
        pp[s] = sz[(kk[s] + 2) * XP + lane];
    }
    if (fabsf(rho) < a.pcg_abs_tol) return 0;
    const float rho0 = fabsf(rho);
    int it = 0;
    for (int i = 0; i < a.max_pcg; i++) {
        it++;
        float Ap[NP], loc = 0.f;
#pragma unroll
        for (int s = 0; s < NP; s++) {
            const float a_ = offdiag(s, pm[s], pp[s], row_bcast_dot<NX>(0.f, p[s], sd[s]));
            Ap[s] = val[s] ? a_ : 0.f;
            loc += p[s] * Ap[s];
            const float q = row_bcast_dot<NX>(0.f, Ap[s], pd[s]);  // Pd Ap, for the neighbours' copies of w
            if (val[s]) sw[(kk[s] + 1) * XP + l] = q;
        }
        // quotients by reciprocal and product, as the reference's fast-math build divides (-use_fast_math)
        const float alpha = rho * __builtin_amdgcn_rcpf(bsum(loc, red + 64 * slot));
        slot ^= 1;
#pragma unroll
        for (int s = 0; s < NP; s++) {
            x[s] += alpha * p[s];
            r[s] -= alpha * Ap[s];
        }
        if (it >= a.max_pcg) break;  // x is final at the cap: its last preconditioner application is not needed
#pragma unroll
        for (int s = 0; s < NP; s++) {
            wm[s] -= alpha * sw[kk[s] * XP + lane];
            wp[s] -= alpha * sw[(kk[s] + 2) * XP + lane];
        }
        const float rn = bsum(precond(z), red + 64 * slot);
        slot ^= 1;
        if (fabsf(rn) < a.pcg_abs_tol + eps * rho0) break;
        const float beta = rn * __builtin_amdgcn_rcpf(rho);
        rho = rn;
#pragma unroll
        for (int s = 0; s < NP; s++) {
            p[s] = z[s] + beta * p[s];
            pm[s] = sz[kk[s] * XP + lane] + beta * pm[s];
            pp[s] = sz[(kk[s] + 2) * XP + lane] + beta * pp[s];
        }
    }
#pragma unroll
    for (int s = 0; s < NP; s++)
        if (val[s]) v.lam[kk[s] * NX + l] = x[s];
    return it;
}

// Work-group sums of two values with one barrier, each summed exactly as block_sum1 sums one. red: 64 floats per
// call, double-buffered by the caller like block_sum1's.
__device__ __forceinline__ float2 block_sum2(float v1, float v2, float* red) {
    const int rows = blockDim.x >> 4;
    v1 = row16_sum(v1);
    v2 = row16_sum(v2);
    if ((threadIdx.x & 15) == 0) {
        red[threadIdx.x >> 4] = v1;
        red[32 + (threadIdx.x >> 4)] = v2;
    }
    __syncthreads();
    const int l = threadIdx.x & 63;
    float s1 = l < rows ? red[l] : 0.f, s2 = l < rows ? red[32 + l] : 0.f;
    s1 = row16_sum(s1);
    s2 = row16_sum(s2);
    return make_float2(lane_f(s1, 0) + lane_f(s1, 16), lane_f(s2, 0) + lane_f(s2, 16));
}

// The backward pass's preconditioner (no reference counterpart): a twisted block-tridiagonal factorization of
// the Schur complement, eliminating from both ends towards the middle knot m = N / 2:
//   top     D_0 = S_00,          G_k = S_{k,k-1} D_{k-1}^-1,    D_k = S_kk - G_k S_{k,k-1}^T   (k = 1 .. m),
//   bottom  E_{N-1} = S_{N-1,N-1}, H_k = S_{k,k+1} E_{k+1}^-1,  E_k = S_kk - H_k S_{k+1,k}     (k = N-2 .. m),
//   middle  C_m = D_m + E_m - S_mm.
// Kept as Gm[k] = G_k (1 <= k <= m), Gm[k + 1] = H_k (m <= k <= N - 2) and Di[k] = D_k^-1 (k < m), C_m^-1 (k = m),
// E_k^-1 (k > m), the inverses by Gauss-Jordan without pivoting (S is definite, and so are D_k, E_k and C_m). The
// factors give S exactly up to fp32 rounding, so a PCG preconditioned with them (btd_apply) stops after one or two
// iterations where the reference's preconditioner P takes 120 to 240 on the same systems at tolerance 1e-6, and
// reaches the same fp32 accuracy. The two eliminations run side by side, each sequential over its half of the
// knots: one problem per 128-thread work-group, the first 16 lanes of wave 0 (top) and wave 1 (bottom) working
// (lane = row), the next knot's rows loaded a step ahead. Gm, Di: B x N blocks of NX x NX floats (row-major).
template <int PL>
__global__ void __launch_bounds__(128) k_btd_factor(Args a, float* Gm, float* Di) {
    constexpr int NX = Plant<PL>::NX, TS = 16;
    __shared__ float tiles[2][NX * TS];  // the bottom side's transpose of So_{k+1}
    __shared__ float pivot[2][NX * TS];  // D_m (top) and E_m (bottom)
    const int b = blockIdx.x, side = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int N = a.N, m = N / 2, r = lane < NX ? lane : 0;
    const size_t blk = (size_t)NX * NX, base = (size_t)b * N * blk;
    const float* Sd = a.Sd + base;
    const float* So = a.So + base;
    float* Gp = Gm + base;
    float* Dp = Di + base;
    float* t2 = tiles[side];
    // in-place Gauss-Jordan inverse, lane = row (lanes 0-15 of the wave: one DPP row), the pivot row broadcast by
    // DPP (gj_pivot_dpp)
    auto invert = [&](float* row) { gj_dpp<NX>(lane, row, std::make_integer_sequence<int, NX>{}); };
    auto store = [&](float* dst, const float* v) {
        if (lane < NX) {
#pragma unroll
            for (int c = 0; c < NX; c++) dst[r * NX + c] = v[c];
        }
    };
    float row[NX];
    if (lane < KG) {
        const int first = side == 0 ? 0 : N - 1, steps = side == 0 ? m : N - 1 - m;
#pragma unroll
        for (int c = 0; c < NX; c++) row[c] = Sd[first * blk + r * NX + c];
        float son[NX], sdn[NX];  // the next step's block rows
        {
            const int k1 = side == 0 ? (steps > 0 ? 1 : 0) : (steps > 0 ? N - 1 : 0);  // top: S_{1,0}; bottom: S_{N-1,N-2}
            const int d1 = side == 0 ? (steps > 0 ? 1 : 0) : (steps > 0 ? N - 2 : 0);
#pragma unroll
            for (int c = 0; c < NX; c++) {
                son[c] = So[k1 * blk + r * NX + c];
                sdn[c] = Sd[d1 * blk + r * NX + c];
            }
        }
        for (int s = 0; s < steps; s++) {
            // top: knot k = s + 1, off-diagonal block S_{k,k-1} = So_k; bottom: knot k = N - 2 - s, S_{k+1,k} = So_{k+1}
            const int k = side == 0 ? s + 1 : N - 2 - s;
            float so[NX], sd[NX], g[NX];
            {
                const int kn = side == 0 ? (k + 1 <= m ? k + 1 : k) : (k - 1 >= m ? k - 1 : k);
                const int kon = side == 0 ? kn : kn + 1;
#pragma unroll
                for (int c = 0; c < NX; c++) {
                    so[c] = son[c];
                    sd[c] = sdn[c];
                    son[c] = So[kon * blk + r * NX + c];
                    sdn[c] = Sd[kn * blk + r * NX + c];
                }
            }
            invert(row);  // D_{k-1}^-1 (top) or E_{k+1}^-1 (bottom)
            store(Dp + (side == 0 ? k - 1 : k + 1) * blk, row);
            // the block products by DPP row broadcasts from the group's lanes (lane j holds row j of the inverse and
            // of So), each sum over j in order as the products were formed before
            float t[NX];
#pragma unroll
            for (int c = 0; c < NX; c++) {
                g[c] = 0.f;
                t[c] = 0.f;
            }
            if (side == 0) {
                row_bcast_dots_all<NX, NX>(g, row, so);  // G_k = So_k D_{k-1}^-1: g[c] = sum_j So_k[r][j] D^-1[j][c]
                lane_dots_all<NX>(t, so, g);             // (G_k So_k^T)[r][c] = sum_j G_k[r][j] So_k[c][j]
            } else {
                // H_k = So_{k+1}^T E_{k+1}^-1 needs column r of So_{k+1}: a transpose through the wave's LDS tile
                if (lane < NX) {
#pragma unroll
                    for (int c = 0; c < NX; c++) t2[lane * TS + c] = so[c];
                }
                wave_sync();
                float sc[NX];
#pragma unroll
                for (int j = 0; j < NX; j++) sc[j] = t2[j * TS + r];
                row_bcast_dots_all<NX, NX>(g, row, sc);  // g[c] = sum_j So_{k+1}[j][r] E^-1[j][c]
                row_bcast_dots_all<NX, NX>(t, so, g);    // (H_k So_{k+1})[r][c] = sum_j H_k[r][j] So_{k+1}[j][c]
            }
            store(Gp + (size_t)(side == 0 ? k : k + 1) * blk, g);
#pragma unroll
            for (int c = 0; c < NX; c++) row[c] = sd[c] - t[c];
            wave_sync();
        }
        if (lane < NX) {
#pragma unroll
            for (int c = 0; c < NX; c++) pivot[side][lane * TS + c] = row[c];
        }
    }
    __syncthreads();
    if (side == 0 && lane < KG) {  // C_m = D_m + E_m - S_mm
#pragma unroll
        for (int c = 0; c < NX; c++) row[c] = pivot[0][r * TS + c] + pivot[1][r * TS + c] - Sd[m * blk + r * NX + c];
        invert(row);
        store(Dp + (size_t)m * blk, row);
    }
}

// z = M^-1 r with the twisted factorization of k_btd_factor, in three phases separated by work-group barriers:
// btd_ends (both ends towards m: top y_k = r_k - G_k y_{k-1}, w_k = D_k^-1 y_k for k < m; bottom u_k = r_k -
// H_k u_{k+1}, v_k = E_k^-1 u_k for k > m; y_m and u_m to LDS), btd_middle (z_m = C_m^-1 (y_m + u_m - r_m)) and
// btd_out (from m outwards: z_k = w_k - G_{k+1}^T z_{k+1} above, z_k = v_k - H_{k-1}^T z_{k-1} below). Each side
// is one 16-lane knot group (lane = row; knot groups are the waves' DPP rows); each step's block row times the vector takes the vector's entries from the group's lanes by DPP row broadcasts
// (row_bcast_dot, summing in column order) and the next step's rows are read a step ahead. rv, zv: padded LDS
// vectors (knot k at (k + 1) * 16); Gp, Dp: the problem's factor blocks (in LDS in the paired PCG); pv: 16
// floats of LDS for the side's vector at m.
template <int PL>
__device__ __forceinline__ void btd_ends(bool top, int lane, int N, const float* Gp, const float* Dp,
                                         const float* rv, float* zv, float* pv) {
    constexpr int NX = Plant<PL>::NX, XP = 16;
    const int blk = NX * NX, r = lane < NX ? lane : 0, m = N / 2;
    const int steps = top ? m + 1 : N - m;
    auto knot = [&](int s) { return top ? s : N - 1 - s; };
    auto mult = [&](int k) { const int g = top ? k : k + 1; return g < N ? g : N - 1; };  // G_k or H_k
    float gc[NX], dc[NX], gn[NX], dn[NX];
    int k = knot(0);
#pragma unroll
    for (int c = 0; c < NX; c++) {
        gc[c] = 0.f;
        dc[c] = Dp[k * blk + r * NX + c];
    }
    float rc = rv[(k + 1) * XP + r], y = 0.f;
    for (int s = 0; s < steps; s++) {
        k = knot(s);
        const int kn = knot(s + 1 < steps ? s + 1 : s);
#pragma unroll
        for (int c = 0; c < NX; c++) {
            gn[c] = Gp[mult(kn) * blk + r * NX + c];
            dn[c] = Dp[kn * blk + r * NX + c];
        }
        const float rn = rv[(kn + 1) * XP + r];
        const float t = row_bcast_dot<NX>(0.f, y, gc);  // gc = 0 at the first step
        y = rc - t;
        if (k != m) {
            const float w = row_bcast_dot<NX>(0.f, y, dc);
            if (lane < NX) zv[(k + 1) * XP + lane] = w;
        } else if (lane < NX) {
            pv[lane] = y;
        }
#pragma unroll
        for (int c = 0; c < NX; c++) {
            gc[c] = gn[c];
            dc[c] = dn[c];
        }
        rc = rn;
    }
}

template <int PL>
__device__ __forceinline__ void btd_middle(int lane, int N, const float* Dp, const float* rv, float* zv,
                                           const float* pt, const float* pb) {
    constexpr int NX = Plant<PL>::NX, XP = 16;
    const int r = lane < NX ? lane : 0, m = N / 2;
    const float t = pt[r] + pb[r] - rv[(m + 1) * XP + r];
    float dm[NX];
#pragma unroll
    for (int c = 0; c < NX; c++) dm[c] = Dp[m * NX * NX + r * NX + c];
    const float z = row_bcast_dot<NX>(0.f, t, dm);
    if (lane < NX) zv[(m + 1) * XP + lane] = z;
}

template <int PL>
__device__ __forceinline__ void btd_out(bool top, int lane, int N, const float* Gp, float* zv) {
    constexpr int NX = Plant<PL>::NX, XP = 16;
    const int blk = NX * NX, r = lane < NX ? lane : 0, m = N / 2;
    const int steps = top ? m : N - 1 - m;
    if (steps <= 0) return;
    auto knot = [&](int s) { return top ? m - 1 - s : m + 1 + s; };
    auto col = [&](int k) { return top ? k + 1 : k; };  // G_{k+1} above, H_{k-1} = Gm[k] below
    float x = zv[(m + 1) * XP + r];
    float gc[NX], gn[NX];
#pragma unroll
    for (int j = 0; j < NX; j++) gc[j] = Gp[col(knot(0)) * blk + j * NX + r];
    for (int s = 0; s < steps; s++) {
        const int k = knot(s), kn = knot(s + 1 < steps ? s + 1 : s);
#pragma unroll
        for (int j = 0; j < NX; j++) gn[j] = Gp[col(kn) * blk + j * NX + r];
        const float t = row_bcast_dot<NX>(0.f, x, gc);
        x = zv[(k + 1) * XP + r] - t;
        if (lane < NX) zv[(k + 1) * XP + lane] = x;
#pragma unroll
        for (int j = 0; j < NX; j++) gc[j] = gn[j];
    }
}

// z = M^-1 r with the twisted factorization: btd_ends on knot groups 0 (top) and 4 (bottom; group 1 in a
// work-group of fewer than eight groups), btd_middle, btd_out, separated by work-group barriers (every thread of the
// work-group calls this). r in the padded LDS vector rv, z to zv, pv: 32 floats of LDS.
template <int PL>
__device__ void btd_apply(int N, const float* Gp, const float* Dp, const float* rv, float* zv, float* pv) {
    const int g = threadIdx.x / KG, lane = threadIdx.x % KG, gb = blockDim.x / KG >= 8 ? 4 : 1;
    if (g == 0) btd_ends<PL>(true, lane, N, Gp, Dp, rv, zv, pv);
    if (g == gb) btd_ends<PL>(false, lane, N, Gp, Dp, rv, zv, pv + 16);
    __syncthreads();
    if (g == 0) btd_middle<PL>(lane, N, Dp, rv, zv, pv, pv + 16);
    __syncthreads();
    if (g == 0) btd_out<PL>(true, lane, N, Gp, zv);
    if (g == gb) btd_out<PL>(false, lane, N, Gp, zv);
    __syncthreads();
}

// Two PCG solves on the same S in lockstep (the backward pass's step and adjoint chains), preconditioned with the
// factorization of S (k_btd_factor; btd_ends, btd_middle, btd_out): the rows of S are loaded once, every iteration
// does both mat-vecs, each pair of dot products shares one barrier, and the four sweep sides (top and bottom of
// each chain) run side by side on knot groups 0, 4, 8, 12 (separate waves; fewer waves in smaller work-groups)
// from a copy of the problem's factor blocks in LDS (2 N NX^2 floats, 49 KB for iiwa14 at N = 32). Solve i has its own right-hand side
// a_i.gam, warm start and result a_i.lam, tolerance a_i.tol[b], absolute tolerance, iteration cap and count, and
// the iteration of phase_pcg_reg; once it stops (its test or its cap) its vectors are left alone while the other
// goes on. run_i = false skips solve i (count 0, lambda untouched). One knot per knot group (N <= 32).
template <int PL>
__device__ __forceinline__ void phase_pcg_reg2(const Args& a1, const Args& a2, int b, bool run1, bool run2,
                                               const float* Gm, const float* Di, float* sm, int* n1, int* n2) {
    constexpr int NX = Plant<PL>::NX, XP = 16;
    View<PL> v1(a1, b), v2(a2, b);
    const int N = a1.N, tid = threadIdx.x, k = tid / KG, lane = tid % KG;
    float* sp1 = sm;
    float* sr1 = sp1 + (N + 2) * XP;
    float* sp2 = sr1 + (N + 2) * XP;
    float* sr2 = sp2 + (N + 2) * XP;
    float* red = sr2 + (N + 2) * XP;  // 2 x 64 floats: two double-buffered pairs of reduction slots
    float* sz1 = red + 128;
    float* sz2 = sz1 + (N + 2) * XP;
    float* sG = sz2 + (N + 2) * XP;  // the problem's G_k and D_k^-1 blocks, copied to LDS for the sweeps
    float* sD = sG + N * NX * NX;
    float* pv = sD + N * NX * NX;  // 4 x 16 floats: each sweep side's vector at the middle knot
    for (int i = tid; i < 4 * (N + 2) * XP; i += blockDim.x) sm[i] = 0.f;
    {
        const size_t pb = (size_t)b * N * NX * NX;
        const float4* g4 = reinterpret_cast<const float4*>(Gm + pb);
        const float4* d4 = reinterpret_cast<const float4*>(Di + pb);
        float4* sg4 = reinterpret_cast<float4*>(sG);
        float4* sd4 = reinterpret_cast<float4*>(sD);
        for (int i = tid; i < N * NX * NX / 4; i += blockDim.x) {
            sg4[i] = g4[i];
            sd4[i] = d4[i];
        }
    }
    const bool val = (lane < NX) && (k < N);
    const int l = lane < NX ? lane : 0;
    const int kk = k < N ? k : 0;
    const int G = blockDim.x / KG;
    // knot groups of the four sweep sides (chain 1 top, bottom, chain 2 top, bottom): one per wave when possible
    const int rg[4] = {0, G >= 16 ? 4 : 1, G >= 16 ? 8 : G >= 8 ? 4 : 2, G >= 16 ? 12 : G >= 8 ? 5 : 3};
    float sd[NX], sl[NX], su[NX];
    {
        const float* D = v1.Sd + (size_t)kk * NX * NX + l * NX;
        const float* L = v1.So + (size_t)kk * NX * NX + l * NX;
        const float* U = v1.So + (size_t)(kk + 1 < N ? kk + 1 : kk) * NX * NX + l;
        const bool hasl = kk > 0, hasu = kk < N - 1;
#pragma unroll
        for (int c = 0; c < NX; c++) {
            const float l0 = L[c], u0 = U[c * NX];  // in bounds: loaded unconditionally, selected below
            sd[c] = D[c];
            sl[c] = hasl ? l0 : 0.f;
            su[c] = hasu ? u0 : 0.f;
        }
    }
    auto blk = [&](const float* M, const float4* xb, float s) {
        float xs[XP];
#pragma unroll
        for (int q = 0; q < XP / 4; q++) {
            const float4 t = xb[q];
            xs[4 * q] = t.x; xs[4 * q + 1] = t.y; xs[4 * q + 2] = t.z; xs[4 * q + 3] = t.w;
        }
#pragma unroll
        for (int c = 0; c < NX; c++) s += M[c] * xs[c];
        return s;
    };
    auto mv = [&](const float* M0, const float* M1, const float* M2, const float* x) {
        const float4* xk = reinterpret_cast<const float4*>(x + (kk + 1) * XP);
        float s = blk(M0, xk, 0.f);
        s = blk(M1, xk - XP / 4, s);
        return blk(M2, xk + XP / 4, s);
    };
    const int o = (kk + 1) * XP + l;
    float x1 = (val && run1) ? v1.lam[kk * NX + l] : 0.f, x2 = (val && run2) ? v2.lam[kk * NX + l] : 0.f;
    __syncthreads();
    if (val) {
        sp1[o] = x1;
        sp2[o] = x2;
    }
    __syncthreads();
    float r1 = (val && run1) ? v1.gam[kk * NX + l] - mv(sd, sl, su, sp1) : 0.f;
    float r2 = (val && run2) ? v2.gam[kk * NX + l] - mv(sd, sl, su, sp2) : 0.f;
    if (val) {
        sr1[o] = r1;
        sr2[o] = r2;
    }
    __syncthreads();
    // z_i = M^-1 r_i: the top and bottom sides of both chains on four knot groups (four waves when there are
    // 16 groups or more) while the others wait at the barriers
    const int role = k == rg[0] ? 0 : k == rg[1] ? 1 : k == rg[2] ? 2 : k == rg[3] ? 3 : -1;
    const bool rtop = (role & 1) == 0;
    const float* rrv = role < 2 ? sr1 : sr2;
    float* rzv = role < 2 ? sz1 : sz2;
    // With 16 groups or more each sweep side has a wave of its own. The sweeps' block products broadcast within the
    // knot group's DPP row, so sides whose groups share a wave (fewer groups) run side by side too.
    auto precond = [&](bool d1, bool d2, float& z1o, float& z2o) {
        const bool mine = role >= 0 && (role < 2 ? d1 : d2);
        if (mine) btd_ends<PL>(rtop, lane, N, sG, sD, rrv, rzv, pv + role * XP);
        __syncthreads();
        if (mine && rtop) btd_middle<PL>(lane, N, sD, rrv, rzv, pv + role * XP, pv + (role + 1) * XP);
        __syncthreads();
        if (mine) btd_out<PL>(rtop, lane, N, sG, rzv);
        __syncthreads();
        z1o = (val && d1) ? sz1[o] : 0.f;
        z2o = (val && d2) ? sz2[o] : 0.f;
    };
    float z1, z2;
    precond(run1, run2, z1, z2);
    float p1 = z1, p2 = z2;
    int slot = 0;
    const float2 rho = block_sum2(r1 * z1, r2 * z2, red + 64 * slot);
    slot ^= 1;
    float rho1 = rho.x, rho2 = rho.y;
    bool act1 = run1 && a1.max_pcg > 0 && !(fabsf(rho1) < a1.pcg_abs_tol);
    bool act2 = run2 && a2.max_pcg > 0 && !(fabsf(rho2) < a2.pcg_abs_tol);
    const float lim1 = a1.pcg_abs_tol + a1.tol[b] * fabsf(rho1), lim2 = a2.pcg_abs_tol + a2.tol[b] * fabsf(rho2);
    int it1 = 0, it2 = 0;
    while (act1 || act2) {  // act1, act2 come from work-group sums: uniform, so are the branches below
        if (act1) it1++;
        if (act2) it2++;
        if (val) {
            if (act1) sp1[o] = p1;
            if (act2) sp2[o] = p2;
        }
        __syncthreads();
        const float Ap1 = (val && act1) ? mv(sd, sl, su, sp1) : 0.f;
        const float Ap2 = (val && act2) ? mv(sd, sl, su, sp2) : 0.f;
        const float2 pap = block_sum2(p1 * Ap1, p2 * Ap2, red + 64 * slot);
        slot ^= 1;
        if (act1) {
            const float alpha = rho1 * __builtin_amdgcn_rcpf(pap.x);
            x1 += alpha * p1;
            r1 -= alpha * Ap1;
            if (val) sr1[o] = r1;
        }
        if (act2) {
            const float alpha = rho2 * __builtin_amdgcn_rcpf(pap.y);
            x2 += alpha * p2;
            r2 -= alpha * Ap2;
            if (val) sr2[o] = r2;
        }
        __syncthreads();
        // a solve at its cap is done: its x is final, so its last preconditioner application and test are skipped
        act1 = act1 && it1 < a1.max_pcg;
        act2 = act2 && it2 < a2.max_pcg;
        if (!act1 && !act2) break;  // the flags come from work-group sums: uniform
        precond(act1, act2, z1, z2);
        const float2 rz = block_sum2(r1 * z1, r2 * z2, red + 64 * slot);
        slot ^= 1;
        if (act1) {
            if (fabsf(rz.x) < lim1) {
                act1 = false;
            } else {
                const float beta = rz.x * __builtin_amdgcn_rcpf(rho1);
                rho1 = rz.x;
                p1 = z1 + beta * p1;
                act1 = it1 < a1.max_pcg;
            }
        }
        if (act2) {
            if (fabsf(rz.y) < lim2) {
                act2 = false;
            } else {
                const float beta = rz.y * __builtin_amdgcn_rcpf(rho2);
                rho2 = rz.y;
                p2 = z2 + beta * p2;
                act2 = it2 < a2.max_pcg;
            }
        }
    }
    if (val && run1) v1.lam[kk * NX + l] = x1;
    if (val && run2) v2.lam[kk * NX + l] = x2;
    *n1 = it1;
    *n2 = it2;
}

// register-resident PCG when every knot group owns a single knot (N <= 32); with two knots per group the problem is
// held on the CU (phase_pcg_res over blockDim.x / 16 <= 32 groups; the caller provides pcg_res_lds_floats(N) of LDS,
// at most 64 KB for N <= 64); else S and P stream
template <int PL, int NPASS>
__device__ __forceinline__ int pcg_dispatch(const Args& a, int b, float eps, float* sm) {
    if constexpr (NPASS == 1) return phase_pcg_reg<PL>(a, b, eps, sm);
    else if constexpr (NPASS == 0) return phase_pcg_res<PL>(a, b, eps, sm);
    else if constexpr (NPASS == 2) return phase_pcg_res<PL, 2, 0>(a, b, eps, sm);
    else return phase_pcg<PL, NPASS>(a, b, eps, sm);
}

// The persistent kernels' PCG: as pcg_dispatch, but with two or four knots per knot group the problem stays resident
// (phase_pcg_res; the caller provides pcg_res_lds_floats(N) of LDS); eight knots per group would not fit in
// registers and stream as phase_pcg.
template <int PL, int NPASS>
__device__ __forceinline__ int pcg_dispatch_fused(const Args& a, int b, float eps, float* sm) {
    if constexpr (NPASS == 2 || NPASS == 4) return phase_pcg_res<PL, NPASS, 0>(a, b, eps, sm);
    else return pcg_dispatch<PL, NPASS>(a, b, eps, sm);
}

template <int PL, typename S = float>
__device__ __forceinline__ S phase_dz(const ArgsT<S>& a, int b, S* sm) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU;
    View<PL, S> v(a, b);
    const int N = a.N, g = threadIdx.x / KG, lane = threadIdx.x % KG, G = blockDim.x / KG;
    S* shv = sm + g * KG;
    S* red = sm + G * KG;
    S res = S(0);
    for (int k = g; k < N; k += G) {
        const S* lam1 = v.lam + (k < N - 1 ? (k + 1) * NX : 0);
        res = fmaxf(res, dz_knot<PL, S>(k, N, lane, v.A + (size_t)k * NX * NX, v.Bm + (size_t)k * NX * NU,
                                     v.Qip + k * NQ * NQ, v.Qiv + k * NQ, v.Ri + k * NU, v.q + k * NX, v.r + k * NU,
                                     v.c + k * NX, v.lam + k * NX, lam1, v.dz + (size_t)k * (NX + NU), shv));
    }
    return block_max(res, red);
}

// merit at nal step lengths alpha_j = 2^-j (merit.cuh:16-92): per-(alpha, knot) partial terms,
// then fixed-order sums over knots. out[j] (LDS) receives phi(alpha_j). The reference's
// initial / final merit (dz = 0, one alpha) is nal = 1. The partial terms go to part (LDS, nal * N values), from
// which the sums read them, and to mpart (global memory, kept for inspection).
template <int PL, typename S = float>
__device__ __forceinline__ void phase_merit(const ArgsT<S>& a, int b, S mu, int nal, S* out, S* part) {
    using P = Plant<PL>;
    constexpr int NX = P::NX, NU = P::NU;
    View<PL, S> v(a, b);
    const int N = a.N;
    for (int it = threadIdx.x; it < N * nal; it += blockDim.x) {
        const int k = it % N, j = it / N;
        const S alpha = S(1) / (S)(1 << j);
        const size_t off = (size_t)k * (NX + NU);
        const S m = merit_item<PL, S>(k, N, alpha, a.dt, a.cw, mu, v.xu + off, v.dz + off, v.ref + 6 * k,
                                      v.xref ? v.xref + k * NX : nullptr, v.Qd, v.Rd, v.fext, v.xu, v.dz, v.xs);
        v.mpart[j * N + k] = m;
        part[j * N + k] = m;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)nal) {
        S s = S(0);
        for (int k = 0; k < N; k++) s += part[threadIdx.x * N + k];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// line search + rho adaptation + update (line_search.cuh:12-98). merits in LDS; returns via shared.
template <int PL, typename S = float>
__device__ __forceinline__ void phase_ls(const ArgsT<S>& a, int b, int it, const S* merits, S* shs) {
    constexpr int NX = Plant<PL>::NX, NU = Plant<PL>::NU;
    View<PL, S> v(a, b);
    if (threadIdx.x == 0) {
        // the problem's scalars read once, together (interleaved with the stores below, each read would be a round
        // trip of its own: the compiler cannot tell that the arrays do not overlap)
        const S phi0 = a.phi[b], drho0 = a.drho[b], rho0 = a.rho[b];
        S best = S(1e38f);
        int jb = 0;
        for (int j = 0; j < NUM_ALPHAS; j++)
            if (merits[j] < best) { best = merits[j]; jb = j; }
        const bool ok = best < phi0;
        S rho = rho0, phi = phi0;
        if (a.adapt_rho) {
            const S mult = ok ? smin(drho0 / S(RHO_FACTOR), S(1) / S(RHO_FACTOR)) : smax(drho0 * S(RHO_FACTOR), S(RHO_FACTOR));
            a.drho[b] = mult;
            rho = smin(smax(rho0 * mult, S(RHO_MIN)), S(RHO_MAX));
        }
        S step = -S(1);
        if (!ok) {
            if (rho > S(RHO_MAX)) rho = S(RHO_INIT);
        } else {
            step = S(1) / (S)(1 << jb);
            phi = best;
        }
        a.rho[b] = rho;
        a.phi[b] = phi;
        a.st_step[it * a.B + b] = step;
        a.st_lsm[it * a.B + b] = phi;
        shs[0] = step;
    }
    __syncthreads();
    const S step = shs[0];
    if (step > S(0))
        for (int i = threadIdx.x; i < a.traj; i += blockDim.x) v.xu[i] += step * v.dz[i];
    (void)NX; (void)NU;
}

// ------------------------------------------------------------------ plain path kernels
// which: 0 initial merit, 1 final merit, 2 line search, 3 line search preceded by the batch-level exit test
// (bsqp.cuh:165-167) on the converged count solved[it] that k_pcg_dz accumulated: every work-group makes the
// same decision from it, and work-group 0 raises the device stop flag for the launches that follow.
template <int PL, typename S = float>
__device__ __forceinline__ void merit_body(const ArgsT<S>& a, int nal, int which, int it) {
    S* sm = dyn_lds<S>();
    const int b = blockIdx.x;
    if (which >= 2 && stopped(a)) return;
    if (which == 3 && (double)a.solved[it] >= (double)a.B * (double)a.solve_ratio) {
        if (b == 0 && threadIdx.x == 0) *a.stop_w = 1;
        return;
    }
    if (which >= 2 && a.exit_rule == GATO_EXIT_KKT && !a.active[b]) return;
    phase_merit<PL, S>(a, b, a.mu[b], nal, sm, sm + 2 * NUM_ALPHAS);
    if (threadIdx.x == 0 && which < 2) {
        a.phi[b] = sm[0];
        (which == 0 ? a.st_init : a.st_final)[b] = sm[0];
    }
    if (which >= 2) phase_ls<PL, S>(a, b, it, sm, sm + NUM_ALPHAS);
}

template <int PL, typename S = float>
__global__ void GATO_PHASE_BOUNDS(S) k_merit(ArgsT<S> a, int nal, int which, int it) {
    merit_body<PL, S>(a, nal, which, it);
}

// The same in a 256-thread work-group (one thread per (step length, knot) item, looping when N * NUM_ALPHAS > 256),
// which leaves room for 256 VGPRs at two work-groups per CU, so the forward dynamics of an item do not spill. Each
// item and each step length's sum over knots are computed as in k_merit: the results are the same bit for bit.
template <int PL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_merit_n(Args a, int nal, int which,
                                                                                         int it) {
    merit_body<PL>(a, nal, which, it);
}

// Solve-start state of the batch (the device stop flag, the converged counts) and of problem b (flags, SQP
// count, the per-iteration records) for the solve of bsqp.cuh:111-114, with the records this build keeps.
template <typename S>
__device__ __forceinline__ void begin_batch(const ArgsT<S>& a) {
    if (a.stop_w) *a.stop_w = 0;
    for (int i = 0; i < a.max_sqp; i++) a.solved[i] = 0;
}

template <typename S>
__device__ __forceinline__ void begin_problem(const ArgsT<S>& a, int b) {
    a.conv[b] = 0;
    a.active[b] = 1;
    a.st_sqp[b] = 0;
    for (int i = 0; i < a.max_sqp; i++) {
        a.st_pcg[i * a.B + b] = 0;
        a.st_kres[i * a.B + b] = S(0);
        a.st_lsm[i * a.B + b] = S(0);
        a.st_step[i * a.B + b] = S(0);
    }
}

// The graph path for N <= 32 runs a solve as k_first_n, then per SQP iteration k_pcg_dz and either
// k_ls_lin_schur_n or, after the last, k_last_n: every launch does one problem's consecutive phases in one
// 256-thread work-group (room for 256 VGPRs at two work-groups per CU, as k_merit_n and k_schur_n), so the
// whole solve takes 2 * iterations + 1 launches.

// Solve start (k_begin, the clear of dz and the initial merit of bsqp.cuh:111-118 for problem b), then the
// linearization and Schur phase of SQP iteration 0.
template <int PL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_first_n(Args a, int ts) {
    extern __shared__ __align__(16) float sm[];
    constexpr int NU = Plant<PL>::NU;
    const int b = blockIdx.x;
    View<PL> v(a, b);
    if (threadIdx.x == 0) {
        if (b == 0) begin_batch(a);
        begin_problem(a, b);
    }
    for (int i = threadIdx.x; i < a.traj + NU; i += blockDim.x) v.dz[i] = 0.f;
    __syncthreads();
    phase_merit<PL>(a, b, a.mu[b], 1, sm, sm + 2 * NUM_ALPHAS);
    if (threadIdx.x == 0) {
        a.phi[b] = sm[0];
        a.st_init[b] = sm[0];
    }
    __syncthreads();
    phase_lin<PL>(a, b, sm, ts);
    __syncthreads();
    phase_schur<PL>(a, b, a.rho[b], sm);
}

// The batch-level exit test and line search of SQP iteration it - 1 (merit_body, which == 3), then the
// linearization and Schur phase of iteration it on the trajectory the line search has just updated.
template <int PL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_ls_lin_schur_n(Args a, int it, int ts) {
    extern __shared__ __align__(16) float sm[];
    const int b = blockIdx.x;
    // the batch's and the problem's flags read together, ahead of the tests that use them
    const int stop = a.stop ? *a.stop : 0, solved = a.solved[it - 1], act = a.active[b];
    const float mu = a.mu[b];
    if (stop) return;
    if ((double)solved >= (double)a.B * (double)a.solve_ratio) {
        if (b == 0 && threadIdx.x == 0) *a.stop_w = 1;
        return;
    }
    if (a.exit_rule == GATO_EXIT_KKT && !act) return;
    phase_merit<PL>(a, b, mu, NUM_ALPHAS, sm, sm + 2 * NUM_ALPHAS);
    phase_ls<PL>(a, b, it - 1, sm, sm + NUM_ALPHAS);
    __syncthreads();  // the updated trajectory and rho are settled, and the LDS is free for the linearization
    phase_lin<PL>(a, b, sm, ts);
    __syncthreads();
    phase_schur<PL>(a, b, a.rho[b], sm);
}

// The exit test and line search of the last SQP iteration it (skipped when the batch has stopped, as in
// k_ls_lin_schur_n), then solve end for problem b: the clear of dz, the final merit (bsqp.cuh:179-182) and the
// reset of the rho multiplier (k_end).
template <int PL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_last_n(Args a, int it) {
    extern __shared__ __align__(16) float sm[];
    constexpr int NU = Plant<PL>::NU;
    const int b = blockIdx.x;
    View<PL> v(a, b);
    if (!stopped(a) && (double)a.solved[it] < (double)a.B * (double)a.solve_ratio &&
        !(a.exit_rule == GATO_EXIT_KKT && !a.active[b])) {
        phase_merit<PL>(a, b, a.mu[b], NUM_ALPHAS, sm, sm + 2 * NUM_ALPHAS);
        phase_ls<PL>(a, b, it, sm, sm + NUM_ALPHAS);
        __syncthreads();  // the line search has read dz
    }
    for (int i = threadIdx.x; i < a.traj + NU; i += blockDim.x) v.dz[i] = 0.f;
    // The final merit (bsqp.cuh:179-182) at the returned trajectory with dz = 0 is phi: the line search stored the merit
    // of the point it accepted, whose every entry it formed as the trajectory update does (z + alpha dz, one fma), and
    // left phi unchanged when it kept the trajectory; the solve start's phi is the merit at the initial trajectory.
    // Reading it back skips a merit pass with the same result.
    if (threadIdx.x == 0) {
        a.st_final[b] = a.phi[b];
        a.drho[b] = a.drho_init[b];  // bsqp.cuh:189
    }
}

template <int PL, typename S = float>
__global__ void GATO_PHASE_BOUNDS(S) k_lin(ArgsT<S> a) {
    S* sm = dyn_lds<S>();
    const int b = blockIdx.x;
    if (stopped(a) || (a.exit_rule == GATO_EXIT_KKT && !a.active[b])) return;
    phase_lin<PL, S>(a, b, sm, blockDim.x / KG);
}

template <int PL, typename S = float>
__global__ void GATO_PHASE_BOUNDS(S) k_schur(ArgsT<S> a) {
    S* sm = dyn_lds<S>();
    const int b = blockIdx.x;
    if (stopped(a) || (a.exit_rule == GATO_EXIT_KKT && !a.active[b])) return;
    phase_schur<PL, S>(a, b, a.rho[b], sm);
}

// The same in 256-thread work-groups (16 knot groups, each looping over N / 16 knots), with room for 256 VGPRs at
// two work-groups per CU, so the Schur products do not spill; every knot's arithmetic is k_schur's.
template <int PL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_schur_n(Args a) {
    extern __shared__ __align__(16) float sm[];
    const int b = blockIdx.x;
    if (stopped(a) || (a.exit_rule == GATO_EXIT_KKT && !a.active[b])) return;
    phase_schur<PL>(a, b, a.rho[b], sm);
}

template <int PL, int NPASS, typename S = float>
__global__ void GATO_PHASE_BOUNDS(S) k_pcg(ArgsT<S> a, int it) {
    S* sm = dyn_lds<S>();
    const int b = blockIdx.x;
    if (stopped(a)) return;
    const bool skip = (a.exit_rule == GATO_EXIT_KKT) ? !a.active[b] : a.conv[b];  // quirk Q9 (pcg.cuh:30-33)
    int n = 0;
    if (!skip) {
        if constexpr (std::is_same_v<S, float>) n = pcg_dispatch<PL, NPASS>(a, b, a.tol[b], sm);
        else n = phase_pcg<PL, NPASS, false, S>(a, b, a.tol[b], sm);  // fp64: the streamed PCG at every shape
    }
    if (threadIdx.x == 0) a.st_pcg[it * a.B + b] = n;
}

// Two PCG solves on one problem's S in lockstep (phase_pcg_reg2), preconditioned with the factorization of S in
// Gm, Di (k_btd_factor), each with the k_pcg skip rule (quirk Q9); iteration counts to a1.st_pcg[b] and
// a2.st_pcg[b] when those are set. Then, per solve i, what k_dz and k_accumulate do for the problem: dz_i and
// the KKT residual (to a_i.st_kres[b]) from the new multipliers, and, when zi is set, z_i += dz_i, l_i += lambda_i
// and lambda_i cleared (the next correction's PCG starts from zero).
template <int PL>
__global__ void __launch_bounds__(512) GATO_TWO_WG_PER_CU k_pcg_dual(Args a1, Args a2, const float* Gm,
                                                                     const float* Di, float* z1, float* l1, float* z2,
                                                                     float* l2) {
    extern __shared__ __align__(16) float sm[];
    const int b = blockIdx.x;
    const bool run1 = !((a1.exit_rule == GATO_EXIT_KKT) ? !a1.active[b] : a1.conv[b]);
    const bool run2 = !((a2.exit_rule == GATO_EXIT_KKT) ? !a2.active[b] : a2.conv[b]);
    int n1 = 0, n2 = 0;
    phase_pcg_reg2<PL>(a1, a2, b, run1, run2, Gm, Di, sm, &n1, &n2);
    if (threadIdx.x == 0) {
        if (a1.st_pcg) a1.st_pcg[b] = n1;
        if (a2.st_pcg) a2.st_pcg[b] = n2;
    }
    auto finish = [&](const Args& a, float* z, float* l) {
        __syncthreads();  // the multipliers are written and the LDS is free
        const float kr = phase_dz<PL>(a, b, sm);  // ends with a barrier: dz is settled for the adds below
        if (threadIdx.x == 0) a.st_kres[b] = kr;
        if (z) {
            const int nz = a.traj + Plant<PL>::NU, nl = a.N * Plant<PL>::NX;
            for (int i = threadIdx.x; i < nz; i += blockDim.x) z[(size_t)b * nz + i] += a.dz[(size_t)b * nz + i];
            for (int i = threadIdx.x; i < nl; i += blockDim.x) {
                l[(size_t)b * nl + i] += a.lam[(size_t)b * nl + i];
                a.lam[(size_t)b * nl + i] = 0.f;
            }
        }
    };
    finish(a1, z1, l1);
    finish(a2, z2, l2);
}

// One PCG solve with S and its rows streamed (several knots per knot group), preconditioned with the twisted
// factorization of S in Gm, Di (k_btd_factor) read from global memory: the backward pass's solves when N > 32.
// Skip rule and iteration count as k_pcg (st_pcg[b]).
template <int PL, int NPASS>
__global__ void __launch_bounds__(512) GATO_TWO_WG_PER_CU k_pcg_btd(Args a, const float* Gm, const float* Di) {
    extern __shared__ __align__(16) float sm[];
    const int b = blockIdx.x;
    const bool skip = (a.exit_rule == GATO_EXIT_KKT) ? !a.active[b] : a.conv[b];  // quirk Q9 (pcg.cuh:30-33)
    const size_t pb = (size_t)b * a.N * Plant<PL>::NX * Plant<PL>::NX;
    int n = 0;
    if (!skip) n = phase_pcg<PL, NPASS, true>(a, b, a.tol[b], sm, Gm + pb, Di + pb);
    if (threadIdx.x == 0 && a.st_pcg) a.st_pcg[b] = n;
}

template <int PL, typename S = float>
__global__ void GATO_PHASE_BOUNDS(S) k_dz(ArgsT<S> a, int it) {
    S* sm = dyn_lds<S>();
    const int b = blockIdx.x;
    if (stopped(a)) return;
    if (a.exit_rule == GATO_EXIT_KKT && !a.active[b]) {
        if (threadIdx.x == 0) a.st_kres[it * a.B + b] = S(0);
        return;
    }
    const S kr = phase_dz<PL, S>(a, b, sm);
    if (threadIdx.x == 0) a.st_kres[it * a.B + b] = kr;
}

// Host-side convergence bookkeeping of bsqp.cuh:142-165 for problem b after SQP iteration it, given its PCG
// iteration count and KKT residual: SQP count, converged and active flags, and the line-search record of a
// frozen problem. Returns the converged flag.
template <typename S>
__device__ __forceinline__ int check_problem(const ArgsT<S>& a, int b, int it, int npcg, S kres) {
    int conv = a.conv[b];
    if (a.exit_rule == GATO_EXIT_PCG0) {
        if (npcg == 0) conv = 1;
        a.st_sqp[b] += 1;  // quirk Q10: batch-uniform count
    } else if (a.active[b]) {
        a.st_sqp[b] += 1;
        if (kres <= a.kkt_tol) { conv = 1; a.active[b] = 0; }
    }
    if (a.exit_rule == GATO_EXIT_KKT && !a.active[b]) {  // frozen: no line search this iteration
        a.st_step[it * a.B + b] = S(0);
        a.st_lsm[it * a.B + b] = a.phi[b];
    }
    a.conv[b] = conv;
    return conv;
}

// k_pcg, k_dz and the per-problem part of k_check in one launch (the graph path's step): after the problem's PCG
// has written lambda, the same work-group computes dz and the KKT residual from it and does the problem's
// convergence bookkeeping, adding it to the converged count solved[it] (zero at the start of the solve). The
// batch-level exit test on that count runs at the start of the line search (merit_body, which == 3).
template <int PL, int NPASS>
__device__ __forceinline__ void pcg_dz_body(const Args& a, int it) {
    extern __shared__ __align__(16) float sm[];
    const int b = blockIdx.x;
    // the problem's flags and tolerance read together, ahead of the tests that use them
    const int stop = a.stop ? *a.stop : 0, act = a.active[b], cnv = a.conv[b];
    const float tol = a.tol[b];
    if (stop) return;
    const bool skip = (a.exit_rule == GATO_EXIT_KKT) ? !act : cnv;  // quirk Q9 (pcg.cuh:30-33)
    const bool frozen = a.exit_rule == GATO_EXIT_KKT && !act;
    int n = 0;
    if (!skip) n = pcg_dispatch<PL, NPASS>(a, b, tol, sm);
    float kr = 0.f;
    if (!frozen) {
        __syncthreads();  // lambda (global) and the LDS of the PCG are settled for the dz phase
        kr = phase_dz<PL>(a, b, sm);
    }
    if (threadIdx.x == 0) {
        a.st_pcg[it * a.B + b] = n;
        a.st_kres[it * a.B + b] = kr;
        if (check_problem(a, b, it, n, kr)) atomicAdd(&a.solved[it], 1);
    }
}

template <int PL, int NPASS>
__global__ void __launch_bounds__(512) GATO_TWO_WG_PER_CU k_pcg_dz(Args a, int it) {
    pcg_dz_body<PL, NPASS>(a, it);
}

// The graph path's step for 64 < N <= 128 with the PCG held on the CU (phase_pcg_res): 1024 lanes, the problem's
// off-diagonal blocks of S in LDS (pcg_res_lds_floats), one work-group per CU.
template <int PL>
__global__ void __launch_bounds__(1024) GATO_TWO_WG_PER_CU k_pcg_dz_res(Args a, int it) {
    pcg_dz_body<PL, 0>(a, it);
}

// phase_pcg_res alone (the PCG of k_pcg_dz_res, for the phase-level tests): lambda and the iteration count.
template <int PL>
__global__ void __launch_bounds__(1024) GATO_TWO_WG_PER_CU k_pcg_res(Args a, int it) {
    extern __shared__ __align__(16) float sm[];
    const int b = blockIdx.x;
    if (stopped(a)) return;
    const bool skip = (a.exit_rule == GATO_EXIT_KKT) ? !a.active[b] : a.conv[b];  // quirk Q9 (pcg.cuh:30-33)
    int n = 0;
    if (!skip) n = phase_pcg_res<PL>(a, b, a.tol[b], sm);
    if (threadIdx.x == 0) a.st_pcg[it * a.B + b] = n;
}

// The convergence bookkeeping (check_problem) for every problem of the batch in one work-group, the converged
// count, and, when a device stop flag is set, the batch-level exit test (bsqp.cuh:165-167) that raises it.
template <typename S = float>
__global__ void __launch_bounds__(1024) k_check(ArgsT<S> a, int it) {
    __shared__ int count;
    if (stopped(a)) return;
    if (threadIdx.x == 0) count = 0;
    __syncthreads();
    int mine = 0;
    for (int b = threadIdx.x; b < a.B; b += blockDim.x)
        mine += check_problem(a, b, it, a.st_pcg[it * a.B + b], a.st_kres[it * a.B + b]);
    if (mine) atomicAdd(&count, mine);
    __syncthreads();
    if (threadIdx.x == 0) {
        a.solved[it] = count;
        if (a.stop_w && (double)count >= (double)a.B * (double)a.solve_ratio) *a.stop_w = 1;
    }
}

template <typename S = float>
__global__ void k_begin(ArgsT<S> a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) begin_batch(a);
    if (b < a.B) begin_problem(a, b);
}

template <typename S = float>
__global__ void k_zero(S* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = S(0);
}

template <typename S = float>
__global__ void k_end(ArgsT<S> a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.B) return;
    // The final merit (bsqp.cuh:179-182) at the returned trajectory with dz = 0 is phi: the line search stored the merit
    // of the point it accepted, whose every entry it formed as the trajectory update does (z + alpha dz, one fma), and
    // left phi unchanged when it kept the trajectory; the solve start's phi is the merit at the initial trajectory.
    a.st_final[b] = a.phi[b];
    a.drho[b] = a.drho_init[b];  // bsqp.cuh:189
}

// The per-problem state a freshly constructed solver starts from (bsqp.cuh:43-59): rho = rho0, drho = 1 (also the
// value it resets to), mu, PCG tolerance, and the joint weights the reference cost implies (Qd = [0, qd_cost],
// Rd = u_cost). Multipliers and wrenches start at zero with the rest of the cleared workspace.
__global__ void k_fresh_state(int B, float* rho, float* drho, float* drho_init, float* mu, float* tol, float* Qd,
                              float* Rd, int nq, float rho0, float mu0, float tol0, float qd_cost, float u_cost) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    rho[b] = rho0;
    drho[b] = 1.f;
    drho_init[b] = 1.f;
    mu[b] = mu0;
    tol[b] = tol0;
    for (int i = 0; i < 2 * nq; i++) Qd[(size_t)b * 2 * nq + i] = i < nq ? 0.f : qd_cost;
    for (int i = 0; i < nq; i++) Rd[(size_t)b * nq + i] = u_cost;
}

// SQP loop iterations run (the largest per-problem count) and iterations that reached the line search (the last one
// with a step recorded for any problem): out[0], out[1], as the host statistics report them.
__global__ void __launch_bounds__(256) k_solve_summary(Args a, int* out) {
    __shared__ int m[2];
    if (threadIdx.x == 0) m[0] = m[1] = 0;
    __syncthreads();
    int n = 0, l = 0;
    for (int b = threadIdx.x; b < a.B; b += blockDim.x) {
        n = max(n, a.st_sqp[b]);
        for (int i = a.max_sqp - 1; i >= 0; i--)
            if (a.st_step[(size_t)i * a.B + b] != 0.f) {
                l = max(l, i + 1);
                break;
            }
    }
    atomicMax(&m[0], n);
    atomicMax(&m[1], l);
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = m[0];
        out[1] = m[1];
    }
}

// Per-problem status bits (GATO_STATUS_*) of the last solve, from its statistics and a scan of the returned
// trajectory; one wavefront per problem. Non-finite values are found from the exponent bits, which no
// floating-point mode folds away.
__device__ __forceinline__ bool not_finite(float x) { return (__float_as_uint(x) & 0x7f800000u) == 0x7f800000u; }
__device__ __forceinline__ bool not_finite(double x) {
    return ((unsigned long long)__double_as_longlong(x) & 0x7ff0000000000000ull) == 0x7ff0000000000000ull;
}

template <typename S>
__global__ void __launch_bounds__(64) k_status(ArgsT<S> a, int* status) {
    const int b = blockIdx.x;
    const S* xu = a.xu + (size_t)b * a.traj;
    bool bad = false;
    for (int i = threadIdx.x; i < a.traj; i += 64) bad |= not_finite(xu[i]);
    bad = __any(bad);
    if (threadIdx.x != 0) return;
    const int n = min(a.st_sqp[b], (int)a.max_sqp);
    const bool conv = a.conv[b] != 0;
    bool pcg_cap = false, ls_fail = false;
    for (int i = 0; i < n; i++) {
        pcg_cap |= a.st_pcg[(size_t)i * a.B + b] >= a.max_pcg;
        ls_fail |= a.st_step[(size_t)i * a.B + b] < S(0);
    }
    bad |= not_finite(a.st_init[b]) || not_finite(a.st_final[b]);
    status[b] = (conv ? GATO_STATUS_KKT_CONVERGED : 0) | (!conv && n >= (int)a.max_sqp ? GATO_STATUS_ITER_CAP : 0) |
                (pcg_cap ? GATO_STATUS_PCG_AT_CAP : 0) | (ls_fail ? GATO_STATUS_LS_FAILED : 0) |
                (bad ? GATO_STATUS_NONFINITE : 0);
}

// ------------------------------------------------------------------ fused persistent kernel
// One whole solve of problem b by the calling work-group: initial merit, the SQP iterations (linearization, Schur
// complement, PCG, dz, merit at every step size, line search) until the problem's own exit rule fires or the
// iteration cap, final merit. Problems are independent (a converged problem stops iterating; the reference instead
// keeps stepping converged problems until the whole batch has converged). shs: 16 floats, shi: 4 ints of LDS.
template <int PL, int NPASS>
__device__ __forceinline__ void fused_problem(const Args& a, int b, float* sm, float* shs, int* shi) {
    constexpr int NU = Plant<PL>::NU;
    {
        View<PL> v(a, b);
        const float mu = a.mu[b], eps = a.tol[b];
        for (int i = threadIdx.x; i < a.traj + NU; i += blockDim.x) v.dz[i] = 0.f;
        __syncthreads();
        phase_merit<PL>(a, b, mu, 1, shs, sm);
        if (threadIdx.x == 0) {
            a.phi[b] = shs[0];
            a.st_init[b] = shs[0];
            a.st_sqp[b] = 0;
        }
        __syncthreads();
        int it = 0;
        for (; it < a.max_sqp; it++) {
            phase_lin<PL>(a, b, sm, fused_lin_slots(a.N, blockDim.x / KG));
            __syncthreads();
            phase_schur<PL>(a, b, a.rho[b], sm);
            __syncthreads();
            const int n = pcg_dispatch_fused<PL, NPASS>(a, b, eps, sm);
            __syncthreads();
            const float kr = phase_dz<PL>(a, b, sm);
            if (threadIdx.x == 0) {
                a.st_pcg[it * a.B + b] = n;
                a.st_kres[it * a.B + b] = kr;
                a.st_sqp[b] = it + 1;
                const int conv = (a.exit_rule == GATO_EXIT_PCG0) ? (n == 0) : (kr <= a.kkt_tol);
                a.conv[b] = conv;
                shi[0] = conv;
            }
            __syncthreads();
            if (shi[0]) break;
            phase_merit<PL>(a, b, mu, NUM_ALPHAS, shs, sm);
            phase_ls<PL>(a, b, it, shs, shs + NUM_ALPHAS);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            for (int j = it + (it < a.max_sqp ? 1 : 0); j < a.max_sqp; j++) {
                a.st_pcg[j * a.B + b] = 0;
                a.st_kres[j * a.B + b] = 0.f;
            }
            for (int j = it; j < a.max_sqp; j++) {
                a.st_step[j * a.B + b] = 0.f;
                a.st_lsm[j * a.B + b] = a.phi[b];
            }
        }
        for (int i = threadIdx.x; i < a.traj + NU; i += blockDim.x) v.dz[i] = 0.f;
        // the final merit at the returned trajectory is phi, as at the end of the graph solve (k_last_n)
        if (threadIdx.x == 0) {
            a.st_final[b] = a.phi[b];
            a.drho[b] = a.drho_init[b];
        }
        __syncthreads();
    }
}

// One work-group per problem, grid-strided over the batch.
template <int PL, int NPASS>
__global__ void __launch_bounds__(512) k_fused(Args a) {
    extern __shared__ __align__(16) float sm[];
    __shared__ float shs[16];
    __shared__ int shi[4];
    for (int b = blockIdx.x; b < a.B; b += gridDim.x) fused_problem<PL, NPASS>(a, b, sm, shs, shi);
}

// Problems of several (plant, horizon) classes in one persistent launch. items lists (class | tag << 8, problem) in
// the order they are to start (longest first), tag = 16 plant + knot passes of the class; cls[class] its arguments.
// Work-groups take the next item from ctr[0] as they finish the last, so the load balances itself; each item is
// the class's begin-of-solve reset of the problem and its whole solve, the same work as k_begin and k_fused for
// it. The last work-group to exit (ctr[1]) sets both counters back to zero for the next launch, so a replay needs
// no clearing of them (they are zero when allocated).
__global__ void __launch_bounds__(512) k_fused_mixed(const Args* __restrict__ cls, const int2* __restrict__ items,
                                                     int nitems, int* __restrict__ ctr) {
    extern __shared__ __align__(16) float sm[];
    __shared__ float shs[16];
    __shared__ int shi[4];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(ctr, 1);
        __syncthreads();
        const int i = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (i >= nitems) break;
        const int2 item = items[i];
        const int c = __builtin_amdgcn_readfirstlane(item.x & 255), tag = __builtin_amdgcn_readfirstlane(item.x >> 8);
        const int b = __builtin_amdgcn_readfirstlane(item.y);
        const Args& a = cls[c];
        if (threadIdx.x == 0) begin_problem(a, b);
        __syncthreads();
        switch (tag) {
            case 0x01: fused_problem<0, 1>(a, b, sm, shs, shi); break;
            case 0x02: fused_problem<0, 2>(a, b, sm, shs, shi); break;
            case 0x04: fused_problem<0, 4>(a, b, sm, shs, shi); break;
            case 0x08: fused_problem<0, 8>(a, b, sm, shs, shi); break;
            case 0x11: fused_problem<1, 1>(a, b, sm, shs, shi); break;
            case 0x12: fused_problem<1, 2>(a, b, sm, shs, shi); break;
            case 0x14: fused_problem<1, 4>(a, b, sm, shs, shi); break;
            case 0x18: fused_problem<1, 8>(a, b, sm, shs, shi); break;
            default: break;
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(ctr + 1, 1) == (int)gridDim.x - 1) {
            ctr[0] = 0;
            ctr[1] = 0;
        }
    }
}


// ------------------------------------------------------------------ adjoint of the SQP step (backward)
// No reference counterpart. The backward differentiates the step map z+ = z + dz(z; theta) at the solution z,
// theta = (x_s, reference xyz, diag Q, diag R), with dz from the step's KKT system K [dz; l] = [-g; c]. The
// adjoint system K [v; v_l] = [dL/dz; 0] has the same Schur complement S and preconditioner P: it is the step
// system with g' = -dL/dz and c' = 0, so it reuses the Schur, PCG and dz phases on separate buffers. Then
//   dL/dx_s = -v_l,0,   dL/dtheta = -v^T dg/dtheta - v^T (dG/dtheta) dz   (G depends on diag Q, diag R, ref).

// Right-hand side of a Schur solve from a KKT right-hand side [-g; c] in block form (g = (q', r') per knot):
//   gamma_0 = c_0 - Q~_0^-1 q'_0,  gamma_{k+1} = c_{k+1} - Q~_{k+1}^-1 q'_{k+1} + phi_k q'_k + B_k R_k^-1 r'_k,
// the same expression as the Schur phase's gamma (schur_linsys.cuh:60-98), for arbitrary (q', r', c').
template <int PL>
__device__ __forceinline__ void gamma_rhs_body(const Args& a, int b, const float* qv, const float* rv,
                                               const float* cv, float* gam) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU;
    View<PL> v(a, b);
    const int N = a.N, g = threadIdx.x / KG, lane = threadIdx.x % KG, G = blockDim.x / KG;
    const float* qb = qv + (size_t)b * N * NX;
    const float* rb = rv + (size_t)b * N * NU;
    const float* cb = cv + (size_t)b * N * NX;
    for (int k = g; k < N; k += G) {
        if (lane >= NX) continue;
        const int r = lane;
        float qi[NX];
        qinv_row<PL>(v.Qip + k * NQ * NQ, v.Qiv + k * NQ, r, qi);
        float gs = cb[k * NX + r];
#pragma unroll
        for (int m = 0; m < NX; m++) gs -= qi[m] * qb[k * NX + m];
        if (k > 0) {
            const float* phi = v.So + (size_t)k * NX * NX + r * NX;
#pragma unroll
            for (int m = 0; m < NX; m++) gs += phi[m] * qb[(k - 1) * NX + m];
            const float* Bk = v.Bm + (size_t)(k - 1) * NX * NU + r * NU;
            const float* Rk = v.Ri + (k - 1) * NU;
#pragma unroll
            for (int l = 0; l < NU; l++) gs += Bk[l] * Rk[l] * rb[(k - 1) * NU + l];
        }
        gam[(size_t)b * N * NX + k * NX + r] = gs;
    }
}

// Residual of the step's KKT system K [z; l] = rhs at a current solution (z, l), from the blocks themselves
// (no inverses), written as the right-hand side of the correction solve: q' = -r_x, r' = -r_u, c' = r_c.
//   r_x = rhs_x - Q~_k z_x + l_k - A_k^T l_{k+1},  r_u = rhs_u - R_k z_u - B_k^T l_{k+1},
//   r_c0 = rhs_c0 + z_x0,  r_c{k+1} = rhs_c{k+1} - A_k z_xk - B_k z_uk + z_x{k+1}.
// rhs_z is rz (trajectory layout) when given, else -(gq, gr); rhs_c is cc when given, else 0.
template <int PL>
__device__ __forceinline__ void kkt_resid_body(const Args& a, int b, const float* rz, const float* gq,
                                               const float* gr, const float* cc, const float* zall,
                                               const float* lall, float* qo, float* ro, float* co) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU, ST = NX + NU;
    View<PL> v(a, b);
    const int N = a.N, g = threadIdx.x / KG, lane = threadIdx.x % KG, G = blockDim.x / KG;
    const float rho = a.rho[b];
    const float* z = zall + (size_t)b * (a.traj + NU);
    const float* l = lall + (size_t)b * N * NX;
    for (int k = g; k < N; k += G) {
        const bool last = (k == N - 1);
        const float* zk = z + (size_t)k * ST;
        const float* l1 = l + (k + 1) * NX;
        if (lane < NX) {
            const int i = lane;
            float hx;
            if (i < NQ) {
                hx = rho * zk[i];
#pragma unroll
                for (int m = 0; m < NQ; m++) hx += v.Qp[k * NQ * NQ + i * NQ + m] * zk[m];
            } else {
                hx = v.Qv[k * NQ + i - NQ] * zk[i];
            }
            hx -= l[k * NX + i];
            if (!last) {
                const float* A = v.A + (size_t)k * NX * NX;
#pragma unroll
                for (int m = 0; m < NX; m++) hx += A[m * NX + i] * l1[m];
            }
            const float rhs = rz ? rz[(size_t)b * a.traj + (size_t)k * ST + i] : -gq[(size_t)b * N * NX + k * NX + i];
            qo[(size_t)b * N * NX + k * NX + i] = -(rhs - hx);
            // constraint row k: c_0 = -z_x0, c_k = A_{k-1} z_x{k-1} + B_{k-1} z_u{k-1} - z_xk
            float cz;
            if (k == 0) {
                cz = -zk[i];
            } else {
                const float* A0 = v.A + (size_t)(k - 1) * NX * NX + i * NX;
                const float* B0 = v.Bm + (size_t)(k - 1) * NX * NU + i * NU;
                const float* zp = z + (size_t)(k - 1) * ST;
                cz = -zk[i];
#pragma unroll
                for (int m = 0; m < NX; m++) cz += A0[m] * zp[m];
#pragma unroll
                for (int m = 0; m < NU; m++) cz += B0[m] * zp[NX + m];
            }
            const float rc = cc ? cc[(size_t)b * N * NX + k * NX + i] : 0.f;
            co[(size_t)b * N * NX + k * NX + i] = rc - cz;
        }
        if (lane < NU) {
            const int i = lane;
            float ru = 0.f;
            if (!last) {
                float hu = v.Rg[k * NU + i] * zk[NX + i];
                const float* Bm = v.Bm + (size_t)k * NX * NU;
#pragma unroll
                for (int m = 0; m < NX; m++) hu += Bm[m * NU + i] * l1[m];
                const float rhs = rz ? rz[(size_t)b * a.traj + (size_t)k * ST + NX + i]
                                     : -gr[(size_t)b * N * NU + k * NU + i];
                ru = -(rhs - hu);
            }
            ro[(size_t)b * N * NU + k * NU + i] = ru;
        }
    }
}

// One backward chain's residual inputs and outputs (kkt_resid_body's and gamma_rhs_body's arguments).
struct ResidChain {
    const float *rz, *gq, *gr, *cc, *z, *l;
    float *q, *r, *c, *gam;
};

// A refinement round's residuals and Schur right-hand sides for both backward chains, problem by problem:
// kkt_resid_body and then gamma_rhs_body for chain 1 (when do1) and chain 2, in one launch.
template <int PL>
__global__ void __launch_bounds__(512) GATO_TWO_WG_PER_CU k_residual2(Args a, ResidChain c1, ResidChain c2, int do1) {
    const int b = blockIdx.x;
    if (do1) kkt_resid_body<PL>(a, b, c1.rz, c1.gq, c1.gr, c1.cc, c1.z, c1.l, c1.q, c1.r, c1.c);
    kkt_resid_body<PL>(a, b, c2.rz, c2.gq, c2.gr, c2.cc, c2.z, c2.l, c2.q, c2.r, c2.c);
    __syncthreads();  // the problem's residual rows are written before the right-hand sides read their neighbours
    if (do1) gamma_rhs_body<PL>(a, b, c1.q, c1.r, c1.c, c1.gam);
    gamma_rhs_body<PL>(a, b, c2.q, c2.r, c2.c, c2.gam);
}

// z += dz, l += dl; dl is cleared (the next correction's PCG starts from zero)
__global__ void k_accumulate(float* z, float* dz, float* l, float* dl, int nz, int nl) {
    const int b = blockIdx.x;
    for (int i = threadIdx.x; i < nz; i += blockDim.x) z[(size_t)b * nz + i] += dz[(size_t)b * nz + i];
    for (int i = threadIdx.x; i < nl; i += blockDim.x) {
        l[(size_t)b * nl + i] += dl[(size_t)b * nl + i];
        dl[(size_t)b * nl + i] = 0.f;
    }
}

// Parameter gradients from the step (a.dz at the solution), the adjoint solution vz and its multipliers lam_a.
// Reference xyz of knot k: w (v_q . J_j) + [rank-1 Hessian] w ((v_q . J_j)(g . dq) + (v_q . g)(J_j . dq)),
// with J, g = J^T (p - ref_k) at the knot's cost state (x_{N-2} for the reference's terminal block).
template <int PL>
__global__ void __launch_bounds__(512) GATO_TWO_WG_PER_CU k_grad(Args a, const float* vzall, const float* lam_a,
                                                                 float* gxs, float* gref, float* gQ, float* gR) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU, ST = NX + NU;
    extern __shared__ __align__(16) float sm[];
    const int b = blockIdx.x;
    View<PL> v(a, b);
    const int N = a.N, g = threadIdx.x / KG, lane = threadIdx.x % KG, G = blockDim.x / KG;
    float* sq = sm;
    float* sr = sm + N * KG;
    const float* vz = vzall + (size_t)b * (a.traj + NU);
    const float h0 = a.cw.hess_mode ? 0.f : 1.f;
    for (int k = g; k < N; k += G) {
        const bool term = (k == N - 1);
        const int src = term ? (a.cw.terminal_mode ? N - 1 : N - 2) : k;
        const float* xsrc = v.xu + (size_t)src * ST;
        const float* vx = vz + (size_t)k * ST;
        const float* dx = v.dz + (size_t)k * ST;
        float cq = 0.f, cr = 0.f;
        if (lane < NX) {
            const float xr = v.xref ? v.xref[k * NX + lane] : 0.f;
            cq = -vx[lane] * (xsrc[lane] - xr) - vx[lane] * dx[lane];
        }
        if (!term && lane < NU) cr = -vx[NX + lane] * (v.xu[(size_t)k * ST + NX + lane] + dx[NX + lane]);
        sq[k * KG + lane] = cq;
        sr[k * KG + lane] = cr;
        float q[NQ];
#pragma unroll
        for (int i = 0; i < NQ; i++) q[i] = xsrc[i];
        Kin<PL> K;
        K.set(q);
        float p[3], Jc[3];
        ee_position<PL>(K, lane < NQ ? lane : -1, p, Jc);
        const float* rk = v.ref + 6 * k;
        const float gl = Jc[0] * (p[0] - rk[0]) + Jc[1] * (p[1] - rk[1]) + Jc[2] * (p[2] - rk[2]);
        const float vq = lane < NQ ? vx[lane] : 0.f, dq = lane < NQ ? dx[lane] : 0.f;
        const float s0 = row16_sum(vq * Jc[0]), s1 = row16_sum(vq * Jc[1]), s2 = row16_sum(vq * Jc[2]);
        const float t0 = row16_sum(Jc[0] * dq), t1 = row16_sum(Jc[1] * dq), t2 = row16_sum(Jc[2] * dq);
        const float gdq = row16_sum(gl * dq), vqg = row16_sum(vq * gl);
        const float w = term ? (a.cw.terminal_mode ? a.cw.N_cost : a.cw.q_cost) : a.cw.q_cost;
        if (lane < 6) {
            const float s = lane == 0 ? s0 : lane == 1 ? s1 : s2;
            const float t = lane == 0 ? t0 : lane == 1 ? t1 : t2;
            gref[(size_t)b * 6 * N + 6 * k + lane] = lane < 3 ? w * (s + h0 * (s * gdq + vqg * t)) : 0.f;
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < NX) {
        float s = 0.f;
        for (int k = 0; k < N; k++) s += sq[k * KG + t];
        gQ[(size_t)b * NX + t] = s;
        gxs[(size_t)b * NX + t] = -lam_a[(size_t)b * N * NX + t];
    } else if (t >= KG && t < KG + NU) {
        float s = 0.f;
        for (int k = 0; k < N; k++) s += sr[k * KG + t - KG];
        gR[(size_t)b * NU + t - KG] = s;
    }
}

// ------------------------------------------------------------------ derivative of the solution (implicit rounds)
// At a solution the step vanishes: F(z, l) = [g(z) + C(z)^T l; c(z)] = 0, and the solution's derivative needs the
// adjoint of the exact Jacobian J = dF/d(z, l), whose upper-left block H is the Hessian of the Lagrangian; the step's
// K has the model G~ there instead. With J = K - D, D = [[G~ - H, 0], [0, 0]], the iteration K w+ = [dL/dz; 0] + D w
// contracts at the SQP step's own local rate (w = K^-1 [dL/dz; 0] is the one-step adjoint), and each of its steps is
// a refinement round of the adjoint chain with the right-hand side dL/dz + (G~ - H) v. H v comes from central
// differences of the stationarity rows along v: the KKT set-up at z +- eps v (k_lin on separate blocks), with the
// step's multipliers l held fixed.

// KKT set-up blocks in fp64 (the implicit rounds' perturbed linearizations)
struct Lin64 {
    double *A, *Bm, *c, *Qp, *Qv, *Rg, *q, *r;
};

// zp = z + eps v, zm = z - eps v for problem b (fp64), with eps = step / |v|_inf (stored per problem)
__global__ void __launch_bounds__(256) k_perturb(const float* z, const float* vall, double* zp, double* zm,
                                                 double* eps, int traj, int ndz, double step) {
    __shared__ float red[8];
    const int b = blockIdx.x;
    const float* v = vall + (size_t)b * ndz;
    float m = 0.f;
    for (int i = threadIdx.x; i < traj; i += blockDim.x) m = fmaxf(m, fabsf(v[i]));
    m = block_max(m, red);
    const double e = step / fmax((double)m, 1e-300);
    if (threadIdx.x == 0) eps[b] = e;
    for (int i = threadIdx.x; i < traj; i += blockDim.x) {
        const size_t j = (size_t)b * traj + i;
        zp[j] = (double)z[j] + e * (double)v[i];
        zm[j] = (double)z[j] - e * (double)v[i];
    }
}

// The KKT set-up of phase_lin in fp64 at the trajectory zall (fp64), the solver's inputs (fp32: x_s, reference,
// wrenches, joint weights and reference) widened in LDS. Knot groups loop over the knots; per group the fp64 lin_knot
// table and the knot's reference.
template <int PL>
__host__ __device__ constexpr int lin64_group_doubles() {
    return lin_lds_floats<PL>() + 8 + Plant<PL>::NX;
}
template <int PL>
__global__ void __launch_bounds__(128) k_lin64(Args a, const double* zall, Lin64 o) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU, ST = NX + NU;
    extern __shared__ __align__(16) double smd[];
    const int b = blockIdx.x, N = a.N, g = threadIdx.x / KG, lane = threadIdx.x % KG, G = blockDim.x / KG;
    View<PL> v(a, b);
    double* sxs = smd;
    double* sQd = sxs + NX;
    double* sRd = sQd + NX;
    double* sfe = sRd + NU;
    double* grp = sfe + 8;
    for (int i = threadIdx.x; i < NX; i += blockDim.x) {
        sxs[i] = v.xs[i];
        sQd[i] = v.Qd[i];
    }
    for (int i = threadIdx.x; i < NU; i += blockDim.x) sRd[i] = v.Rd[i];
    for (int i = threadIdx.x; i < 6; i += blockDim.x) sfe[i] = v.fext[i];
    __syncthreads();
    const double* z = zall + (size_t)b * a.traj;
    double* shg = grp + g * lin64_group_doubles<PL>();
    double* sref = shg + lin_lds_floats<PL>();
    double* sxr = sref + 8;
    const double* xt = z + (size_t)(a.cw.terminal_mode ? N - 1 : N - 2) * ST;
    for (int k = g; k < N; k += G) {
        if (lane < 6) sref[lane] = v.ref[6 * k + lane];
        if (lane < NX && v.xref) sxr[lane] = v.xref[k * NX + lane];
        wave_sync();
        const size_t bk = (size_t)b * N + k;
        const double* xk = z + (size_t)(k < N - 1 ? k : 0) * ST;
        lin_knot<PL, double>(k, N, opaque(lane), (double)a.dt, a.cw, xk, xt, sref, v.xref ? sxr : nullptr, sQd, sRd,
                             sfe, z, sxs, o.A + bk * NX * NX, o.Bm + bk * NX * NU,
                             o.c + (size_t)b * N * NX + (k + 1) * NX, o.c + (size_t)b * N * NX, o.Qp + bk * NQ * NQ,
                             o.Qv + bk * NQ, o.Rg + bk * NU, o.q + bk * NX, o.r + bk * NU, shg);
    }
}

// rz = dL/dz + G~ v - H v for problem b (trajectory layout), H v = [F_z(z + eps v) - F_z(z - eps v)] / (2 eps) in fp64
// from the two perturbed set-ups (p, m: q, r, A, B) and the step's multipliers lam; G~ from the blocks at z (a)
template <int PL>
__global__ void __launch_bounds__(512) GATO_TWO_WG_PER_CU k_hvp_rhs(Args a, Lin64 p, Lin64 m, const float* dL,
                                                                    const float* vall, const float* lamall,
                                                                    const double* eps, float* rz) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, NX = P::NX, NU = P::NU, ST = NX + NU;
    const int b = blockIdx.x;
    View<PL> v0(a, b);
    const int N = a.N, g = threadIdx.x / KG, lane = threadIdx.x % KG, G = blockDim.x / KG;
    const float rho = a.rho[b];
    const double inv2e = 0.5 / eps[b];
    const float* v = vall + (size_t)b * (a.traj + NU);
    const float* l = lamall + (size_t)b * N * NX;
    const float* dl = dL + (size_t)b * a.traj;
    float* out = rz + (size_t)b * a.traj;
    for (int k = g; k < N; k += G) {
        const bool last = (k == N - 1);
        const size_t bk = (size_t)b * N + k;
        const float* vk = v + (size_t)k * ST;
        const float* l1 = l + (k + 1) * NX;
        if (lane < NX) {
            const int i = lane;
            float gv;
            if (i < NQ) {
                gv = rho * vk[i];
#pragma unroll
                for (int j = 0; j < NQ; j++) gv += v0.Qp[k * NQ * NQ + i * NQ + j] * vk[j];
            } else {
                gv = v0.Qv[k * NQ + i - NQ] * vk[i];
            }
            double dF = p.q[bk * NX + i] - m.q[bk * NX + i];
            if (!last) {
                const double* Ap = p.A + bk * NX * NX;
                const double* Am = m.A + bk * NX * NX;
#pragma unroll
                for (int j = 0; j < NX; j++) dF += (Ap[j * NX + i] - Am[j * NX + i]) * (double)l1[j];
            }
            out[(size_t)k * ST + i] = (float)((double)dl[(size_t)k * ST + i] + (double)gv - dF * inv2e);
        }
        if (!last && lane < NU) {
            const int i = lane;
            const float gv = v0.Rg[k * NU + i] * vk[NX + i];
            double dF = p.r[bk * NU + i] - m.r[bk * NU + i];
            const double* Bp = p.Bm + bk * NX * NU;
            const double* Bn = m.Bm + bk * NX * NU;
#pragma unroll
            for (int j = 0; j < NX; j++) dF += (Bp[j * NU + i] - Bn[j * NU + i]) * (double)l1[j];
            out[(size_t)k * ST + NX + i] = (float)((double)dl[(size_t)k * ST + NX + i] + (double)gv - dF * inv2e);
        }
    }
}

// backward set-up: flags cleared, the step's multipliers warm-started from the solve's, the correction and
// adjoint buffers (lam_a, dz_a, dz_b; dz_v, lam_v, lam_w) zeroed, one work-group per problem
__global__ void k_bwd_begin(Args a, const float* lam_keep, float* lam_b, float* lam_a, float* dz_b, float* dz_a,
                            float* dz_v, float* lam_v, float* lam_w, int nlam, int ndz) {
    const int b = blockIdx.x;
    if (threadIdx.x == 0) {
        a.conv[b] = 0;
        a.active[b] = 1;
    }
    for (int i = threadIdx.x; i < nlam; i += blockDim.x) {
        const size_t j = (size_t)b * nlam + i;
        lam_b[j] = lam_keep[j];
        lam_a[j] = 0.f;
        lam_v[j] = 0.f;
        lam_w[j] = 0.f;
    }
    for (int i = threadIdx.x; i < ndz; i += blockDim.x) {
        const size_t j = (size_t)b * ndz + i;
        dz_b[j] = 0.f;
        dz_a[j] = 0.f;
        dz_v[j] = 0.f;
    }
}


// ------------------------------------------------------------------ MPC hypothesis selection (no reference kernel)
// The MPC caller's per-step host work (mpc_controller.py:294-309): step every wrench hypothesis b from
// (x_last, u_last) with its f_ext_b, err_b = |x_next_b - x_curr|_2, best = argmin (first on ties, as numpy's
// argmin), then every problem's warm start becomes problem best's trajectory (mpc_controller.py:240-242).
template <int PL>
__global__ void __launch_bounds__(1024) k_select(const float* x, const float* u, const float* xcur,
                                                 const float* fext, float dt, int B, float* err, int* best) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ;
    __shared__ float se[1024];
    __shared__ int si[1024];
    // numpy.argmin's order (mpc_controller.py:303-304): the first NaN if there is one, else the first minimum
    auto better = [](float e2, int i2, float e1, int i1) {
        const bool n2 = e2 != e2, n1 = e1 != e1;
        if (n2 != n1) return n2;
        return (!n2 && e2 < e1) || ((n2 || e2 == e1) && i2 < i1);
    };
    float emin = 3.4e38f;
    int imin = B;
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        float q[NQ], qd[NQ], uu[NQ], qdd[NQ], L[NQ * (NQ + 1) / 2];
#pragma unroll
        for (int i = 0; i < NQ; i++) { q[i] = x[i]; qd[i] = x[NQ + i]; uu[i] = u[i]; }
        Kin<PL> K;
        K.set(q);
        forward_dynamics<PL>(K, qd, uu, fext + 6 * b, qdd, L);
        float e = 0.f;
#pragma unroll
        for (int i = 0; i < NQ; i++) {
            const float dq = q[i] + dt * qd[i] + 0.5f * qdd[i] * dt * dt - xcur[i];
            const float dv = qd[i] + dt * qdd[i] - xcur[NQ + i];
            e += dq * dq + dv * dv;
        }
        e = sqrtf(e);
        if (err) err[b] = e;
        if (better(e, b, emin, imin)) { emin = e; imin = b; }
    }
    se[threadIdx.x] = emin;
    si[threadIdx.x] = imin;
    __syncthreads();
    // Tree over n = blockDim.x slots, which need not be a power of two: the first step starts at half the next
    // power of two and folds only the slots that exist (i + h < n) into [0, h); every later step halves a power
    // of two. `better` is a total order, so the fold order does not change the winner.
    const int n = blockDim.x;
    int h = 1;
    while (2 * h < n) h *= 2;
    for (; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h && (int)threadIdx.x + h < n) {
            const float e2 = se[threadIdx.x + h];
            const int i2 = si[threadIdx.x + h];
            if (better(e2, i2, se[threadIdx.x], si[threadIdx.x])) {
                se[threadIdx.x] = e2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *best = si[0] < B ? si[0] : 0;  // B = 0 aside, si[0] < B always holds here
}

__global__ void k_broadcast_row(float* xu, int traj, int B, const int* best) {
    const int src = *best;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)B * traj; i += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(i / traj);
        if (b != src) xu[i] = xu[(size_t)src * traj + i % traj];
    }
}


// ------------------------------------------------------------------ plant simulation and reference windows (MPC)
// One RK4 step of length h of qdd = FD(q, qd, u; f_ext) for one system (q, v updated in place): the integration the
// reference's MPC harness runs on the host with pinocchio (common.py:49-91: k1..k4 from ABA, revolute joints
// integrated as q + v dt). fe may be null (no wrench).
template <int PL>
__device__ __forceinline__ void rk4_step(float* q, float* v, const float* uu, const float* fe, float h) {
    constexpr int NQ = Plant<PL>::NQ;
    auto accel = [&](const float* qq, const float* vv, float* a) {
        Kin<PL> K;
        K.set(qq);
        float L[NQ * (NQ + 1) / 2];
        forward_dynamics<PL>(K, vv, uu, fe, a, L);
    };
    float k1v[NQ], k2v[NQ], k3v[NQ], k4v[NQ], qt[NQ], vt[NQ];
    accel(q, v, k1v);
#pragma unroll
    for (int i = 0; i < NQ; i++) { qt[i] = q[i] + 0.5f * h * v[i]; vt[i] = v[i] + 0.5f * h * k1v[i]; }
    accel(qt, vt, k2v);
    float v2[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) { v2[i] = vt[i]; qt[i] = q[i] + 0.5f * h * v2[i]; vt[i] = v[i] + 0.5f * h * k2v[i]; }
    accel(qt, vt, k3v);
    float v3[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) { v3[i] = vt[i]; qt[i] = q[i] + h * v3[i]; vt[i] = v[i] + h * k3v[i]; }
    accel(qt, vt, k4v);
#pragma unroll
    for (int i = 0; i < NQ; i++) {
        q[i] += h / 6.f * (v[i] + 2.f * v2[i] + 2.f * v3[i] + vt[i]);
        v[i] += h / 6.f * (k1v[i] + 2.f * k2v[i] + 2.f * k3v[i] + k4v[i]);
    }
}

// `steps` RK4 sub-steps of h for every problem b with its own control u_b and wrench f_ext_b (x_b in place).
template <int PL>
__global__ void __launch_bounds__(64) k_rk4(float* x, const float* u, const float* fext, float h, int steps, int B) {
    constexpr int NQ = Plant<PL>::NQ;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float q[NQ], v[NQ], uu[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) { q[i] = x[b * 2 * NQ + i]; v[i] = x[b * 2 * NQ + NQ + i]; uu[i] = u[b * NQ + i]; }
    for (int st = 0; st < steps; st++) rk4_step<PL>(q, v, uu, fext + 6 * b, h);
#pragma unroll
    for (int i = 0; i < NQ; i++) { x[b * 2 * NQ + i] = q[i]; x[b * 2 * NQ + NQ + i] = v[i]; }
}

// The MPC harness's plant advance (mpc_controller.py:194-213): sub-steps i = first .. first + steps - 1 of length h,
// sub-step i applying the control of knot min(int(i / (knot_dt / h)), knots - 1) of system b's trajectory xu_b
// (the quotient in double, as the harness computes it in Python), under wrench fext_b (null: none).
template <int PL>
__global__ void __launch_bounds__(64) k_plant_rk4(float* x, const float* xu, int traj, int knots, const float* fext,
                                                  double h, int first, int steps, double knot_dt, int n) {
    constexpr int NQ = Plant<PL>::NQ, NX = 2 * NQ;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    float q[NQ], v[NQ], uu[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) { q[i] = x[b * NX + i]; v[i] = x[b * NX + NQ + i]; }
    const float* fe = fext ? fext + 6 * b : nullptr;
    const double per_knot = knot_dt / h;
    for (int i = first; i < first + steps; i++) {
        const int k = min((int)((double)i / per_knot), knots - 1);
        const float* uk = xu + (size_t)b * traj + (size_t)k * (NX + NQ) + NX;
#pragma unroll
        for (int j = 0; j < NQ; j++) uu[j] = uk[j];
        rk4_step<PL>(q, v, uu, fe, (float)h);
    }
#pragma unroll
    for (int i = 0; i < NQ; i++) { x[b * NX + i] = q[i]; x[b * NX + NQ + i] = v[i]; }
}

// ref_b[k] = path[offset + k] (6 floats per point) for every problem: the reference window the MPC caller
// tiles on the host each step (mpc_controller.py:226-229), from a device-resident path.
__global__ void k_window_ref(float* ref, const float* path, int npoints, int offset, int N, int B) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)B * N * 6; i += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(i % (6 * N));
        const int k = j / 6;
        const int pt = std::min(offset + k, npoints - 1);
        ref[i] = path[(size_t)pt * 6 + j % 6];
    }
}

// Figure-8 end-effector path of the MPC examples (common.py:10-46), generated on the device: point j of a cycle
// of m points sits at t_j = 2 pi j / (m - 1) (the last one exactly at 2 pi, as numpy.linspace places it) on
// (ax sin t, 0, az sin(2t)/2 + az/2) + o, rotated by theta about z; rpy = 0; the cycle repeats `cycles` times.
// Evaluated in double and rounded once, as the callers round the float64 path.
__global__ void k_figure8(float* path, int m, int cycles, double ax, double az, double ox, double oy, double oz,
                          double theta) {
    const double c = cos(theta), s = sin(theta);
    const double step = m > 1 ? 2.0 * M_PI / (double)(m - 1) : 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)m * cycles; i += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(i % m);
        const double t = (m > 1 && j == m - 1) ? 2.0 * M_PI : (double)j * step;
        const double xu = ox + ax * sin(t), yu = oy, zu = oz + az * sin(2.0 * t) / 2.0 + az / 2.0;
        float* pt = path + 6 * i;
        pt[0] = (float)(c * xu - s * yu);
        pt[1] = (float)(s * xu + c * yu);
        pt[2] = (float)zu;
        pt[3] = pt[4] = pt[5] = 0.f;
    }
}

// ------------------------------------------------------------------ force hypotheses for MPC (force_estimator.py)
// State of the MPC examples' wrench estimator (force_estimator.py:4-31) kept on the device. Wrenches here are
// world-frame [force; moment]. radius and confidence stay in double as the reference's Python floats do.
struct ForceEst {
    float est[6], mom[6], smooth[6];
    float R[9];      // rotation applied to the exploration directions (row-major)
    double radius, confidence;
    double hist[5];  // last five minimum prediction errors, oldest first
    int count;       // updates so far (the length of the reference's error history)
};

// Hypothesis i of the batch (force_estimator.py:82-97): 0 the smoothed estimate, 1 zero, 2 the estimate pushed
// along its momentum, i >= 3 the blended estimate plus radius times the rotated exploration direction i - 3.
// Float arithmetic is rounded per operation like numpy's float32 arrays (no contraction).
__device__ __forceinline__ void force_hypothesis(const ForceEst& st, const float* dirs, int i, float* f) {
#pragma clang fp contract(off)
    if (i == 0) {
        for (int c = 0; c < 6; c++) f[c] = st.smooth[c];
    } else if (i == 1) {
        for (int c = 0; c < 6; c++) f[c] = 0.f;
    } else if (i == 2) {
        for (int c = 0; c < 6; c++) f[c] = st.smooth[c] + 0.5f * st.mom[c];
    } else {
        const float* d = dirs + 3 * (i - 3);
        for (int r = 0; r < 3; r++) {
            const float dir = (st.R[3 * r] * d[0] + st.R[3 * r + 1] * d[1]) + st.R[3 * r + 2] * d[2];
            const float base = 0.7f * st.smooth[r] + 0.3f * st.est[r];
            f[r] = (float)((double)base + st.radius * (double)dir);
        }
        for (int c = 3; c < 6; c++) f[c] = st.smooth[c];
    }
}

// World placements (R row-major, origin) of the last two joint frames at q, in double: the frames whose relative
// placement the wrench transform of the MPC caller uses (the end-effector's parent joint and the joint before it).
template <int PL>
__device__ void last_joint_frames(const float* q, double* R1, double* o1, double* R0, double* o0) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ;
    double W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, o[3] = {0, 0, 0};
    Unroll<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr auto& T = P::J::T[i];  // R_tree (row-major), p_tree of joint i in double
#pragma unroll
        for (int r = 0; r < 3; r++) o[r] += W[3 * r] * T[9] + W[3 * r + 1] * T[10] + W[3 * r + 2] * T[11];
        double WT[9];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                WT[3 * r + k] = W[3 * r] * T[k] + W[3 * r + 1] * T[3 + k] + W[3 * r + 2] * T[6 + k];
        const double s = sin((double)q[i]), c = cos((double)q[i]);
#pragma unroll
        for (int r = 0; r < 3; r++) {
            W[3 * r] = WT[3 * r] * c + WT[3 * r + 1] * s;
            W[3 * r + 1] = -WT[3 * r] * s + WT[3 * r + 1] * c;
            W[3 * r + 2] = WT[3 * r + 2];
        }
        if constexpr (i == NQ - 2) {
#pragma unroll
            for (int k = 0; k < 9; k++) R0[k] = W[k];
#pragma unroll
            for (int k = 0; k < 3; k++) o0[k] = o[k];
        }
    });
    for (int k = 0; k < 9; k++) R1[k] = W[k];
    for (int k = 0; k < 3; k++) o1[k] = o[k];
}

// Inverse action of the placement (R, p) on a wrench [f; n]: (R^T f, R^T (n - p x f)).
__device__ __forceinline__ void wrench_act_inv(const double* R, const double* p, const double* w, double* out) {
    const double n0 = w[3] - (p[1] * w[2] - p[2] * w[1]);
    const double n1 = w[4] - (p[2] * w[0] - p[0] * w[2]);
    const double n2 = w[5] - (p[0] * w[1] - p[1] * w[0]);
    for (int k = 0; k < 3; k++) {
        out[k] = R[k] * w[0] + R[3 + k] * w[1] + R[6 + k] * w[2];
        out[3 + k] = R[k] * n0 + R[3 + k] * n1 + R[6 + k] * n2;
    }
}

// Every hypothesis b < B in the world frame (optional output) and mapped to the wrench the solver takes, the way
// mpc_controller.py:311-337 maps it: inverse action of the last joint's world placement, then inverse action of
// that joint's placement relative to the joint before it. Frames at q are computed once per work-group.
template <int PL>
__global__ void __launch_bounds__(256) k_force_hypotheses(const ForceEst* st, const float* dirs, const float* q,
                                                          int B, float* fext, float* world) {
    __shared__ double fr[12];  // R' (relative rotation), p' (relative origin) of the last joint
    __shared__ double M1[12];  // world rotation and origin of the last joint
    __shared__ ForceEst S;
    if (threadIdx.x == 0) {
        double R1[9], o1[3], R0[9], o0[3];
        last_joint_frames<PL>(q, R1, o1, R0, o0);
        const double d[3] = {o1[0] - o0[0], o1[1] - o0[1], o1[2] - o0[2]};
        for (int r = 0; r < 3; r++) {
            for (int k = 0; k < 3; k++)
                fr[3 * r + k] = R0[r] * R1[k] + R0[3 + r] * R1[3 + k] + R0[6 + r] * R1[6 + k];  // R0^T R1
            fr[9 + r] = R0[r] * d[0] + R0[3 + r] * d[1] + R0[6 + r] * d[2];                 // R0^T (o1 - o0)
        }
        for (int k = 0; k < 9; k++) M1[k] = R1[k];
        for (int k = 0; k < 3; k++) M1[9 + k] = o1[k];
        S = *st;
    }
    __syncthreads();
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) {
        float f[6];
        force_hypothesis(S, dirs, b, f);
        if (world)
            for (int c = 0; c < 6; c++) world[6 * b + c] = f[c];
        double w[6], wl[6], wj[6];
        for (int c = 0; c < 6; c++) w[c] = f[c];
        wrench_act_inv(M1, M1 + 9, w, wl);
        wrench_act_inv(fr, fr + 9, wl, wj);
        for (int c = 0; c < 6; c++) fext[6 * b + c] = (float)wj[c];
    }
}

// Estimator update after a solve (force_estimator.py:99-136) by one thread: the selected hypothesis (*best) is
// regenerated from the unchanged state, momentum / estimate / smoothed estimate move toward it, the search
// radius shrinks when an exploitation hypothesis (< 3) won and grows otherwise, the last five minimum
// prediction errors adjust it further, and the exploration directions get the rotation of the unit quaternion
// drawn from the uniforms u (the reference draws them with numpy.random.rand).
__global__ void k_force_update(ForceEst* st, const float* dirs, const int* best, const float* err, int B, float alpha,
                               float beta, double smoothing, double r_min, double r_max, double u1, double u2,
                               double u3) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ForceEst S = *st;
    const int bi = *best;
    if (bi < 0 || bi >= B) return;  // an index outside the batch leaves the estimator unchanged
    const double emin = (double)err[bi];  // the minimum error is the selected hypothesis's (first argmin)
    if (S.count < 5) {
        S.hist[S.count] = emin;
    } else {
        for (int k = 0; k < 4; k++) S.hist[k] = S.hist[k + 1];
        S.hist[4] = emin;
    }
    S.count++;
    float bf[6];
    force_hypothesis(*st, dirs, bi, bf);
    const float b32 = beta, omb = (float)(1.0 - (double)beta);
    const float a32 = alpha, oma = (float)(1.0 - (double)alpha);
    const float sm = (float)smoothing, oms = (float)(1.0 - smoothing);
    for (int c = 0; c < 6; c++) {
        const float delta = bf[c] - S.est[c];
        S.mom[c] = b32 * S.mom[c] + omb * delta;
        const float raw = a32 * bf[c] + oma * S.est[c];
        S.est[c] = 0.8f * S.est[c] + 0.2f * (raw + 0.5f * S.mom[c]);
        S.smooth[c] = oms * S.smooth[c] + sm * S.est[c];
    }
    if (bi < 3) {
        S.radius *= 0.95;
        S.confidence = fmin(1.0, S.confidence + 0.05);
    } else {
        S.radius *= 1.05;
        S.confidence = fmax(0.0, S.confidence - 0.1);
    }
    S.radius = fmin(fmax(S.radius, r_min), r_max);
    if (S.count > 5) {
        double mean = 0.0;
        for (int k = 0; k < 5; k++) mean += S.hist[k];
        mean /= 5.0;
        double var = 0.0;
        for (int k = 0; k < 5; k++) var += (S.hist[k] - mean) * (S.hist[k] - mean);
        const double sd = sqrt(var / 5.0);
        const double mean4 = (((S.hist[0] + S.hist[1]) + S.hist[2]) + S.hist[3]) / 4.0;
        if (sd < 0.01) {
            S.radius *= 0.9;
        } else if (S.hist[4] > 1.5 * mean4) {
            S.radius *= 1.3;
            S.confidence *= 0.5;
        }
        S.radius = fmin(fmax(S.radius, r_min), r_max);
    }
    // unit quaternion (x, y, z, w) from three uniforms, then its rotation matrix
    const double x = sqrt(1.0 - u1) * sin(2.0 * M_PI * u2), y = sqrt(1.0 - u1) * cos(2.0 * M_PI * u2);
    const double z = sqrt(u1) * sin(2.0 * M_PI * u3), w = sqrt(u1) * cos(2.0 * M_PI * u3);
    const double R[9] = {1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y),
                         2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x),
                         2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)};
    for (int k = 0; k < 9; k++) S.R[k] = (float)R[k];
    *st = S;
}

// ------------------------------------------------------------------ sim_forward (sim.cuh:14-48)
// End-effector position p(q_k) (origin of the last joint frame) of every knot of every problem, and optionally its
// Jacobian dp/dq (row-major 3 x nq per knot): one thread per (problem, knot).
template <int PL>
__global__ void __launch_bounds__(256) k_ee_positions(const float* xu, int traj, int N, int B, float* pos, float* jac) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ, ST = P::NX + P::NU;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * N) return;
    const int b = t / N, k = t % N;
    const float* x = xu + (size_t)b * traj + (size_t)k * ST;
    float q[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) q[i] = x[i];
    Kin<PL> K;
    K.set(q);
    float p[3], Jc[3];
    ee_position<PL>(K, -1, p, static_cast<float*>(nullptr));
    for (int i = 0; i < 3; i++) pos[(size_t)t * 3 + i] = p[i];
    if (!jac) return;
    for (int j = 0; j < NQ; j++) {
        ee_position<PL>(K, j, p, Jc);
        for (int i = 0; i < 3; i++) jac[(size_t)t * 3 * NQ + i * NQ + j] = Jc[i];
    }
}

template <int PL, typename S = float>
__global__ void k_sim(S* xn, const S* x, const S* u, const S* fext, S dt, int B) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    S q[NQ], qd[NQ], uu[NQ], qdd[NQ], L[NQ * (NQ + 1) / 2];
#pragma unroll
    for (int i = 0; i < NQ; i++) { q[i] = x[i]; qd[i] = x[NQ + i]; uu[i] = u[i]; }
    Kin<PL, S> K;
    K.set(q);
    forward_dynamics<PL>(K, qd, uu, fext + 6 * b, qdd, L);
#pragma unroll
    for (int i = 0; i < NQ; i++) {
        xn[b * 2 * NQ + NQ + i] = qd[i] + dt * qdd[i];
        xn[b * 2 * NQ + i] = q[i] + dt * qd[i] + S(0.5) * qdd[i] * dt * dt;
    }
}

}  // namespace gato

// Workspace of the allocation-free solve (gato_solve_noalloc): one static array per GPU, resident from the library's
// load until the process ends. Host code never passes its name to a kernel or a copy (in host code the name is not a
// device address): it uses the address hipGetSymbolAddress returns for the current device (scratch_address).
static __device__ __align__(256) unsigned char scratch[GATO_NOALLOC_SCRATCH_BYTES];

// =================================================================== host side
namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail((int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int plant_nq(int plant) { return plant == GATO_PLANT_IIWA14 ? 7 : plant == GATO_PLANT_INDY7 ? 6 : -1; }

// Device address of `scratch` on the current device, looked up once per device (each GPU holds its own copy).
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[MAX_DEVICES] = {};

int scratch_address(void** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(GATO_E_ARG, "device index beyond the allocation-free solve's table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

}  // namespace

struct gato_solver {
    int plant, N, B, nq, nx, nu, traj;
    gato_params prm;
    int adapt_rho = 1;
    char* pool = nullptr;
    size_t pool_bytes = 0;
    gato::Args a{};
    float *d_fext, *d_Qd, *d_Rd, *d_xref;
    float *d_rho, *d_drho, *d_drho_init, *d_mu, *d_tol;
    std::vector<float> h_rho_init, h_drho_init;
    bool joint_cost = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int* d_stop = nullptr;
    int* d_status = nullptr;  // per-problem GATO_STATUS_* bits of the last solve
    int G = 4;  // knot groups per work-group (launch shape; see grid_threads)
    int gkey_G = 0;
    float *lam_save = nullptr, *xu_tune = nullptr;  // launch-shape tuner: saved multipliers, scratch trajectory
    hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;
    // backward workspace: the step at the solution (lam_b, dz_b) and its correction solves (q_a, r_a, c_a,
    // gam_a, lam_a, dz_a); the adjoint solution (dz_v, lam_v) and its correction solves (q_w, r_w, c_w, gam_w,
    // lam_w, dz_w); the factorization of S that preconditions the backward's PCG (bwd_G, bwd_Di); the
    // backward PCG settings and the iteration counts of the two chains
    float *bwd_G = nullptr, *bwd_Di = nullptr;
    float *lam_b = nullptr, *dz_b = nullptr, *q_a = nullptr, *r_a = nullptr, *c_a = nullptr, *gam_a = nullptr,
          *lam_a = nullptr, *dz_a = nullptr, *dz_v = nullptr, *lam_v = nullptr, *d_bwd_tol = nullptr,
          *bst_kres = nullptr, *q_w = nullptr, *r_w = nullptr, *c_w = nullptr, *gam_w = nullptr, *lam_w = nullptr,
          *dz_w = nullptr;
    int *bst_pcg = nullptr, *bst_pcg_s = nullptr;
    int bwd_max_pcg = -1;  // -1: one iteration per round with the factorization preconditioner, 1000 with P
    int bwd_refine = 2;
    float bwd_tol = 1e-6f;
    // derivative of the solution (gato_set_backward_implicit): rounds, finite-difference step, and their workspace
    // (allocated on first use): per-problem steps, z +- eps v, the augmented right-hand side, and the KKT set-up
    // blocks at z + eps v (imp_p) and z - eps v (imp_m)
    int bwd_implicit = 0;
    float bwd_fd_step = 1e-4f;
    char* imp_pool = nullptr;
    double *imp_eps = nullptr, *imp_zp = nullptr, *imp_zm = nullptr;
    float* imp_rz = nullptr;
    gato::Lin64 imp_p{}, imp_m{};
    // graph mode: the captured solve, valid for the launch arguments it was captured with
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    gato::Args gkey{};
    double last_kernel_us = 0.0;
    int last_launches = 0;
    // host copies of device stats
    std::vector<int> h_pcg, h_sqp, h_conv;
    std::vector<float> h_kres, h_lsm, h_step, h_init, h_final;
};

namespace {

template <typename T>
T* carve(gato_solver* s, size_t& off, size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(s->pool + off);
    off += count * sizeof(T);
    return p;
}

// The solver's workspace, carved from s->pool in 256-byte-aligned arrays. The forward solve's arrays come first;
// with forward_only the layout stops after them (the allocation-free solve carves only those).
size_t layout(gato_solver* s, bool assign, bool forward_only = false) {
    const size_t N = s->N, B = s->B, nx = s->nx, nu = s->nu, nq = s->nq;
    size_t off = 0;
    auto F = [&](size_t n) -> float* { float* p = carve<float>(s, off, n); return assign ? p : nullptr; };
    auto I = [&](size_t n) -> int* { int* p = carve<int>(s, off, n); return assign ? p : nullptr; };
    gato::Args& a = s->a;
    float* t;
    t = F(B * 6); if (assign) s->d_fext = t;
    t = F(B * nx); if (assign) s->d_Qd = t;
    t = F(B * nu); if (assign) s->d_Rd = t;
    t = F(B * N * nx); if (assign) s->d_xref = t;
    t = F(B * N * nx); if (assign) a.lam = t;
    t = F(B); if (assign) s->d_rho = t;
    t = F(B); if (assign) s->d_drho = t;
    t = F(B); if (assign) s->d_drho_init = t;
    t = F(B); if (assign) s->d_mu = t;
    t = F(B); if (assign) s->d_tol = t;
    t = F(B * N * nx * nx); if (assign) a.A = t;
    t = F(B * N * nx * nu); if (assign) a.Bm = t;
    t = F(B * N * nx); if (assign) a.c = t;
    t = F(B * N * nq * nq); if (assign) a.Qp = t;
    t = F(B * N * nq); if (assign) a.Qv = t;
    t = F(B * N * nu); if (assign) a.Rg = t;
    t = F(B * N * nx); if (assign) a.q = t;
    t = F(B * N * nu); if (assign) a.r = t;
    t = F(B * N * nq * nq); if (assign) a.Qip = t;
    t = F(B * N * nq); if (assign) a.Qiv = t;
    t = F(B * N * nu); if (assign) a.Ri = t;
    t = F(B * N * nx * nx); if (assign) a.Sd = t;
    t = F(B * N * nx * nx); if (assign) a.So = t;
    t = F(B * N * nx * nx); if (assign) a.Pd = t;
    t = F(B * N * nx * nx); if (assign) a.Po = t;
    t = F(B * N * nx); if (assign) a.gam = t;
    t = F(B * (s->traj + nu)); if (assign) a.dz = t;
    t = F(B * gato::NUM_ALPHAS * N); if (assign) a.mpart = t;
    t = F(B); if (assign) a.phi = t;
    int* ip;
    ip = I(B); if (assign) a.conv = ip;
    ip = I(B); if (assign) a.active = ip;
    ip = I(std::max<size_t>(B, s->prm.max_sqp_iters)); if (assign) a.solved = ip;
    const size_t it = std::max<uint32_t>(1u, s->prm.max_sqp_iters);
    ip = I(it * B); if (assign) a.st_pcg = ip;
    t = F(it * B); if (assign) a.st_kres = t;
    t = F(it * B); if (assign) a.st_lsm = t;
    t = F(it * B); if (assign) a.st_step = t;
    ip = I(B); if (assign) a.st_sqp = ip;
    // [0] graph-mode stop flag, [1] selected hypothesis, [2] SQP loop iterations run, [3] line-search iterations
    ip = I(4); if (assign) s->d_stop = ip;
    t = F(B); if (assign) a.st_init = t;
    t = F(B); if (assign) a.st_final = t;
    ip = I(B); if (assign) s->d_status = ip;
    if (forward_only) return off + 256;
    t = F(B * N * nx); if (assign) s->lam_save = t;
    t = F(B * s->traj); if (assign) s->xu_tune = t;
    t = F(B * N * nx); if (assign) s->lam_b = t;
    t = F(B * (s->traj + nu)); if (assign) s->dz_b = t;
    t = F(B * N * nx); if (assign) s->q_a = t;
    t = F(B * N * nu); if (assign) s->r_a = t;
    t = F(B * N * nx); if (assign) s->c_a = t;
    t = F(B * (s->traj + nu)); if (assign) s->dz_v = t;
    t = F(B * N * nx); if (assign) s->lam_v = t;
    t = F(B * N * nx); if (assign) s->gam_a = t;
    t = F(B * N * nx); if (assign) s->lam_a = t;
    t = F(B * (s->traj + nu)); if (assign) s->dz_a = t;
    t = F(B); if (assign) s->d_bwd_tol = t;
    t = F(B); if (assign) s->bst_kres = t;
    ip = I(B); if (assign) s->bst_pcg = ip;
    ip = I(B); if (assign) s->bst_pcg_s = ip;
    t = F(B * N * nx); if (assign) s->q_w = t;
    t = F(B * N * nu); if (assign) s->r_w = t;
    t = F(B * N * nx); if (assign) s->c_w = t;
    t = F(B * N * nx); if (assign) s->gam_w = t;
    t = F(B * N * nx); if (assign) s->lam_w = t;
    t = F(B * (s->traj + nu)); if (assign) s->dz_w = t;
    t = F(B * N * nx * nx); if (assign) s->bwd_G = t;  // the backward's factorization of S
    t = F(B * N * nx * nx); if (assign) s->bwd_Di = t;
    return off + 256;
}

// Launch shape: G 16-lane knot groups per work-group (a multiple of 4: whole wavefronts, since the DPP
// reductions read all 64 lanes of every wave; surplus groups own no knot). Each group owns knots
// g, g + G, ...: ceil(N / G) of them, at most 8, and the kernels are instantiated for 1, 2, 4 and 8 passes over
// them (PCG keeps a row per pass and lane), so ceil(N / G) is rounded up to the next of those (a pass whose knot
// is beyond N owns nothing).
int passes(int N, int G) {
    const int np = (N + G - 1) / G;
    return np <= 2 ? np : np <= 4 ? 4 : 8;
}
int default_groups(int N) { return 4 * ((std::min(N, 32) + 3) / 4); }
// Backward refinement rounds by default: each round cuts the relative gradient error by about 8x down to the fp32
// floor; two reach it at N <= 64, while at N = 128 two leave about 1e-2 and four about 2e-4 (measured against the
// FP64 oracle on the scale workload).
int default_refine(int N) { return N > 64 ? 4 : 2; }
bool valid_groups(int N, int G) {
    if (G < 4 || G > 32 || G % 4) return false;
    return (N + G - 1) / G <= 8 && G <= default_groups(N);
}
int grid_threads(const gato_solver* s) { return gato::KG * s->G; }
int check_threads(int B) { return std::min(1024, 64 * ((B + 63) / 64)); }
int npass(const gato_solver* s) { return passes(s->N, s->G); }

template <int PL>
void launch_schur(int B, dim3 blk, size_t lds, hipStream_t st, const gato::Args& a) {
    if (blk.x >= 256)
        hipLaunchKernelGGL((gato::k_schur_n<PL>), dim3(B), dim3(256), lds, st, a);
    else
        hipLaunchKernelGGL((gato::k_schur<PL>), dim3(B), blk, lds, st, a);
}

template <int PL>
void launch_merit(int B, dim3 blk, size_t lds, hipStream_t st, const gato::Args& a, int nal, int which, int it) {
    if (a.N * gato::NUM_ALPHAS <= 256 || blk.x >= 256)
        hipLaunchKernelGGL((gato::k_merit_n<PL>), dim3(B), dim3(256), lds, st, a, nal, which, it);
    else
        hipLaunchKernelGGL((gato::k_merit<PL>), dim3(B), blk, lds, st, a, nal, which, it);
}

// The backward's streamed PCG with the factorization preconditioner (k_pcg_btd), np knots per knot group.
template <int PL>
void launch_pcg_btd(int np, int B, dim3 blk, size_t lds, hipStream_t st, const gato::Args& a, const float* Gm,
                    const float* Di) {
    if (np == 1) hipLaunchKernelGGL((gato::k_pcg_btd<PL, 1>), dim3(B), blk, lds, st, a, Gm, Di);
    else if (np == 2) hipLaunchKernelGGL((gato::k_pcg_btd<PL, 2>), dim3(B), blk, lds, st, a, Gm, Di);
    else if (np == 4) hipLaunchKernelGGL((gato::k_pcg_btd<PL, 4>), dim3(B), blk, lds, st, a, Gm, Di);
    else hipLaunchKernelGGL((gato::k_pcg_btd<PL, 8>), dim3(B), blk, lds, st, a, Gm, Di);
}

template <int PL>
void launch_pcg(int np, int B, dim3 blk, size_t lds, hipStream_t st, const gato::Args& a, int it) {
    if (np == 1) hipLaunchKernelGGL((gato::k_pcg<PL, 1>), dim3(B), blk, lds, st, a, it);
    else if (np == 2) hipLaunchKernelGGL((gato::k_pcg<PL, 2>), dim3(B), blk, lds, st, a, it);
    else if (np == 4) hipLaunchKernelGGL((gato::k_pcg<PL, 4>), dim3(B), blk, lds, st, a, it);
    else hipLaunchKernelGGL((gato::k_pcg<PL, 8>), dim3(B), blk, lds, st, a, it);
}

template <int PL>
void launch_pcg_dz(int np, int B, dim3 blk, size_t lds, hipStream_t st, const gato::Args& a, int it) {
    if (np == 1) hipLaunchKernelGGL((gato::k_pcg_dz<PL, 1>), dim3(B), blk, lds, st, a, it);
    else if (np == 2) hipLaunchKernelGGL((gato::k_pcg_dz<PL, 2>), dim3(B), blk, lds, st, a, it);
    else if (np == 4) hipLaunchKernelGGL((gato::k_pcg_dz<PL, 4>), dim3(B), blk, lds, st, a, it);
    else hipLaunchKernelGGL((gato::k_pcg_dz<PL, 8>), dim3(B), blk, lds, st, a, it);
}

template <int PL>
void launch_fused(int np, int B, dim3 blk, size_t lds, hipStream_t st, const gato::Args& a) {
    if (np == 1) hipLaunchKernelGGL((gato::k_fused<PL, 1>), dim3(B), blk, lds, st, a);
    else if (np == 2) hipLaunchKernelGGL((gato::k_fused<PL, 2>), dim3(B), blk, lds, st, a);
    else if (np == 4) hipLaunchKernelGGL((gato::k_fused<PL, 4>), dim3(B), blk, lds, st, a);
    else hipLaunchKernelGGL((gato::k_fused<PL, 8>), dim3(B), blk, lds, st, a);
}

// dynamic LDS of the per-phase kernels and k_fused for G knot groups per work-group
size_t lds_bytes_g(const gato_solver* s, int G) {
    const int nx = s->nx, nq = s->nq;
    size_t f = 0;
    f = std::max<size_t>(f, (size_t)G * (24 * nq + 9 * (nq * (nq + 1) / 2) + 8));  // lin (lin_lds_floats)
    const size_t tile = s->plant == GATO_PLANT_IIWA14 ? gato::SchurTile<0>::SIZE : gato::SchurTile<1>::SIZE;
    f = std::max<size_t>(f, (size_t)G * tile);                                  // schur (SchurTile)
    f = std::max<size_t>(f, (size_t)2 * (s->N + 2) * 16 + 64);      // pcg
    f = std::max<size_t>(f, (size_t)G * gato::KG + 16);             // dz
    f = std::max<size_t>(f, (size_t)(2 + s->N) * gato::NUM_ALPHAS);  // merit + ls (out, shs, partial terms)
    return f * sizeof(float);
}

size_t lds_bytes(const gato_solver* s) { return lds_bytes_g(s, s->G); }

// LDS of the plain and graph paths' PCG launches (k_pcg, k_pcg_dz): with two knots per knot group the PCG is held on
// the CU (pcg_dispatch), which needs pcg_res_lds_floats(N), at most 64 KB for N <= 64
size_t pcg_lds_bytes(const gato_solver* s) {
    size_t f = lds_bytes(s);
    if (npass(s) == 2)
        f = std::max(f, (size_t)(s->plant == GATO_PLANT_IIWA14 ? gato::pcg_res_lds_floats<0>(s->N)
                                                             : gato::pcg_res_lds_floats<1>(s->N)) * sizeof(float));
    return f;
}

// LDS of k_fused and k_fused_mixed for G knot groups: the phases', or with two or four knots per group the resident
// PCG's (pcg_dispatch_fused) where that is larger
size_t fused_lds_bytes(const gato_solver* s, int G) {
    size_t f = lds_bytes_g(s, G);
    const int lin = s->plant == GATO_PLANT_IIWA14 ? gato::lin_lds_floats<0>() : gato::lin_lds_floats<1>();
    f = std::max(f, (size_t)gato::fused_lin_slots(s->N, G) * lin * sizeof(float));  // the linearization's chunk
    const int np = passes(s->N, G);
    if (np == 2 || np == 4) {
        const int r = s->plant == GATO_PLANT_IIWA14 ? gato::pcg_res_lds_floats<0>(s->N) : gato::pcg_res_lds_floats<1>(s->N);
        f = std::max(f, (size_t)r * sizeof(float));
    }
    return f;
}

void fill_args(gato_solver* s, float* d_xu, const float* d_xs, const float* d_ref, float dt) {
    gato::Args& a = s->a;
    a.N = s->N;
    a.B = s->B;
    a.traj = s->traj;
    a.max_sqp = s->prm.max_sqp_iters;
    a.max_pcg = s->prm.max_pcg_iters;
    a.exit_rule = s->prm.exit_rule;
    a.adapt_rho = s->adapt_rho;
    a.dt = dt;
    a.kkt_tol = s->prm.kkt_tol;
    a.solve_ratio = s->prm.solve_ratio;
    a.pcg_abs_tol = (float)gato::PCG_ABS_TOL;
    a.skip_po = 0;
    a.cw = gato::CostW{s->prm.q_cost, s->prm.N_cost, s->prm.q_lim_cost, s->prm.vel_lim_cost, s->prm.ctrl_lim_cost,
                       s->prm.terminal_mode, s->prm.hess_mode};
    a.xu = d_xu;
    a.xs = d_xs;
    a.ref = d_ref;
    a.fext = s->d_fext;
    a.Qd = s->d_Qd;
    a.Rd = s->d_Rd;
    a.xref = s->joint_cost ? s->d_xref : nullptr;
    a.rho = s->d_rho;
    a.drho = s->d_drho;
    a.drho_init = s->d_drho_init;
    a.mu = s->d_mu;
    a.tol = s->d_tol;
    a.stop = nullptr;
    a.stop_w = nullptr;
}

// Every SQP iteration of the plain path enqueued without a host round trip: PCG with dz and the convergence
// bookkeeping (k_pcg_dz), then the line search, whose batch-level exit test raises the device stop flag so that
// the remaining launches return at once. For N <= 32 the line search shares a launch with the next iteration's
// linearization and Schur phase, and the solve start and end with the first and last of those (k_first_n,
// k_ls_lin_schur_n, k_last_n); otherwise the iteration is k_lin, k_schur, k_pcg_dz and k_merit.
template <int PL>
void enqueue_graph_solve(gato_solver* s, hipStream_t st) {
    const int B = s->B, N = s->N;
    const dim3 blk(grid_threads(s));
    const size_t lds = lds_bytes(s);
    const gato::Args& a = s->a;
    const int sb = (B + 255) / 256;
    const int np = npass(s);
    const size_t nz = (size_t)B * (s->traj + s->nu);
    const int iters = (int)s->prm.max_sqp_iters;
    if (iters >= 1 && N <= 32 && N * gato::NUM_ALPHAS <= 256 && blk.x >= 256) {
        const int ts = std::min(N, s->G);  // LDS tables of the linearization: lds_bytes holds s->G of them
        hipLaunchKernelGGL((gato::k_first_n<PL>), dim3(B), dim3(256), lds, st, a, ts);
        for (int it = 0; it < iters; it++) {
            launch_pcg_dz<PL>(np, B, blk, pcg_lds_bytes(s), st, a, it);
            if (it + 1 < iters)
                hipLaunchKernelGGL((gato::k_ls_lin_schur_n<PL>), dim3(B), dim3(256), lds, st, a, it + 1, ts);
            else
                hipLaunchKernelGGL((gato::k_last_n<PL>), dim3(B), dim3(256), lds, st, a, it);
        }
        return;
    }
    hipLaunchKernelGGL(gato::k_begin, dim3(sb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(gato::k_zero, dim3(256), dim3(256), 0, st, a.dz, nz);
    launch_merit<PL>(B, blk, lds, st, a, 1, 0, 0);
    // 64 < N <= 128: the PCG held on the CU in 1024-thread work-groups, two knots per knot group (k_pcg_dz_res)
    const bool wide = N > 64 && N <= 128;
    const size_t lds_res = std::max(lds, (size_t)gato::pcg_res_lds_floats<PL>(N) * sizeof(float));
    gato::Args as = a;  // k_pcg_dz_res applies P in factored form: its off-diagonal blocks are not formed
    as.skip_po = wide ? 1 : 0;
    for (int it = 0; it < iters; it++) {
        hipLaunchKernelGGL((gato::k_lin<PL>), dim3(B), blk, lds, st, a);
        launch_schur<PL>(B, blk, lds, st, as);
        if (wide)
            hipLaunchKernelGGL((gato::k_pcg_dz_res<PL>), dim3(B), dim3(1024), lds_res, st, a, it);
        else
            launch_pcg_dz<PL>(np, B, blk, pcg_lds_bytes(s), st, a, it);
        launch_merit<PL>(B, blk, lds, st, a, gato::NUM_ALPHAS, 3, it);
    }
    hipLaunchKernelGGL(gato::k_zero, dim3(256), dim3(256), 0, st, a.dz, nz);
    hipLaunchKernelGGL(gato::k_end, dim3(sb), dim3(256), 0, st, a);  // with the final merit (phi)
}

void drop_graph(gato_solver* s) {
    if (s->gexec) hipGraphExecDestroy(s->gexec);
    if (s->graph) hipGraphDestroy(s->graph);
    s->gexec = nullptr;
    s->graph = nullptr;
}

template <int PL>
int launch_graph(gato_solver* s, hipStream_t st) {
    gato::Args& a = s->a;
    a.stop = s->d_stop;
    a.stop_w = s->d_stop;
    if (!s->gexec || s->gkey_G != s->G || std::memcmp(&s->gkey, &a, sizeof(gato::Args)) != 0) {
        drop_graph(s);
        if (!s->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(s->cap_stream, hipStreamCaptureModeThreadLocal));
        enqueue_graph_solve<PL>(s, s->cap_stream);
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(s->cap_stream, &s->graph);
        if (le != hipSuccess) { drop_graph(s); return fail((int)le, std::string("graph capture launch: ") + hipGetErrorString(le)); }
        if (ce != hipSuccess) { drop_graph(s); return fail((int)ce, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce)); }
        const hipError_t ie = hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) { drop_graph(s); return fail((int)ie, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie)); }
        std::memcpy(&s->gkey, &a, sizeof(gato::Args));
        s->gkey_G = s->G;
    }
    HIP_TRY(hipEventRecord(s->ev0, st));
    HIP_TRY(hipGraphLaunch(s->gexec, st));
    HIP_TRY(hipEventRecord(s->ev1, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_kernel_us = ms * 1e3;
    s->last_launches = 1;
    return 0;
}

template <int PL>
int launch_solve(gato_solver* s, hipStream_t st) {
    if (s->prm.mode == GATO_MODE_GRAPH) return launch_graph<PL>(s, st);
    const int B = s->B, N = s->N;
    const dim3 blk(grid_threads(s));
    const size_t lds = lds_bytes(s);
    gato::Args& a = s->a;
    const int sb = (B + 255) / 256;
    hipLaunchKernelGGL(gato::k_begin, dim3(sb), dim3(256), 0, st, a);
    HIP_TRY(hipGetLastError());
    const int np = npass(s);
    if (s->prm.mode == GATO_MODE_FUSED) {
        HIP_TRY(hipEventRecord(s->ev0, st));
        launch_fused<PL>(np, B, blk, fused_lds_bytes(s, s->G), st, a);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(s->ev1, st));
        HIP_TRY(hipStreamSynchronize(st));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        s->last_kernel_us = ms * 1e3;
        s->last_launches = 1;
        return 0;
    }
    // plain: host-driven loop (bsqp.cuh:103-197)
    {
        const size_t n = (size_t)B * (s->traj + s->nu);
        hipLaunchKernelGGL(gato::k_zero, dim3(256), dim3(256), 0, st, a.dz, n);
    }
    launch_merit<PL>(B, blk, lds, st, a, 1, 0, 0);
    HIP_TRY(hipGetLastError());
    double pcg_ms = 0.0;
    int pcg_launches = 0;
    int it = 0;
    for (; it < (int)s->prm.max_sqp_iters; it++) {
        hipLaunchKernelGGL((gato::k_lin<PL>), dim3(B), blk, lds, st, a);
        launch_schur<PL>(B, blk, lds, st, a);
        HIP_TRY(hipEventRecord(s->ev0, st));
        launch_pcg<PL>(np, B, blk, pcg_lds_bytes(s), st, a, it);
        HIP_TRY(hipEventRecord(s->ev1, st));
        hipLaunchKernelGGL((gato::k_dz<PL>), dim3(B), blk, lds, st, a, it);
        hipLaunchKernelGGL(gato::k_check, dim3(1), dim3(check_threads(B)), 0, st, a, it);
        HIP_TRY(hipGetLastError());
        int solved = 0;
        HIP_TRY(hipMemcpyAsync(&solved, a.solved + it, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        pcg_ms += ms;
        pcg_launches++;
        if ((double)solved >= (double)B * s->prm.solve_ratio) { it++; break; }
        launch_merit<PL>(B, blk, lds, st, a, gato::NUM_ALPHAS, 2, it);
        HIP_TRY(hipGetLastError());
    }
    {
        const size_t n = (size_t)B * (s->traj + s->nu);
        hipLaunchKernelGGL(gato::k_zero, dim3(256), dim3(256), 0, st, a.dz, n);
    }
    hipLaunchKernelGGL(gato::k_end, dim3(sb), dim3(256), 0, st, a);  // with the final merit (phi)
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    s->last_kernel_us = pcg_launches ? pcg_ms * 1e3 / pcg_launches : 0.0;
    s->last_launches = pcg_launches;
    return 0;
}


// The backward pass: the step at the solution and the adjoint system, each solved with the step's Schur
// complement, preconditioned with its twisted block-tridiagonal factorization (k_btd_factor; k_pcg_dual for both
// chains with the rows of S in registers when N <= 32, k_pcg_btd per chain otherwise), and improved by
// `bwd_refine` rounds of iterative refinement (the residual is evaluated with the KKT blocks themselves, so fp32
// error in S and Q~^-1 does not limit the result).
template <int PL>
int launch_backward(gato_solver* s, const float* dL, float* gxs, float* gref, float* gQ, float* gR, hipStream_t st) {
    const int B = s->B, N = s->N, np = npass(s);
    const dim3 blk(grid_threads(s));
    const size_t lds = std::max(lds_bytes(s), (size_t)2 * N * gato::KG * sizeof(float));
    // the dual PCG keeps six padded vectors, two pairs of reduction slots, the problem's factor blocks and the
    // sweeps' middle vectors in LDS
    const size_t lds_dual =
        std::max(lds, (size_t)(6 * (N + 2) * 16 + 128 + 2 * N * s->nx * s->nx + 64) * sizeof(float));
    // the streamed PCG with the factorization keeps p, r, z, the reduction slots and the middle vectors in LDS
    const size_t lds_btd = std::max(lds, (size_t)(3 * (N + 2) * 16 + 16 + 32) * sizeof(float));
    const int nlam = N * s->nx, ndz = s->traj + s->nu;
    gato::Args ab = s->a;  // the step at the solution, on its own multipliers and dz
    ab.exit_rule = GATO_EXIT_PCG0;
    ab.stop = nullptr;
    ab.stop_w = nullptr;
    ab.lam = s->lam_b;
    ab.dz = s->dz_b;
    ab.st_pcg = s->bst_pcg;
    ab.st_kres = s->bst_kres;
    hipLaunchKernelGGL(gato::k_bwd_begin, dim3(B), dim3(256), 0, st, ab, s->a.lam, s->lam_b, s->lam_a, s->dz_b,
                       s->dz_a, s->dz_v, s->lam_v, s->lam_w, nlam, ndz);
    hipLaunchKernelGGL((gato::k_lin<PL>), dim3(B), blk, lds, st, ab);
    launch_schur<PL>(B, blk, lds, st, ab);
    hipLaunchKernelGGL((gato::k_btd_factor<PL>), dim3(B), dim3(128), 0, st, ab, s->bwd_G, s->bwd_Di);
    // correction solves of the step: same blocks, right-hand side (q_a, r_a, c_a) -> gam_a, solution (dz_a, lam_a)
    gato::Args ac = ab;
    ac.q = s->q_a;
    ac.r = s->r_a;
    ac.gam = s->gam_a;
    ac.lam = s->lam_a;
    ac.dz = s->dz_a;
    ac.max_pcg = s->bwd_max_pcg > 0 ? s->bwd_max_pcg : 1;
    ab.max_pcg = ac.max_pcg;  // the step's own solve too: the refinement rounds correct it like its corrections
    ac.tol = s->d_bwd_tol;
    ac.pcg_abs_tol = 1e-30f;
    ac.st_pcg = s->bst_pcg_s;
    ab.st_pcg = s->bst_pcg_s;
    // solves of the adjoint K [v; v_l] = [dL/dz; 0] (from zero, then corrections): (q_w, r_w, c_w) -> gam_w,
    // solution (dz_w, lam_w), accumulated into (dz_v, lam_v)
    gato::Args aw = ac;
    aw.q = s->q_w;
    aw.r = s->r_w;
    aw.gam = s->gam_w;
    aw.lam = s->lam_w;
    aw.dz = s->dz_w;
    aw.st_pcg = s->bst_pcg;
    // residuals of the step's correction (from refinement round 1 on) and of the adjoint: (q, r, c) and gamma
    const gato::ResidChain rs{nullptr, s->a.q, s->a.r, s->a.c, s->dz_b, s->lam_b, s->q_a, s->r_a, s->c_a, s->gam_a};
    const gato::ResidChain rw{dL, nullptr, nullptr, nullptr, s->dz_v, s->lam_v, s->q_w, s->r_w, s->c_w, s->gam_w};
    // the two chains are independent and share K's Schur complement and preconditioner: round i solves the
    // step's correction i (refinement round 0: the step itself) and the adjoint's correction i side by side
    for (int i = 0; i <= s->bwd_refine; i++) {
        const gato::Args& as = i == 0 ? ab : ac;
        hipLaunchKernelGGL((gato::k_residual2<PL>), dim3(B), blk, lds, st, ab, rs, rw, i > 0 ? 1 : 0);
        if (np == 1) {  // the solves, their dz and the accumulation in one launch
            hipLaunchKernelGGL((gato::k_pcg_dual<PL>), dim3(B), blk, lds_dual, st, as, aw, s->bwd_G, s->bwd_Di,
                               i > 0 ? s->dz_b : nullptr, i > 0 ? s->lam_b : nullptr, s->dz_v, s->lam_v);
            continue;
        }
        // S streams: the chains are separate launches (a paired streamed PCG measured slower than two launches)
        launch_pcg_btd<PL>(np, B, blk, lds_btd, st, as, s->bwd_G, s->bwd_Di);
        launch_pcg_btd<PL>(np, B, blk, lds_btd, st, aw, s->bwd_G, s->bwd_Di);
        hipLaunchKernelGGL((gato::k_dz<PL>), dim3(B), blk, lds, st, as, 0);
        if (i > 0)
            hipLaunchKernelGGL(gato::k_accumulate, dim3(B), dim3(256), 0, st, s->dz_b, s->dz_a, s->lam_b, s->lam_a, ndz,
                               nlam);
        hipLaunchKernelGGL((gato::k_dz<PL>), dim3(B), blk, lds, st, aw, 0);
        hipLaunchKernelGGL(gato::k_accumulate, dim3(B), dim3(256), 0, st, s->dz_v, s->dz_w, s->lam_v, s->lam_w, ndz,
                           nlam);
    }
    if (s->bwd_implicit > 0) {
        // derivative of the solution: rounds of K w+ = [dL/dz; 0] + (G~ - H) v on the adjoint chain (k_perturb,
        // the set-up at z +- eps v, k_hvp_rhs, then a refinement round with that right-hand side), H from the step's
        // multipliers lam_b; the step term of the parameter gradients is dropped (dz = 0 at a solution)
        const gato::ResidChain ri{s->imp_rz, nullptr, nullptr, nullptr, s->dz_v, s->lam_v, s->q_w, s->r_w, s->c_w,
                                  s->gam_w};
        constexpr int NXp = gato::Plant<PL>::NX, NUp = gato::Plant<PL>::NU;
        const size_t lds64 = (size_t)(2 * NXp + NUp + 8 + 8 * gato::lin64_group_doubles<PL>()) * sizeof(double);
        for (int it = 0; it < s->bwd_implicit; it++) {
            hipLaunchKernelGGL(gato::k_perturb, dim3(B), dim3(256), 0, st, ab.xu, s->dz_v, s->imp_zp, s->imp_zm,
                               s->imp_eps, s->traj, ndz, (double)s->bwd_fd_step);
            hipLaunchKernelGGL((gato::k_lin64<PL>), dim3(B), dim3(8 * gato::KG), lds64, st, ab, s->imp_zp, s->imp_p);
            hipLaunchKernelGGL((gato::k_lin64<PL>), dim3(B), dim3(8 * gato::KG), lds64, st, ab, s->imp_zm, s->imp_m);
            hipLaunchKernelGGL((gato::k_hvp_rhs<PL>), dim3(B), blk, 0, st, ab, s->imp_p, s->imp_m, dL, s->dz_v,
                               s->lam_b, s->imp_eps, s->imp_rz);
            hipLaunchKernelGGL((gato::k_residual2<PL>), dim3(B), blk, lds, st, ab, rs, ri, 0);
            launch_pcg_btd<PL>(np, B, blk, lds_btd, st, aw, s->bwd_G, s->bwd_Di);
            hipLaunchKernelGGL((gato::k_dz<PL>), dim3(B), blk, lds, st, aw, 0);
            hipLaunchKernelGGL(gato::k_accumulate, dim3(B), dim3(256), 0, st, s->dz_v, s->dz_w, s->lam_v, s->lam_w,
                               ndz, nlam);
        }
        hipLaunchKernelGGL(gato::k_zero, dim3(256), dim3(256), 0, st, s->dz_b, (size_t)B * ndz);
    }
    hipLaunchKernelGGL((gato::k_grad<PL>), dim3(B), blk, lds, st, ab, s->dz_v, s->lam_v, gxs, gref, gQ, gR);
    HIP_TRY(hipGetLastError());
    return 0;
}

// The implicit rounds' workspace (allocated on the first gato_set_backward_implicit with rounds > 0): per-problem
// steps and the perturbed trajectories z +- eps v in fp64, the augmented right-hand side, and the fp64 KKT set-up
// blocks at the two perturbed trajectories.
int alloc_implicit(gato_solver* s) {
    if (s->imp_pool) return 0;
    const size_t N = s->N, B = s->B, nx = s->nx, nu = s->nu, nq = s->nq, traj = s->traj;
    const size_t set = N * (nx * nx + nx * nu + nx + nq * nq + nq + nu + nx + nu);  // A Bm c Qp Qv Rg q r
    const size_t per = 1 + 2 * traj + 2 * set;                                     // doubles
    const size_t bytes = (B * per + 32 * 16) * sizeof(double) + (B * traj + 64) * sizeof(float);
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(GATO_E_ALLOC, std::string("hipMalloc implicit workspace: ") + hipGetErrorString(e));
    s->imp_pool = static_cast<char*>(p);
    size_t off = 0;  // in doubles, 256-byte aligned arrays
    auto D = [&](size_t n) {
        off = (off + 31) & ~size_t(31);
        double* q = reinterpret_cast<double*>(s->imp_pool) + off;
        off += n;
        return q;
    };
    s->imp_eps = D(B);
    s->imp_zp = D(B * traj);
    s->imp_zm = D(B * traj);
    for (gato::Lin64* a : {&s->imp_p, &s->imp_m}) {
        a->A = D(B * N * nx * nx);
        a->Bm = D(B * N * nx * nu);
        a->c = D(B * N * nx);
        a->Qp = D(B * N * nq * nq);
        a->Qv = D(B * N * nq);
        a->Rg = D(B * N * nu);
        a->q = D(B * N * nx);
        a->r = D(B * N * nu);
    }
    s->imp_rz = reinterpret_cast<float*>(D((B * traj + 1) / 2 + 1));
    return 0;
}

int upload_defaults(gato_solver* s) {
    const int B = s->B;
    std::vector<float> Qd((size_t)B * s->nx), Rd((size_t)B * s->nu);
    for (int b = 0; b < B; b++) {
        for (int i = 0; i < s->nx; i++) Qd[(size_t)b * s->nx + i] = i < s->nq ? 0.f : s->prm.qd_cost;
        for (int i = 0; i < s->nu; i++) Rd[(size_t)b * s->nu + i] = s->prm.u_cost;
    }
    HIP_TRY(hipMemcpy(s->d_Qd, Qd.data(), Qd.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->d_Rd, Rd.data(), Rd.size() * 4, hipMemcpyHostToDevice));
    return 0;
}


// Device statistics of the last solve -> host gato_stats (bindings.cu:96-147 result fields).
int collect_stats(gato_solver* s, hipStream_t st, gato_stats* stats, double t_us) {
    const int B = s->B;
    const size_t it = s->prm.max_sqp_iters;
    gato::Args& a = s->a;
    HIP_TRY(hipMemcpyAsync(s->h_pcg.data(), a.st_pcg, it * B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_kres.data(), a.st_kres, it * B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_lsm.data(), a.st_lsm, it * B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_step.data(), a.st_step, it * B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_sqp.data(), a.st_sqp, B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_conv.data(), a.conv, B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_init.data(), a.st_init, B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_final.data(), a.st_final, B * 4, hipMemcpyDeviceToHost, st));
    if (stats->status) {
        hipLaunchKernelGGL(gato::k_status<float>, dim3(B), dim3(64), 0, st, a, s->d_status);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(stats->status, s->d_status, B * 4, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    stats->solve_time_us = t_us;
    int n_iters = 0;
    for (int b = 0; b < B; b++) n_iters = std::max(n_iters, s->h_sqp[b]);
    int ls_iters = 0;
    for (size_t i = 0; i < it; i++)
        for (int b = 0; b < B; b++)
            if (s->h_step[i * B + b] != 0.f) ls_iters = std::max<int>(ls_iters, (int)i + 1);
    stats->n_iters = n_iters;
    stats->ls_num_iters = ls_iters;
    auto cp = [&](auto* dst, const auto& src, size_t n) { if (dst) std::memcpy(dst, src.data(), n * sizeof(src[0])); };
    cp(stats->sqp_iters, s->h_sqp, B);
    cp(stats->kkt_converged, s->h_conv, B);
    cp(stats->initial_merit, s->h_init, B);
    cp(stats->final_merit, s->h_final, B);
    cp(stats->pcg_iters, s->h_pcg, it * B);
    cp(stats->ls_min_merit, s->h_lsm, it * B);
    cp(stats->ls_step_size, s->h_step, it * B);
    cp(stats->kkt_residual, s->h_kres, it * B);
    return 0;
}

// Launch-shape tuner: times one whole solve (after a warm-up) for every valid knot-group count on the
// caller's inputs, with the persistent state (lambda, rho, drho) saved before and restored after each
// trial and the trajectory solved in a scratch copy; keeps the fastest shape.
int autotune(gato_solver* s, float* d_xu, const float* d_xs, const float* d_ref, float dt, hipStream_t st,
             int32_t* chosen) {
    const size_t nlam = (size_t)s->B * s->N * s->nx, nxu = (size_t)s->B * s->traj;
    std::vector<float> rho(s->B), drho(s->B);
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMemcpy(s->lam_save, s->a.lam, nlam * 4, hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(rho.data(), s->d_rho, s->B * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(drho.data(), s->d_drho, s->B * 4, hipMemcpyDeviceToHost));
    auto restore = [&]() -> int {
        HIP_TRY(hipMemcpy(s->a.lam, s->lam_save, nlam * 4, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(s->d_rho, rho.data(), s->B * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->d_drho, drho.data(), s->B * 4, hipMemcpyHostToDevice));
        return 0;
    };
    const int G0 = s->G;
    int best = G0;
    float best_ms = 1e30f;
    for (int G = 4; G <= 32; G *= 2) {
        if (!valid_groups(s->N, G)) continue;
        s->G = G;
        float ms = 0.f;
        for (int rep = 0; rep < 2; rep++) {
            int rc = restore();
            if (rc) { s->G = G0; return rc; }
            HIP_TRY(hipMemcpyAsync(s->xu_tune, d_xu, nxu * 4, hipMemcpyDeviceToDevice, st));
            fill_args(s, s->xu_tune, d_xs, d_ref, dt);
            HIP_TRY(hipEventRecord(s->ev_t0, st));
            rc = s->plant == GATO_PLANT_IIWA14 ? launch_solve<0>(s, st) : launch_solve<1>(s, st);
            if (rc) { s->G = G0; restore(); return rc; }
            HIP_TRY(hipEventRecord(s->ev_t1, st));
            HIP_TRY(hipStreamSynchronize(st));
            HIP_TRY(hipEventElapsedTime(&ms, s->ev_t0, s->ev_t1));
        }
        if (ms < best_ms) { best_ms = ms; best = G; }
    }
    s->G = best;
    fill_args(s, d_xu, d_xs, d_ref, dt);
    const int rc = restore();
    if (rc) return rc;
    if (chosen) *chosen = best;
    return 0;
}

}  // namespace

// =================================================================== FP64 mode
// The reference's USE_DOUBLES build (gato/settings.h:7-11, bindings.cu:244-252): the same SQP iteration in double
// precision. It runs the plain path's kernels instantiated at S = double (one kernel per phase, a host loop reading
// the converged count every iteration, the streamed PCG), with library sin, cos, division and square root.
struct gato_solver64 {
    int plant, N, B, nq, nx, nu, traj;
    gato_params prm;
    int adapt_rho = 1;
    int G = 4;
    char* pool = nullptr;
    gato::ArgsT<double> a{};
    double *d_fext, *d_Qd, *d_Rd, *d_xref, *d_rho, *d_drho, *d_drho_init, *d_mu, *d_tol;
    int* d_status = nullptr;
    std::vector<double> h_rho_init, h_drho_init;
    bool joint_cost = false;
    std::vector<int> h_pcg, h_sqp, h_conv;
    std::vector<double> h_kres, h_lsm, h_step, h_init, h_final;
};

namespace {

// The FP64 solver's workspace (the fp32 solver's forward arrays, in doubles).
size_t layout64(gato_solver64* s, bool assign) {
    const size_t N = s->N, B = s->B, nx = s->nx, nu = s->nu, nq = s->nq;
    size_t off = 0;
    auto D = [&](size_t n) -> double* {
        off = (off + 255) & ~size_t(255);
        double* p = reinterpret_cast<double*>(s->pool + off);
        off += n * sizeof(double);
        return assign ? p : nullptr;
    };
    auto I = [&](size_t n) -> int* {
        off = (off + 255) & ~size_t(255);
        int* p = reinterpret_cast<int*>(s->pool + off);
        off += n * sizeof(int);
        return assign ? p : nullptr;
    };
    gato::ArgsT<double>& a = s->a;
    const size_t it = std::max<uint32_t>(1u, s->prm.max_sqp_iters);
    s->d_fext = D(B * 6); s->d_Qd = D(B * nx); s->d_Rd = D(B * nu); s->d_xref = D(B * N * nx);
    a.lam = D(B * N * nx);
    s->d_rho = D(B); s->d_drho = D(B); s->d_drho_init = D(B); s->d_mu = D(B); s->d_tol = D(B);
    a.A = D(B * N * nx * nx); a.Bm = D(B * N * nx * nu); a.c = D(B * N * nx);
    a.Qp = D(B * N * nq * nq); a.Qv = D(B * N * nq); a.Rg = D(B * N * nu); a.q = D(B * N * nx); a.r = D(B * N * nu);
    a.Qip = D(B * N * nq * nq); a.Qiv = D(B * N * nq); a.Ri = D(B * N * nu);
    a.Sd = D(B * N * nx * nx); a.So = D(B * N * nx * nx); a.Pd = D(B * N * nx * nx); a.Po = D(B * N * nx * nx);
    a.gam = D(B * N * nx); a.dz = D(B * (s->traj + nu)); a.mpart = D(B * gato::NUM_ALPHAS * N); a.phi = D(B);
    a.conv = I(B); a.active = I(B); a.solved = I(std::max<size_t>(B, s->prm.max_sqp_iters));
    a.st_pcg = I(it * B); a.st_kres = D(it * B); a.st_lsm = D(it * B); a.st_step = D(it * B); a.st_sqp = I(B);
    a.st_init = D(B); a.st_final = D(B);
    s->d_status = I(B);
    return off + 256;
}

// knot groups per work-group: one per knot up to 16 (the PCG then makes at most 8 passes for N <= 128)
int groups64(int N) { return std::min(16, default_groups(N)); }

int upload_defaults64(gato_solver64* s) {
    std::vector<double> Qd((size_t)s->B * s->nx), Rd((size_t)s->B * s->nu);
    for (int b = 0; b < s->B; b++) {
        for (int i = 0; i < s->nx; i++) Qd[(size_t)b * s->nx + i] = i < s->nq ? 0.0 : (double)s->prm.qd_cost;
        for (int i = 0; i < s->nu; i++) Rd[(size_t)b * s->nu + i] = (double)s->prm.u_cost;
    }
    HIP_TRY(hipMemcpy(s->d_Qd, Qd.data(), Qd.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->d_Rd, Rd.data(), Rd.size() * 8, hipMemcpyHostToDevice));
    return 0;
}

template <int PL, int NP>
void launch_pcg64(int B, dim3 blk, size_t lds, hipStream_t st, const gato::ArgsT<double>& a, int it) {
    hipLaunchKernelGGL((gato::k_pcg<PL, NP, double>), dim3(B), blk, lds, st, a, it);
}

// bsqp.cuh:103-197 in double: the plain path's host loop
template <int PL>
int launch_solve64(gato_solver64* s, hipStream_t st) {
    const int B = s->B, N = s->N;
    const dim3 blk(gato::KG * s->G);
    const int np = (N + s->G - 1) / s->G;
    const size_t tile = s->plant == GATO_PLANT_IIWA14 ? gato::SchurTile<0>::SIZE : gato::SchurTile<1>::SIZE;
    size_t f = 0;
    f = std::max<size_t>(f, (size_t)s->G * gato::lin_lds_floats<PL>());
    f = std::max<size_t>(f, (size_t)s->G * tile);
    f = std::max<size_t>(f, (size_t)2 * (N + 2) * 16 + 64);
    f = std::max<size_t>(f, (size_t)s->G * gato::KG + 16);
    f = std::max<size_t>(f, (size_t)(2 + N) * gato::NUM_ALPHAS);
    const size_t lds = f * sizeof(double);
    gato::ArgsT<double>& a = s->a;
    const int sb = (B + 255) / 256;
    const size_t nz = (size_t)B * (s->traj + s->nu);
    hipLaunchKernelGGL(gato::k_begin<double>, dim3(sb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(gato::k_zero<double>, dim3(256), dim3(256), 0, st, a.dz, nz);
    hipLaunchKernelGGL((gato::k_merit<PL, double>), dim3(B), blk, lds, st, a, 1, 0, 0);
    HIP_TRY(hipGetLastError());
    for (int it = 0; it < (int)s->prm.max_sqp_iters; it++) {
        hipLaunchKernelGGL((gato::k_lin<PL, double>), dim3(B), blk, lds, st, a);
        hipLaunchKernelGGL((gato::k_schur<PL, double>), dim3(B), blk, lds, st, a);
        if (np == 1) launch_pcg64<PL, 1>(B, blk, lds, st, a, it);
        else if (np == 2) launch_pcg64<PL, 2>(B, blk, lds, st, a, it);
        else if (np == 4) launch_pcg64<PL, 4>(B, blk, lds, st, a, it);
        else launch_pcg64<PL, 8>(B, blk, lds, st, a, it);
        hipLaunchKernelGGL((gato::k_dz<PL, double>), dim3(B), blk, lds, st, a, it);
        hipLaunchKernelGGL(gato::k_check<double>, dim3(1), dim3(check_threads(B)), 0, st, a, it);
        HIP_TRY(hipGetLastError());
        int solved = 0;
        HIP_TRY(hipMemcpyAsync(&solved, a.solved + it, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if ((double)solved >= (double)B * s->prm.solve_ratio) break;
        hipLaunchKernelGGL((gato::k_merit<PL, double>), dim3(B), blk, lds, st, a, gato::NUM_ALPHAS, 2, it);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(gato::k_zero<double>, dim3(256), dim3(256), 0, st, a.dz, nz);
    hipLaunchKernelGGL(gato::k_end<double>, dim3(sb), dim3(256), 0, st, a);  // with the final merit (phi)
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

}  // namespace

extern "C" {

const char* gato_last_error(void) { return g_err.c_str(); }

int gato_create_f64(gato_solver64** out, int32_t plant, int32_t knots, int32_t batch, const gato_params* params) {
    if (!out || !params) return fail(GATO_E_ARG, "null argument");
    const int nq = plant_nq(plant);
    if (nq < 0) return fail(GATO_E_ARG, "unknown plant");
    if (knots < 2 || knots > 128) return fail(GATO_E_ARG, "knots must be in [2, 128]");
    if (batch < 1) return fail(GATO_E_ARG, "batch must be >= 1");
    if (params->max_sqp_iters < 1 || params->max_sqp_iters > 1000) return fail(GATO_E_ARG, "max_sqp_iters must be in [1, 1000]");
    if (params->exit_rule != GATO_EXIT_PCG0 && params->exit_rule != GATO_EXIT_KKT) return fail(GATO_E_ARG, "unknown exit rule");
    gato_solver64* s = new gato_solver64();
    s->plant = plant; s->N = knots; s->B = batch; s->nq = nq; s->nx = 2 * nq; s->nu = nq;
    s->traj = (s->nx + s->nu) * knots - s->nu;
    s->prm = *params;
    s->prm.mode = GATO_MODE_PLAIN;
    s->G = groups64(knots);
    const size_t bytes = layout64(s, false);
    hipError_t e = hipMalloc(&s->pool, bytes);
    if (e != hipSuccess) {
        delete s;
        return fail(GATO_E_ALLOC, std::string("hipMalloc workspace: ") + hipGetErrorString(e));
    }
    layout64(s, true);
    auto cleanup = [&](int code) { gato_destroy_f64(s); return code; };
    if ((e = hipMemset(s->pool, 0, bytes)) != hipSuccess) return cleanup(fail((int)e, "hipMemset workspace"));
    s->h_rho_init.assign(batch, (double)params->rho);
    s->h_drho_init.assign(batch, 1.0);
    std::vector<double> mu(batch, (double)params->mu), tol(batch, (double)params->pcg_tol);
    if ((e = hipMemcpy(s->d_rho, s->h_rho_init.data(), batch * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(s->d_drho, s->h_drho_init.data(), batch * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(s->d_drho_init, s->h_drho_init.data(), batch * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(s->d_mu, mu.data(), batch * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(s->d_tol, tol.data(), batch * 8, hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(fail((int)e, "hipMemcpy initial state"));
    int rc = upload_defaults64(s);
    if (rc) return cleanup(rc);
    const size_t it = params->max_sqp_iters;
    s->h_pcg.resize(it * batch); s->h_kres.resize(it * batch); s->h_lsm.resize(it * batch); s->h_step.resize(it * batch);
    s->h_sqp.resize(batch); s->h_conv.resize(batch); s->h_init.resize(batch); s->h_final.resize(batch);
    *out = s;
    return 0;
}

int gato_destroy_f64(gato_solver64* s) {
    if (!s) return 0;
    if (s->pool) hipFree(s->pool);
    delete s;
    return 0;
}

int gato_solve_f64(gato_solver64* s, double* d_xu, const double* d_xs, const double* d_ref, double dt, void* stream,
                   gato_stats* stats) {
    if (!s || !d_xu || !d_xs || !d_ref) return fail(GATO_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    gato::ArgsT<double>& a = s->a;
    a.N = s->N; a.B = s->B; a.traj = s->traj; a.max_sqp = s->prm.max_sqp_iters; a.max_pcg = s->prm.max_pcg_iters;
    a.exit_rule = s->prm.exit_rule; a.adapt_rho = s->adapt_rho; a.dt = dt; a.kkt_tol = s->prm.kkt_tol;
    a.solve_ratio = s->prm.solve_ratio; a.pcg_abs_tol = gato::PCG_ABS_TOL;
    a.cw = gato::CostW{s->prm.q_cost, s->prm.N_cost, s->prm.q_lim_cost, s->prm.vel_lim_cost, s->prm.ctrl_lim_cost,
                       s->prm.terminal_mode, s->prm.hess_mode};
    a.xu = d_xu; a.xs = d_xs; a.ref = d_ref; a.fext = s->d_fext; a.Qd = s->d_Qd; a.Rd = s->d_Rd;
    a.xref = s->joint_cost ? s->d_xref : nullptr;
    a.rho = s->d_rho; a.drho = s->d_drho; a.drho_init = s->d_drho_init; a.mu = s->d_mu; a.tol = s->d_tol;
    a.stop = nullptr; a.stop_w = nullptr;
    const auto t0 = std::chrono::high_resolution_clock::now();
    const int rc = s->plant == GATO_PLANT_IIWA14 ? launch_solve64<0>(s, st) : launch_solve64<1>(s, st);
    if (rc) return rc;
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (!stats) return 0;
    const int B = s->B;
    const size_t it = s->prm.max_sqp_iters;
    HIP_TRY(hipMemcpyAsync(s->h_pcg.data(), a.st_pcg, it * B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_kres.data(), a.st_kres, it * B * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_lsm.data(), a.st_lsm, it * B * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_step.data(), a.st_step, it * B * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_sqp.data(), a.st_sqp, B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_conv.data(), a.conv, B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_init.data(), a.st_init, B * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(s->h_final.data(), a.st_final, B * 8, hipMemcpyDeviceToHost, st));
    if (stats->status) {
        hipLaunchKernelGGL(gato::k_status<double>, dim3(B), dim3(64), 0, st, a, s->d_status);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(stats->status, s->d_status, B * 4, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    stats->solve_time_us = std::chrono::duration<double, std::micro>(t1 - t0).count();
    int n_iters = 0, ls_iters = 0;
    for (int b = 0; b < B; b++) n_iters = std::max(n_iters, s->h_sqp[b]);
    for (size_t i = 0; i < it; i++)
        for (int b = 0; b < B; b++)
            if (s->h_step[i * B + b] != 0.0) ls_iters = std::max<int>(ls_iters, (int)i + 1);
    stats->n_iters = n_iters;
    stats->ls_num_iters = ls_iters;
    auto cpi = [&](int32_t* dst, const std::vector<int>& v) { if (dst) std::memcpy(dst, v.data(), v.size() * 4); };
    auto cpf = [&](float* dst, const std::vector<double>& v) {
        if (dst) for (size_t i = 0; i < v.size(); i++) dst[i] = (float)v[i];
    };
    cpi(stats->sqp_iters, s->h_sqp); cpi(stats->kkt_converged, s->h_conv); cpi(stats->pcg_iters, s->h_pcg);
    cpf(stats->initial_merit, s->h_init); cpf(stats->final_merit, s->h_final); cpf(stats->ls_min_merit, s->h_lsm);
    cpf(stats->ls_step_size, s->h_step); cpf(stats->kkt_residual, s->h_kres);
    return 0;
}

int gato_last_merits_f64(gato_solver64* s, double* h_initial, double* h_final) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (h_initial) std::memcpy(h_initial, s->h_init.data(), s->B * 8);
    if (h_final) std::memcpy(h_final, s->h_final.data(), s->B * 8);
    return 0;
}

int gato_set_joint_cost_f64(gato_solver64* s, const double* d_Qd, const double* d_Rd, const double* d_xref,
                            void* stream) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    hipStream_t st = (hipStream_t)stream;
    if (!d_Qd && !d_Rd && !d_xref) {
        s->joint_cost = false;
        return upload_defaults64(s);
    }
    if (d_Qd) HIP_TRY(hipMemcpyAsync(s->d_Qd, d_Qd, (size_t)s->B * s->nx * 8, hipMemcpyDeviceToDevice, st));
    if (d_Rd) HIP_TRY(hipMemcpyAsync(s->d_Rd, d_Rd, (size_t)s->B * s->nu * 8, hipMemcpyDeviceToDevice, st));
    if (d_xref) {
        HIP_TRY(hipMemcpyAsync(s->d_xref, d_xref, (size_t)s->B * s->N * s->nx * 8, hipMemcpyDeviceToDevice, st));
        s->joint_cost = true;
    }
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int gato_set_f_ext_batch_f64(gato_solver64* s, const double* d_f_ext, void* stream) {
    if (!s || !d_f_ext) return fail(GATO_E_ARG, "null argument");
    HIP_TRY(hipMemcpyAsync(s->d_fext, d_f_ext, (size_t)s->B * 6 * 8, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int gato_set_batch_f64(gato_solver64* s, int32_t what, const double* h_values, int32_t set_as_reset_default) {
    if (!s || !h_values) return fail(GATO_E_ARG, "null argument");
    const size_t n = (size_t)s->B * 8;
    switch (what) {
        case GATO_SET_RHO:
            if (set_as_reset_default) std::memcpy(s->h_rho_init.data(), h_values, n);
            HIP_TRY(hipMemcpy(s->d_rho, h_values, n, hipMemcpyHostToDevice));
            return 0;
        case GATO_SET_DRHO:
            if (set_as_reset_default) {
                std::memcpy(s->h_drho_init.data(), h_values, n);
                HIP_TRY(hipMemcpy(s->d_drho_init, h_values, n, hipMemcpyHostToDevice));
            }
            HIP_TRY(hipMemcpy(s->d_drho, h_values, n, hipMemcpyHostToDevice));
            return 0;
        case GATO_SET_MU: HIP_TRY(hipMemcpy(s->d_mu, h_values, n, hipMemcpyHostToDevice)); return 0;
        case GATO_SET_PCG_TOL: HIP_TRY(hipMemcpy(s->d_tol, h_values, n, hipMemcpyHostToDevice)); return 0;
        default: return fail(GATO_E_ARG, "unknown batch setting");
    }
}

int gato_reset_f64(gato_solver64* s, int32_t what) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (what == GATO_RESET_DUAL) {
        HIP_TRY(hipMemset(s->a.lam, 0, (size_t)s->B * s->N * s->nx * 8));
        return 0;
    }
    if (what == GATO_RESET_RHO) {
        HIP_TRY(hipMemcpy(s->d_rho, s->h_rho_init.data(), s->B * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->d_drho, s->h_drho_init.data(), s->B * 8, hipMemcpyHostToDevice));
        return 0;
    }
    return fail(GATO_E_ARG, "unknown reset");
}

int gato_set_rho_adaptation_f64(gato_solver64* s, int32_t enabled) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    s->adapt_rho = enabled ? 1 : 0;
    return 0;
}

int gato_get_state_f64(gato_solver64* s, double* d_lambda, double* h_rho, double* h_drho) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (d_lambda) HIP_TRY(hipMemcpy(d_lambda, s->a.lam, (size_t)s->B * s->N * s->nx * 8, hipMemcpyDeviceToDevice));
    if (h_rho) HIP_TRY(hipMemcpy(h_rho, s->d_rho, s->B * 8, hipMemcpyDeviceToHost));
    if (h_drho) HIP_TRY(hipMemcpy(h_drho, s->d_drho, s->B * 8, hipMemcpyDeviceToHost));
    return 0;
}

int gato_sim_forward_f64(gato_solver64* s, double* d_xnext, const double* d_x, const double* d_u, double dt,
                         void* stream) {
    if (!s || !d_xnext || !d_x || !d_u) return fail(GATO_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int nb = (s->B + 63) / 64;
    if (s->plant == GATO_PLANT_IIWA14)
        hipLaunchKernelGGL((gato::k_sim<0, double>), dim3(nb), dim3(64), 0, st, d_xnext, d_x, d_u, s->d_fext, dt, s->B);
    else
        hipLaunchKernelGGL((gato::k_sim<1, double>), dim3(nb), dim3(64), 0, st, d_xnext, d_x, d_u, s->d_fext, dt, s->B);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}


int gato_plant_dims(int32_t plant, int32_t* nq, int32_t* nx) {
    const int n = plant_nq(plant);
    if (n < 0) return fail(GATO_E_ARG, "unknown plant");
    if (nq) *nq = n;
    if (nx) *nx = 2 * n;
    return 0;
}

// k_pcg_dz_res holds a problem's off-diagonal Schur blocks in LDS (more than the default 64 KB of dynamic LDS at
// N = 128): allow each instantiation its largest size once per process, before any launch or graph capture.
static int allow_res_lds() {
    static std::once_flag once;
    static hipError_t err = hipSuccess;
    std::call_once(once, [] {
        const int b0 = gato::pcg_res_lds_floats<0>(128) * (int)sizeof(float);
        const int b1 = gato::pcg_res_lds_floats<1>(128) * (int)sizeof(float);
        const void* fn[9] = {reinterpret_cast<const void*>(&gato::k_pcg_dz_res<0>),
                             reinterpret_cast<const void*>(&gato::k_pcg_dz_res<1>),
                             reinterpret_cast<const void*>(&gato::k_pcg_res<0>),
                             reinterpret_cast<const void*>(&gato::k_pcg_res<1>),
                             reinterpret_cast<const void*>(&gato::k_fused<0, 2>),
                             reinterpret_cast<const void*>(&gato::k_fused<0, 4>),
                             reinterpret_cast<const void*>(&gato::k_fused<1, 2>),
                             reinterpret_cast<const void*>(&gato::k_fused<1, 4>),
                             reinterpret_cast<const void*>(&gato::k_fused_mixed)};
        const int bytes[9] = {b0, b1, b0, b1, b0, b0, b1, b1, b0};
        for (int i = 0; i < 9 && err == hipSuccess; i++)
            err = hipFuncSetAttribute(fn[i], hipFuncAttributeMaxDynamicSharedMemorySize, bytes[i]);
    });
    return err == hipSuccess ? 0 : fail((int)err, std::string("hipFuncSetAttribute (PCG LDS): ") + hipGetErrorString(err));
}

int gato_create(gato_solver** out, int32_t plant, int32_t knots, int32_t batch, const gato_params* params) {
    if (!out || !params) return fail(GATO_E_ARG, "null argument");
    const int nq = plant_nq(plant);
    if (nq < 0) return fail(GATO_E_ARG, "unknown plant");
    if (knots < 2 || knots > 128) return fail(GATO_E_ARG, "knots must be in [2, 128]");
    if (batch < 1) return fail(GATO_E_ARG, "batch must be >= 1");
    if (params->max_sqp_iters < 1 || params->max_sqp_iters > 1000) return fail(GATO_E_ARG, "max_sqp_iters must be in [1, 1000]");
    if (params->mode < GATO_MODE_PLAIN || params->mode > GATO_MODE_GRAPH) return fail(GATO_E_ARG, "unknown mode");
    if (params->exit_rule != GATO_EXIT_PCG0 && params->exit_rule != GATO_EXIT_KKT) return fail(GATO_E_ARG, "unknown exit rule");
    if (knots > 64)
        if (int rc0 = allow_res_lds()) return rc0;
    gato_solver* s = new gato_solver();
    s->plant = plant;
    s->N = knots;
    s->B = batch;
    s->nq = nq;
    s->nx = 2 * nq;
    s->nu = nq;
    s->traj = (s->nx + s->nu) * knots - s->nu;
    s->prm = *params;
    s->G = default_groups(knots);
    s->bwd_refine = default_refine(knots);
    s->pool_bytes = layout(s, false);
    hipError_t e = hipMalloc(&s->pool, s->pool_bytes);
    if (e != hipSuccess) {
        delete s;
        return fail(GATO_E_ALLOC, std::string("hipMalloc workspace: ") + hipGetErrorString(e));
    }
    layout(s, true);
    int rc = 0;
    auto cleanup = [&](int code) { gato_destroy(s); return code; };
    if ((e = hipMemset(s->pool, 0, s->pool_bytes)) != hipSuccess) return cleanup(fail((int)e, "hipMemset workspace"));
    if ((e = hipEventCreate(&s->ev0)) != hipSuccess || (e = hipEventCreate(&s->ev1)) != hipSuccess ||
        (e = hipEventCreate(&s->ev_t0)) != hipSuccess || (e = hipEventCreate(&s->ev_t1)) != hipSuccess)
        return cleanup(fail((int)e, "hipEventCreate"));
    s->h_rho_init.assign(batch, params->rho);
    s->h_drho_init.assign(batch, 1.f);
    std::vector<float> mu(batch, params->mu), tol(batch, params->pcg_tol);
    if ((e = hipMemcpy(s->d_rho, s->h_rho_init.data(), batch * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(s->d_drho, s->h_drho_init.data(), batch * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(s->d_drho_init, s->h_drho_init.data(), batch * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(s->d_mu, mu.data(), batch * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(s->d_tol, tol.data(), batch * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(fail((int)e, "hipMemcpy initial state"));
    if ((rc = upload_defaults(s)) != 0) return cleanup(rc);
    const size_t it = params->max_sqp_iters;
    s->h_pcg.resize(it * batch);
    s->h_kres.resize(it * batch);
    s->h_lsm.resize(it * batch);
    s->h_step.resize(it * batch);
    s->h_sqp.resize(batch);
    s->h_conv.resize(batch);
    s->h_init.resize(batch);
    s->h_final.resize(batch);
    *out = s;
    return 0;
}

int gato_destroy(gato_solver* s) {
    if (!s) return 0;
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
    if (s->ev_t0) hipEventDestroy(s->ev_t0);
    if (s->ev_t1) hipEventDestroy(s->ev_t1);
    drop_graph(s);
    if (s->cap_stream) hipStreamDestroy(s->cap_stream);
    if (s->pool) hipFree(s->pool);
    if (s->imp_pool) hipFree(s->imp_pool);
    delete s;
    return 0;
}

int gato_solve(gato_solver* s, float* d_xu, const float* d_xs, const float* d_ref, float dt, void* stream,
               gato_stats* stats) {
    if (!s || !d_xu || !d_xs || !d_ref) return fail(GATO_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    fill_args(s, d_xu, d_xs, d_ref, dt);
    const auto t0 = std::chrono::high_resolution_clock::now();
    int rc = s->plant == GATO_PLANT_IIWA14 ? launch_solve<0>(s, st) : launch_solve<1>(s, st);
    if (rc) return rc;
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (!stats) return 0;
    return collect_stats(s, st, stats, std::chrono::duration<double, std::micro>(t1 - t0).count());
}

int gato_set_joint_cost(gato_solver* s, const float* d_Qd, const float* d_Rd, const float* d_xref, void* stream) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    hipStream_t st = (hipStream_t)stream;
    if (!d_Qd && !d_Rd && !d_xref) {
        s->joint_cost = false;
        return upload_defaults(s);
    }
    if (d_Qd) HIP_TRY(hipMemcpyAsync(s->d_Qd, d_Qd, (size_t)s->B * s->nx * 4, hipMemcpyDeviceToDevice, st));
    if (d_Rd) HIP_TRY(hipMemcpyAsync(s->d_Rd, d_Rd, (size_t)s->B * s->nu * 4, hipMemcpyDeviceToDevice, st));
    if (d_xref) {
        HIP_TRY(hipMemcpyAsync(s->d_xref, d_xref, (size_t)s->B * s->N * s->nx * 4, hipMemcpyDeviceToDevice, st));
        s->joint_cost = true;
    }
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int gato_set_f_ext_batch(gato_solver* s, const float* d_f_ext, void* stream) {
    if (!s || !d_f_ext) return fail(GATO_E_ARG, "null argument");
    HIP_TRY(hipMemcpyAsync(s->d_fext, d_f_ext, (size_t)s->B * 6 * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int gato_set_rho_penalty_batch(gato_solver* s, const float* h_rho, int32_t set_as_reset_default) {
    if (!s || !h_rho) return fail(GATO_E_ARG, "null argument");
    if (set_as_reset_default) std::memcpy(s->h_rho_init.data(), h_rho, s->B * 4);
    HIP_TRY(hipMemcpy(s->d_rho, h_rho, s->B * 4, hipMemcpyHostToDevice));
    return 0;
}

int gato_set_drho_batch(gato_solver* s, const float* h_drho, int32_t set_as_reset_default) {
    if (!s || !h_drho) return fail(GATO_E_ARG, "null argument");
    if (set_as_reset_default) {
        std::memcpy(s->h_drho_init.data(), h_drho, s->B * 4);
        HIP_TRY(hipMemcpy(s->d_drho_init, h_drho, s->B * 4, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(s->d_drho, h_drho, s->B * 4, hipMemcpyHostToDevice));
    return 0;
}

int gato_set_mu_batch(gato_solver* s, const float* h_mu) {
    if (!s || !h_mu) return fail(GATO_E_ARG, "null argument");
    HIP_TRY(hipMemcpy(s->d_mu, h_mu, s->B * 4, hipMemcpyHostToDevice));
    return 0;
}

int gato_set_pcg_tol_batch(gato_solver* s, const float* h_tol) {
    if (!s || !h_tol) return fail(GATO_E_ARG, "null argument");
    HIP_TRY(hipMemcpy(s->d_tol, h_tol, s->B * 4, hipMemcpyHostToDevice));
    return 0;
}

int gato_reset_dual(gato_solver* s) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    HIP_TRY(hipMemset(s->a.lam, 0, (size_t)s->B * s->N * s->nx * 4));
    return 0;
}

int gato_reset_rho(gato_solver* s) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    HIP_TRY(hipMemcpy(s->d_rho, s->h_rho_init.data(), s->B * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->d_drho, s->h_drho_init.data(), s->B * 4, hipMemcpyHostToDevice));
    return 0;
}

int gato_set_rho_adaptation(gato_solver* s, int32_t enabled) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    s->adapt_rho = enabled ? 1 : 0;
    return 0;
}

int gato_set_mode(gato_solver* s, int32_t mode, int32_t exit_rule) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (mode < GATO_MODE_PLAIN || mode > GATO_MODE_GRAPH) return fail(GATO_E_ARG, "unknown mode");
    if (exit_rule != GATO_EXIT_PCG0 && exit_rule != GATO_EXIT_KKT) return fail(GATO_E_ARG, "unknown exit rule");
    s->prm.mode = mode;
    s->prm.exit_rule = exit_rule;
    return 0;
}

int gato_get_state(gato_solver* s, float* d_lambda, float* h_rho, float* h_drho) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (d_lambda) HIP_TRY(hipMemcpy(d_lambda, s->a.lam, (size_t)s->B * s->N * s->nx * 4, hipMemcpyDeviceToDevice));
    if (h_rho) HIP_TRY(hipMemcpy(h_rho, s->d_rho, s->B * 4, hipMemcpyDeviceToHost));
    if (h_drho) HIP_TRY(hipMemcpy(h_drho, s->d_drho, s->B * 4, hipMemcpyDeviceToHost));
    return 0;
}

int gato_set_state(gato_solver* s, const float* d_lambda, const float* h_rho, const float* h_drho) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (d_lambda) HIP_TRY(hipMemcpy(s->a.lam, d_lambda, (size_t)s->B * s->N * s->nx * 4, hipMemcpyDeviceToDevice));
    if (h_rho) HIP_TRY(hipMemcpy(s->d_rho, h_rho, s->B * 4, hipMemcpyHostToDevice));
    if (h_drho) HIP_TRY(hipMemcpy(s->d_drho, h_drho, s->B * 4, hipMemcpyHostToDevice));
    return 0;
}

int gato_sim_forward(gato_solver* s, float* d_xnext, const float* d_x, const float* d_u, float dt, void* stream) {
    if (!s || !d_xnext || !d_x || !d_u) return fail(GATO_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int nb = (s->B + 63) / 64;
    if (s->plant == GATO_PLANT_IIWA14)
        hipLaunchKernelGGL(gato::k_sim<0>, dim3(nb), dim3(64), 0, st, d_xnext, d_x, d_u, s->d_fext, dt, s->B);
    else
        hipLaunchKernelGGL(gato::k_sim<1>, dim3(nb), dim3(64), 0, st, d_xnext, d_x, d_u, s->d_fext, dt, s->B);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int gato_ee_positions(gato_solver* s, const float* d_xu, float* d_pos, float* d_jac, void* stream) {
    if (!s || !d_xu || !d_pos) return fail(GATO_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int n = s->B * s->N;
    if (s->plant == GATO_PLANT_IIWA14)
        hipLaunchKernelGGL(gato::k_ee_positions<0>, dim3((n + 255) / 256), dim3(256), 0, st, d_xu, s->traj, s->N, s->B,
                           d_pos, d_jac);
    else
        hipLaunchKernelGGL(gato::k_ee_positions<1>, dim3((n + 255) / 256), dim3(256), 0, st, d_xu, s->traj, s->N, s->B,
                           d_pos, d_jac);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gato_select_hypothesis(gato_solver* s, const float* d_x_last, const float* d_u_last, const float* d_x_curr,
                           float dt, float* d_xu, float* d_errors, int32_t* h_best, void* stream) {
    if (!s || !d_x_last || !d_u_last || !d_x_curr || !d_xu) return fail(GATO_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int threads = std::min(1024, 64 * ((s->B + 63) / 64));
    int* d_best = s->d_stop + 1;
    if (s->plant == GATO_PLANT_IIWA14)
        hipLaunchKernelGGL(gato::k_select<0>, dim3(1), dim3(threads), 0, st, d_x_last, d_u_last, d_x_curr, s->d_fext, dt,
                           s->B, d_errors, d_best);
    else
        hipLaunchKernelGGL(gato::k_select<1>, dim3(1), dim3(threads), 0, st, d_x_last, d_u_last, d_x_curr, s->d_fext, dt,
                           s->B, d_errors, d_best);
    hipLaunchKernelGGL(gato::k_broadcast_row, dim3(256), dim3(256), 0, st, d_xu, s->traj, s->B, d_best);
    HIP_TRY(hipGetLastError());
    if (h_best) {  // only a host copy of the index waits for the device
        int32_t best = 0;
        HIP_TRY(hipMemcpyAsync(&best, d_best, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        *h_best = best;
    }
    return 0;
}

int gato_simulate_rk4(gato_solver* s, float* d_x, const float* d_u, const float* d_f_ext, float h, int32_t steps,
                      void* stream) {
    if (!s || !d_x || !d_u || !d_f_ext) return fail(GATO_E_ARG, "null argument");
    if (steps < 0 || !(h > 0.f)) return fail(GATO_E_ARG, "need steps >= 0 and h > 0");
    hipStream_t st = (hipStream_t)stream;
    const int nb = (s->B + 63) / 64;
    if (s->plant == GATO_PLANT_IIWA14)
        hipLaunchKernelGGL(gato::k_rk4<0>, dim3(nb), dim3(64), 0, st, d_x, d_u, d_f_ext, h, steps, s->B);
    else
        hipLaunchKernelGGL(gato::k_rk4<1>, dim3(nb), dim3(64), 0, st, d_x, d_u, d_f_ext, h, steps, s->B);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gato_simulate_plant(int32_t model, int32_t n, float* d_x, const float* d_xu, int32_t knots, const float* d_f_ext,
                        double h, int32_t first_substep, int32_t steps, double knot_dt, void* stream) {
    if (!d_x || !d_xu) return fail(GATO_E_ARG, "null argument");
    if (model < 0 || model > 2) return fail(GATO_E_ARG, "unknown plant model");
    if (n <= 0 || knots <= 0 || steps < 0 || first_substep < 0 || !(h > 0.0) || !(knot_dt > 0.0))
        return fail(GATO_E_ARG, "need n, knots > 0, steps, first_substep >= 0 and h, knot_dt > 0");
    const int nq = model == GATO_PLANT_IIWA14 ? 7 : 6;
    const int traj = (3 * nq) * knots - nq;
    hipStream_t st = (hipStream_t)stream;
    const int nb = (n + 63) / 64;
    if (model == GATO_PLANT_IIWA14)
        hipLaunchKernelGGL(gato::k_plant_rk4<0>, dim3(nb), dim3(64), 0, st, d_x, d_xu, traj, knots, d_f_ext, h,
                           first_substep, steps, knot_dt, n);
    else if (model == GATO_PLANT_INDY7)
        hipLaunchKernelGGL(gato::k_plant_rk4<1>, dim3(nb), dim3(64), 0, st, d_x, d_xu, traj, knots, d_f_ext, h,
                           first_substep, steps, knot_dt, n);
    else
        hipLaunchKernelGGL(gato::k_plant_rk4<2>, dim3(nb), dim3(64), 0, st, d_x, d_xu, traj, knots, d_f_ext, h,
                           first_substep, steps, knot_dt, n);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gato_window_reference(gato_solver* s, float* d_ref, const float* d_path, int32_t npoints, int32_t offset,
                          void* stream) {
    if (!s || !d_ref || !d_path) return fail(GATO_E_ARG, "null argument");
    if (npoints < 1 || offset < 0) return fail(GATO_E_ARG, "need npoints >= 1 and offset >= 0");
    hipLaunchKernelGGL(gato::k_window_ref, dim3(64), dim3(256), 0, (hipStream_t)stream, d_ref, d_path, npoints, offset,
                       s->N, s->B);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gato_figure8_path(float* d_path, int32_t points_per_cycle, int32_t cycles, double amp_x, double amp_z,
                      const double* offset, double theta, void* stream) {
    if (!d_path || !offset) return fail(GATO_E_ARG, "null argument");
    if (points_per_cycle < 1 || cycles < 1) return fail(GATO_E_ARG, "need points_per_cycle >= 1 and cycles >= 1");
    hipLaunchKernelGGL(gato::k_figure8, dim3(64), dim3(256), 0, (hipStream_t)stream, d_path, points_per_cycle, cycles,
                       amp_x, amp_z, offset[0], offset[1], offset[2], theta);
    HIP_TRY(hipGetLastError());
    return 0;
}

struct gato_force_est {
    int B = 0;
    double smoothing = 0.3, r_min = 1.0, r_max = 100.0;
    uint64_t rng = 0;
    gato::ForceEst* d_st = nullptr;
    float* d_dirs = nullptr;
};

namespace {
// SplitMix64 (Steele, Lea and Flood, 2014): the library's own uniforms for the exploration rotations
uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

gato::ForceEst initial_force_state(double radius) {
    gato::ForceEst st{};
    st.R[0] = st.R[4] = st.R[8] = 1.f;
    st.radius = radius;
    return st;
}
}  // namespace

int gato_force_est_create(gato_force_est** out, int32_t batch, float initial_radius, float min_radius,
                          float max_radius, float smoothing_factor, uint64_t seed) {
    if (!out) return fail(GATO_E_ARG, "null argument");
    *out = nullptr;
    if (batch <= 3) return fail(GATO_E_ARG, "batch must exceed 3 (three exploitation hypotheses plus exploration)");
    if (!(min_radius > 0.f) || !(max_radius >= min_radius) || !(initial_radius > 0.f))
        return fail(GATO_E_ARG, "need 0 < min_radius <= max_radius and initial_radius > 0");
    if (!(smoothing_factor >= 0.f && smoothing_factor <= 1.f)) return fail(GATO_E_ARG, "smoothing_factor outside [0, 1]");
    // exploration directions: Fibonacci spiral on the unit sphere (force_estimator.py:33-61), double then float
    const int n = batch - 3;
    std::vector<float> dirs(3 * (size_t)n);
    const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
    for (int i = 0; i < n; i++) {
        const double y = n > 1 ? 1.0 - 2.0 * i / (double)(n - 1) : 0.0;
        const double r = std::sqrt(1.0 - y * y), th = 2.0 * M_PI * i / phi;
        dirs[3 * i] = (float)(r * std::cos(th));
        dirs[3 * i + 1] = (float)y;
        dirs[3 * i + 2] = (float)(r * std::sin(th));
    }
    auto* e = new gato_force_est();
    e->B = batch;
    e->smoothing = smoothing_factor;
    e->r_min = min_radius;
    e->r_max = max_radius;
    e->rng = seed;
    const gato::ForceEst st = initial_force_state(initial_radius);
    if (hipMalloc(&e->d_st, sizeof(gato::ForceEst)) != hipSuccess ||
        hipMalloc(&e->d_dirs, sizeof(float) * std::max<size_t>(dirs.size(), 1)) != hipSuccess) {
        if (e->d_st) (void)hipFree(e->d_st);
        delete e;
        return fail(GATO_E_ALLOC, "device allocation of the force estimator failed");
    }
    if (hipMemcpy(e->d_st, &st, sizeof(st), hipMemcpyHostToDevice) != hipSuccess ||
        (n > 0 && hipMemcpy(e->d_dirs, dirs.data(), sizeof(float) * dirs.size(), hipMemcpyHostToDevice) != hipSuccess)) {
        (void)hipFree(e->d_st);
        (void)hipFree(e->d_dirs);
        delete e;
        return fail(GATO_E_LAUNCH, "upload of the force estimator state failed");
    }
    *out = e;
    return 0;
}

int gato_force_est_destroy(gato_force_est* e) {
    if (!e) return fail(GATO_E_ARG, "null handle");
    (void)hipFree(e->d_st);
    (void)hipFree(e->d_dirs);
    delete e;
    return 0;
}

int gato_force_est_reset(gato_force_est* e, float radius, void* stream) {
    if (!e) return fail(GATO_E_ARG, "null handle");
    if (!(radius > 0.f)) return fail(GATO_E_ARG, "radius must be positive");
    const gato::ForceEst st = initial_force_state(radius);
    hipStream_t sm = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(e->d_st, &st, sizeof(st), hipMemcpyHostToDevice, sm));
    HIP_TRY(hipStreamSynchronize(sm));
    return 0;
}

int gato_force_est_generate(gato_force_est* e, gato_solver* s, const float* d_q, float* d_world, void* stream) {
    if (!e || !s || !d_q) return fail(GATO_E_ARG, "null argument");
    if (e->B != s->B) return fail(GATO_E_ARG, "estimator and solver batch sizes differ");
    const dim3 grid((e->B + 255) / 256);
    hipStream_t st = (hipStream_t)stream;
    if (s->plant == GATO_PLANT_IIWA14)
        hipLaunchKernelGGL(gato::k_force_hypotheses<0>, grid, dim3(256), 0, st, e->d_st, e->d_dirs, d_q, e->B, s->d_fext,
                           d_world);
    else
        hipLaunchKernelGGL(gato::k_force_hypotheses<1>, grid, dim3(256), 0, st, e->d_st, e->d_dirs, d_q, e->B, s->d_fext,
                           d_world);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gato_force_est_update(gato_force_est* e, gato_solver* s, const int32_t* d_best, const float* d_errors, float alpha,
                          float beta, const double* h_uniforms, void* stream) {
    if (!e || !s || !d_errors) return fail(GATO_E_ARG, "null argument");
    if (e->B != s->B) return fail(GATO_E_ARG, "estimator and solver batch sizes differ");
    double u[3];
    for (int k = 0; k < 3; k++) u[k] = h_uniforms ? h_uniforms[k] : uniform01(e->rng);
    for (int k = 0; k < 3; k++)
        if (!(u[k] >= 0.0 && u[k] <= 1.0)) return fail(GATO_E_ARG, "uniforms must lie in [0, 1]");
    const int* best = d_best ? d_best : s->d_stop + 1;  // the solver's last gato_select_hypothesis result
    hipLaunchKernelGGL(gato::k_force_update, dim3(1), dim3(64), 0, (hipStream_t)stream, e->d_st, e->d_dirs, best,
                       d_errors, e->B, alpha, beta, e->smoothing, e->r_min, e->r_max, u[0], u[1], u[2]);
    HIP_TRY(hipGetLastError());
    return 0;
}

int gato_force_est_state(gato_force_est* e, double* h_state, void* stream) {
    if (!e || !h_state) return fail(GATO_E_ARG, "null argument");
    gato::ForceEst st;
    hipStream_t sm = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(&st, e->d_st, sizeof(st), hipMemcpyDeviceToHost, sm));
    HIP_TRY(hipStreamSynchronize(sm));
    for (int c = 0; c < 6; c++) {
        h_state[c] = st.est[c];
        h_state[6 + c] = st.mom[c];
        h_state[12 + c] = st.smooth[c];
    }
    for (int k = 0; k < 9; k++) h_state[18 + k] = st.R[k];
    h_state[27] = st.radius;
    h_state[28] = st.confidence;
    h_state[29] = st.count;
    return 0;
}

int gato_debug_phase(gato_solver* s, int32_t phase, float* d_xu, const float* d_xs, const float* d_ref, float dt,
                     void* stream) {
    if (!s || !d_xu || !d_xs || !d_ref) return fail(GATO_E_ARG, "null argument");
    if (phase < 0 || phase > 5) return fail(GATO_E_ARG, "unknown phase");
    if (phase == 5 && (s->N <= 64 || s->N > 128)) return fail(GATO_E_ARG, "phase 5 needs 64 < N <= 128");
    hipStream_t st = (hipStream_t)stream;
    fill_args(s, d_xu, d_xs, d_ref, dt);
    const dim3 blk(grid_threads(s));
    const size_t lds = lds_bytes(s);
    gato::Args a = s->a;
    a.exit_rule = GATO_EXIT_PCG0;  // run every problem
    const int sb = (s->B + 255) / 256;
    hipLaunchKernelGGL(gato::k_begin, dim3(sb), dim3(256), 0, st, a);
    const int np = npass(s);
    const bool ii = s->plant == GATO_PLANT_IIWA14;
    switch (phase) {
        case 0:
            if (ii) hipLaunchKernelGGL(gato::k_lin<0>, dim3(s->B), blk, lds, st, a);
            else hipLaunchKernelGGL(gato::k_lin<1>, dim3(s->B), blk, lds, st, a);
            break;
        case 1:
            if (ii) launch_schur<0>(s->B, blk, lds, st, a);
            else launch_schur<1>(s->B, blk, lds, st, a);
            break;
        case 2:
            if (ii) launch_pcg<0>(np, s->B, blk, pcg_lds_bytes(s), st, a, 0);
            else launch_pcg<1>(np, s->B, blk, pcg_lds_bytes(s), st, a, 0);
            break;
        case 3:
            if (ii) hipLaunchKernelGGL(gato::k_dz<0>, dim3(s->B), blk, lds, st, a, 0);
            else hipLaunchKernelGGL(gato::k_dz<1>, dim3(s->B), blk, lds, st, a, 0);
            break;
        case 4:
            // merit terms only: reuse the line-search launch with a large phi so nothing is applied
            if (ii) launch_merit<0>(s->B, blk, lds, st, a, gato::NUM_ALPHAS, 3, 0);
            else launch_merit<1>(s->B, blk, lds, st, a, gato::NUM_ALPHAS, 3, 0);
            break;
        case 5: {  // the graph path's PCG for 64 < N <= 128 (held on the CU, P in factored form)
            const size_t lr = std::max(lds, (size_t)(ii ? gato::pcg_res_lds_floats<0>(s->N)
                                                        : gato::pcg_res_lds_floats<1>(s->N)) * sizeof(float));
            if (ii) hipLaunchKernelGGL(gato::k_pcg_res<0>, dim3(s->B), dim3(1024), lr, st, a, 0);
            else hipLaunchKernelGGL(gato::k_pcg_res<1>, dim3(s->B), dim3(1024), lr, st, a, 0);
            break;
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int gato_debug_copy(gato_solver* s, const char* name, void* d_dst, int64_t cap, int64_t* count) {
    if (!s || !name || !d_dst) return fail(GATO_E_ARG, "null argument");
    const size_t N = s->N, B = s->B, nx = s->nx, nu = s->nu, nq = s->nq, it = s->prm.max_sqp_iters;
    const gato::Args& a = s->a;
    struct E { const char* n; const void* p; size_t c; };
    const E tab[] = {{"A", a.A, B * N * nx * nx},         {"B", a.Bm, B * N * nx * nu},   {"c", a.c, B * N * nx},
                     {"Qp", a.Qp, B * N * nq * nq},       {"Qv", a.Qv, B * N * nq},       {"Rg", a.Rg, B * N * nu},
                     {"q", a.q, B * N * nx},              {"r", a.r, B * N * nu},         {"Qip", a.Qip, B * N * nq * nq},
                     {"Qiv", a.Qiv, B * N * nq},          {"Ri", a.Ri, B * N * nu},       {"Sd", a.Sd, B * N * nx * nx},
                     {"So", a.So, B * N * nx * nx},       {"Pd", a.Pd, B * N * nx * nx},  {"Po", a.Po, B * N * nx * nx},
                     {"gamma", a.gam, B * N * nx},        {"lambda", a.lam, B * N * nx},  {"dz", a.dz, B * (s->traj + nu)},
                     {"merit_parts", a.mpart, B * gato::NUM_ALPHAS * N}, {"pcg_iters", a.st_pcg, it * B},
                     {"kkt_residual", a.st_kres, it * B}, {"f_ext", s->d_fext, B * 6},
                     {"Qd", s->d_Qd, B * nx},             {"Rd", s->d_Rd, B * nu}};
    for (const E& e : tab) {
        if (std::strcmp(e.n, name) != 0) continue;
        if (count) *count = (int64_t)e.c;
        if ((int64_t)e.c > cap) return fail(GATO_E_ARG, std::string("buffer too small for ") + name);
        HIP_TRY(hipMemcpy(d_dst, e.p, e.c * 4, hipMemcpyDeviceToDevice));
        return 0;
    }
    return fail(GATO_E_ARG, std::string("unknown workspace array ") + name);
}

int gato_set_backward_implicit(gato_solver* s, int32_t rounds, float fd_step) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (rounds < 0 || rounds > 32 || !(fd_step > 0.f) || !(fd_step <= 1.f))
        return fail(GATO_E_ARG, "need rounds in [0, 32] and fd_step in (0, 1]");
    if (rounds > 0) {
        const int rc = alloc_implicit(s);
        if (rc) return rc;
    }
    s->bwd_implicit = rounds;
    s->bwd_fd_step = fd_step;
    return 0;
}

int gato_set_backward_params(gato_solver* s, int32_t max_pcg_iters, float pcg_tol, int32_t refine) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (refine == -1) refine = default_refine(s->N);
    if (!(max_pcg_iters >= 1 || max_pcg_iters == -1) || !(pcg_tol >= 0.f) || refine < 0 || refine > 8)
        return fail(GATO_E_ARG, "need max_pcg_iters >= 1 (or -1), pcg_tol >= 0 and refine in [0, 8]");
    s->bwd_max_pcg = max_pcg_iters;
    s->bwd_tol = pcg_tol;
    s->bwd_refine = refine;
    return 0;
}

int gato_get_backward_params(gato_solver* s, int32_t* max_pcg_iters, float* pcg_tol, int32_t* refine) {
    if (!s || !max_pcg_iters || !pcg_tol || !refine) return fail(GATO_E_ARG, "null argument");
    *max_pcg_iters = s->bwd_max_pcg;
    *pcg_tol = s->bwd_tol;
    *refine = s->bwd_refine;
    return 0;
}

int gato_backward(gato_solver* s, const float* d_xu, const float* d_xs, const float* d_ref, float dt,
                  const float* d_dL_dxu, float* d_grad_xs, float* d_grad_ref, float* d_grad_Qd, float* d_grad_Rd,
                  void* stream, int32_t* h_adj_pcg_iters) {
    if (!s || !d_xu || !d_xs || !d_ref || !d_dL_dxu || !d_grad_xs || !d_grad_ref || !d_grad_Qd || !d_grad_Rd)
        return fail(GATO_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    fill_args(s, const_cast<float*>(d_xu), d_xs, d_ref, dt);
    std::vector<float> tol(s->B, s->bwd_tol);
    HIP_TRY(hipMemcpyAsync(s->d_bwd_tol, tol.data(), s->B * 4, hipMemcpyHostToDevice, st));
    const int rc = s->plant == GATO_PLANT_IIWA14
                       ? launch_backward<0>(s, d_dL_dxu, d_grad_xs, d_grad_ref, d_grad_Qd, d_grad_Rd, st)
                       : launch_backward<1>(s, d_dL_dxu, d_grad_xs, d_grad_ref, d_grad_Qd, d_grad_Rd, st);
    if (rc) return rc;
    if (h_adj_pcg_iters) HIP_TRY(hipMemcpyAsync(h_adj_pcg_iters, s->bst_pcg, s->B * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int gato_set_launch_shape(gato_solver* s, int32_t groups) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (!valid_groups(s->N, groups))
        return fail(GATO_E_ARG, "groups must be a multiple of 4 in [4, 32], at most the default, with ceil(N / groups) <= 8");
    s->G = groups;
    return 0;
}

int gato_get_launch_shape(gato_solver* s, int32_t* groups, int32_t* threads) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (groups) *groups = s->G;
    if (threads) *threads = grid_threads(s);
    return 0;
}

int gato_autotune(gato_solver* s, float* d_xu, const float* d_xs, const float* d_ref, float dt, void* stream,
                  int32_t* chosen_groups) {
    if (!s || !d_xu || !d_xs || !d_ref) return fail(GATO_E_ARG, "null argument");
    return autotune(s, d_xu, d_xs, d_ref, dt, (hipStream_t)stream, chosen_groups);
}

struct gato_mixed {
    std::vector<gato_solver*> cls;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> joins;
    hipStream_t cap = nullptr;
    hipEvent_t fork = nullptr, ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<gato::Args> keys;
    std::vector<int> key_groups;
    int tune = 1;
    bool tuned = false;
    double last_us = 0.0;
    // GATO_MODE_FUSED: k_fused_mixed over a work list of every class's problems, longest class first
    int mode = GATO_MODE_GRAPH, fgroups = 32;
    bool fgroups_fixed = false, ftuned = false;  // persistent mode: caller's knot-group count, or tuned once
    std::vector<int2> items;  // (class, problem) in start order; the class's tag is added by gato_mixed_set_mode
    std::vector<gato::Args> h_cls;
    gato::Args* d_cls = nullptr;
    int2* d_items = nullptr;
    int* d_ctr = nullptr;
};

static void mixed_drop_graph(gato_mixed* m) {
    if (m->exec) hipGraphExecDestroy(m->exec);
    if (m->graph) hipGraphDestroy(m->graph);
    m->exec = nullptr;
    m->graph = nullptr;
}

int gato_mixed_destroy(gato_mixed* m) {
    if (!m) return 0;
    mixed_drop_graph(m);
    for (auto* s : m->cls) gato_destroy(s);
    for (auto st : m->streams) hipStreamDestroy(st);
    for (auto e : m->joins) hipEventDestroy(e);
    if (m->cap) hipStreamDestroy(m->cap);
    if (m->fork) hipEventDestroy(m->fork);
    if (m->ev0) hipEventDestroy(m->ev0);
    if (m->ev1) hipEventDestroy(m->ev1);
    if (m->d_cls) hipFree(m->d_cls);
    if (m->d_items) hipFree(m->d_items);
    if (m->d_ctr) hipFree(m->d_ctr);
    delete m;
    return 0;
}

int gato_mixed_create(gato_mixed** out, int32_t nclass, const int32_t* plants, const int32_t* knots,
                      const int32_t* batches, const gato_params* params) {
    if (!out || !plants || !knots || !batches || !params || nclass < 1 || nclass > 64)
        return fail(GATO_E_ARG, "need 1..64 classes and non-null arrays");
    gato_mixed* m = new gato_mixed();
    gato_params p = *params;
    p.mode = GATO_MODE_GRAPH;
    for (int i = 0; i < nclass; i++) {
        gato_solver* s = nullptr;
        const int rc = gato_create(&s, plants[i], knots[i], batches[i], &p);
        if (rc) { gato_destroy(s); gato_mixed_destroy(m); return rc; }
        m->cls.push_back(s);
        hipStream_t st;
        hipEvent_t ev;
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) { gato_mixed_destroy(m); return fail((int)e, "hipStreamCreate"); }
        m->streams.push_back(st);
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) { gato_mixed_destroy(m); return fail((int)e, "hipEventCreate"); }
        m->joins.push_back(ev);
    }
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&m->cap, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&m->fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreate(&m->ev0)) != hipSuccess || (e = hipEventCreate(&m->ev1)) != hipSuccess) {
        gato_mixed_destroy(m);
        return fail((int)e, "stream/event creation");
    }
    m->keys.resize(nclass);
    m->key_groups.assign(nclass, 0);
    // work list of the persistent mode: classes in decreasing order of per-iteration work (N nx^2), each class's
    // problems in order
    std::vector<int> order(nclass);
    for (int i = 0; i < nclass; i++) order[i] = i;
    auto work = [&](int i) { return (double)knots[i] * (2.0 * plant_nq(plants[i])) * (2.0 * plant_nq(plants[i])); };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
    for (int i : order)
        for (int b = 0; b < batches[i]; b++) m->items.push_back(int2{i, b});
    m->h_cls.resize(nclass);
    if ((e = hipMalloc(&m->d_cls, nclass * sizeof(gato::Args))) != hipSuccess ||
        (e = hipMalloc(&m->d_items, m->items.size() * sizeof(int2))) != hipSuccess ||
        (e = hipMalloc(&m->d_ctr, 2 * sizeof(int))) != hipSuccess ||
        (e = hipMemset(m->d_ctr, 0, 2 * sizeof(int))) != hipSuccess) {
        gato_mixed_destroy(m);
        return fail((int)e, std::string("mixed work list: ") + hipGetErrorString(e));
    }
    *out = m;
    return 0;
}

// Whether the persistent kernel can take every class at G knot groups per work-group: at most 8 knots per group for
// each (N / G rounded up), G a multiple of 4 in [4, 32].
static bool mixed_groups_ok(const gato_mixed* m, int G) {
    if (G < 4 || G > 32 || G % 4) return false;
    for (auto* c : m->cls)
        if ((c->N + G - 1) / G > 8) return false;
    return true;
}

// The work list with every item's tag (16 plant + knot passes) for G knot groups, uploaded; G becomes the kernel's.
static int mixed_use_groups(gato_mixed* m, int G) {
    std::vector<int2> items(m->items);
    for (auto& it : items) {
        const gato_solver* c = m->cls[it.x];
        it.x |= ((c->plant == GATO_PLANT_IIWA14 ? 0 : 16) + passes(c->N, G)) << 8;
    }
    HIP_TRY(hipMemcpy(m->d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
    m->fgroups = G;
    return 0;
}

int gato_mixed_set_mode(gato_mixed* m, int32_t mode, int32_t groups) {
    if (!m) return fail(GATO_E_ARG, "null solver");
    if (mode != GATO_MODE_GRAPH && mode != GATO_MODE_FUSED)
        return fail(GATO_E_ARG, "mixed mode must be GATO_MODE_GRAPH or GATO_MODE_FUSED");
    if (mode == GATO_MODE_FUSED) {
        const int G = groups ? groups : 32;
        if (!mixed_groups_ok(m, G))
            return fail(GATO_E_ARG, "groups " + std::to_string(G) + ": need a multiple of 4 in [4, 32] (0: tuned, else "
                                    "32) that gives every class at most 8 knots per knot group");
        if (int rc = allow_res_lds()) return rc;
        if (int rc = mixed_use_groups(m, G)) return rc;
        m->fgroups_fixed = groups != 0;
        m->ftuned = false;
    }
    m->mode = mode;
    return 0;
}

int gato_mixed_get_mode(gato_mixed* m, int32_t* mode, int32_t* groups) {
    if (!m || !mode || !groups) return fail(GATO_E_ARG, "null argument");
    *mode = m->mode;
    *groups = m->fgroups;
    return 0;
}

// One launch of k_fused_mixed over every class's problems (fresh solves from the classes' persistent state, as
// GATO_MODE_FUSED solves each class), on as many work-groups as fit on the device at once.
static int mixed_solve_fused(gato_mixed* m, hipStream_t st) {
    const int nc = (int)m->cls.size(), threads = gato::KG * m->fgroups;
    size_t lds = 0;
    for (auto* s : m->cls) lds = std::max(lds, fused_lds_bytes(s, m->fgroups));
    int dev = 0, cus = 0, per = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gato::k_fused_mixed, threads, lds));
    const int nitems = (int)m->items.size();
    const int grid = std::max(1, std::min(nitems, std::max(1, per) * cus));
    bool same = true;
    for (int i = 0; i < nc; i++) same = same && std::memcmp(&m->h_cls[i], &m->cls[i]->a, sizeof(gato::Args)) == 0;
    if (!same) {  // the classes' arguments changed (new tensors, dt): upload them before the launch
        for (int i = 0; i < nc; i++) m->h_cls[i] = m->cls[i]->a;
        HIP_TRY(hipMemcpyAsync(m->d_cls, m->h_cls.data(), nc * sizeof(gato::Args), hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipEventRecord(m->ev0, st));
    hipLaunchKernelGGL(gato::k_fused_mixed, dim3(grid), dim3(threads), lds, st, m->d_cls, m->d_items, nitems, m->d_ctr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(m->ev1, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    m->last_us = ms * 1e3;
    return 0;
}

// Knot groups per work-group of the persistent kernel, tuned on the caller's inputs: every valid count of 8, 16 and
// 32 runs the whole mixed solve twice (the second timed) on scratch copies of the trajectories, each run from the
// classes' saved multipliers and rho; the fastest is kept, and the classes' state is restored.
static int mixed_tune_fused(gato_mixed* m, float* const* d_xu, const float* const* d_xs, const float* const* d_ref,
                            float dt, hipStream_t st) {
    const int nc = (int)m->cls.size();
    std::vector<std::vector<float>> rho(nc), drho(nc);
    HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < nc; i++) {
        gato_solver* c = m->cls[i];
        rho[i].resize(c->B);
        drho[i].resize(c->B);
        HIP_TRY(hipMemcpy(c->lam_save, c->a.lam, (size_t)c->B * c->N * c->nx * 4, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(rho[i].data(), c->d_rho, c->B * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(drho[i].data(), c->d_drho, c->B * 4, hipMemcpyDeviceToHost));
    }
    auto restore = [&]() -> int {
        for (int i = 0; i < nc; i++) {
            gato_solver* c = m->cls[i];
            HIP_TRY(hipMemcpy(c->a.lam, c->lam_save, (size_t)c->B * c->N * c->nx * 4, hipMemcpyDeviceToDevice));
            HIP_TRY(hipMemcpy(c->d_rho, rho[i].data(), c->B * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(c->d_drho, drho[i].data(), c->B * 4, hipMemcpyHostToDevice));
        }
        return 0;
    };
    int best = m->fgroups;
    double best_us = 1e300;
    for (int G = 8; G <= 32; G *= 2) {
        if (!mixed_groups_ok(m, G)) continue;
        int rc = mixed_use_groups(m, G);
        for (int rep = 0; rep < 2 && !rc; rep++) {
            rc = restore();
            for (int i = 0; i < nc && !rc; i++) {
                gato_solver* c = m->cls[i];
                const hipError_t e = hipMemcpyAsync(c->xu_tune, d_xu[i], (size_t)c->B * c->traj * 4,
                                                    hipMemcpyDeviceToDevice, st);
                if (e != hipSuccess) rc = fail((int)e, "mixed tuner copy");
                fill_args(c, c->xu_tune, d_xs[i], d_ref[i], dt);
            }
            if (!rc) rc = mixed_solve_fused(m, st);
        }
        if (rc) { restore(); return rc; }
        if (m->last_us < best_us) { best_us = m->last_us; best = G; }
    }
    if (int rc = mixed_use_groups(m, best)) return rc;
    return restore();
}

int gato_mixed_class(gato_mixed* m, int32_t i, gato_solver** out) {
    if (!m || !out || i < 0 || i >= (int)m->cls.size()) return fail(GATO_E_ARG, "bad class index");
    *out = m->cls[i];
    return 0;
}

int gato_mixed_set_autotune(gato_mixed* m, int32_t enabled) {
    if (!m) return fail(GATO_E_ARG, "null solver");
    m->tune = enabled ? 1 : 0;
    m->tuned = false;
    return 0;
}

int gato_mixed_solve(gato_mixed* m, float* const* d_xu, const float* const* d_xs, const float* const* d_ref, float dt,
                     void* stream, gato_stats* stats) {
    if (!m || !d_xu || !d_xs || !d_ref) return fail(GATO_E_ARG, "null argument");
    const int nc = (int)m->cls.size();
    for (int i = 0; i < nc; i++)
        if (!d_xu[i] || !d_xs[i] || !d_ref[i]) return fail(GATO_E_ARG, "null class pointer");
    hipStream_t st = (hipStream_t)stream;
    const auto t0 = std::chrono::high_resolution_clock::now();
    if (m->mode == GATO_MODE_FUSED) {
        if (m->tune && !m->fgroups_fixed && !m->ftuned) {
            const int rt = mixed_tune_fused(m, d_xu, d_xs, d_ref, dt, st);
            if (rt) return rt;
            m->ftuned = true;
        }
        for (int i = 0; i < nc; i++) fill_args(m->cls[i], d_xu[i], d_xs[i], d_ref[i], dt);
        const int rc = mixed_solve_fused(m, st);
        if (rc) return rc;
        const double us = std::chrono::duration<double, std::micro>(std::chrono::high_resolution_clock::now() - t0).count();
        for (int i = 0; i < nc && stats; i++) {
            const int rs = collect_stats(m->cls[i], st, &stats[i], us);
            if (rs) return rs;
        }
        return 0;
    }
    if (m->tune && !m->tuned) {
        for (int i = 0; i < nc; i++) {
            const int rc = autotune(m->cls[i], d_xu[i], d_xs[i], d_ref[i], dt, st, nullptr);
            if (rc) return rc;
        }
        m->tuned = true;
    }
    bool same = m->exec != nullptr;
    for (int i = 0; i < nc; i++) {
        gato_solver* s = m->cls[i];
        fill_args(s, d_xu[i], d_xs[i], d_ref[i], dt);
        s->a.stop = s->d_stop;
        s->a.stop_w = s->d_stop;
        same = same && m->key_groups[i] == s->G && std::memcmp(&m->keys[i], &s->a, sizeof(gato::Args)) == 0;
    }
    if (!same) {
        mixed_drop_graph(m);
        HIP_TRY(hipStreamBeginCapture(m->cap, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipEventRecord(m->fork, m->cap);
        for (int i = 0; i < nc && e == hipSuccess; i++) {
            e = hipStreamWaitEvent(m->streams[i], m->fork, 0);
            if (e != hipSuccess) break;
            if (m->cls[i]->plant == GATO_PLANT_IIWA14) enqueue_graph_solve<0>(m->cls[i], m->streams[i]);
            else enqueue_graph_solve<1>(m->cls[i], m->streams[i]);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipEventRecord(m->joins[i], m->streams[i]);
            if (e == hipSuccess) e = hipStreamWaitEvent(m->cap, m->joins[i], 0);
        }
        const hipError_t ce = hipStreamEndCapture(m->cap, &m->graph);
        if (e != hipSuccess) { mixed_drop_graph(m); return fail((int)e, std::string("mixed capture: ") + hipGetErrorString(e)); }
        if (ce != hipSuccess) { mixed_drop_graph(m); return fail((int)ce, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce)); }
        const hipError_t ie = hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) { mixed_drop_graph(m); return fail((int)ie, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie)); }
        for (int i = 0; i < nc; i++) {
            std::memcpy(&m->keys[i], &m->cls[i]->a, sizeof(gato::Args));
            m->key_groups[i] = m->cls[i]->G;
        }
    }
    HIP_TRY(hipEventRecord(m->ev0, st));
    HIP_TRY(hipGraphLaunch(m->exec, st));
    HIP_TRY(hipEventRecord(m->ev1, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    m->last_us = ms * 1e3;
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (stats) {
        const double us = std::chrono::duration<double, std::micro>(t1 - t0).count();
        for (int i = 0; i < nc; i++) {
            const int rc = collect_stats(m->cls[i], st, &stats[i], us);
            if (rc) return rc;
        }
    }
    return 0;
}

int gato_mixed_last_time(gato_mixed* m, double* us) {
    if (!m || !us) return fail(GATO_E_ARG, "null argument");
    *us = m->last_us;
    return 0;
}

int gato_solve_noalloc(int32_t plant, int32_t knots, int32_t batch, const gato_params* params, float* d_xu,
                       const float* d_xs, const float* d_ref, float dt, void* stream, gato_stats* stats) {
    if (!params || !d_xu || !d_xs || !d_ref) return fail(GATO_E_ARG, "null argument");
    const int nq = plant_nq(plant);
    if (nq < 0) return fail(GATO_E_ARG, "unknown plant");
    if (knots < 2 || knots > 128) return fail(GATO_E_ARG, "knots must be in [2, 128]");
    if (batch < 1) return fail(GATO_E_ARG, "batch must be >= 1");
    if (params->max_sqp_iters < 1 || params->max_sqp_iters > 1000) return fail(GATO_E_ARG, "max_sqp_iters must be in [1, 1000]");
    if (params->exit_rule != GATO_EXIT_PCG0 && params->exit_rule != GATO_EXIT_KKT) return fail(GATO_E_ARG, "unknown exit rule");
    gato_solver s;  // a fresh solver whose workspace is the static buffer: no pool of its own, nothing to free
    s.plant = plant;
    s.N = knots;
    s.B = batch;
    s.nq = nq;
    s.nx = 2 * nq;
    s.nu = nq;
    s.traj = (s.nx + s.nu) * knots - s.nu;
    s.prm = *params;
    s.prm.mode = GATO_MODE_GRAPH;
    s.G = default_groups(knots);
    const size_t need = layout(&s, false, true);
    if (need > (size_t)GATO_NOALLOC_SCRATCH_BYTES)
        return fail(GATO_E_ARG, "gato_solve_noalloc: " + std::to_string(batch) + " problems of " + std::to_string(knots) +
                                    " knots need " + std::to_string(need) + " bytes of workspace, more than the static " +
                                    std::to_string((size_t)GATO_NOALLOC_SCRATCH_BYTES) + "-byte buffer holds");
    if (knots > 64)
        if (int rc0 = allow_res_lds()) return rc0;
    void* base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    s.pool = static_cast<char*>(base);
    s.pool_bytes = need;
    layout(&s, true, true);
    hipStream_t st = (hipStream_t)stream;
    const auto t0 = std::chrono::high_resolution_clock::now();
    // the previous call's contents: cleared, then the fresh solver's state (gato_create clears its pool the same way)
    hipLaunchKernelGGL(gato::k_zero, dim3(256), dim3(256), 0, st, reinterpret_cast<float*>(s.pool), need / sizeof(float));
    hipLaunchKernelGGL(gato::k_fresh_state, dim3((batch + 255) / 256), dim3(256), 0, st, batch, s.d_rho, s.d_drho,
                       s.d_drho_init, s.d_mu, s.d_tol, s.d_Qd, s.d_Rd, nq, params->rho, params->mu, params->pcg_tol,
                       params->qd_cost, params->u_cost);
    fill_args(&s, d_xu, d_xs, d_ref, dt);
    s.a.stop = s.d_stop;
    s.a.stop_w = s.d_stop;
    if (plant == GATO_PLANT_IIWA14) enqueue_graph_solve<0>(&s, st);
    else enqueue_graph_solve<1>(&s, st);
    HIP_TRY(hipGetLastError());
    if (!stats) return 0;
    hipLaunchKernelGGL(gato::k_solve_summary, dim3(1), dim3(256), 0, st, s.a, s.d_stop + 2);
    HIP_TRY(hipGetLastError());
    if (stats->status) {
        hipLaunchKernelGGL(gato::k_status<float>, dim3(batch), dim3(64), 0, st, s.a, s.d_status);
        HIP_TRY(hipGetLastError());
    }
    const size_t B = batch, it = params->max_sqp_iters;
    int32_t counts[2] = {0, 0};
    auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
        if (dst) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        return 0;
    };
    if ((rc = cp(counts, s.d_stop + 2, sizeof(counts))) || (rc = cp(stats->sqp_iters, s.a.st_sqp, B * 4)) ||
        (rc = cp(stats->kkt_converged, s.a.conv, B * 4)) || (rc = cp(stats->initial_merit, s.a.st_init, B * 4)) ||
        (rc = cp(stats->final_merit, s.a.st_final, B * 4)) || (rc = cp(stats->pcg_iters, s.a.st_pcg, it * B * 4)) ||
        (rc = cp(stats->ls_min_merit, s.a.st_lsm, it * B * 4)) || (rc = cp(stats->ls_step_size, s.a.st_step, it * B * 4)) ||
        (rc = cp(stats->kkt_residual, s.a.st_kres, it * B * 4)) || (rc = cp(stats->status, s.d_status, B * 4)))
        return rc;
    HIP_TRY(hipStreamSynchronize(st));
    const auto t1 = std::chrono::high_resolution_clock::now();
    stats->solve_time_us = std::chrono::duration<double, std::micro>(t1 - t0).count();
    stats->n_iters = counts[0];
    stats->ls_num_iters = counts[1];
    return 0;
}

int gato_last_kernel_time(gato_solver* s, double* avg_us, int32_t* launches) {
    if (!s) return fail(GATO_E_ARG, "null solver");
    if (avg_us) *avg_us = s->last_kernel_us;
    if (launches) *launches = s->last_launches;
    return 0;
}

}  // extern "C"
