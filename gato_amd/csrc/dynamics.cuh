// Rigid-body dynamics of the serial arms (iiwa14, indy7) for CDNA4, one lane per evaluation.
//
// Written from Featherstone's spatial-algebra equations over the URDF tables in model_tables.h
// (not from GRiD's generated code). Same functions as the reference's device dynamics:
//   RNEA bias c(q, qd; f_ext) with f_ext subtracted from the last link (iiwa14_fext.cuh:29-268),
//   M(q)^-1 (iiwa14_grid.cuh:4742-5175), qdd = M^-1 (u - c) (iiwa14_grid.cuh:5341-5352),
//   d qdd / d(q, qd) = -M^-1 dID/d(q, qd) (iiwa14_plant.cuh:230-268),
//   end-effector position = origin of the last joint frame (iiwa14_grid.cuh:2596-2664).
//
// Spatial vectors are [angular; linear]. Joint i rotates about its local z:
// R_i = R_tree_i Rz(q_i) (child orientation in parent), p_i = p_tree_i.
// Every loop over joints has a compile-time trip count, so the table look-ups fold into
// immediates and the many exact 0 / +-1 rotation entries vanish from the instruction stream.
// Scalar type S: float for the solver (hardware sine/cosine and reciprocals, as the reference's fast-math build
// evaluates them), double where FP64 is asked for (the library's sin, cos, division and square root).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "model_tables.h"

namespace gato {

constexpr double kGravity = 9.81;  // iiwa14_plant.cuh:25-28

template <int PL>
struct Plant {
    using J = gato_model::Joints<PL>;
    static constexpr int NQ = J::NQ;
    static constexpr int NX = 2 * NQ;
    static constexpr int NU = NQ;
    __host__ __device__ static constexpr float T(int i, int k) { return static_cast<float>(J::T[i][k]); }
    // limits tightened by the 0.1 margin: {lo - m, hi + m} with m = -0.1 (iiwa14_plant.cuh:31-70)
    template <typename S = float>
    __host__ __device__ static constexpr S qlo(int i) { return static_cast<S>(-J::T[i][22] + 0.1); }
    template <typename S = float>
    __host__ __device__ static constexpr S qhi(int i) { return static_cast<S>(J::T[i][22] - 0.1); }
    template <typename S = float>
    __host__ __device__ static constexpr S vlo(int i) { return static_cast<S>(-J::T[i][23] + 0.1); }
    template <typename S = float>
    __host__ __device__ static constexpr S vhi(int i) { return static_cast<S>(J::T[i][23] - 0.1); }
    template <typename S = float>
    __host__ __device__ static constexpr S ulo(int i) { return static_cast<S>(-J::T[i][24] + 0.1); }
    template <typename S = float>
    __host__ __device__ static constexpr S uhi(int i) { return static_cast<S>(J::T[i][24] - 0.1); }
};

// joint table entry in scalar type S
template <int PL, typename S>
__host__ __device__ constexpr S tb(int i, int k) {
    return static_cast<S>(gato_model::Joints<PL>::T[i][k]);
}

template <typename S>
struct V6T {
    S w0, w1, w2, l0, l1, l2;
};
using V6 = V6T<float>;

template <typename S = float>
__device__ __forceinline__ V6T<S> v6zero() { return V6T<S>{S(0), S(0), S(0), S(0), S(0), S(0)}; }
template <typename S>
__device__ __forceinline__ V6T<S> operator+(V6T<S> a, V6T<S> b) {
    return V6T<S>{a.w0 + b.w0, a.w1 + b.w1, a.w2 + b.w2, a.l0 + b.l0, a.l1 + b.l1, a.l2 + b.l2};
}
template <typename S>
__device__ __forceinline__ V6T<S> operator-(V6T<S> a, V6T<S> b) {
    return V6T<S>{a.w0 - b.w0, a.w1 - b.w1, a.w2 - b.w2, a.l0 - b.l0, a.l1 - b.l1, a.l2 - b.l2};
}
template <typename S>
__device__ __forceinline__ V6T<S> operator*(S s, V6T<S> a) {
    return V6T<S>{s * a.w0, s * a.w1, s * a.w2, s * a.l0, s * a.l1, s * a.l2};
}

// motion cross product v x m
template <typename S>
__device__ __forceinline__ V6T<S> crm(const V6T<S>& v, const V6T<S>& m) {
    return V6T<S>{v.w1 * m.w2 - v.w2 * m.w1, v.w2 * m.w0 - v.w0 * m.w2, v.w0 * m.w1 - v.w1 * m.w0,
                  v.l1 * m.w2 - v.l2 * m.w1 + v.w1 * m.l2 - v.w2 * m.l1,
                  v.l2 * m.w0 - v.l0 * m.w2 + v.w2 * m.l0 - v.w0 * m.l2,
                  v.l0 * m.w1 - v.l1 * m.w0 + v.w0 * m.l1 - v.w1 * m.l0};
}
// v x (S qd) with S = e_z (angular): [w x qd e_z; l x qd e_z]
template <typename S>
__device__ __forceinline__ V6T<S> crm_s(const V6T<S>& v, S qd) {
    return V6T<S>{v.w1 * qd, -v.w0 * qd, S(0), v.l1 * qd, -v.l0 * qd, S(0)};
}
// force cross product v x* f
template <typename S>
__device__ __forceinline__ V6T<S> crf(const V6T<S>& v, const V6T<S>& f) {
    return V6T<S>{v.w1 * f.w2 - v.w2 * f.w1 + v.l1 * f.l2 - v.l2 * f.l1,
                  v.w2 * f.w0 - v.w0 * f.w2 + v.l2 * f.l0 - v.l0 * f.l2,
                  v.w0 * f.w1 - v.w1 * f.w0 + v.l0 * f.l1 - v.l1 * f.l0,
                  v.w1 * f.l2 - v.w2 * f.l1, v.w2 * f.l0 - v.w0 * f.l2, v.w0 * f.l1 - v.w1 * f.l0};
}
// S x m = e_z (angular) x m
template <typename S>
__device__ __forceinline__ V6T<S> s_crm(const V6T<S>& m) { return V6T<S>{-m.w1, m.w0, S(0), -m.l1, m.l0, S(0)}; }
// S x* f
template <typename S>
__device__ __forceinline__ V6T<S> s_crf(const V6T<S>& f) { return V6T<S>{-f.w1, f.w0, S(0), -f.l1, f.l0, S(0)}; }

// motion parent -> child through joint i: w' = R^T w, l' = R^T (l - p x w), R^T = Rz^T R_tree^T
template <int PL, int I, typename S>
__device__ __forceinline__ V6T<S> xmot(S s, S c, const V6T<S>& m) {
    constexpr S p0 = tb<PL, S>(I, 9), p1 = tb<PL, S>(I, 10), p2 = tb<PL, S>(I, 11);
    const S t0 = m.l0 - (p1 * m.w2 - p2 * m.w1);
    const S t1 = m.l1 - (p2 * m.w0 - p0 * m.w2);
    const S t2 = m.l2 - (p0 * m.w1 - p1 * m.w0);
    // u = R_tree^T x  (R_tree row-major, T(I, 3r + c))
    const S a0 = tb<PL, S>(I, 0) * m.w0 + tb<PL, S>(I, 3) * m.w1 + tb<PL, S>(I, 6) * m.w2;
    const S a1 = tb<PL, S>(I, 1) * m.w0 + tb<PL, S>(I, 4) * m.w1 + tb<PL, S>(I, 7) * m.w2;
    const S a2 = tb<PL, S>(I, 2) * m.w0 + tb<PL, S>(I, 5) * m.w1 + tb<PL, S>(I, 8) * m.w2;
    const S b0 = tb<PL, S>(I, 0) * t0 + tb<PL, S>(I, 3) * t1 + tb<PL, S>(I, 6) * t2;
    const S b1 = tb<PL, S>(I, 1) * t0 + tb<PL, S>(I, 4) * t1 + tb<PL, S>(I, 7) * t2;
    const S b2 = tb<PL, S>(I, 2) * t0 + tb<PL, S>(I, 5) * t1 + tb<PL, S>(I, 8) * t2;
    return V6T<S>{c * a0 + s * a1, -s * a0 + c * a1, a2, c * b0 + s * b1, -s * b0 + c * b1, b2};
}

// force child -> parent through joint i (X^T f): f = R f', n = R n' + p x f
template <int PL, int I, typename S>
__device__ __forceinline__ V6T<S> xfrc(S s, S c, const V6T<S>& f) {
    constexpr S p0 = tb<PL, S>(I, 9), p1 = tb<PL, S>(I, 10), p2 = tb<PL, S>(I, 11);
    const S a0 = c * f.w0 - s * f.w1, a1 = s * f.w0 + c * f.w1, a2 = f.w2;
    const S b0 = c * f.l0 - s * f.l1, b1 = s * f.l0 + c * f.l1, b2 = f.l2;
    const S n0 = tb<PL, S>(I, 0) * a0 + tb<PL, S>(I, 1) * a1 + tb<PL, S>(I, 2) * a2;
    const S n1 = tb<PL, S>(I, 3) * a0 + tb<PL, S>(I, 4) * a1 + tb<PL, S>(I, 5) * a2;
    const S n2 = tb<PL, S>(I, 6) * a0 + tb<PL, S>(I, 7) * a1 + tb<PL, S>(I, 8) * a2;
    const S g0 = tb<PL, S>(I, 0) * b0 + tb<PL, S>(I, 1) * b1 + tb<PL, S>(I, 2) * b2;
    const S g1 = tb<PL, S>(I, 3) * b0 + tb<PL, S>(I, 4) * b1 + tb<PL, S>(I, 5) * b2;
    const S g2 = tb<PL, S>(I, 6) * b0 + tb<PL, S>(I, 7) * b1 + tb<PL, S>(I, 8) * b2;
    return V6T<S>{n0 + (p1 * g2 - p2 * g1), n1 + (p2 * g0 - p0 * g2), n2 + (p0 * g1 - p1 * g0), g0, g1, g2};
}

// spatial inertia of link i about its joint origin times motion: [Ib w + h x l; m l - h x w]
template <int PL, int I, typename S>
__device__ __forceinline__ V6T<S> imul(const V6T<S>& v) {
    constexpr S m = tb<PL, S>(I, 12), h0 = tb<PL, S>(I, 13), h1 = tb<PL, S>(I, 14), h2 = tb<PL, S>(I, 15);
    constexpr S xx = tb<PL, S>(I, 16), yy = tb<PL, S>(I, 17), zz = tb<PL, S>(I, 18), xy = tb<PL, S>(I, 19),
                xz = tb<PL, S>(I, 20), yz = tb<PL, S>(I, 21);
    return V6T<S>{xx * v.w0 + xy * v.w1 + xz * v.w2 + (h1 * v.l2 - h2 * v.l1),
                  xy * v.w0 + yy * v.w1 + yz * v.w2 + (h2 * v.l0 - h0 * v.l2),
                  xz * v.w0 + yz * v.w1 + zz * v.w2 + (h0 * v.l1 - h1 * v.l0),
                  m * v.l0 - (h1 * v.w2 - h2 * v.w1), m * v.l1 - (h2 * v.w0 - h0 * v.w2),
                  m * v.l2 - (h0 * v.w1 - h1 * v.w0)};
}

// compile-time loop helper: body(std::integral_constant<int, I>) for I in [0, N)
template <int I, int N>
struct Unroll {
    template <typename F>
    __device__ __forceinline__ static void run(F&& f) {
        f(std::integral_constant<int, I>{});
        Unroll<I + 1, N>::run(f);
    }
};
template <int N>
struct Unroll<N, N> {
    template <typename F>
    __device__ __forceinline__ static void run(F&&) {}
};
template <int I, int N>
struct UnrollDown {  // I = N-1 .. 0
    template <typename F>
    __device__ __forceinline__ static void run(F&& f) {
        f(std::integral_constant<int, N - 1 - I>{});
        UnrollDown<I + 1, N>::run(f);
    }
};
template <int N>
struct UnrollDown<N, N> {
    template <typename F>
    __device__ __forceinline__ static void run(F&&) {}
};

// 1 / x: the hardware reciprocal in fp32 (the reference's fast-math quotient), a division in fp64
template <typename S>
__device__ __forceinline__ S rcp(S x) {
    if constexpr (std::is_same_v<S, float>) return __builtin_amdgcn_rcpf(x);
    else return S(1) / x;
}

template <int PL, typename S = float>
struct Kin {
    static constexpr int NQ = Plant<PL>::NQ;
    S s[NQ], c[NQ];
    // fp32: the hardware sine and cosine, as the reference's fast-math build evaluates them (-use_fast_math)
    __device__ __forceinline__ void set(const S* q) {
#pragma unroll
        for (int i = 0; i < NQ; i++) {
            if constexpr (std::is_same_v<S, float>) __sincosf(q[i], &s[i], &c[i]);
            else sincos(q[i], &s[i], &c[i]);
        }
    }
};

// Where the RNEA keeps its per-link forces between the forward and backward sweeps.
template <int NQ, typename S = float>
struct FRegs {  // registers (6 NQ scalars per lane)
    V6T<S> f[NQ];
    __device__ __forceinline__ V6T<S> get(int i) const { return f[i]; }
    __device__ __forceinline__ void set(int i, const V6T<S>& v) { f[i] = v; }
};
template <typename S = float>
struct FLdsT {  // LDS: component c of link i at p[(6 i + c) * cs]
    S* p;
    int cs;
    __device__ __forceinline__ V6T<S> get(int i) const {
        const S* q = p + 6 * i * cs;
        return V6T<S>{q[0], q[cs], q[2 * cs], q[3 * cs], q[4 * cs], q[5 * cs]};
    }
    __device__ __forceinline__ void set(int i, const V6T<S>& v) {
        S* q = p + 6 * i * cs;
        q[0] = v.w0; q[cs] = v.w1; q[2 * cs] = v.w2; q[3 * cs] = v.l0; q[4 * cs] = v.l1; q[5 * cs] = v.l2;
    }
};
using FLds = FLdsT<float>;

// RNEA: tau = ID(q, qd, qdd) with gravity as base acceleration, f_ext (link-local) on the last link.
template <int PL, typename S, typename FS>
__device__ __forceinline__ void rnea_fs(const Kin<PL, S>& K, const S* qd, const S* qdd, const S* fext, S* tau,
                                        FS& fs) {
    constexpr int NQ = Plant<PL>::NQ;
    V6T<S> v = v6zero<S>(), a = V6T<S>{S(0), S(0), S(0), S(0), S(0), S(kGravity)};
    Unroll<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        v = xmot<PL, i>(K.s[i], K.c[i], v);
        a = xmot<PL, i>(K.s[i], K.c[i], a);
        v.w2 += qd[i];
        a = a + crm_s(v, qd[i]);
        if (qdd) a.w2 += qdd[i];
        V6T<S> f = imul<PL, i>(a) + crf(v, imul<PL, i>(v));
        if (i == NQ - 1 && fext) {
            f.w0 -= fext[0]; f.w1 -= fext[1]; f.w2 -= fext[2];
            f.l0 -= fext[3]; f.l1 -= fext[4]; f.l2 -= fext[5];
        }
        fs.set(i, f);
    });
    V6T<S> fa = fs.get(NQ - 1);
    UnrollDown<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        tau[i] = fa.w2;
        if constexpr (i > 0) fa = fs.get(i - 1) + xfrc<PL, i>(K.s[i], K.c[i], fa);
    });
}

template <int PL, bool KEEP, typename S>
__device__ __forceinline__ void rnea(const Kin<PL, S>& K, const S* qd, const S* qdd, const S* fext, S* tau) {
    FRegs<Plant<PL>::NQ, S> fs;
    rnea_fs<PL>(K, qd, qdd, fext, tau, fs);
}

// Rigid-body inertia (m, h, Ib) transported from joint frame I to its parent frame I-1.
template <typename S>
struct RBIT {
    S m, h0, h1, h2, xx, yy, zz, xy, xz, yz;
};

template <int PL, int I, typename S>
__device__ __forceinline__ RBIT<S> rbi_local() {
    return RBIT<S>{tb<PL, S>(I, 12), tb<PL, S>(I, 13), tb<PL, S>(I, 14), tb<PL, S>(I, 15), tb<PL, S>(I, 16),
                   tb<PL, S>(I, 17), tb<PL, S>(I, 18), tb<PL, S>(I, 19), tb<PL, S>(I, 20), tb<PL, S>(I, 21)};
}

// Ib' = R Ib R^T + m(|p|^2 1 - p p^T) + 2 (p.h~) 1 - p h~^T - h~ p^T,  h' = h~ + m p,  h~ = R h
template <int PL, int I, typename S>
__device__ __forceinline__ RBIT<S> rbi_to_parent(S s, S c, const RBIT<S>& X) {
    constexpr S p0 = tb<PL, S>(I, 9), p1 = tb<PL, S>(I, 10), p2 = tb<PL, S>(I, 11);
    // R = R_tree Rz: first Rz (rotate the inertia about z), then R_tree
    // Rz Ib Rz^T entries
    const S cc = c * c, ss = s * s, cs = c * s;
    const S zxx = cc * X.xx - S(2) * cs * X.xy + ss * X.yy;
    const S zyy = ss * X.xx + S(2) * cs * X.xy + cc * X.yy;
    const S zxy = cs * (X.xx - X.yy) + (cc - ss) * X.xy;
    const S zxz = c * X.xz - s * X.yz;
    const S zyz = s * X.xz + c * X.yz;
    const S zzz = X.zz;
    const S hz0 = c * X.h0 - s * X.h1, hz1 = s * X.h0 + c * X.h1, hz2 = X.h2;
    // R_tree (row-major T(I, 3r+k)) applied: M' = T M T^T
    S Mz[3][3] = {{zxx, zxy, zxz}, {zxy, zyy, zyz}, {zxz, zyz, zzz}};
    S TM[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            TM[r][k] = tb<PL, S>(I, 3 * r) * Mz[0][k] + tb<PL, S>(I, 3 * r + 1) * Mz[1][k] +
                       tb<PL, S>(I, 3 * r + 2) * Mz[2][k];
    S Mp[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            Mp[r][k] = TM[r][0] * tb<PL, S>(I, 3 * k) + TM[r][1] * tb<PL, S>(I, 3 * k + 1) +
                       TM[r][2] * tb<PL, S>(I, 3 * k + 2);
    const S g0 = tb<PL, S>(I, 0) * hz0 + tb<PL, S>(I, 1) * hz1 + tb<PL, S>(I, 2) * hz2;
    const S g1 = tb<PL, S>(I, 3) * hz0 + tb<PL, S>(I, 4) * hz1 + tb<PL, S>(I, 5) * hz2;
    const S g2 = tb<PL, S>(I, 6) * hz0 + tb<PL, S>(I, 7) * hz1 + tb<PL, S>(I, 8) * hz2;
    const S pp = p0 * p0 + p1 * p1 + p2 * p2;
    const S ph = p0 * g0 + p1 * g1 + p2 * g2;
    RBIT<S> o;
    o.m = X.m;
    o.h0 = g0 + X.m * p0;
    o.h1 = g1 + X.m * p1;
    o.h2 = g2 + X.m * p2;
    o.xx = Mp[0][0] + X.m * (pp - p0 * p0) + S(2) * ph - S(2) * p0 * g0;
    o.yy = Mp[1][1] + X.m * (pp - p1 * p1) + S(2) * ph - S(2) * p1 * g1;
    o.zz = Mp[2][2] + X.m * (pp - p2 * p2) + S(2) * ph - S(2) * p2 * g2;
    o.xy = Mp[0][1] - X.m * p0 * p1 - p0 * g1 - g0 * p1;
    o.xz = Mp[0][2] - X.m * p0 * p2 - p0 * g2 - g0 * p2;
    o.yz = Mp[1][2] - X.m * p1 * p2 - p1 * g2 - g1 * p2;
    return o;
}

template <typename S>
__device__ __forceinline__ RBIT<S> operator+(const RBIT<S>& a, const RBIT<S>& b) {
    return RBIT<S>{a.m + b.m,   a.h0 + b.h0, a.h1 + b.h1, a.h2 + b.h2, a.xx + b.xx,
                   a.yy + b.yy, a.zz + b.zz, a.xy + b.xy, a.xz + b.xz, a.yz + b.yz};
}

// Packed lower-triangular index of (r, c), r >= c
__host__ __device__ constexpr int tri(int r, int c) { return r * (r + 1) / 2 + c; }

// Joint-space inertia matrix by composite rigid bodies; M stored packed lower (tri(r, c)).
template <int PL, typename S>
__device__ __forceinline__ void crba(const Kin<PL, S>& K, S* M) {
    constexpr int NQ = Plant<PL>::NQ;
    RBIT<S> Ic = rbi_local<PL, NQ - 1, S>();
    UnrollDown<0, NQ>::run([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        // F = Ic_j S = [Ixz, Iyz, Izz, -h_y, h_x, 0]
        V6T<S> F{Ic.xz, Ic.yz, Ic.zz, -Ic.h1, Ic.h0, S(0)};
        M[tri(j, j)] = Ic.zz;
        Unroll<0, j>::run([&](auto kc) {
            constexpr int i = j - 1 - decltype(kc)::value;  // i = j-1 .. 0
            F = xfrc<PL, i + 1>(K.s[i + 1], K.c[i + 1], F);
            M[tri(j, i)] = F.w2;
        });
        if constexpr (j > 0) Ic = rbi_local<PL, j - 1, S>() + rbi_to_parent<PL, j>(K.s[j], K.c[j], Ic);
    });
}

// In-place Cholesky of packed lower M (SPD): M = L L^T
template <int N, typename S>
__device__ __forceinline__ void chol_packed(S* L) {
#pragma unroll
    for (int j = 0; j < N; j++) {
        S d = L[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; k++) d -= L[tri(j, k)] * L[tri(j, k)];
        S r;
        if constexpr (std::is_same_v<S, float>) r = rsqrtf(d);
        else r = S(1) / sqrt(d);
        L[tri(j, j)] = d * r;  // sqrt(d)
#pragma unroll
        for (int i = j + 1; i < N; i++) {
            S v = L[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; k++) v -= L[tri(i, k)] * L[tri(j, k)];
            L[tri(i, j)] = v * r;
        }
    }
}

// b <- (L L^T)^-1 b
template <int N, typename S>
__device__ __forceinline__ void chol_solve(const S* L, S* b) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        S v = b[i];
#pragma unroll
        for (int k = 0; k < i; k++) v -= L[tri(i, k)] * b[k];
        b[i] = v * rcp(L[tri(i, i)]);  // fp32: fast-math quotient, as the reference divides
    }
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        S v = b[i];
#pragma unroll
        for (int k = i + 1; k < N; k++) v -= L[tri(k, i)] * b[k];
        b[i] = v * rcp(L[tri(i, i)]);
    }
}

// qdd = M^-1 (u - c); returns the Cholesky factor in L (packed). FS keeps the RNEA link forces.
template <int PL, typename S, typename FS>
__device__ __forceinline__ void forward_dynamics_fs(const Kin<PL, S>& K, const S* qd, const S* u, const S* fext,
                                                    S* qdd, S* L, FS& fs) {
    constexpr int NQ = Plant<PL>::NQ;
    S cb[NQ];
    rnea_fs<PL>(K, qd, static_cast<const S*>(nullptr), fext, cb, fs);
    __builtin_amdgcn_sched_barrier(0);  // keep the RNEA and CRBA live ranges apart
    crba<PL>(K, L);
    chol_packed<NQ>(L);
#pragma unroll
    for (int i = 0; i < NQ; i++) qdd[i] = u[i] - cb[i];
    chol_solve<NQ>(L, qdd);
}

template <int PL, typename S>
__device__ __forceinline__ void forward_dynamics(const Kin<PL, S>& K, const S* qd, const S* u, const S* fext,
                                                 S* qdd, S* L) {
    constexpr int NQ = Plant<PL>::NQ;
    S cb[NQ];
    rnea<PL, false>(K, qd, static_cast<const S*>(nullptr), fext, cb);
    __builtin_amdgcn_sched_barrier(0);  // keep the RNEA and CRBA live ranges apart
    crba<PL>(K, L);
    chol_packed<NQ>(L);
#pragma unroll
    for (int i = 0; i < NQ; i++) qdd[i] = u[i] - cb[i];
    chol_solve<NQ>(L, qdd);
}

// Per-link RNEA table shared by the lanes of one knot (stride 8 scalars per spatial vector):
// velocity v_i at [8 i], acceleration a_i at [8 (NQ + i)], force transmitted across joint i at [8 (2 NQ + i)].
template <int PL>
struct RneaTab {
    static constexpr int NQ = Plant<PL>::NQ;
    static constexpr int SIZE = 8 * 3 * NQ;
    template <typename S>
    __device__ static void put(S* t, int slot, const V6T<S>& x) {
        S* p = t + 8 * slot;
        p[0] = x.w0; p[1] = x.w1; p[2] = x.w2; p[3] = x.l0; p[4] = x.l1; p[5] = x.l2;
    }
    template <typename S>
    __device__ static V6T<S> get(const S* t, int slot) {
        const S* p = t + 8 * slot;
        return V6T<S>{p[0], p[1], p[2], p[3], p[4], p[5]};
    }
};

// RNEA at (q, qd, qdd) with f_ext; the `writer` lane publishes v_i, a_i and the transmitted forces.
// Value barrier: the optimiser may not assume x is unchanged, so sub-expressions computed from it
// are not merged with earlier ones (which would keep those earlier values alive in registers).
__device__ __forceinline__ float opaque(float x) {
    asm volatile("" : "+v"(x));
    return x;
}
__device__ __forceinline__ double opaque(double x) {
    asm volatile("" : "+v"(x));
    return x;
}
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Every lane of the knot group computes identical values and stores them to the same addresses,
// so the table needs no ordering beyond the lane's own program order.
template <int PL, typename S>
__device__ __forceinline__ void rnea_publish(const Kin<PL, S>& K0, const S* qd0, const S* qdd, const S* fext, S* tab) {
    constexpr int NQ = Plant<PL>::NQ;
    using T = RneaTab<PL>;
    // recompute from opaque copies instead of reusing the forward-dynamics sweep's values
    Kin<PL, S> K;
    S qd[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) {
        K.s[i] = opaque(K0.s[i]);
        K.c[i] = opaque(K0.c[i]);
        qd[i] = opaque(qd0[i]);
    }
    V6T<S> v = v6zero<S>(), a = V6T<S>{S(0), S(0), S(0), S(0), S(0), S(kGravity)};
    Unroll<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        v = xmot<PL, i>(K.s[i], K.c[i], v);
        a = xmot<PL, i>(K.s[i], K.c[i], a);
        v.w2 += qd[i];
        a = a + crm_s(v, qd[i]);
        a.w2 += qdd[i];
        T::put(tab, i, v);
        T::put(tab, NQ + i, a);
        V6T<S> f = imul<PL, i>(a) + crf(v, imul<PL, i>(v));
        if (i == NQ - 1 && fext) {
            f.w0 -= fext[0]; f.w1 -= fext[1]; f.w2 -= fext[2];
            f.l0 -= fext[3]; f.l1 -= fext[4]; f.l2 -= fext[5];
        }
        T::put(tab, 2 * NQ + i, f);
    });
    V6T<S> fa = T::get(tab, 3 * NQ - 1);
    UnrollDown<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (i > 0) {
            fa = T::get(tab, 2 * NQ + i - 1) + xfrc<PL, i>(K.s[i], K.c[i], fa);
            T::put(tab, 2 * NQ + i - 1, fa);
        }
    });
}

// Joint axes carried into every later link frame: S_i^(k) = X_k ... X_{i+1} S_i for i <= k, stored at
// slot k(k+1)/2 + i of an 8-scalar-stride table (identical in every lane of a knot group).
template <int PL, typename S>
__device__ __forceinline__ void publish_axes(const Kin<PL, S>& K, S* ax) {
    constexpr int NQ = Plant<PL>::NQ;
    using T = RneaTab<PL>;
    Unroll<0, NQ>::run([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        T::put(ax, tri(k, k), V6T<S>{S(0), S(0), S(1), S(0), S(0), S(0)});
        Unroll<0, k>::run([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            T::put(ax, tri(k, i), xmot<PL, k>(K.s[k], K.c[k], T::get(ax, tri(k - 1, i))));
        });
    });
}

template <typename S>
__device__ __forceinline__ S dot6(const V6T<S>& m, const V6T<S>& f) {
    return m.w0 * f.w0 + m.w1 * f.w1 + m.w2 * f.w2 + m.l0 * f.l0 + m.l1 * f.l1 + m.l2 * f.l2;
}

// Tangent of ID along q_j (dir 0) or qd_j (dir 1), qdd fixed, as a single forward sweep:
//   d tau_i = sum_{k >= i} (S_i^(k))^T g_k,  g_k = dI_k-force of link k  (+ S x* F_j at k = j for dir 0),
// which equals the backward force accumulation of the differentiated RNEA, dF_{k-1} += X_k^T (dF_k + [k = j]
// S x* F_j), unrolled (the extra term contributes (S_j)^T (S x* F_j) = 0 to d tau_j itself).
template <int PL, typename S>
__device__ __forceinline__ void rnea_tangent_ax(const Kin<PL, S>& K, const S* qd, const S* qdd, const S* tab0,
                                                const S* ax0, int j, int dir, S* dtau) {
    constexpr int NQ = Plant<PL>::NQ;
    using T = RneaTab<PL>;
    // read the tables through an offset the optimiser cannot see through: the values are reloaded
    // link by link instead of being forwarded from the sweep that stored them
    int off = 0;
    asm volatile("" : "+v"(off));
    const S* tab = tab0 + off;
    const S* ax = ax0 + off;
    const V6T<S> fj = T::get(tab, 2 * NQ + j);
    // the seed enters through 0/1 factors (no per-link branches: a run-time j would otherwise make the
    // compiler carry both variants of every link's terms)
    const S wq = dir == 0 ? S(1) : S(0), wv = S(1) - wq;
    V6T<S> dv = v6zero<S>(), da = v6zero<S>();
#pragma unroll
    for (int i = 0; i < NQ; i++) dtau[i] = S(0);
    Unroll<0, NQ>::run([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        asm volatile("" ::: "memory");  // one link at a time: table loads are neither hoisted nor merged
        const V6T<S> v = T::get(tab, k);
        const V6T<S> a = T::get(tab, NQ + k);
        V6T<S> xv = v;  // X_k v_{k-1}
        xv.w2 -= qd[k];
        V6T<S> xa = a - crm_s(v, qd[k]);  // X_k a_{k-1}
        xa.w2 -= qdd[k];
        const S h = (k == j) ? S(1) : S(0);
        const S hq = h * wq, hv = h * wv;
        dv = xmot<PL, k>(K.s[k], K.c[k], dv) - hq * s_crm(xv);
        dv.w2 += hv;
        da = xmot<PL, k>(K.s[k], K.c[k], da) - hq * s_crm(xa) + crm_s(dv, qd[k]) +
             hv * V6T<S>{v.w1, -v.w0, S(0), v.l1, -v.l0, S(0)};  // v_k x S
        const V6T<S> g = imul<PL, k>(da) + crf(dv, imul<PL, k>(v)) + crf(v, imul<PL, k>(dv)) + hq * s_crf(fj);
        Unroll<0, k + 1>::run([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            dtau[i] += dot6(T::get(ax, tri(k, i)), g);
        });
    });
}

// End-effector position (origin of the last joint frame) and, for lane-owned column jc < NQ,
// the Jacobian column dp/dq_jc = z_jc x (p - o_jc). Pass jc < 0 to skip the Jacobian.
template <int PL, typename S>
__device__ __forceinline__ void ee_position(const Kin<PL, S>& K, int jc, S* p, S* Jcol) {
    constexpr int NQ = Plant<PL>::NQ;
    S W[3][3] = {{S(1), S(0), S(0)}, {S(0), S(1), S(0)}, {S(0), S(0), S(1)}};
    S o[3] = {S(0), S(0), S(0)};
    S zj[3] = {S(0), S(0), S(0)}, oj[3] = {S(0), S(0), S(0)};
    Unroll<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
#pragma unroll
        for (int r = 0; r < 3; r++)
            o[r] += W[r][0] * tb<PL, S>(i, 9) + W[r][1] * tb<PL, S>(i, 10) + W[r][2] * tb<PL, S>(i, 11);
        // W <- W R_tree Rz
        S WT[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                WT[r][k] = W[r][0] * tb<PL, S>(i, k) + W[r][1] * tb<PL, S>(i, 3 + k) + W[r][2] * tb<PL, S>(i, 6 + k);
        const S s = K.s[i], c = K.c[i];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            W[r][0] = WT[r][0] * c + WT[r][1] * s;
            W[r][1] = -WT[r][0] * s + WT[r][1] * c;
            W[r][2] = WT[r][2];
        }
        if (i == jc) {
#pragma unroll
            for (int r = 0; r < 3; r++) { zj[r] = W[r][2]; oj[r] = o[r]; }
        }
    });
    p[0] = o[0]; p[1] = o[1]; p[2] = o[2];
    if (Jcol) {
        const S d0 = o[0] - oj[0], d1 = o[1] - oj[1], d2 = o[2] - oj[2];
        Jcol[0] = zj[1] * d2 - zj[2] * d1;
        Jcol[1] = zj[2] * d0 - zj[0] * d2;
        Jcol[2] = zj[0] * d1 - zj[1] * d0;
    }
}

}  // namespace gato
