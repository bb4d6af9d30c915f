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
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "model_tables.h"

namespace gato {

constexpr float kGravity = 9.81f;  // iiwa14_plant.cuh:25-28

template <int PL>
struct Plant {
    using J = gato_model::Joints<PL>;
    static constexpr int NQ = J::NQ;
    static constexpr int NX = 2 * NQ;
    static constexpr int NU = NQ;
    __host__ __device__ static constexpr float T(int i, int k) { return static_cast<float>(J::T[i][k]); }
    // limits tightened by the 0.1 margin: {lo - m, hi + m} with m = -0.1 (iiwa14_plant.cuh:31-70)
    __host__ __device__ static constexpr float qlo(int i) { return static_cast<float>(-J::T[i][22] + 0.1); }
    __host__ __device__ static constexpr float qhi(int i) { return static_cast<float>(J::T[i][22] - 0.1); }
    __host__ __device__ static constexpr float vlo(int i) { return static_cast<float>(-J::T[i][23] + 0.1); }
    __host__ __device__ static constexpr float vhi(int i) { return static_cast<float>(J::T[i][23] - 0.1); }
    __host__ __device__ static constexpr float ulo(int i) { return static_cast<float>(-J::T[i][24] + 0.1); }
    __host__ __device__ static constexpr float uhi(int i) { return static_cast<float>(J::T[i][24] - 0.1); }
};

struct V6 {
    float w0, w1, w2, l0, l1, l2;
};

__device__ __forceinline__ V6 v6zero() { return V6{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}; }
__device__ __forceinline__ V6 operator+(V6 a, V6 b) {
    return V6{a.w0 + b.w0, a.w1 + b.w1, a.w2 + b.w2, a.l0 + b.l0, a.l1 + b.l1, a.l2 + b.l2};
}
__device__ __forceinline__ V6 operator-(V6 a, V6 b) {
    return V6{a.w0 - b.w0, a.w1 - b.w1, a.w2 - b.w2, a.l0 - b.l0, a.l1 - b.l1, a.l2 - b.l2};
}
__device__ __forceinline__ V6 operator*(float s, V6 a) {
    return V6{s * a.w0, s * a.w1, s * a.w2, s * a.l0, s * a.l1, s * a.l2};
}

// motion cross product v x m
__device__ __forceinline__ V6 crm(const V6& v, const V6& m) {
    return V6{v.w1 * m.w2 - v.w2 * m.w1, v.w2 * m.w0 - v.w0 * m.w2, v.w0 * m.w1 - v.w1 * m.w0,
              v.l1 * m.w2 - v.l2 * m.w1 + v.w1 * m.l2 - v.w2 * m.l1,
              v.l2 * m.w0 - v.l0 * m.w2 + v.w2 * m.l0 - v.w0 * m.l2,
              v.l0 * m.w1 - v.l1 * m.w0 + v.w0 * m.l1 - v.w1 * m.l0};
}
// v x (S qd) with S = e_z (angular): [w x qd e_z; l x qd e_z]
__device__ __forceinline__ V6 crm_s(const V6& v, float qd) {
    return V6{v.w1 * qd, -v.w0 * qd, 0.f, v.l1 * qd, -v.l0 * qd, 0.f};
}
// force cross product v x* f
__device__ __forceinline__ V6 crf(const V6& v, const V6& f) {
    return V6{v.w1 * f.w2 - v.w2 * f.w1 + v.l1 * f.l2 - v.l2 * f.l1,
              v.w2 * f.w0 - v.w0 * f.w2 + v.l2 * f.l0 - v.l0 * f.l2,
              v.w0 * f.w1 - v.w1 * f.w0 + v.l0 * f.l1 - v.l1 * f.l0,
              v.w1 * f.l2 - v.w2 * f.l1, v.w2 * f.l0 - v.w0 * f.l2, v.w0 * f.l1 - v.w1 * f.l0};
}
// S x m = e_z (angular) x m
__device__ __forceinline__ V6 s_crm(const V6& m) { return V6{-m.w1, m.w0, 0.f, -m.l1, m.l0, 0.f}; }
// S x* f
__device__ __forceinline__ V6 s_crf(const V6& f) { return V6{-f.w1, f.w0, 0.f, -f.l1, f.l0, 0.f}; }

// motion parent -> child through joint i: w' = R^T w, l' = R^T (l - p x w), R^T = Rz^T R_tree^T
template <int PL, int I>
__device__ __forceinline__ V6 xmot(float s, float c, const V6& m) {
    using P = Plant<PL>;
    constexpr float p0 = P::T(I, 9), p1 = P::T(I, 10), p2 = P::T(I, 11);
    const float t0 = m.l0 - (p1 * m.w2 - p2 * m.w1);
    const float t1 = m.l1 - (p2 * m.w0 - p0 * m.w2);
    const float t2 = m.l2 - (p0 * m.w1 - p1 * m.w0);
    // u = R_tree^T x  (R_tree row-major, T(I, 3r + c))
    const float a0 = P::T(I, 0) * m.w0 + P::T(I, 3) * m.w1 + P::T(I, 6) * m.w2;
    const float a1 = P::T(I, 1) * m.w0 + P::T(I, 4) * m.w1 + P::T(I, 7) * m.w2;
    const float a2 = P::T(I, 2) * m.w0 + P::T(I, 5) * m.w1 + P::T(I, 8) * m.w2;
    const float b0 = P::T(I, 0) * t0 + P::T(I, 3) * t1 + P::T(I, 6) * t2;
    const float b1 = P::T(I, 1) * t0 + P::T(I, 4) * t1 + P::T(I, 7) * t2;
    const float b2 = P::T(I, 2) * t0 + P::T(I, 5) * t1 + P::T(I, 8) * t2;
    return V6{c * a0 + s * a1, -s * a0 + c * a1, a2, c * b0 + s * b1, -s * b0 + c * b1, b2};
}

// force child -> parent through joint i (X^T f): f = R f', n = R n' + p x f
template <int PL, int I>
__device__ __forceinline__ V6 xfrc(float s, float c, const V6& f) {
    using P = Plant<PL>;
    constexpr float p0 = P::T(I, 9), p1 = P::T(I, 10), p2 = P::T(I, 11);
    const float a0 = c * f.w0 - s * f.w1, a1 = s * f.w0 + c * f.w1, a2 = f.w2;
    const float b0 = c * f.l0 - s * f.l1, b1 = s * f.l0 + c * f.l1, b2 = f.l2;
    const float n0 = P::T(I, 0) * a0 + P::T(I, 1) * a1 + P::T(I, 2) * a2;
    const float n1 = P::T(I, 3) * a0 + P::T(I, 4) * a1 + P::T(I, 5) * a2;
    const float n2 = P::T(I, 6) * a0 + P::T(I, 7) * a1 + P::T(I, 8) * a2;
    const float g0 = P::T(I, 0) * b0 + P::T(I, 1) * b1 + P::T(I, 2) * b2;
    const float g1 = P::T(I, 3) * b0 + P::T(I, 4) * b1 + P::T(I, 5) * b2;
    const float g2 = P::T(I, 6) * b0 + P::T(I, 7) * b1 + P::T(I, 8) * b2;
    return V6{n0 + (p1 * g2 - p2 * g1), n1 + (p2 * g0 - p0 * g2), n2 + (p0 * g1 - p1 * g0), g0, g1, g2};
}

// spatial inertia of link i about its joint origin times motion: [Ib w + h x l; m l - h x w]
template <int PL, int I>
__device__ __forceinline__ V6 imul(const V6& v) {
    using P = Plant<PL>;
    constexpr float m = P::T(I, 12), h0 = P::T(I, 13), h1 = P::T(I, 14), h2 = P::T(I, 15);
    constexpr float xx = P::T(I, 16), yy = P::T(I, 17), zz = P::T(I, 18), xy = P::T(I, 19), xz = P::T(I, 20),
                    yz = P::T(I, 21);
    return V6{xx * v.w0 + xy * v.w1 + xz * v.w2 + (h1 * v.l2 - h2 * v.l1),
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

template <int PL>
struct Kin {
    static constexpr int NQ = Plant<PL>::NQ;
    float s[NQ], c[NQ];
    // the hardware sine and cosine, as the reference's fast-math build evaluates them (-use_fast_math)
    __device__ __forceinline__ void set(const float* q) {
#pragma unroll
        for (int i = 0; i < NQ; i++) __sincosf(q[i], &s[i], &c[i]);
    }
};

// Where the RNEA keeps its per-link forces between the forward and backward sweeps.
template <int NQ>
struct FRegs {  // registers (6 NQ floats per lane)
    V6 f[NQ];
    __device__ __forceinline__ V6 get(int i) const { return f[i]; }
    __device__ __forceinline__ void set(int i, const V6& v) { f[i] = v; }
};
struct FLds {  // LDS: component c of link i at p[(6 i + c) * cs]
    float* p;
    int cs;
    __device__ __forceinline__ V6 get(int i) const {
        const float* q = p + 6 * i * cs;
        return V6{q[0], q[cs], q[2 * cs], q[3 * cs], q[4 * cs], q[5 * cs]};
    }
    __device__ __forceinline__ void set(int i, const V6& v) {
        float* q = p + 6 * i * cs;
        q[0] = v.w0; q[cs] = v.w1; q[2 * cs] = v.w2; q[3 * cs] = v.l0; q[4 * cs] = v.l1; q[5 * cs] = v.l2;
    }
};

// RNEA: tau = ID(q, qd, qdd) with gravity as base acceleration, f_ext (link-local) on the last link.
template <int PL, typename FS>
__device__ __forceinline__ void rnea_fs(const Kin<PL>& K, const float* qd, const float* qdd, const float* fext,
                                        float* tau, FS& fs) {
    constexpr int NQ = Plant<PL>::NQ;
    V6 v = v6zero(), a = V6{0.f, 0.f, 0.f, 0.f, 0.f, kGravity};
    Unroll<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        v = xmot<PL, i>(K.s[i], K.c[i], v);
        a = xmot<PL, i>(K.s[i], K.c[i], a);
        v.w2 += qd[i];
        a = a + crm_s(v, qd[i]);
        if (qdd) a.w2 += qdd[i];
        V6 f = imul<PL, i>(a) + crf(v, imul<PL, i>(v));
        if (i == NQ - 1 && fext) {
            f.w0 -= fext[0]; f.w1 -= fext[1]; f.w2 -= fext[2];
            f.l0 -= fext[3]; f.l1 -= fext[4]; f.l2 -= fext[5];
        }
        fs.set(i, f);
    });
    V6 fa = fs.get(NQ - 1);
    UnrollDown<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        tau[i] = fa.w2;
        if constexpr (i > 0) fa = fs.get(i - 1) + xfrc<PL, i>(K.s[i], K.c[i], fa);
    });
}

template <int PL, bool KEEP>
__device__ __forceinline__ void rnea(const Kin<PL>& K, const float* qd, const float* qdd, const float* fext,
                                     float* tau) {
    FRegs<Plant<PL>::NQ> fs;
    rnea_fs<PL>(K, qd, qdd, fext, tau, fs);
}

// Rigid-body inertia (m, h, Ib) transported from joint frame I to its parent frame I-1.
struct RBI {
    float m, h0, h1, h2, xx, yy, zz, xy, xz, yz;
};

template <int PL, int I>
__device__ __forceinline__ RBI rbi_local() {
    using P = Plant<PL>;
    return RBI{P::T(I, 12), P::T(I, 13), P::T(I, 14), P::T(I, 15), P::T(I, 16),
               P::T(I, 17), P::T(I, 18), P::T(I, 19), P::T(I, 20), P::T(I, 21)};
}

// Ib' = R Ib R^T + m(|p|^2 1 - p p^T) + 2 (p.h~) 1 - p h~^T - h~ p^T,  h' = h~ + m p,  h~ = R h
template <int PL, int I>
__device__ __forceinline__ RBI rbi_to_parent(float s, float c, const RBI& X) {
    using P = Plant<PL>;
    constexpr float p0 = P::T(I, 9), p1 = P::T(I, 10), p2 = P::T(I, 11);
    // R = R_tree Rz: first Rz (rotate the inertia about z), then R_tree
    // Rz Ib Rz^T entries
    const float cc = c * c, ss = s * s, cs = c * s;
    const float zxx = cc * X.xx - 2.f * cs * X.xy + ss * X.yy;
    const float zyy = ss * X.xx + 2.f * cs * X.xy + cc * X.yy;
    const float zxy = cs * (X.xx - X.yy) + (cc - ss) * X.xy;
    const float zxz = c * X.xz - s * X.yz;
    const float zyz = s * X.xz + c * X.yz;
    const float zzz = X.zz;
    const float hz0 = c * X.h0 - s * X.h1, hz1 = s * X.h0 + c * X.h1, hz2 = X.h2;
    // R_tree (row-major T(I, 3r+k)) applied: M' = T M T^T
    float Mz[3][3] = {{zxx, zxy, zxz}, {zxy, zyy, zyz}, {zxz, zyz, zzz}};
    float TM[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            TM[r][k] = P::T(I, 3 * r) * Mz[0][k] + P::T(I, 3 * r + 1) * Mz[1][k] + P::T(I, 3 * r + 2) * Mz[2][k];
    float Mp[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            Mp[r][k] = TM[r][0] * P::T(I, 3 * k) + TM[r][1] * P::T(I, 3 * k + 1) + TM[r][2] * P::T(I, 3 * k + 2);
    const float g0 = P::T(I, 0) * hz0 + P::T(I, 1) * hz1 + P::T(I, 2) * hz2;
    const float g1 = P::T(I, 3) * hz0 + P::T(I, 4) * hz1 + P::T(I, 5) * hz2;
    const float g2 = P::T(I, 6) * hz0 + P::T(I, 7) * hz1 + P::T(I, 8) * hz2;
    const float pp = p0 * p0 + p1 * p1 + p2 * p2;
    const float ph = p0 * g0 + p1 * g1 + p2 * g2;
    RBI o;
    o.m = X.m;
    o.h0 = g0 + X.m * p0;
    o.h1 = g1 + X.m * p1;
    o.h2 = g2 + X.m * p2;
    o.xx = Mp[0][0] + X.m * (pp - p0 * p0) + 2.f * ph - 2.f * p0 * g0;
    o.yy = Mp[1][1] + X.m * (pp - p1 * p1) + 2.f * ph - 2.f * p1 * g1;
    o.zz = Mp[2][2] + X.m * (pp - p2 * p2) + 2.f * ph - 2.f * p2 * g2;
    o.xy = Mp[0][1] - X.m * p0 * p1 - p0 * g1 - g0 * p1;
    o.xz = Mp[0][2] - X.m * p0 * p2 - p0 * g2 - g0 * p2;
    o.yz = Mp[1][2] - X.m * p1 * p2 - p1 * g2 - g1 * p2;
    return o;
}

__device__ __forceinline__ RBI operator+(const RBI& a, const RBI& b) {
    return RBI{a.m + b.m,   a.h0 + b.h0, a.h1 + b.h1, a.h2 + b.h2, a.xx + b.xx,
               a.yy + b.yy, a.zz + b.zz, a.xy + b.xy, a.xz + b.xz, a.yz + b.yz};
}

// Packed lower-triangular index of (r, c), r >= c
__host__ __device__ constexpr int tri(int r, int c) { return r * (r + 1) / 2 + c; }

// Joint-space inertia matrix by composite rigid bodies; M stored packed lower (tri(r, c)).
template <int PL>
__device__ __forceinline__ void crba(const Kin<PL>& K, float* M) {
    constexpr int NQ = Plant<PL>::NQ;
    RBI Ic = rbi_local<PL, NQ - 1>();
    UnrollDown<0, NQ>::run([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        // F = Ic_j S = [Ixz, Iyz, Izz, -h_y, h_x, 0]
        V6 F{Ic.xz, Ic.yz, Ic.zz, -Ic.h1, Ic.h0, 0.f};
        M[tri(j, j)] = Ic.zz;
        Unroll<0, j>::run([&](auto kc) {
            constexpr int i = j - 1 - decltype(kc)::value;  // i = j-1 .. 0
            F = xfrc<PL, i + 1>(K.s[i + 1], K.c[i + 1], F);
            M[tri(j, i)] = F.w2;
        });
        if constexpr (j > 0) Ic = rbi_local<PL, j - 1>() + rbi_to_parent<PL, j>(K.s[j], K.c[j], Ic);
    });
}

// In-place Cholesky of packed lower M (SPD): M = L L^T
template <int N>
__device__ __forceinline__ void chol_packed(float* L) {
#pragma unroll
    for (int j = 0; j < N; j++) {
        float d = L[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; k++) d -= L[tri(j, k)] * L[tri(j, k)];
        const float r = rsqrtf(d);
        L[tri(j, j)] = d * r;  // sqrt(d)
#pragma unroll
        for (int i = j + 1; i < N; i++) {
            float v = L[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; k++) v -= L[tri(i, k)] * L[tri(j, k)];
            L[tri(i, j)] = v * r;
        }
    }
}

// b <- (L L^T)^-1 b
template <int N>
__device__ __forceinline__ void chol_solve(const float* L, float* b) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        float v = b[i];
#pragma unroll
        for (int k = 0; k < i; k++) v -= L[tri(i, k)] * b[k];
        b[i] = v * __builtin_amdgcn_rcpf(L[tri(i, i)]);  // fast-math quotient, as the reference divides
    }
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        float v = b[i];
#pragma unroll
        for (int k = i + 1; k < N; k++) v -= L[tri(k, i)] * b[k];
        b[i] = v * __builtin_amdgcn_rcpf(L[tri(i, i)]);
    }
}

// qdd = M^-1 (u - c); returns the Cholesky factor in L (packed). FS keeps the RNEA link forces.
template <int PL, typename FS>
__device__ __forceinline__ void forward_dynamics_fs(const Kin<PL>& K, const float* qd, const float* u,
                                                    const float* fext, float* qdd, float* L, FS& fs) {
    constexpr int NQ = Plant<PL>::NQ;
    float cb[NQ];
    rnea_fs<PL>(K, qd, nullptr, fext, cb, fs);
    __builtin_amdgcn_sched_barrier(0);  // keep the RNEA and CRBA live ranges apart
    crba<PL>(K, L);
    chol_packed<NQ>(L);
#pragma unroll
    for (int i = 0; i < NQ; i++) qdd[i] = u[i] - cb[i];
    chol_solve<NQ>(L, qdd);
}

template <int PL>
__device__ __forceinline__ void forward_dynamics(const Kin<PL>& K, const float* qd, const float* u,
                                                 const float* fext, float* qdd, float* L) {
    constexpr int NQ = Plant<PL>::NQ;
    float cb[NQ];
    rnea<PL, false>(K, qd, nullptr, fext, cb);
    __builtin_amdgcn_sched_barrier(0);  // keep the RNEA and CRBA live ranges apart
    crba<PL>(K, L);
    chol_packed<NQ>(L);
#pragma unroll
    for (int i = 0; i < NQ; i++) qdd[i] = u[i] - cb[i];
    chol_solve<NQ>(L, qdd);
}

// Per-link RNEA table shared by the lanes of one knot (stride 8 floats per spatial vector):
// velocity v_i at [8 i], acceleration a_i at [8 (NQ + i)], force transmitted across joint i at [8 (2 NQ + i)].
template <int PL>
struct RneaTab {
    static constexpr int NQ = Plant<PL>::NQ;
    static constexpr int SIZE = 8 * 3 * NQ;
    __device__ static void put(float* t, int slot, const V6& x) {
        float* p = t + 8 * slot;
        p[0] = x.w0; p[1] = x.w1; p[2] = x.w2; p[3] = x.l0; p[4] = x.l1; p[5] = x.l2;
    }
    __device__ static V6 get(const float* t, int slot) {
        const float* p = t + 8 * slot;
        return V6{p[0], p[1], p[2], p[3], p[4], p[5]};
    }
};

// RNEA at (q, qd, qdd) with f_ext; the `writer` lane publishes v_i, a_i and the transmitted forces.
// Value barrier: the optimiser may not assume x is unchanged, so sub-expressions computed from it
// are not merged with earlier ones (which would keep those earlier values alive in registers).
__device__ __forceinline__ float opaque(float x) {
    asm volatile("" : "+v"(x));
    return x;
}
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Every lane of the knot group computes identical values and stores them to the same addresses,
// so the table needs no ordering beyond the lane's own program order.
template <int PL>
__device__ __forceinline__ void rnea_publish(const Kin<PL>& K0, const float* qd0, const float* qdd, const float* fext,
                                             float* tab) {
    constexpr int NQ = Plant<PL>::NQ;
    using T = RneaTab<PL>;
    // recompute from opaque copies instead of reusing the forward-dynamics sweep's values
    Kin<PL> K;
    float qd[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) {
        K.s[i] = opaque(K0.s[i]);
        K.c[i] = opaque(K0.c[i]);
        qd[i] = opaque(qd0[i]);
    }
    V6 v = v6zero(), a = V6{0.f, 0.f, 0.f, 0.f, 0.f, kGravity};
    Unroll<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        v = xmot<PL, i>(K.s[i], K.c[i], v);
        a = xmot<PL, i>(K.s[i], K.c[i], a);
        v.w2 += qd[i];
        a = a + crm_s(v, qd[i]);
        a.w2 += qdd[i];
        T::put(tab, i, v);
        T::put(tab, NQ + i, a);
        V6 f = imul<PL, i>(a) + crf(v, imul<PL, i>(v));
        if (i == NQ - 1 && fext) {
            f.w0 -= fext[0]; f.w1 -= fext[1]; f.w2 -= fext[2];
            f.l0 -= fext[3]; f.l1 -= fext[4]; f.l2 -= fext[5];
        }
        T::put(tab, 2 * NQ + i, f);
    });
    V6 fa = T::get(tab, 3 * NQ - 1);
    UnrollDown<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if constexpr (i > 0) {
            fa = T::get(tab, 2 * NQ + i - 1) + xfrc<PL, i>(K.s[i], K.c[i], fa);
            T::put(tab, 2 * NQ + i - 1, fa);
        }
    });
}

// Joint axes carried into every later link frame: S_i^(k) = X_k ... X_{i+1} S_i for i <= k, stored at
// slot k(k+1)/2 + i of an 8-float-stride table (identical in every lane of a knot group).
template <int PL>
__device__ __forceinline__ void publish_axes(const Kin<PL>& K, float* ax) {
    constexpr int NQ = Plant<PL>::NQ;
    using T = RneaTab<PL>;
    Unroll<0, NQ>::run([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        T::put(ax, tri(k, k), V6{0.f, 0.f, 1.f, 0.f, 0.f, 0.f});
        Unroll<0, k>::run([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            T::put(ax, tri(k, i), xmot<PL, k>(K.s[k], K.c[k], T::get(ax, tri(k - 1, i))));
        });
    });
}

__device__ __forceinline__ float dot6(const V6& m, const V6& f) {
    return m.w0 * f.w0 + m.w1 * f.w1 + m.w2 * f.w2 + m.l0 * f.l0 + m.l1 * f.l1 + m.l2 * f.l2;
}

// Tangent of ID along q_j (dir 0) or qd_j (dir 1), qdd fixed, as a single forward sweep:
//   d tau_i = sum_{k >= i} (S_i^(k))^T g_k,  g_k = dI_k-force of link k  (+ S x* F_j at k = j for dir 0),
// which equals the backward force accumulation of the differentiated RNEA, dF_{k-1} += X_k^T (dF_k + [k = j]
// S x* F_j), unrolled (the extra term contributes (S_j)^T (S x* F_j) = 0 to d tau_j itself).
template <int PL>
__device__ __forceinline__ void rnea_tangent_ax(const Kin<PL>& K, const float* qd, const float* qdd,
                                                const float* tab0, const float* ax0, int j, int dir, float* dtau) {
    constexpr int NQ = Plant<PL>::NQ;
    using T = RneaTab<PL>;
    // read the tables through an offset the optimiser cannot see through: the values are reloaded
    // link by link instead of being forwarded from the sweep that stored them
    int off = 0;
    asm volatile("" : "+v"(off));
    const float* tab = tab0 + off;
    const float* ax = ax0 + off;
    const V6 fj = T::get(tab, 2 * NQ + j);
    // the seed enters through 0/1 factors (no per-link branches: a run-time j would otherwise make the
    // compiler carry both variants of every link's terms)
    const float wq = dir == 0 ? 1.f : 0.f, wv = 1.f - wq;
    V6 dv = v6zero(), da = v6zero();
#pragma unroll
    for (int i = 0; i < NQ; i++) dtau[i] = 0.f;
    Unroll<0, NQ>::run([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        asm volatile("" ::: "memory");  // one link at a time: table loads are neither hoisted nor merged
        const V6 v = T::get(tab, k);
        const V6 a = T::get(tab, NQ + k);
        V6 xv = v;  // X_k v_{k-1}
        xv.w2 -= qd[k];
        V6 xa = a - crm_s(v, qd[k]);  // X_k a_{k-1}
        xa.w2 -= qdd[k];
        const float h = (k == j) ? 1.f : 0.f;
        const float hq = h * wq, hv = h * wv;
        dv = xmot<PL, k>(K.s[k], K.c[k], dv) - hq * s_crm(xv);
        dv.w2 += hv;
        da = xmot<PL, k>(K.s[k], K.c[k], da) - hq * s_crm(xa) + crm_s(dv, qd[k]) +
             hv * V6{v.w1, -v.w0, 0.f, v.l1, -v.l0, 0.f};  // v_k x S
        const V6 g = imul<PL, k>(da) + crf(dv, imul<PL, k>(v)) + crf(v, imul<PL, k>(dv)) + hq * s_crf(fj);
        Unroll<0, k + 1>::run([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            dtau[i] += dot6(T::get(ax, tri(k, i)), g);
        });
    });
}

// End-effector position (origin of the last joint frame) and, for lane-owned column jc < NQ,
// the Jacobian column dp/dq_jc = z_jc x (p - o_jc). Pass jc < 0 to skip the Jacobian.
template <int PL>
__device__ __forceinline__ void ee_position(const Kin<PL>& K, int jc, float* p, float* Jcol) {
    using P = Plant<PL>;
    constexpr int NQ = P::NQ;
    float W[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    float o[3] = {0.f, 0.f, 0.f};
    float zj[3] = {0.f, 0.f, 0.f}, oj[3] = {0.f, 0.f, 0.f};
    Unroll<0, NQ>::run([&](auto ic) {
        constexpr int i = decltype(ic)::value;
#pragma unroll
        for (int r = 0; r < 3; r++)
            o[r] += W[r][0] * P::T(i, 9) + W[r][1] * P::T(i, 10) + W[r][2] * P::T(i, 11);
        // W <- W R_tree Rz
        float WT[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                WT[r][k] = W[r][0] * P::T(i, k) + W[r][1] * P::T(i, 3 + k) + W[r][2] * P::T(i, 6 + k);
        const float s = K.s[i], c = K.c[i];
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
        const float d0 = o[0] - oj[0], d1 = o[1] - oj[1], d2 = o[2] - oj[2];
        Jcol[0] = zj[1] * d2 - zj[2] * d1;
        Jcol[1] = zj[2] * d0 - zj[0] * d2;
        Jcol[2] = zj[0] * d1 - zj[1] * d0;
    }
}

}  // namespace gato
