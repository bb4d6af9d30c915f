/* CPU FP64 restatement of GATO's batched SQP iteration (Mo0dy/GATO).
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load this library, and only as the checker / the timed CPU baseline. The product
 * (gato_amd/) never links or calls it.
 *
 * Every phase cites the reference file:line whose behaviour it restates. The rigid-body
 * dynamics are written from Featherstone's spatial-algebra equations over the URDF-derived
 * tables in gato_amd/csrc/model_tables.h, not from GRiD's generated code; they compute the
 * same functions (RNEA with f_ext on the last link, M^-1, d(FD)/d(q,qd), end-effector FK).
 * How each function is pinned (closed forms, identities, finite differences, dense KKT
 * solves, GRiD's inertia tables) is listed in tests/test_oracle_*.py.
 */
#include "gato_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../gato_amd/csrc/model_tables.h"

#define MAXQ 7
#define MAXX 14
#define NUM_ALPHAS 8

/* ρ schedule constants: gato/settings.h:18-22 */
#define RHO_INIT 1e-3
#define RHO_FACTOR 1.2
#define RHO_MIN 1e-8
#define RHO_MAX 10.0
#define GRAVITY 9.81              /* iiwa14_plant.cuh:25-28 */
#define LIMIT_MARGIN (-0.1)       /* iiwa14_plant.cuh:31-34 */
#define PCG_ABS_TOL 1e-6          /* pcg.cuh:27 */

typedef struct {
    int nq, indy;
    double Rt[MAXQ][9], pt[MAXQ][3], m[MAXQ], h[MAXQ][3], Ib[MAXQ][6];
    double qlo[MAXQ], qhi[MAXQ], vlo[MAXQ], vhi[MAXQ], ulo[MAXQ], uhi[MAXQ];
} plant_t;

static plant_t g_plants[3];
static int g_init = 0;

static void init_plant(plant_t* P, const double (*tab)[GATO_JOINT_REC], int nq, int indy) {
    P->nq = nq;
    P->indy = indy;
    for (int i = 0; i < nq; i++) {
        const double* t = tab[i];
        memcpy(P->Rt[i], t, 9 * sizeof(double));
        memcpy(P->pt[i], t + 9, 3 * sizeof(double));
        P->m[i] = t[12];
        memcpy(P->h[i], t + 13, 3 * sizeof(double));
        memcpy(P->Ib[i], t + 16, 6 * sizeof(double));
        /* {lo - margin, hi + margin} with margin = -0.1: iiwa14_plant.cuh:36-70 */
        P->qlo[i] = -t[22] - LIMIT_MARGIN;
        P->qhi[i] = t[22] + LIMIT_MARGIN;
        P->vlo[i] = -t[23] - LIMIT_MARGIN;
        P->vhi[i] = t[23] + LIMIT_MARGIN;
        P->ulo[i] = -t[24] - LIMIT_MARGIN;
        P->uhi[i] = t[24] + LIMIT_MARGIN;
    }
}

static const plant_t* get_plant(int plant) {
    if (!g_init) {
#pragma omp critical(orc_init)
        {
            if (!g_init) {
                init_plant(&g_plants[0], gato_iiwa14_joints, GATO_IIWA14_NQ, 0);
                init_plant(&g_plants[1], gato_indy7_joints, GATO_INDY7_NQ, 1);
                init_plant(&g_plants[2], gato_indy7_urdf_joints, GATO_INDY7_URDF_NQ, 1);
                g_init = 1;
            }
        }
    }
    return &g_plants[plant == ORC_PLANT_INDY7 ? 1 : plant == ORC_PLANT_INDY7_URDF ? 2 : 0];
}

int orc_nq(int plant) { return get_plant(plant)->nq; }

void orc_limits(int plant, double* qlo, double* qhi, double* vlo, double* vhi, double* ulo, double* uhi) {
    const plant_t* P = get_plant(plant);
    for (int i = 0; i < P->nq; i++) {
        qlo[i] = P->qlo[i]; qhi[i] = P->qhi[i];
        vlo[i] = P->vlo[i]; vhi[i] = P->vhi[i];
        ulo[i] = P->ulo[i]; uhi[i] = P->uhi[i];
    }
}

/* ------------------------------------------------------------------ spatial algebra */
/* spatial vectors are [angular(3); linear(3)] (Featherstone; GRiD uses the same order). */

static void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* orientation R (child in parent) and origin p of joint frame i at angle q: R = R_tree Rz(q) */
static void joint_pose(const plant_t* P, int i, double q, double R[9], double p[3]) {
    double c = cos(q), s = sin(q);
    const double* T = P->Rt[i];
    for (int r = 0; r < 3; r++) {
        R[3 * r + 0] = T[3 * r + 0] * c + T[3 * r + 1] * s;
        R[3 * r + 1] = -T[3 * r + 0] * s + T[3 * r + 1] * c;
        R[3 * r + 2] = T[3 * r + 2];
    }
    p[0] = P->pt[i][0]; p[1] = P->pt[i][1]; p[2] = P->pt[i][2];
}

/* motion vector parent -> child coordinates: w' = R^T w, v' = R^T (v - p x w) */
static void xmot(const double R[9], const double p[3], const double* m, double* o) {
    double t[3], pw[3];
    cross3(p, m, pw);
    t[0] = m[3] - pw[0]; t[1] = m[4] - pw[1]; t[2] = m[5] - pw[2];
    for (int r = 0; r < 3; r++) {
        o[r] = R[r] * m[0] + R[3 + r] * m[1] + R[6 + r] * m[2];
        o[3 + r] = R[r] * t[0] + R[3 + r] * t[1] + R[6 + r] * t[2];
    }
}

/* force vector child -> parent coordinates (X^T f): f = R f', n = R n' + p x f */
static void xfrcT(const double R[9], const double p[3], const double* f, double* o) {
    double fl[3], pn[3];
    for (int r = 0; r < 3; r++) {
        o[r] = R[3 * r] * f[0] + R[3 * r + 1] * f[1] + R[3 * r + 2] * f[2];
        fl[r] = R[3 * r] * f[3] + R[3 * r + 1] * f[4] + R[3 * r + 2] * f[5];
    }
    cross3(p, fl, pn);
    for (int r = 0; r < 3; r++) {
        o[r] += pn[r];
        o[3 + r] = fl[r];
    }
}

/* spatial inertia about the frame origin times motion: [Ib w + h x v; m v - h x w] */
static void imul(const plant_t* P, int i, const double* v, double* o) {
    const double* I = P->Ib[i];
    const double* h = P->h[i];
    double hv[3], hw[3];
    cross3(h, v + 3, hv);
    cross3(h, v, hw);
    o[0] = I[0] * v[0] + I[3] * v[1] + I[4] * v[2] + hv[0];
    o[1] = I[3] * v[0] + I[1] * v[1] + I[5] * v[2] + hv[1];
    o[2] = I[4] * v[0] + I[5] * v[1] + I[2] * v[2] + hv[2];
    for (int r = 0; r < 3; r++) o[3 + r] = P->m[i] * v[3 + r] - hw[r];
}

/* motion cross product v x m */
static void crm(const double* v, const double* m, double* o) {
    double a[3], b[3];
    cross3(v, m, o);
    cross3(v + 3, m, a);
    cross3(v, m + 3, b);
    for (int r = 0; r < 3; r++) o[3 + r] = a[r] + b[r];
}

/* force cross product v x* f */
static void crf(const double* v, const double* f, double* o) {
    double a[3], b[3];
    cross3(v, f, a);
    cross3(v + 3, f + 3, b);
    for (int r = 0; r < 3; r++) o[r] = a[r] + b[r];
    cross3(v, f + 3, o + 3);
}

/* ------------------------------------------------------------------ RNEA (+ tangent) */
/* ID(q, qd, qdd) with gravity as base acceleration and f_ext subtracted from the last link's
 * force in its own frame (iiwa14_fext.cuh:29-268; f_ext: :164-171).
 * kind = 0: plain; kind = 1: also d tau / d q_j; kind = 2: also d tau / d qd_j (qdd fixed). */
static void rnea_t(const plant_t* P, const double* q, const double* qd, const double* qdd, const double* fext,
                   int gravity, int kind, int jd, double* tau, double* dtau) {
    const int n = P->nq;
    double R[MAXQ][9], p[MAXQ][3];
    double v[MAXQ][6], a[MAXQ][6], f[MAXQ][6];
    double dv[MAXQ][6], da[MAXQ][6], df[MAXQ][6];
    const double a0[6] = {0, 0, 0, 0, 0, gravity ? GRAVITY : 0.0};
    const double zero[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        joint_pose(P, i, q[i], R[i], p[i]);
        const double* vp = i ? v[i - 1] : zero;
        const double* ap = i ? a[i - 1] : a0;
        const double* dvp = i ? dv[i - 1] : zero;
        const double* dap = i ? da[i - 1] : zero;
        double xv[6], xa[6];
        xmot(R[i], p[i], vp, xv);
        xmot(R[i], p[i], ap, xa);
        xmot(R[i], p[i], dvp, dv[i]);
        xmot(R[i], p[i], dap, da[i]);
        if (kind == 1 && i == jd) { /* dX/dq = -crm(S) X */
            const double S[6] = {0, 0, 1, 0, 0, 0};
            double t[6];
            crm(S, xv, t);
            for (int r = 0; r < 6; r++) dv[i][r] -= t[r];
            crm(S, xa, t);
            for (int r = 0; r < 6; r++) da[i][r] -= t[r];
        }
        for (int r = 0; r < 6; r++) { v[i][r] = xv[r]; a[i][r] = xa[r]; }
        v[i][2] += qd[i];
        a[i][2] += qdd[i];
        const double Sqd[6] = {0, 0, qd[i], 0, 0, 0};
        double t[6];
        crm(v[i], Sqd, t);
        for (int r = 0; r < 6; r++) a[i][r] += t[r];
        if (kind) {
            if (kind == 2 && i == jd) dv[i][2] += 1.0;
            crm(dv[i], Sqd, t);
            for (int r = 0; r < 6; r++) da[i][r] += t[r];
            if (kind == 2 && i == jd) {
                const double S[6] = {0, 0, 1, 0, 0, 0};
                crm(v[i], S, t);
                for (int r = 0; r < 6; r++) da[i][r] += t[r];
            }
        }
        double Iv[6], Ia[6];
        imul(P, i, v[i], Iv);
        imul(P, i, a[i], Ia);
        crf(v[i], Iv, t);
        for (int r = 0; r < 6; r++) f[i][r] = Ia[r] + t[r];
        if (kind) {
            double Idv[6], Ida[6], t2[6];
            imul(P, i, dv[i], Idv);
            imul(P, i, da[i], Ida);
            crf(dv[i], Iv, t);
            crf(v[i], Idv, t2);
            for (int r = 0; r < 6; r++) df[i][r] = Ida[r] + t[r] + t2[r];
        }
    }
    if (fext)
        for (int r = 0; r < 6; r++) f[n - 1][r] -= fext[r];
    for (int i = n - 1; i >= 0; i--) {
        tau[i] = f[i][2];
        if (kind) dtau[i] = df[i][2];
        if (i > 0) {
            double t[6];
            xfrcT(R[i], p[i], f[i], t);
            for (int r = 0; r < 6; r++) f[i - 1][r] += t[r];
            if (kind) {
                xfrcT(R[i], p[i], df[i], t);
                for (int r = 0; r < 6; r++) df[i - 1][r] += t[r];
                if (kind == 1 && i == jd) { /* d(X^T)/dq f = X^T crf(S) f */
                    const double S[6] = {0, 0, 1, 0, 0, 0};
                    double u[6];
                    crf(S, f[i], u);
                    xfrcT(R[i], p[i], u, t);
                    for (int r = 0; r < 6; r++) df[i - 1][r] += t[r];
                }
            }
        }
    }
}

void orc_rnea(int plant, const double* q, const double* qd, const double* qdd, const double* fext, int gravity,
              double* tau) {
    rnea_t(get_plant(plant), q, qd, qdd, fext, gravity, 0, 0, tau, NULL);
}

/* M(q) column j = ID(q, 0, e_j) without gravity or f_ext */
static void mass_matrix(const plant_t* P, const double* q, double* M) {
    const int n = P->nq;
    double z[MAXQ] = {0}, e[MAXQ];
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) e[i] = (i == j);
        rnea_t(P, q, z, e, NULL, 0, 0, 0, M + j * n, NULL);
    }
    for (int i = 0; i < n; i++) /* symmetrise exactly */
        for (int j = i + 1; j < n; j++) {
            double s = 0.5 * (M[i + j * n] + M[j + i * n]);
            M[i + j * n] = M[j + i * n] = s;
        }
}

void orc_mass(int plant, const double* q, double* M) { mass_matrix(get_plant(plant), q, M); }

/* Cholesky of an SPD n x n (column-major) in place (lower), returns 0 on success */
static int chol(double* A, int n) {
    for (int j = 0; j < n; j++) {
        double d = A[j + j * n];
        for (int k = 0; k < j; k++) d -= A[j + k * n] * A[j + k * n];
        if (!(d > 0)) return -1;
        d = sqrt(d);
        A[j + j * n] = d;
        for (int i = j + 1; i < n; i++) {
            double s = A[i + j * n];
            for (int k = 0; k < j; k++) s -= A[i + k * n] * A[j + k * n];
            A[i + j * n] = s / d;
        }
    }
    return 0;
}

static void chol_solve(const double* L, int n, double* b) {
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= L[i + k * n] * b[k];
        b[i] = s / L[i + i * n];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int k = i + 1; k < n; k++) s -= L[k + i * n] * b[k];
        b[i] = s / L[i + i * n];
    }
}

/* qdd = M^-1 (u - c(q, qd; f_ext)) (iiwa14_fext.cuh:504-508) */
static void fd(const plant_t* P, const double* q, const double* qd, const double* u, const double* fext,
               double* qdd) {
    const int n = P->nq;
    double M[MAXQ * MAXQ], z[MAXQ] = {0}, c[MAXQ];
    rnea_t(P, q, qd, z, fext, 1, 0, 0, c, NULL);
    mass_matrix(P, q, M);
    chol(M, n);
    for (int i = 0; i < n; i++) qdd[i] = u[i] - c[i];
    chol_solve(M, n, qdd);
}

void orc_fd(int plant, const double* q, const double* qd, const double* u, const double* fext, double* qdd) {
    fd(get_plant(plant), q, qd, u, fext, qdd);
}

/* d qdd / d(q, qd) = -M^-1 dID/d(q, qd) at qdd = FD (iiwa14_plant.cuh:230-268);
 * dqdd is nq x 2nq column-major ([d/dq | d/dqd]); Minv nq x nq. */
static void fd_grad(const plant_t* P, const double* q, const double* qd, const double* u, const double* fext,
                    double* qdd, double* dqdd, double* Minv) {
    const int n = P->nq;
    double M[MAXQ * MAXQ], z[MAXQ] = {0}, c[MAXQ], tau[MAXQ];
    rnea_t(P, q, qd, z, fext, 1, 0, 0, c, NULL);
    mass_matrix(P, q, M);
    chol(M, n);
    for (int i = 0; i < n; i++) qdd[i] = u[i] - c[i];
    chol_solve(M, n, qdd);
    for (int kind = 1; kind <= 2; kind++)
        for (int j = 0; j < n; j++) {
            double* col = dqdd + n * ((kind - 1) * n + j);
            rnea_t(P, q, qd, qdd, fext, 1, kind, j, tau, col);
            for (int i = 0; i < n; i++) col[i] = -col[i];
            chol_solve(M, n, col);
        }
    if (Minv)
        for (int j = 0; j < n; j++) {
            double* col = Minv + j * n;
            for (int i = 0; i < n; i++) col[i] = (i == j);
            chol_solve(M, n, col);
        }
}

void orc_fd_grad(int plant, const double* q, const double* qd, const double* u, const double* fext, double* qdd,
                 double* dqdd, double* Minv) {
    fd_grad(get_plant(plant), q, qd, u, fext, qdd, dqdd, Minv);
}

/* end-effector = origin of the last joint frame, no tool offset (iiwa14_grid.cuh:2596-2664);
 * J (3 x nq, column-major) = d p / d q. */
static void fk(const plant_t* P, const double* q, double* pee, double* J) {
    const int n = P->nq;
    double Rw[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, o[3] = {0, 0, 0};
    double zj[MAXQ][3], oj[MAXQ][3];
    for (int i = 0; i < n; i++) {
        double R[9], p[3], Rn[9];
        joint_pose(P, i, q[i], R, p);
        for (int r = 0; r < 3; r++) o[r] += Rw[3 * r] * p[0] + Rw[3 * r + 1] * p[1] + Rw[3 * r + 2] * p[2];
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++)
                Rn[3 * r + cc] = Rw[3 * r] * R[cc] + Rw[3 * r + 1] * R[3 + cc] + Rw[3 * r + 2] * R[6 + cc];
        memcpy(Rw, Rn, sizeof(Rw));
        for (int r = 0; r < 3; r++) { zj[i][r] = Rw[3 * r + 2]; oj[i][r] = o[r]; }
    }
    pee[0] = o[0]; pee[1] = o[1]; pee[2] = o[2];
    if (J)
        for (int j = 0; j < n; j++) {
            double d[3] = {o[0] - oj[j][0], o[1] - oj[j][1], o[2] - oj[j][2]};
            cross3(zj[j], d, J + 3 * j);
        }
}

void orc_fk(int plant, const double* q, double* p, double* J) { fk(get_plant(plant), q, p, J); }

/* world orientation (row-major) and origin of every joint frame: R [nq][9], o [nq][3] */
void orc_frames(int plant, const double* q, double* Rout, double* oout) {
    const plant_t* P = get_plant(plant);
    double Rw[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, o[3] = {0, 0, 0};
    for (int i = 0; i < P->nq; i++) {
        double R[9], p[3], Rn[9];
        joint_pose(P, i, q[i], R, p);
        for (int r = 0; r < 3; r++) o[r] += Rw[3 * r] * p[0] + Rw[3 * r + 1] * p[1] + Rw[3 * r + 2] * p[2];
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++)
                Rn[3 * r + cc] = Rw[3 * r] * R[cc] + Rw[3 * r + 1] * R[3 + cc] + Rw[3 * r + 2] * R[6 + cc];
        memcpy(Rw, Rn, sizeof(Rw));
        memcpy(Rout + 9 * i, Rw, sizeof(Rw));
        memcpy(oout + 3 * i, o, sizeof(o));
    }
}

/* ------------------------------------------------------------------ cost */
/* log barrier (iiwa14_plant.cuh:104-111; identical in indy7_plant.cuh) */
static double barrier(double v, double lo, double hi) {
    double a = v - lo, b = hi - v;
    a = a <= 1e-10 ? 1e-10 : a;
    b = b <= 1e-10 ? 1e-10 : b;
    return -log(a) - log(b);
}

/* barrier gradient: iiwa14 sign-preserving clamp (iiwa14_plant.cuh:114-136);
 * indy7 plain clamp (indy7_plant.cuh:141-148) */
static double barrier_g(const plant_t* P, double v, double lo, double hi) {
    double a = v - lo, b = hi - v;
    const double eps = 1e-6;
    if (P->indy) {
        a = a <= eps ? eps : a;
        b = b <= eps ? eps : b;
    } else {
        if (a >= 0) { if (a < eps) a = eps; } else { if (a > -eps) a = -eps; }
        if (b >= 0) { if (b < eps) b = eps; } else { if (b > -eps) b = -eps; }
    }
    return -1.0 / a + 1.0 / b;
}

/* iiwa14 barrier curvature (iiwa14_plant.cuh:141-155) */
static double barrier_h(double v, double lo, double hi) {
    double a = fabs(v - lo), b = fabs(hi - v);
    if (a < 1e-6) a = 1e-6;
    if (b < 1e-6) b = 1e-6;
    return 1.0 / (a * a) + 1.0 / (b * b);
}

static void default_weights(const orc_cfg* cfg, int nq, const double* Qd, const double* Rd, double* Qw, double* Rw) {
    for (int i = 0; i < 2 * nq; i++) Qw[i] = Qd ? Qd[i] : (i < nq ? 0.0 : cfg->qd_cost);
    for (int i = 0; i < nq; i++) Rw[i] = Rd ? Rd[i] : cfg->u_cost;
}

/* stage cost of the merit (iiwa14_plant.cuh:276-327); joint-space diagonal weights Qw/Rw and
 * joint reference xr generalise qd_cost/u_cost (reference: Qw = [0, qd_cost], xr = 0, Rw = u_cost). */
static double stage_cost(const orc_cfg* cfg, const plant_t* P, int k, const double* x, const double* u,
                         const double* rk, const double* Qw, const double* Rw, const double* xr) {
    const int nq = P->nq, N = cfg->N;
    double c = 0.0, pee[3];
    for (int i = 0; i < nq; i++) {
        c += cfg->q_lim_cost * barrier(x[i], P->qlo[i], P->qhi[i]);
        c += cfg->vel_lim_cost * barrier(x[nq + i], P->vlo[i], P->vhi[i]);
    }
    for (int i = 0; i < 2 * nq; i++) {
        double e = x[i] - (xr ? xr[i] : 0.0);
        c += 0.5 * Qw[i] * e * e;
    }
    if (k < N - 1)
        for (int i = 0; i < nq; i++) {
            c += 0.5 * Rw[i] * u[i] * u[i];
            c += cfg->ctrl_lim_cost * barrier(u[i], P->ulo[i], P->uhi[i]);
        }
    fk(P, x, pee, NULL);
    double w = (k == N - 1) ? cfg->N_cost : cfg->q_cost;
    for (int i = 0; i < 3; i++) {
        double e = pee[i] - rk[i];
        c += 0.5 * w * e * e;
    }
    return c;
}

/* quadratic model of one knot (iiwa14_plant.cuh:339-424, indy7_plant.cuh:335-415).
 * Q (nx x nx column-major), q (nx); R diagonal returned as nu x nu, r (nu) when R != NULL. */
static void cost_model(const orc_cfg* cfg, const plant_t* P, double w, const double* x, const double* u,
                       const double* rk, const double* Qw, const double* Rw, const double* xr, double* Q, double* q,
                       double* R, double* r) {
    const int nq = P->nq, nx = 2 * nq;
    double pee[3], J[3 * MAXQ], g[MAXQ], bg[MAXQ];
    fk(P, x, pee, J);
    for (int j = 0; j < nq; j++)
        g[j] = J[3 * j] * (pee[0] - rk[0]) + J[3 * j + 1] * (pee[1] - rk[1]) + J[3 * j + 2] * (pee[2] - rk[2]);
    for (int i = 0; i < nq; i++) bg[i] = barrier_g(P, x[i], P->qlo[i], P->qhi[i]);
    memset(Q, 0, sizeof(double) * nx * nx);
    for (int i = 0; i < nq; i++) {
        q[i] = w * g[i] + cfg->q_lim_cost * bg[i] + Qw[i] * (x[i] - (xr ? xr[i] : 0.0));
        double bv = barrier_g(P, x[nq + i], P->vlo[i], P->vhi[i]);
        q[nq + i] = Qw[nq + i] * (x[nq + i] - (xr ? xr[nq + i] : 0.0)) + cfg->vel_lim_cost * bv;
        for (int j = 0; j < nq; j++) {
            double v = cfg->hess_mode ? w * (J[3 * i] * J[3 * j] + J[3 * i + 1] * J[3 * j + 1] + J[3 * i + 2] * J[3 * j + 2])
                                      : w * g[i] * g[j]; /* rank-1 EE term (quirk Q5) */
            if (P->indy)
                v += cfg->q_lim_cost * bg[i] * bg[j];
            else if (i == j)
                v += cfg->q_lim_cost * barrier_h(x[i], P->qlo[i], P->qhi[i]);
            Q[i + j * nx] = v;
        }
        Q[i + i * nx] += Qw[i];
        Q[(nq + i) * (nx + 1)] = Qw[nq + i] + cfg->vel_lim_cost * (P->indy ? bv * bv : barrier_h(x[nq + i], P->vlo[i], P->vhi[i]));
    }
    if (R) {
        memset(R, 0, sizeof(double) * nq * nq);
        for (int i = 0; i < nq; i++) {
            double bu = barrier_g(P, u[i], P->ulo[i], P->uhi[i]);
            r[i] = Rw[i] * u[i] + cfg->ctrl_lim_cost * bu;
            R[i * (nq + 1)] = Rw[i] + cfg->ctrl_lim_cost * (P->indy ? bu * bu : barrier_h(u[i], P->ulo[i], P->uhi[i]));
        }
    }
}

/* ------------------------------------------------------------------ KKT setup */
/* setupKKTSystemBatchedKernel (setup_kkt.cuh:14-108) for one problem:
 * A_k (nx x nx), B_k (nx x nu) column-major, c_{k+1} = x_{k+1} - x+(x_k,u_k) for k <= N-2, c_0 = x_0 - x_s,
 * trapezoidal step (integrator.cuh:34-37, 143-184). */
void orc_setup_kkt(const orc_cfg* cfg, const double* xu, const double* xs, const double* ref, const double* fext,
                   const double* Qd, const double* Rd, const double* xref, double* A, double* Bm, double* c,
                   double* Q, double* R, double* q, double* r) {
    const plant_t* P = get_plant(cfg->plant);
    const int nq = P->nq, nx = 2 * nq, nu = nq, N = cfg->N, st = nx + nu;
    const double dt = cfg->dt, h2 = 0.5 * dt * dt;
    double Qw[MAXX], Rw[MAXQ];
    default_weights(cfg, nq, Qd, Rd, Qw, Rw);
    for (int k = 0; k < N - 1; k++) {
        const double* x = xu + k * st;
        const double* u = x + nx;
        const double* xn = x + st;
        double qdd[MAXQ], dq[MAXQ * 2 * MAXQ], Mi[MAXQ * MAXQ];
        fd_grad(P, x, x + nq, u, fext, qdd, dq, Mi);
        double* Ak = A + k * nx * nx;
        double* Bk = Bm + k * nx * nu;
        for (int col = 0; col < nx; col++)
            for (int row = 0; row < nx; row++) {
                int rr = row % nq;
                double d = dq[rr + nq * col];
                double v = (row == col);
                if (row < nq) {
                    if (col >= nq && row == col - nq) v += dt;
                    v += h2 * d;
                } else {
                    v += dt * d;
                }
                Ak[row + col * nx] = v;
            }
        for (int col = 0; col < nu; col++)
            for (int row = 0; row < nx; row++) {
                int rr = row % nq;
                double d = Mi[rr + nq * col];
                Bk[row + col * nx] = row < nq ? h2 * d : dt * d;
            }
        double* ck = c + (k + 1) * nx;
        for (int i = 0; i < nq; i++) {
            ck[i] = xn[i] - (x[i] + dt * x[nq + i] + 0.5 * qdd[i] * dt * dt);
            ck[nq + i] = xn[nq + i] - (x[nq + i] + dt * qdd[i]);
        }
        const double* xr = xref ? xref + k * nx : NULL;
        cost_model(cfg, P, cfg->q_cost, x, u, ref + 6 * k, Qw, Rw, xr, Q + k * nx * nx, q + k * nx,
                   R + k * nu * nu, r + k * nu);
    }
    {
        /* terminal block: quirk Q4 (iiwa14_plant.cuh:427-450) evaluates it at x_{N-2} with ref_{N-1}
         * and q_cost (quirk Q2); terminal_mode 1 uses x_{N-1} and N_cost. */
        int k = N - 1;
        const double* x = cfg->terminal_mode ? xu + k * st : xu + (k - 1) * st;
        double w = cfg->terminal_mode ? cfg->N_cost : cfg->q_cost;
        const double* xr = xref ? xref + k * nx : NULL;
        cost_model(cfg, P, w, x, NULL, ref + 6 * k, Qw, Rw, xr, Q + k * nx * nx, q + k * nx, NULL, NULL);
        memset(R + k * nu * nu, 0, sizeof(double) * nu * nu);
        memset(r + k * nu, 0, sizeof(double) * nu);
    }
    for (int i = 0; i < nx; i++) c[i] = xu[i] - xs[i];
}

/* ------------------------------------------------------------------ Schur system */
/* Gauss-Jordan inverse without pivoting (linalg.cuh:365-519); A n x n column-major -> Ai */
static void gj_inverse(const double* A, double* Ai, int n) {
    double W[MAXX * 2 * MAXX];
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++) {
            W[r + c * n] = A[r + c * n];
            W[r + (c + n) * n] = (r == c);
        }
    for (int p = 0; p < n; p++) {
        double piv = 1.0 / W[p + p * n];
        for (int c = 0; c < 2 * n; c++) W[p + c * n] *= piv;
        for (int r = 0; r < n; r++) {
            if (r == p) continue;
            double f = W[r + p * n];
            if (f == 0.0) continue;
            for (int c = 0; c < 2 * n; c++) W[r + c * n] -= f * W[p + c * n];
        }
    }
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++) Ai[r + c * n] = W[r + (c + n) * n];
}

/* C = A B (m x l)(l x n), column-major */
static void mm(const double* A, const double* B, double* C, int m, int l, int n) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double s = 0;
            for (int k = 0; k < l; k++) s += A[i + k * m] * B[k + j * l];
            C[i + j * m] = s;
        }
}

/* formSchurSystemBatchedKernel1/2 (schur_linsys.cuh:13-260) for one problem.
 * Q~_k = Q_k + rho diag(1_nq, 0) (quirk Q6, linalg.cuh:85-96). Outputs (column-major blocks):
 *   Qi[k] = Q~_k^-1, Ri[k] = R_k^-1 (k <= N-2),
 *   S: Sd[0] = -Q~_0^-1, Sd[k+1] = -theta_k, So[k+1] = S_{k+1,k} = phi_k = A_k Q~_k^-1,
 *   P: Pd[0] = -Q~_0, Pd[k+1] = -(theta_k + rho diag(1_nq,0))^-1 (quirk Q8), Po[k+1] = -Pd[k+1] phi_k Pd[k],
 *   gamma_0 = c_0 - Q~_0^-1 q_0, gamma_{k+1} = c_{k+1} - Q~_{k+1}^-1 q_{k+1} + phi_k q_k + B_k R_k^-1 r_k. */
void orc_schur(const orc_cfg* cfg, double rho, const double* A, const double* Bm, const double* Q, const double* R,
               const double* q, const double* r, const double* c, double* Qi, double* Ri, double* Sd, double* So,
               double* Pd, double* Po, double* gamma) {
    const plant_t* P = get_plant(cfg->plant);
    const int nq = P->nq, nx = 2 * nq, nu = nq, N = cfg->N, n2 = nx * nx;
    double Qt[MAXX * MAXX];
    for (int k = 0; k < N; k++) {
        memcpy(Qt, Q + k * n2, sizeof(double) * n2);
        for (int i = 0; i < nq; i++) Qt[i * (nx + 1)] += rho;
        gj_inverse(Qt, Qi + k * n2, nx);
        if (k == 0) {
            for (int i = 0; i < n2; i++) { Pd[i] = -Qt[i]; Sd[i] = -Qi[i]; }
        }
        if (k < N - 1) gj_inverse(R + k * nu * nu, Ri + k * nu * nu, nu);
        else memset(Ri + k * nu * nu, 0, sizeof(double) * nu * nu);
    }
    memset(So, 0, sizeof(double) * n2); /* block row 0 has no left block */
    memset(Po, 0, sizeof(double) * n2);
    for (int i = 0; i < nx; i++) {
        double s = c[i];
        for (int j = 0; j < nx; j++) s -= Qi[i + j * nx] * q[j];
        gamma[i] = s;
    }
    for (int k = 0; k < N - 1; k++) {
        const double* Ak = A + k * n2;
        const double* Bk = Bm + k * nx * nu;
        double phi[MAXX * MAXX], BRi[MAXX * MAXQ], th[MAXX * MAXX], thr[MAXX * MAXX], thi[MAXX * MAXX];
        mm(Ak, Qi + k * n2, phi, nx, nx, nx);
        mm(Bk, Ri + k * nu * nu, BRi, nx, nu, nu);
        for (int j = 0; j < nx; j++)
            for (int i = 0; i < nx; i++) {
                double s = Qi[(k + 1) * n2 + i + j * nx];
                for (int l = 0; l < nx; l++) s += phi[i + l * nx] * Ak[j + l * nx];
                for (int l = 0; l < nu; l++) s += BRi[i + l * nx] * Bk[j + l * nx];
                th[i + j * nx] = s;
            }
        double* g = gamma + (k + 1) * nx;
        for (int i = 0; i < nx; i++) {
            double s = c[(k + 1) * nx + i];
            for (int j = 0; j < nx; j++) s += -Qi[(k + 1) * n2 + i + j * nx] * q[(k + 1) * nx + j] + phi[i + j * nx] * q[k * nx + j];
            for (int j = 0; j < nu; j++) s += BRi[i + j * nx] * r[k * nu + j];
            g[i] = s;
        }
        for (int i = 0; i < n2; i++) {
            Sd[(k + 1) * n2 + i] = -th[i];
            So[(k + 1) * n2 + i] = phi[i];
            thr[i] = th[i];
        }
        for (int i = 0; i < nq; i++) thr[i * (nx + 1)] += rho;
        gj_inverse(thr, thi, nx);
        for (int i = 0; i < n2; i++) Pd[(k + 1) * n2 + i] = -thi[i];
    }
    for (int k = 0; k < N - 1; k++) {
        double t[MAXX * MAXX], t2[MAXX * MAXX];
        mm(So + (k + 1) * n2, Pd + k * n2, t, nx, nx, nx);
        mm(Pd + (k + 1) * n2, t, t2, nx, nx, nx);
        for (int i = 0; i < n2; i++) Po[(k + 1) * n2 + i] = -t2[i];
    }
}

/* y = M x for the symmetric block-tridiagonal (diag blocks D[k], sub-diagonal blocks L[k] = M_{k,k-1}) */
static void btd_mv(int nx, int N, const double* D, const double* L, const double* x, double* y) {
    const int n2 = nx * nx;
    for (int k = 0; k < N; k++) {
        for (int i = 0; i < nx; i++) {
            double s = 0;
            for (int j = 0; j < nx; j++) s += D[k * n2 + i + j * nx] * x[k * nx + j];
            if (k > 0)
                for (int j = 0; j < nx; j++) s += L[k * n2 + i + j * nx] * x[(k - 1) * nx + j];
            if (k < N - 1)
                for (int j = 0; j < nx; j++) s += L[(k + 1) * n2 + j + i * nx] * x[(k + 1) * nx + j];
            y[k * nx + i] = s;
        }
    }
}

static double dotv(const double* a, const double* b, int n) {
    double s = 0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* solvePCGBatchedKernel (pcg.cuh:13-148): warm-started PCG on S lambda = gamma with
 * |rho| < 1e-6 early exit and |rho_new| < 1e-6 + eps |rho_0| stop. Returns the iteration count. */
int orc_pcg(int nx, int N, const double* Sd, const double* So, const double* Pd, const double* Po,
            const double* gamma, double* lambda, double eps, int max_iters) {
    const int n = nx * N;
    double* w = (double*)malloc(sizeof(double) * 4 * n);
    double *r = w, *z = w + n, *p = w + 2 * n, *Ap = w + 3 * n;
    btd_mv(nx, N, Sd, So, lambda, r);
    for (int i = 0; i < n; i++) r[i] = gamma[i] - r[i];
    btd_mv(nx, N, Pd, Po, r, z);
    memcpy(p, z, sizeof(double) * n);
    double rho = dotv(r, z, n);
    int it = 0;
    if (fabs(rho) < PCG_ABS_TOL) {
        free(w);
        return 0;
    }
    double rho0 = fabs(rho);
    for (int i = 0; i < max_iters; i++) {
        it++;
        btd_mv(nx, N, Sd, So, p, Ap);
        double alpha = rho / dotv(p, Ap, n);
        for (int j = 0; j < n; j++) {
            lambda[j] += alpha * p[j];
            r[j] -= alpha * Ap[j];
        }
        btd_mv(nx, N, Pd, Po, r, z);
        double rn = dotv(r, z, n);
        if (fabs(rn) < PCG_ABS_TOL + eps * rho0) break;
        double beta = rn / rho;
        rho = rn;
        for (int j = 0; j < n; j++) p[j] = z[j] + beta * p[j];
    }
    free(w);
    return it;
}

/* computeDzBatchedKernel (schur_linsys.cuh:315-431): dz and the stationarity residuals
 * qres_k = q_k - lambda_k + A_k^T lambda_{k+1}, rres_k = r_k + B_k^T lambda_{k+1} (rres_{N-1} = 0). */
void orc_dz(int nx, int nu, int N, const double* Qi, const double* Ri, const double* A, const double* Bm,
            const double* q, const double* r, const double* lambda, double* dz, double* qres, double* rres) {
    const int n2 = nx * nx, st = nx + nu;
    for (int k = 0; k < N; k++) {
        double v[MAXX];
        for (int i = 0; i < nx; i++) {
            double s = q[k * nx + i] - lambda[k * nx + i];
            if (k < N - 1)
                for (int j = 0; j < nx; j++) s += A[k * n2 + j + i * nx] * lambda[(k + 1) * nx + j];
            v[i] = s;
            qres[k * nx + i] = s;
        }
        for (int i = 0; i < nx; i++) {
            double s = 0;
            for (int j = 0; j < nx; j++) s += Qi[k * n2 + i + j * nx] * v[j];
            dz[k * st + i] = -s;
        }
        if (k < N - 1) {
            double w[MAXQ];
            for (int i = 0; i < nu; i++) {
                double s = r[k * nu + i];
                for (int j = 0; j < nx; j++) s += Bm[k * nx * nu + j + i * nx] * lambda[(k + 1) * nx + j];
                w[i] = s;
                rres[k * nu + i] = s;
            }
            for (int i = 0; i < nu; i++) {
                double s = 0;
                for (int j = 0; j < nu; j++) s += Ri[k * nu * nu + i + j * nu] * w[j];
                dz[k * st + nx + i] = -s;
            }
        } else {
            for (int i = 0; i < nu; i++) rres[k * nu + i] = 0.0;
        }
    }
}

/* computeMeritBatchedKernel (merit.cuh:16-92): phi = sum_k l_k + mu (sum_k |c_{k+1}|_1 + |x_0 - x_s|_1)
 * at z + alpha dz (knot order summation). */
double orc_merit(const orc_cfg* cfg, const double* xu, const double* dz, double alpha, const double* xs,
                 const double* ref, const double* fext, const double* Qd, const double* Rd, const double* xref,
                 double mu) {
    const plant_t* P = get_plant(cfg->plant);
    const int nq = P->nq, nx = 2 * nq, nu = nq, N = cfg->N, st = nx + nu;
    double Qw[MAXX], Rw[MAXQ];
    default_weights(cfg, nq, Qd, Rd, Qw, Rw);
    double total = 0.0;
    for (int k = 0; k < N; k++) {
        double z[3 * MAXQ + MAXX];
        int cnt = (k < N - 1) ? st + nx : nx;
        for (int i = 0; i < cnt; i++) z[i] = xu[k * st + i] + alpha * dz[k * st + i];
        double cost = stage_cost(cfg, P, k, z, z + nx, ref + 6 * k, Qw, Rw, xref ? xref + k * nx : NULL);
        double cons = 0.0;
        if (k < N - 1) {
            double qdd[MAXQ];
            fd(P, z, z + nq, z + nx, fext, qdd);
            const double* xn = z + st;
            for (int i = 0; i < nq; i++) {
                cons += fabs(xn[i] - (z[i] + cfg->dt * z[nq + i] + 0.5 * qdd[i] * cfg->dt * cfg->dt));
                cons += fabs(xn[nq + i] - (z[nq + i] + cfg->dt * qdd[i]));
            }
        } else {
            for (int i = 0; i < nx; i++) cons += fabs(xu[i] + alpha * dz[i] - xs[i]);
        }
        total += cost + mu * cons;
    }
    return total;
}

/* ------------------------------------------------------------------ batched SQP */
typedef struct {
    double *A, *Bm, *c, *Q, *R, *q, *r, *Qi, *Ri, *Sd, *So, *Pd, *Po, *gamma, *dz, *qres, *rres;
} work_t;

static void work_alloc(work_t* W, int nx, int nu, int N) {
    size_t n2 = (size_t)nx * nx * N, nb = (size_t)nx * nu * N, nuu = (size_t)nu * nu * N;
    W->A = malloc(sizeof(double) * n2); W->Bm = malloc(sizeof(double) * nb);
    W->c = malloc(sizeof(double) * nx * N); W->Q = malloc(sizeof(double) * n2);
    W->R = malloc(sizeof(double) * nuu); W->q = malloc(sizeof(double) * nx * N);
    W->r = malloc(sizeof(double) * nu * N); W->Qi = malloc(sizeof(double) * n2);
    W->Ri = malloc(sizeof(double) * nuu); W->Sd = malloc(sizeof(double) * n2);
    W->So = malloc(sizeof(double) * n2); W->Pd = malloc(sizeof(double) * n2);
    W->Po = malloc(sizeof(double) * n2); W->gamma = malloc(sizeof(double) * nx * N);
    W->dz = malloc(sizeof(double) * (nx + nu) * N); W->qres = malloc(sizeof(double) * nx * N);
    W->rres = malloc(sizeof(double) * nu * N);
}

static void work_free(work_t* W) {
    free(W->A); free(W->Bm); free(W->c); free(W->Q); free(W->R); free(W->q); free(W->r); free(W->Qi);
    free(W->Ri); free(W->Sd); free(W->So); free(W->Pd); free(W->Po); free(W->gamma); free(W->dz);
    free(W->qres); free(W->rres);
}

static double absmax(const double* v, int n) {
    double m = 0;
    for (int i = 0; i < n; i++) m = fmax(m, fabs(v[i]));
    return m;
}

/* BSQP::solve (bsqp.cuh:103-197) over B independent problems. Batch coupling is only the
 * break test (#converged >= B * solve_ratio, bsqp.cuh:165) and, under the reference exit rule,
 * the batch-uniform iteration count (quirk Q10). lambda (B x N*nx, unpadded), rho, drho persist. */
int orc_solve(const orc_cfg* cfg, int B, double* xu, const double* xs, const double* ref, const double* fext,
              const double* Qd, const double* Rd, const double* xref, double* lambda, double* rho, double* drho,
              const double* drho_init, const double* mu, const double* pcg_tol, orc_stats* st, int nthreads) {
    const plant_t* P = get_plant(cfg->plant);
    const int nq = P->nq, nx = 2 * nq, nu = nq, N = cfg->N;
    const int traj = (nx + nu) * N - nu;
    const int nxN = nx * N;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int* conv = (int*)calloc(B, sizeof(int));
    int* active = (int*)malloc(sizeof(int) * B); /* KKT rule: problems still iterating */
    double* phi_cur = (double*)malloc(sizeof(double) * B);
    double** dzs = (double**)malloc(sizeof(double*) * B);
    for (int b = 0; b < B; b++) {
        dzs[b] = (double*)calloc(traj + nu, sizeof(double));
        active[b] = 1;
        st->sqp_iters[b] = 0;
    }
#define PB(arr, w) ((arr) ? (arr) + (size_t)b * (w) : NULL)
#pragma omp parallel for schedule(static)
    for (int b = 0; b < B; b++) {
        phi_cur[b] = orc_merit(cfg, xu + (size_t)b * traj, dzs[b], 0.0, xs + b * nx, ref + (size_t)b * 6 * N,
                               PB(fext, 6), PB(Qd, nx), PB(Rd, nu), PB(xref, nxN), mu[b]);
        st->initial_merit[b] = phi_cur[b];
    }
    st->n_iters = 0;
    st->n_ls_iters = 0;
    for (int it = 0; it < cfg->max_sqp_iters; it++) {
        st->n_iters = it + 1;
#pragma omp parallel
        {
            work_t W;
            work_alloc(&W, nx, nu, N);
#pragma omp for schedule(static)
            for (int b = 0; b < B; b++) {
                double* z = xu + (size_t)b * traj;
                int skip = (cfg->exit_rule == ORC_EXIT_PCG0) ? conv[b] : !active[b];
                if (cfg->exit_rule == ORC_EXIT_KKT && skip) {
                    st->pcg_iters[it * B + b] = 0;
                    st->kkt_res[it * B + b] = 0.0;
                    continue;
                }
                orc_setup_kkt(cfg, z, xs + b * nx, ref + (size_t)b * 6 * N, PB(fext, 6), PB(Qd, nx), PB(Rd, nu),
                              PB(xref, nxN), W.A, W.Bm, W.c, W.Q, W.R, W.q, W.r);
                orc_schur(cfg, rho[b], W.A, W.Bm, W.Q, W.R, W.q, W.r, W.c, W.Qi, W.Ri, W.Sd, W.So, W.Pd, W.Po,
                          W.gamma);
                int pit = 0;
                if (!skip) /* quirk Q9: converged problems skip the PCG (pcg.cuh:30-33) */
                    pit = orc_pcg(nx, N, W.Sd, W.So, W.Pd, W.Po, W.gamma, lambda + (size_t)b * nxN, pcg_tol[b],
                                  cfg->max_pcg_iters);
                st->pcg_iters[it * B + b] = pit;
                orc_dz(nx, nu, N, W.Qi, W.Ri, W.A, W.Bm, W.q, W.r, lambda + (size_t)b * nxN, W.dz, W.qres, W.rres);
                memcpy(dzs[b], W.dz, sizeof(double) * traj);
                double kr = fmax(fmax(absmax(W.qres, nxN), absmax(W.rres, nu * N)), absmax(W.c, nxN));
                st->kkt_res[it * B + b] = kr;
            }
            work_free(&W);
        }
        /* host convergence bookkeeping (bsqp.cuh:142-165) */
        int solved = 0;
        for (int b = 0; b < B; b++) {
            if (cfg->exit_rule == ORC_EXIT_PCG0) {
                if (st->pcg_iters[it * B + b] == 0) { conv[b] = 1; }
                st->sqp_iters[b] += 1; /* quirk Q10: +1 for every problem */
                solved += conv[b];
            } else {
                if (active[b]) {
                    st->sqp_iters[b] += 1;
                    if (st->kkt_res[it * B + b] <= cfg->kkt_tol) { conv[b] = 1; active[b] = 0; }
                }
                solved += conv[b];
            }
        }
        if ((double)solved >= (double)B * cfg->solve_ratio) break;
        st->n_ls_iters = it + 1;
        /* merit at 8 step lengths + line search (merit.cuh, line_search.cuh:12-98) */
#pragma omp parallel for schedule(static)
        for (int b = 0; b < B; b++) {
            double* z = xu + (size_t)b * traj;
            if (cfg->exit_rule == ORC_EXIT_KKT && !active[b]) {
                st->ls_min_merit[it * B + b] = phi_cur[b];
                st->ls_step[it * B + b] = 0.0;
                continue;
            }
            double best = 1e38;
            int jbest = 0;
            for (int j = 0; j < NUM_ALPHAS; j++) {
                double a = 1.0 / (double)(1 << j);
                double m = orc_merit(cfg, z, dzs[b], a, xs + b * nx, ref + (size_t)b * 6 * N, PB(fext, 6),
                                     PB(Qd, nx), PB(Rd, nu), PB(xref, nxN), mu[b]);
                if (m < best) { best = m; jbest = j; }
            }
            int ok = best < phi_cur[b];
            if (cfg->adapt_rho) {
                double mult = ok ? fmin(drho[b] / RHO_FACTOR, 1.0 / RHO_FACTOR) : fmax(drho[b] * RHO_FACTOR, RHO_FACTOR);
                drho[b] = mult;
                rho[b] = fmin(fmax(rho[b] * mult, RHO_MIN), RHO_MAX);
            }
            if (!ok) {
                if (rho[b] > RHO_MAX) rho[b] = RHO_INIT;
                st->ls_step[it * B + b] = -1.0;
            } else {
                double a = 1.0 / (double)(1 << jbest);
                phi_cur[b] = best;
                st->ls_step[it * B + b] = a;
                for (int i = 0; i < traj; i++) z[i] += a * dzs[b][i];
            }
            st->ls_min_merit[it * B + b] = phi_cur[b];
        }
    }
#pragma omp parallel for schedule(static)
    for (int b = 0; b < B; b++) {
        memset(dzs[b], 0, sizeof(double) * traj);
        st->final_merit[b] = orc_merit(cfg, xu + (size_t)b * traj, dzs[b], 0.0, xs + b * nx,
                                       ref + (size_t)b * 6 * N, PB(fext, 6), PB(Qd, nx), PB(Rd, nu), PB(xref, nxN),
                                       mu[b]);
        st->kkt_converged[b] = conv[b];
        drho[b] = drho_init[b];
    }
#undef PB
    for (int b = 0; b < B; b++) free(dzs[b]);
    free(dzs);
    free(phi_cur);
    free(active);
    free(conv);
    return 0;
}

/* simForwardBatchedKernel (sim.cuh:14-48): one trapezoidal step of (x, u) under B wrenches */
void orc_sim_forward(int plant, int B, const double* x, const double* u, const double* fext, double dt,
                     double* xout) {
    const plant_t* P = get_plant(plant);
    const int nq = P->nq, nx = 2 * nq;
    for (int b = 0; b < B; b++) {
        double qdd[MAXQ];
        fd(P, x, x + nq, u, fext ? fext + 6 * b : NULL, qdd);
        for (int i = 0; i < nq; i++) {
            xout[b * nx + nq + i] = x[nq + i] + dt * qdd[i];
            xout[b * nx + i] = x[i] + dt * x[nq + i] + 0.5 * qdd[i] * dt * dt;
        }
    }
}
