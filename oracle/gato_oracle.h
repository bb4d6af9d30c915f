/* CPU restatement of GATO's batched SQP iteration. TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (gato_amd/) never links or calls it.
 */
#ifndef GATO_ORACLE_H
#define GATO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_PLANT_IIWA14 0
#define ORC_PLANT_INDY7 1
/* indy7 as the reference's MPC examples simulate it (pinocchio on the URDF, inertial origins kept): plant dynamics
 * only (orc_rnea, orc_mass, orc_fd, orc_fk, orc_sim_forward); the solver's indy7 is ORC_PLANT_INDY7 */
#define ORC_PLANT_INDY7_URDF 2

#define ORC_EXIT_PCG0 0 /* reference rule: converged <=> PCG took 0 iterations (bsqp.cuh:149-163) */
#define ORC_EXIT_KKT 1  /* converged <=> max(|stationarity|_inf, |defect|_inf) <= kkt_tol, per problem */

typedef struct orc_cfg {
    int plant;
    int N;
    double dt;
    int max_sqp_iters;
    double kkt_tol;
    int max_pcg_iters;
    double solve_ratio;
    double q_cost, qd_cost, u_cost, N_cost, q_lim_cost, vel_lim_cost, ctrl_lim_cost;
    int adapt_rho;
    int exit_rule;     /* ORC_EXIT_* */
    int terminal_mode; /* 0: reference (terminal block from x_{N-2}, q_cost); 1: consistent */
    int hess_mode;     /* 0: reference rank-1 EE Hessian w g g^T (quirk Q5); 1: Gauss-Newton w J^T J */
} orc_cfg;

typedef struct orc_stats {
    int* sqp_iters;        /* [B] */
    int* kkt_converged;    /* [B] */
    double* initial_merit; /* [B] */
    double* final_merit;   /* [B] */
    int* pcg_iters;        /* [max_sqp_iters][B] */
    double* ls_min_merit;  /* [max_sqp_iters][B] */
    double* ls_step;       /* [max_sqp_iters][B] */
    double* kkt_res;       /* [max_sqp_iters][B] */
    int n_iters;           /* loop iterations run */
    int n_ls_iters;        /* iterations that reached the line search */
} orc_stats;

int orc_nq(int plant);
void orc_limits(int plant, double* qlo, double* qhi, double* vlo, double* vhi, double* ulo, double* uhi);

/* dynamics (one configuration) */
void orc_fk(int plant, const double* q, double* p, double* J);
void orc_frames(int plant, const double* q, double* R, double* o);
void orc_rnea(int plant, const double* q, const double* qd, const double* qdd, const double* fext, int gravity,
              double* tau);
void orc_mass(int plant, const double* q, double* M);
void orc_fd(int plant, const double* q, const double* qd, const double* u, const double* fext, double* qdd);
void orc_fd_grad(int plant, const double* q, const double* qd, const double* u, const double* fext, double* qdd,
                 double* dqdd, double* Minv);

/* one problem, one phase at a time */
void orc_setup_kkt(const orc_cfg* cfg, const double* xu, const double* xs, const double* ref, const double* fext,
                   const double* Qd, const double* Rd, const double* xref, double* A, double* Bm, double* c, double* Q,
                   double* R, double* q, double* r);
void orc_schur(const orc_cfg* cfg, double rho, const double* A, const double* Bm, const double* Q, const double* R,
               const double* q, const double* r, const double* c, double* Qi, double* Ri, double* Sd, double* So,
               double* Pd, double* Po, double* gamma);
int orc_pcg(int nx, int N, const double* Sd, const double* So, const double* Pd, const double* Po,
            const double* gamma, double* lambda, double eps, int max_iters);
void orc_dz(int nx, int nu, int N, const double* Qi, const double* Ri, const double* A, const double* Bm,
            const double* q, const double* r, const double* lambda, double* dz, double* qres, double* rres);
double orc_merit(const orc_cfg* cfg, const double* xu, const double* dz, double alpha, const double* xs,
                 const double* ref, const double* fext, const double* Qd, const double* Rd, const double* xref,
                 double mu);

/* full batched solve (bsqp.cuh:103-197) */
int orc_solve(const orc_cfg* cfg, int B, double* xu, const double* xs, const double* ref, const double* fext,
              const double* Qd, const double* Rd, const double* xref, double* lambda, double* rho, double* drho,
              const double* drho_init, const double* mu, const double* pcg_tol, orc_stats* st, int nthreads);

void orc_sim_forward(int plant, int B, const double* x, const double* u, const double* fext, double dt,
                     double* xout);

#ifdef __cplusplus
}
#endif
#endif
