/* CPU restatement of Tahoe's continuum-solid Newton step -- TEST INFRASTRUCTURE.
 *
 * This library is the checker for the HIP path: only tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it.  It is never part of the
 * shipped path.  Every routine cites the reference file:line it restates
 * (paths relative to the Tahoe source tree).
 *
 * Conventions (all 0-based unless stated):
 *   X, u      [nnd][3] node-major coordinates / displacements
 *   conn      [nel][8] hexahedron connectivity (HexahedronT vertex order)
 *   eqnos     [nnd][3] 1-based equation numbers, <= 0 for prescribed dofs
 *   bindx/val Aztec MSR, upper triangle only (MSRMatrixT.h:68-71):
 *             bindx[0] = n+1, bindx[i+1]-bindx[i] = #off-diagonals of row i,
 *             bindx[n+1..] 0-based column numbers (> row, ascending),
 *             val[0..n-1] diagonal, val[n] unused, val[k] for k > n
 *             the off-diagonal matching bindx[k].
 *   element matrices are 24x24 row-major, dof index 3*a+i.
 *   Status codes follow ExceptionT::CodeT (toolbox/src/common/ExceptionT.h:17-30):
 *             0 ok, 1 general fail, 7 bad Jacobian determinant.
 */
#ifndef TAHOE_ORACLE_H
#define TAHOE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_TL = 0, OR_Q1P0 = 1, OR_TL_STVENANT = 2, OR_Q1P0_STVENANT = 3 };

/* ---- material (SimoIso3D) ---- */
void or_mu_kappa(double E, double nu, double* mu, double* kappa);
int or_simo_cauchy(double mu, double kappa, const double F[9], double sig[6]);
int or_simo_moduli(double mu, double kappa, const double F[9], double c[36]);
double or_simo_energy(double mu, double kappa, const double F[9]);
/* ---- material (large_strain_StVenant) ---- */
int or_stvenant_cauchy(double mu, double kappa, const double F[9], double sig[6]);
int or_stvenant_moduli(double mu, double kappa, const double F[9], double c[36]);
double or_stvenant_energy(double mu, double kappa, const double F[9]);

/* ---- parent domain ---- */
void or_hex8_parent(double Na[8][8], double DNa[8][3][8], double w[8]);

/* ---- element kernels ---- */
int or_tl_element(const double Xe[24], const double ue[24], double mu, double kappa,
                  double Re[24], double Ke[576]);
int or_q1p0_element(const double Xe[24], const double ue[24], double mu, double kappa,
                    double Re[24], double Ke[576]);
/* the same element with the large_strain_StVenant material (reference deck Q1P0.3D.a) */
int or_q1p0_stvenant_element(const double Xe[24], const double ue[24], double mu, double kappa, double Re[24],
                             double Ke[576]);
int or_tl_stvenant_element(const double Xe[24], const double ue[24], double mu, double kappa,
                           double Re[24], double Ke[576]);

/* ---- equation system ---- */
int or_equation_numbers(int nnd, int nkbc, const int* kbc_nodes, const int* kbc_dofs, int* eqnos);
long or_msr_pattern(int n_eq, int nnd, int nel, const int* conn, const int* eqnos,
                    int* bindx, long cap);

/* ---- assembly ---- */
int or_residual(int type, int nnd, int nel, const double* X, const int* conn, const int* mat,
                const double* mu, const double* kappa, const double* u, const int* eqnos,
                int n_eq, double* R);
int or_stiffness(int type, int nnd, int nel, const double* X, const int* conn, const int* mat,
                 const double* mu, const double* kappa, const double* u, const int* eqnos,
                 int n_eq, const int* bindx, double* val);

/* Element colouring (no two elements of a colour share a node) and the
 * residual/stiffness passes over it on nthreads OpenMP threads: the
 * multi-threaded CPU baseline.  color_ptr has room for 65 entries. */
int or_color_elements(int nnd, int nel, const int* conn, int* color_ptr, int* color_elems);
int or_element_passes_parallel(int type, int nel, const double* X, const int* conn, const int* mat,
                               const double* mu, const double* kappa, const double* u, const int* eqnos,
                               int n_eq, const int* bindx, double* R, double* val, int ncolors,
                               const int* color_ptr, const int* color_elems, int nthreads);

/* ---- linear algebra ---- */
void or_msr_multx(int n, const int* bindx, const double* val, const double* x, double* b);
/* restarted GMRES(m), left Jacobi (NLSolver_NK::GMRES); maxit <= 0: 3 n */
int or_gmres_dense(int n, const double* A, const double* b, double* x, int m, double tol, int maxit, int* steps);
int or_gmres_msr(int n, const int* bindx, const double* val, const double* b, double* x, int m, double tol,
                 int maxit, int* steps);
int or_pcg_jacobi(int n, const int* bindx, const double* val, const double* b, double* x,
                  double rtol, int maxit, int* iters, double* relres);

/* ---- Newton time loop ---- */
typedef struct {
    double rel_tol, abs_tol, div_tol; /* NLSolver rel_tolerance, abs_tolerance, divergence_tolerance */
    int max_iterations;
    double lin_rtol;                  /* Jacobi-PCG relative residual */
    int lin_maxit;
} or_newton_opts;

int or_static_solve(int type, int nnd, int nel, const double* X, const int* conn, const int* mat,
                    const double* mu, const double* kappa, int nkbc, const int* kbc_nodes,
                    const int* kbc_dofs, const double* kbc_values, int nforce, const int* f_nodes,
                    const int* f_dofs, const double* f_values, int num_steps,
                    const or_newton_opts* opts, double* u /* in: start, out: final */,
                    double* step_u /* [num_steps][nnd][3] or NULL */, int* step_iters,
                    int* lin_iters_total);

#ifdef __cplusplus
}
#endif
#endif
