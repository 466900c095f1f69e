// FP64 CPU oracle for the periodic homogenization solve path (test infrastructure only).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
// library; the product library never does.
//
// Parity pins (what fixes each function without running the reference; the tests that
// check them live in tests/test_oracle_*.py):
//   orc_ke / orc_ke_lame      KE(lam=1,mu=0)*72 and KE(0,1)*72 equal the reference's integer
//                             tables kelam72/kemu72 (tests/golden/ke_tables.json); symmetric;
//                             null space = 6 rigid modes.
//   orc_level_sizes           closed-form sizes (32->8,4; 48->6; 80->5; 256->7 levels).
//   orc_macro_rhs             uniform density gives f = 0; f = K chi0 on the unpinned operator.
//   orc_apply_K / residual    dense assembly brute force at n<=6; symmetry; constants in kernel.
//   orc_gs_fine               equals a dense block Gauss-Seidel on the same colour order (n=4).
//   orc_restrict / prolong    R = P^T exactly; P exact on trilinear fields.
//   orc_galerkin_*            equals dense P^T A P at n<=8 (pin and clamp included).
//   orc_mg_solve_*            dense FP64 direct solve at n=8; solid -> C0; laminate -> analytic.
//   orc_elastic_matrix        solid -> C0; uniform -> (Emin+x) C0; laminate -> analytic; cubic
//                             symmetry under reflect6; SPD.
//   orc_sensitivity           Euler homogeneity sum_e w_e dC/dx_e = C; central differences.
//   orc_symmetrize_*          projection: idempotent, invariant under every group element.
//   orc_oc_filter/update      uniform-field invariants; volume within 1e-4; box/move limits;
//                             the power law rho'/rho = (-s/g)^damp; the B = 0.01 branch for
//                             positive sensitivities; the hand-computed filter impulse table.
//   orc_mma_update            KKT of the convex toy min sum x s.t. sum 1/x <= c; closed-form
//                             asymptotes (0.5 span, then 1.2 / 0.7 / 1.0 with their clamps);
//                             the subproblem optimum against an independent SLSQP solve.
//   orc_philox4x32_10         Random123 known-answer vectors.
//   orc_randc_coeffs/field    coefficients: standard-normal moments (mean, variance, fourth
//                             moment 3), seed determinism; field: a single-coefficient design
//                             against the closed form sigmoid(3 cos sin cos / 8) scaled to the
//                             volume by an independent root finder.
//   orc_randc_density         the composition of the two above plus symmetry projection.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void orc_ke(double E, double nu, double* ke);
void orc_ke_lame(double lam, double mu, double* ke);
int orc_level_sizes(int n, int* sizes, int max_levels);
void orc_macro_rhs(const int n[3], const double* ke, const double* w, int s, double* f);
void orc_apply_K(const int n[3], const double* ke, const double* w, const double* u, double* y, int pin);
void orc_residual(const int n[3], const double* ke, const double* w, const double* u, const double* f, double* r);
void orc_gs_fine(const int n[3], const double* ke, const double* w, double* u, const double* f, double omega,
                 int reverse);
void orc_stencil_apply(const int n[3], const double* st, const double* u, double* y);
void orc_stencil_residual(const int n[3], const double* st, const double* u, const double* f, double* r);
void orc_stencil_gs(const int n[3], const double* st, double* u, const double* f, double omega, int reverse);
void orc_restrict(const int nf[3], const double* rf, double* fc);
void orc_prolong_add(const int nf[3], const double* uc, double* uf, int zero_pin);
void orc_galerkin_fine(const int nf[3], const double* ke, const double* w, double* stc, double clamp);
void orc_galerkin_coarse(const int nf[3], const double* stf, double* stc, double clamp);
void orc_dense_from_stencil(const int n[3], const double* st, double* A);
void orc_dense_fine(const int n[3], const double* ke, const double* w, double* A);
int orc_cholesky(int m, double* A);
void orc_chol_solve(int m, const double* L, double* b);
void orc_elastic_matrix(const int n[3], const double* ke, const double* w, const double* chi, double* C);
void orc_sensitivity(const int n[3], const double* ke, const double* chi, const double* dC, double* sens);

void* orc_mg_create(const int n[3], double E, double nu, double emin);
void orc_mg_destroy(void* h);
int orc_mg_levels(void* h, int* sizes);
const double* orc_mg_ke(void* h);
int orc_mg_set_density(void* h, const double* x);
const double* orc_mg_weights(void* h);
void orc_mg_get_stencil(void* h, int level, double* out);
double orc_mg_vcycle(void* h, double* u, const double* f, double omega, int symmetric);
int orc_mg_solve_mg(void* h, double* u, const double* f, double tol, int maxit, double omega, double* relres,
                    int* iters);
int orc_mg_solve_pcg(void* h, double* u, const double* f, double tol, int maxit, double omega, double* relres,
                     int* iters);
int orc_mg_homogenize(void* h, double tol, int maxit, int mode, int warm, double* chi, double* C, int* iters);

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_symmetrize_f32(int n, int sym, const float* in, float* out);
void orc_symmetrize_f64(int n, int sym, const double* in, double* out);
void orc_oc_filter(const int n[3], const double* sens, const double* rho, double r, int periodic, double* out);
int orc_oc_update(long ne, const float* sens, float* rho, double vol, double minRho, double step, double damp);
int orc_mma_update(int m, long n, int iter, double* x, const double* xmin, const double* xmax, double* xold1,
                   double* xold2, const double* df0dx, const double* fval, const double* dfdx, double* low,
                   double* upp, double a0, const double* a, const double* c, const double* d, double move);
void orc_randc_density(int n, uint64_t seed, double vol, int sym, float* out);
void orc_randc_coeffs(uint64_t seed, double* alpha);
void orc_randc_field(int n, const double* alpha, double vol, int sym, float* out);

#ifdef __cplusplus
}
#endif
