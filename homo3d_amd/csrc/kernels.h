// Launch wrappers of the gfx950 kernels (kernels.hip).  All fields carry 6 load cases.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace homo3d {

void launch_fine_apply(hipStream_t s, int n, const float* ke, const float* w, const float* u, const float* f,
                       float* out);
void launch_fine_gs(hipStream_t s, int n, int colour, float omega, int reverse, const float* ke, const float* w,
                    float* u, const float* f);
void launch_macro_rhs(hipStream_t s, int n, const float* kchi, const float* w, float* f);
void launch_fine_residual64(hipStream_t s, int n, const double* ke, const double* kchi, const float* xc, double emin,
                            const double* u, float* r32, double* partial, double* out6, int do_f);
void launch_rhs_norm64(hipStream_t s, int n, const double* kchi, const float* xc, double emin, double* partial,
                       double* out6);
void launch_add_correction(hipStream_t s, size_t M, const float* d, double* u);
void launch_round_f32(hipStream_t s, size_t M, const double* u, float* x);
void launch_widen_f64(hipStream_t s, size_t M, const float* x, double* u);
void launch_zero_pin(hipStream_t s, size_t N, float* u);
void launch_zero_pin64(hipStream_t s, size_t N, double* u);
void launch_stencil_apply(hipStream_t s, int n, const float* st, const float* u, const float* f, float* out);
void launch_stencil_gs(hipStream_t s, int n, int colour, float omega, int reverse, const float* st, float* u,
                       const float* f);
void launch_restrict(hipStream_t s, int nf, const float* rf, float* fc);
void launch_prolong(hipStream_t s, int nf, const float* uc, float* uf, int pin);
void launch_galerkin_fine(hipStream_t s, int nf, const float* G, const float* ke, const float* w, float* stc,
                          float clamp);
void launch_galerkin_coarse(hipStream_t s, int nf, const float* stf, float* stc, float clamp);
void launch_coarse_solve(hipStream_t s, int nc, const double* Ainv, const float* f, float* u);
void launch_dot6(hipStream_t s, size_t M, const float* a, const float* b, double* partial, int nblocks, double* out);
void launch_sum_partials(hipStream_t s, int R, int nblocks, const double* partial, double* out);
void launch_pcg_alpha(hipStream_t s, double* sc);
void launch_pcg_beta(hipStream_t s, double* sc);
void launch_pcg_update_xr(hipStream_t s, size_t M, const double* sc, const float* p, const float* q, float* x,
                          float* r);
void launch_pcg_update_p(hipStream_t s, size_t M, const double* sc, const float* z, float* p);
void launch_elastic_matrix(hipStream_t s, int n, const float* ke, const float* w, const float* chi, double* partial,
                           int nblocks, double* out21);
void launch_sensitivity(hipStream_t s, int n, int pitch, const float* ke, const float* chi, const float* wdc,
                        float* sens);
void launch_symmetrize(hipStream_t s, int n, int pitch, int ng, const int* gperm, const int* gflip, const float* in,
                       float* out);
void launch_oc_filter(hipStream_t s, int n, int pitch, int h, int periodic, const float* wk, const float* sens,
                      const float* rho, float* out);
void launch_oc_trial(hipStream_t s, int n, int pitch, double g, double min_rho, double step, double damp,
                     const float* sens, const float* rho, float* out, double* partial, int nblocks, double* sum);
void launch_absmax(hipStream_t s, int n, int pitch, const float* a, float* partial, int nblocks);
void launch_set_weights(hipStream_t s, int n, int pitch, double emin, const float* x, float* w, float* xc);
void launch_copy_pitched(hipStream_t s, int n, int pitch_in, int pitch_out, const float* in, float* out);

}  // namespace homo3d
