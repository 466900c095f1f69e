// Launch wrappers of the gfx950 kernels (kernels.hip).  All fields carry 6 load cases.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace homo3d {

// Fine-level kernels take u as FP32 planes ([6][3][n^3]) or, with u_half, as packed FP16 load-case
// pairs ([3][3][n^3] words of two halves, cases 2p and 2p + 1; only with n % 8 == 0).
// Element matrix for the matrix-core form of the fine operator, built from the integer Lame tables
// Kl = 72 KE(1, 0), Km = 72 KE(0, 1): either one FP16-exact table T = (lambda / mu) Kl + Km with
// KE = c0 T (nt = 1), or Kl and Km with KE = c0 Kl + c1 Km (nt = 2), laid out as the A operands of
// 16x16x32 tiles (kKeMfmaBytes).  ke_mfma_table returns nt.
constexpr int kKeMfmaBytes = 3 * 64 * 8 * 2;
int ke_mfma_table(const double* kl, const double* km, double lam, double mu, void* table, float* c0, float* c1);
struct FineMfma {
    const void* table;  // device copy of the ke_mfma_table output
    float c0, c1;
    int nt;
};
// n for which packed-FP16 fine operators run in element form on the matrix cores
int fine_apply_mfma_ok(int n);
// out = K u (f == nullptr) or f - K u.  With mf, packed-FP16 u and fine_apply_mfma_ok(n) the operator
// runs in element form on the matrix cores.
void launch_fine_apply(hipStream_t s, int n, const float* ke, const FineMfma* mf, const float* w, const void* u,
                       int u_half, const float* f, float* out);
// One CG step's operator pass with the vector updates fused (z the preconditioned residual, FP32 or
// packed FP16 by z_half; sc the device CG scalars, alpha at sc + 12, beta at sc + 24):
//   first: p = z, q = K p;  else: x += alpha p_old, p = z + beta p_old, q = K z + beta q_old (= K p),
// and the per-load-case FP64 dot p . q -> out6.  partial holds 6 * n^3 / 512 doubles (and at least
// 6 * dot_blocks).
// staged (packed FP16 z, fine_cg_staged_ok(n)): each workgroup stages its z node box in LDS instead of
// gathering the neighbours from memory; the results are the same bit for bit.
int fine_cg_staged_ok(int n);
void launch_fine_apply_cg(hipStream_t s, int n, const float* ke, const FineMfma* mf, const float* w, const void* z,
                          int z_half, int staged, float* p, float* x, const double* sc, int first, float* q,
                          double* partial, int dot_blocks, double* out6);
// zero_start: colour pass of a forward sweep that starts from u = 0 (u need not be cleared)
void launch_fine_gs(hipStream_t s, int n, int colour, float omega, int reverse, int zero_start, const float* ke,
                    const float* w, void* u, int u_half, const float* f);
// Colours cfirst and cfirst + 1 (cfirst even; REVERSE: cfirst and cfirst - 1, cfirst odd) in one launch,
// identical to two launch_fine_gs calls; only when fine_gs_pairs_ok(n).  With partial (reverse sweeps of
// a packed-FP16 u only), each of the (n/2)^3 / 256 workgroups also writes its FP64 sums of f . u_new per
// load case to partial[r * pstride + workgroup].
int fine_gs_pairs_ok(int n);
// rown (zero-start forward passes; ignored otherwise): both colours also store, per vertex, f minus the
// K u terms of the colours swept so far (the vertex itself included); after the sweep,
// launch_fine_residual_later turns these into the residual f - K u.
void launch_fine_gs2(hipStream_t s, int n, int cfirst, float omega, int reverse, int zero_start, const float* ke,
                     const float* w, void* u, int u_half, const float* f, double* partial = nullptr,
                     int pstride = 0, float* rown = nullptr);
// r -= the K u terms of later-colour neighbours, colours 0..6 (one launch): with the own-colour
// residuals of a complete zero-start forward sweep (launch_fine_gs2 rown) in r, leaves r = f - K u.
void launch_fine_residual_later(hipStream_t s, int n, const float* ke, const float* w, const void* u, int u_half,
                                float* r);
void launch_macro_rhs(hipStream_t s, int n, const float* kchi, const float* w, float* f);
void launch_fine_residual64(hipStream_t s, int n, const double* ke, const double* kchi, const float* xc, double emin,
                            const double* u, float* r32, double* partial, double* out6, int do_k);
void launch_rhs_norm64(hipStream_t s, int n, const double* kchi, const float* xc, double emin, double* partial,
                       double* out6);
// The FP64 refinement iterate u of N vertices is nine planes of component pairs: component c = 3 r + i
// (component i of load case r) of vertex v is u[u64_index(c, v, N)].  x, p are FP32 planes [18][N].
__host__ __device__ inline size_t u64_index(int c, size_t v, size_t N) {
    return (size_t)(c >> 1) * 2 * N + 2 * v + (size_t)(c & 1);
}
void launch_round_f32(hipStream_t s, size_t N, const double* u, float* x);
// u (+)= x + alpha p for the 6 load cases (N % 2 == 0; alpha = sc + 12, only when pending): the last
// CG step's x update fused into the FP64 refinement iterate (add) or its first value.
void launch_refine_accumulate(hipStream_t s, size_t N, const double* sc, int pending, const float* x, const float* p,
                              double* u, int add);
void launch_widen_f64(hipStream_t s, size_t N, const float* x, double* u);
// zero load case r (all its components, every vertex) of the FP64 iterate
void launch_zero_case64(hipStream_t s, size_t N, int r, double* u);
void launch_zero_pin(hipStream_t s, size_t N, void* u, int u_half);
// FP32 field [6][3][N] <-> packed FP16 field [3][3][N] words (round to nearest)
void launch_pack_half(hipStream_t s, size_t N, const float* u, void* z);
void launch_unpack_half(hipStream_t s, size_t N, const void* z, float* u);
void launch_zero_pin64(hipStream_t s, size_t N, double* u);
// The coarsest level's dense FP64 inverse on the GPU: assembled from its FP32 stencil (lexicographic
// n_c^3 level, m = 3 n_c^3), symmetrised and inverted in place in ainv (m x m) by Gauss-Jordan
// elimination on the diagonal, one launch per step; piv is 4 m doubles of pivot rows and columns.
// status (device) = 0, or 1 + the first step whose pivot is not positive (the system is not positive
// definite; ainv is then not an inverse).
void launch_coarse_factor(hipStream_t s, int nc, const float* st, double* ainv, double* piv, int* status);
// Coarse stencils: FP32 [243][n^3] or, with *_half, packed FP16 [122][n^3] words (entries 2w, 2w + 1).
// u FP32 planes, or (u_half, only where stencil_u_half_ok: six load cases per launch) packed FP16 pairs.
int stencil_u_half_ok(int n, int lex);
void launch_stencil_apply(hipStream_t s, int n, int lex, const void* st, int st_half, const void* u, int u_half,
                          const float* f, float* out);
void launch_stencil_gs(hipStream_t s, int n, int colour, float omega, int reverse, int zero_start, const void* st,
                       int st_half, void* u, int u_half, const float* f);
// Colours cfirst and cfirst + 1 (cfirst even; reverse: cfirst and cfirst - 1, cfirst odd) of a coarse
// parity-blocked level in one launch, identical to two launch_stencil_gs calls; only when
// stencil_gs_pairs_ok(n, lex) (levels above 32^3 whose grid rows fit a workgroup).
// rown (zero-start forward passes; ignored otherwise): the own-colour residuals, as launch_fine_gs2
// stores them; launch_stencil_residual_later completes them to f - A u after the sweep.
int stencil_gs_pairs_ok(int n, int lex);
void launch_stencil_gs2(hipStream_t s, int n, int cfirst, float omega, int reverse, int zero_start, const void* st,
                        int st_half, void* u, int u_half, const float* f, float* rown = nullptr);
void launch_stencil_residual_later(hipStream_t s, int n, const void* st, int st_half, const void* u, int u_half,
                                   float* r);
void launch_restrict(hipStream_t s, int nf, int lexc, const float* rf, float* fc);
// uf += P uc.  uc FP32, or packed FP16 (uc_half: parity-blocked coarse level, packed FP16 fine u).
void launch_prolong(hipStream_t s, int nf, int lexc, const void* uc, int uc_half, void* uf, int u_half, int pin);
void launch_galerkin_fine(hipStream_t s, int nf, int lexc, const float* G, const float* ke, const float* w, void* stc,
                          int out_half, float clamp);
void launch_galerkin_coarse(hipStream_t s, int nf, int lexc, const void* stf, int in_half, void* stc, int out_half,
                            float clamp);
void launch_coarse_solve(hipStream_t s, int nc, const double* Ainv, const float* f, float* u);
void launch_dot6(hipStream_t s, size_t M, const float* a, const float* b, double* partial, int nblocks, double* out);
void launch_sum_partials(hipStream_t s, int R, int nblocks, const double* partial, double* out);
void launch_pcg_alpha(hipStream_t s, double* sc);
void launch_pcg_beta(hipStream_t s, double* sc);
// CG updates over 6 load cases of M floats each (M % 4 == 0): r -= alpha q with |r|^2 -> out6;
// x += alpha p and, when do_p, p = z + beta p for the active load cases.
void launch_pcg_update_r(hipStream_t s, size_t M, const double* sc, const float* q, float* r, double* partial,
                         int nblocks, double* out6);
// z FP32 or packed FP16 (z_half); p and x FP32
void launch_pcg_update_px(hipStream_t s, size_t M, const double* sc, const void* z, int z_half, float* p, float* x,
                          int do_p, int nblocks);
// p = z
hipError_t launch_set_p(hipStream_t s, size_t M, const void* z, int z_half, float* p, int nblocks);
// per-load-case r . z -> out6 (z FP32 or packed FP16)
void launch_dot_rz(hipStream_t s, size_t M, const float* r, const void* z, int z_half, double* partial, int nblocks,
                   double* out6);
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
// x = rho^p; drho = p rho^(p - 1) dx (drho may alias dx); FP64 block sums of a (nblocks partials); a = value
void launch_density_pow(hipStream_t s, int n, int pitch, float p, const float* rho, float* x);
void launch_density_pow_grad(hipStream_t s, int n, int pitch, float p, const float* rho, const float* dx, float* drho);
void launch_field_sum(hipStream_t s, int n, int pitch, const float* a, double* partial, int nblocks);
void launch_field_fill(hipStream_t s, int n, int pitch, float value, float* a);
void launch_set_weights(hipStream_t s, int n, int pitch, double emin, const float* x, float* w, float* xc);
void launch_copy_pitched(hipStream_t s, int n, int pitch_in, int pitch_out, const float* in, float* out);

}  // namespace homo3d

namespace homo3d {
constexpr int kMmaMaxM = 4;
constexpr int kMmaNewtonSums = kMmaMaxM + kMmaMaxM * (kMmaMaxM + 1) / 2;
void launch_mma_setup(hipStream_t s, int m, long n, int iter, double move, const double* x, const double* xmin,
                      const double* xmax, const double* xold1, const double* xold2, const double* df0,
                      const double* dfdx, double* low, double* upp, double* alfa, double* beta, double* p0, double* q0,
                      double* P, double* Q, double* part, int nb, double* out);
void launch_mma_resid(hipStream_t s, int m, long n, const double* lam, double epsi, const double* x,
                      const double* xsi, const double* eta, const double* low, const double* upp, const double* alfa,
                      const double* beta, const double* p0, const double* q0, const double* P, const double* Q,
                      double* part, double* partmax, int nb, double* out);
void launch_mma_newton(hipStream_t s, int m, long n, const double* lam, double epsi, const double* x,
                       const double* xsi, const double* eta, const double* low, const double* upp, const double* alfa,
                       const double* beta, const double* p0, const double* q0, const double* P, const double* Q,
                       double* delx, double* diagx, double* part, int nb, double* out);
void launch_mma_dir(hipStream_t s, int m, long n, double epsi, const double* dlam, const double* x, const double* xsi,
                    const double* eta, const double* low, const double* upp, const double* alfa, const double* beta,
                    const double* P, const double* Q, const double* delx, const double* diagx, double* dx,
                    double* dxsi, double* deta, double* partmax, int nb, double* out);
void launch_mma_step(hipStream_t s, long n, double steg, const double* xo, const double* xsio, const double* etao,
                     const double* dx, const double* dxsi, const double* deta, double* x, double* xsi, double* eta);
void launch_mma_init(hipStream_t s, long n, const double* alfa, const double* beta, double* x, double* xsi,
                     double* eta);
void launch_f32_to_f64(hipStream_t s, long n, const float* a, double* b);
void launch_f64_to_f32(hipStream_t s, long n, const double* a, float* b);
// Density filter with Spline4/linear radial weights (wk: normalised (2h+1)^3 box) and its adjoint (wsum: the
// forward pass's in-grid weight sums); sgm projection and its backward.  Pitched lexicographic arrays.
void launch_conv_filter(hipStream_t s, int n, int pitch, int h, int periodic, const float* wk, const float* in,
                        float* out, float* wsum);
void launch_conv_filter_back(hipStream_t s, int n, int pitch, int h, int periodic, const float* wk, const float* dout,
                             const float* wsum, float* din);
void launch_sgm(hipStream_t s, int n, int pitch, double k, double c, const float* x, float* y);
void launch_sgm_back(hipStream_t s, int n, int pitch, double k, const float* y, const float* gy, float* gx);

// Bridges to the learned warm-start network (g | n; network tensors bf16 channels-last, see net.hip):
// input channel 0 = mean of x^(1/3) over each coarse cell; target = scale (u - mean) / n sampled at the
// coarse vertices, the mean weighted by the input's rho^3 (mean18: 18 doubles of device scratch); warm start = periodic trilinear upsampling of
// the prediction P ([g^3][ldp]) minus its value at vertex 0, times n / scale, into U64 and x.
void launch_net_input(hipStream_t s, int n, int g, const float* x, uint16_t* in);
void launch_net_target(hipStream_t s, int n, int g, const double* u64, const uint16_t* in, double scale, double* mean18,
                       uint16_t* tg);
void launch_net_warm_start(hipStream_t s, int n, int g, const float* P, int ldp, double scale, double* u64, float* x);

}  // namespace homo3d
