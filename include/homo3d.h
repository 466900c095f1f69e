/*
 * homo3d.h - C ABI of the MI355X-native periodic homogenization solver.
 *
 * Plain C types only (pointers, sizes, doubles).  Every entry point returns an int status
 * (HOMO3D_OK = 0) and records a message retrievable with homo3d_last_error().
 *
 * Reference interfaces each entry point replaces (lavenklau/homo3d):
 *   homo3d_create            homo::Homogenization::Homogenization(cfg::HomoConfig) / build
 *                            (homogenization/homogenization.h:20-22, homogenization.cpp:40-66)
 *                            and MG::build (homogenization/MG.h:22, MG.cpp:13-44)
 *   homo3d_update_density    Homogenization::update(float* rho, int pitchT)
 *                            (homogenization.h:28, homogenization.cpp:79-85) incl. MG::updateStencils
 *                            (MG.cpp:341-354)
 *   homo3d_solve             the six-load-case loop of Homogenization::elasticMatrix
 *                            (homogenization.cu:471-477): Grid::useFchar / useUchar /
 *                            MG::solveEquation / Grid::setUchar (grid.h:257-262, MG.h:30)
 *   homo3d_elastic_matrix    Homogenization::elasticMatrix(double C[6][6]) (homogenization.h:24,
 *                            homogenization.cu:478-575)
 *   homo3d_sensitivity       Homogenization::Sensitivity(float dC[6][6], float* sens, int pitchT,
 *                            bool lexiOrder) (homogenization.h:38, homogenization.cu:860-948)
 *   homo3d_macro_rhs         Grid::enforce_unit_macro_strain (grid.h:255, grid.cu:1921-1935)
 *   homo3d_residual          Grid::update_residual (grid.h:253, grid.cu:1566-1593)
 *   homo3d_apply_K           Grid::v3_stencilOnLeft (grid.h:355, grid.cu:3686-3857)
 *   homo3d_smooth            Grid::gs_relaxation (grid.h:241, grid.cu:1192-1236)
 *   homo3d_restrict          Grid::restrict_residual (grid.h:249, grid.cu:1257-1275)
 *   homo3d_prolong           Grid::prolongate_correction (grid.h:247, grid.cu:1238-1255)
 *   homo3d_coarse_solve      Grid::solveHostEquation (grid.h:236, grid.cpp:349-374)
 *   homo3d_vcycle            MG::v_cycle (MG.h:26, MG.cpp:46-70); homo3d_vcycle_sweeps with its
 *                            w_SOR / pre / post arguments
 *   homo3d_field_copy        Grid::v3_upload / v3_download (grid.h:339-340)
 *   homo3d_solution_copy     Grid::uchar_h read/write, the stored load-case solutions used as warm start
 *                            (grid.h:142, grid.cpp:341-345, 459-462), here the FP64 refinement iterate
 *   homo3d_get_stencil       Grid::stencil_g read-back (grid.h:126)
 *   homo3d_oc_filter         OCOptimizer::filterSens (optimization/OCOptimizer.h:17-18)
 *   homo3d_oc_update         OCOptimizer::update (optimization/OCOptimizer.h:19-20)
 *   homo3d_symmetrize        symmetrizeField / Tensor::symmetrize (homogenization/Framework.cuh:221-235,
 *                            AutoDiff/TensorExpression.h:425-447)
 *   homo3d_mma_*             homo::MMAOptimizer (optimization/mmaOptimizer.h:12-34) and
 *                            mmasubDevice (optimization/mmaOpt.h:17-19)
 *   homo3d_randc_density     initDensity "randc" (Framework.cu:61-134, Framework.cuh:237-310), replaced
 *                            by a seeded generator
 *
 * Field layout: a displacement/force/residual field of one level holds 6 load cases, each three
 * components of n^3 floats: index ((rhs * 3 + comp) * n^3 + v), v = x + n*(y + n*z).  Densities and
 * sensitivities cross the boundary as the reference's lexicographic row-pitched arrays:
 * element (x,y,z) at x + (y + z*n) * pitch (AutoDiff/TensorExpression.h:137-150).
 * C is a row-major double[36] in Voigt order xx, yy, zz, yz, xz, xy.
 *
 * Arguments named on_device take 1 for device (or managed) memory, 0 for host memory and -1 to let
 * the library inspect the pointer (hipPointerGetAttributes).
 */
#ifndef HOMO3D_H
#define HOMO3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HOMO3D_OK = 0,
    HOMO3D_NOT_CONVERGED = 1,
    HOMO3D_DIVERGED = 2,
    HOMO3D_BAD_ARG = 3,
    HOMO3D_HIP_ERROR = 4,
    HOMO3D_NO_SUPPORT = 5
};

/* HOMO3D_SOLVER_NONE creates a density-only handle (symmetry, filter, OC) without a solver hierarchy. */
enum { HOMO3D_SOLVER_MG = 0, HOMO3D_SOLVER_PCG = 1, HOMO3D_SOLVER_NONE = 2 };
enum { HOMO3D_FIELD_U = 0, HOMO3D_FIELD_F = 1, HOMO3D_FIELD_R = 2 };
enum { HOMO3D_SYM_REFLECT3 = 0, HOMO3D_SYM_REFLECT6 = 1, HOMO3D_SYM_ROTATE3 = 2, HOMO3D_SYM_NONE = 3 };

typedef struct homo3d_handle homo3d_handle;

/* Mirrors cfg::HomoConfig (cmdline.h:39-60) for the fields the solve path reads. */
typedef struct {
    int reso[3];            /* cells per axis; cubic grids, 8 <= n <= 256 and n = c * 2^k with 4 <= c <= 7
                               (coarsest level 4^3..7^3); larger n would overflow the 32-bit byte offsets
                               of the FP64 field gathers and is rejected (HOMO3D_NO_SUPPORT) */
    double youngs_modulus;  /* E (reference default 10) */
    double poisson_ratio;   /* nu (0.3) */
    double emin;            /* stiffness floor added to x_e (reference rhoPenalMin = 1e-9) */
    double fem_rel_thres;   /* default stop tolerance (reference femRelThres, 1e-2) */
    int max_iter;           /* max V-cycles / PCG iterations per solve (200) */
    int solver;             /* HOMO3D_SOLVER_MG (reference stand-alone V-cycles) or HOMO3D_SOLVER_PCG */
    int device;             /* HIP device ordinal */
    double sor_omega;       /* smoother relaxation factor (1.0) */
    double diag_clamp;      /* coarse-stencil diagonal floor (1e-4) */
    int precond_half;       /* HOMO3D_SOLVER_PCG: keep the preconditioner's finest-level iterate in FP16
                               (accumulation stays FP32; needs n % 8 == 0, else FP32 is used).  Default 1. */
    int coarse_half;        /* store the Galerkin stencils of the intermediate levels in FP16, as the reference
                               does (grid_ex.cu:936-1160); the coarsest level stays FP32.  Default 1. */
    int fine_mfma;          /* apply the finest-level operator to FP16 fields in element form on the matrix cores
                               (integer Lame tables, exact in FP16; FP32 sums; needs n % 32 == 0 and
                               72 n^3 < 2^32).  Default 0. */
} homo3d_config;

typedef struct {
    int status;             /* worst status over the six load cases */
    int iters[6];           /* iterations per load case */
    double relres[6];       /* final ||f - K u|| / ||f|| per load case (FP64, unique DOFs) */
    double seconds;         /* wall time of the solve call */
} homo3d_solve_stats;

void homo3d_default_config(homo3d_config* cfg);
int homo3d_create(const homo3d_config* cfg, homo3d_handle** out);
int homo3d_destroy(homo3d_handle* h);
const char* homo3d_last_error(const homo3d_handle* h);
int homo3d_levels(const homo3d_handle* h, int* sizes, int max_levels);
/* Opaque hipStream_t the handle launches on (for event timing by the caller). */
void* homo3d_stream(homo3d_handle* h);

/* Weights and Galerkin operators of a new density.  The coarsest level's FP64 inverse is formed on the
 * GPU (on a second stream of the handle, overlapping the finest-level work of the next solve); a
 * coarsest system that is not positive definite is reported (HOMO3D_DIVERGED) by the first coarse
 * solve after the update. */
int homo3d_update_density(homo3d_handle* h, const float* x, int pitch, int on_device);
int homo3d_solve(homo3d_handle* h, double tol, int warm, homo3d_solve_stats* stats);
int homo3d_elastic_matrix(homo3d_handle* h, double C[36]);
/* Autotuner (no reference counterpart): times the finest-level operator's compiled tile forms on this
 * handle's grid and density (row form: 512-thread workgroups over 128 x 2 x 2 parity tiles; element form
 * on the matrix cores: 32 x 4 x 4 vertex tiles, needs n % 32 == 0) and keeps the faster for every later
 * CG pass and V-cycle residual.  Times of -1 mean the form cannot run at this n. */
typedef struct {
    double row_ms;   /* average r = K z (packed FP16, six load cases) in the row form */
    double mfma_ms;  /* the same in element form on the matrix cores */
    int chosen_mfma; /* 1 when the element form was chosen */
} homo3d_tune_report;
int homo3d_autotune(homo3d_handle* h, homo3d_tune_report* report);
/* CG pass of the refined solver (no reference counterpart): on = 1 (the default) stages each workgroup's
 * node box of the packed-FP16 z in LDS where the grid allows it (parity-blocked, n % 128 == 0) and reads
 * the 27 neighbours from there; on = 0 gathers them from memory.  Both give the same results bit for bit;
 * the switch exists for timing and testing. */
int homo3d_stage_tiles(homo3d_handle* h, int on);
/* FP64 relres ||f - K u|| / ||f|| per load case of the refined solver's current iterate, without solving
 * (Grid::update_residual with the norms of MG::solveEquation, MG.cpp:113-203, in FP64). */
int homo3d_residual_norms(homo3d_handle* h, double relres[6]);
int homo3d_sensitivity(homo3d_handle* h, const float dC[36], float* sens, int pitch, int on_device);
/* Homogenization::Sensitivity(dC, sens, pitchT, lexiOrder) (homogenization.h:38): lexi_order = 1 is
 * homo3d_sensitivity; 0 writes the library's own cell order (parity-blocked: eight (n/2)^3 blocks,
 * b = (x&1) | (y&1)<<1 | (z&1)<<2, lexicographic in (x>>1, y>>1, z>>1) inside), n^3 floats, pitch ignored. */
int homo3d_sensitivity_order(homo3d_handle* h, const float dC[36], float* sens, int pitch, int lexi_order,
                             int on_device);

/* Grid / MG level operations (all six load cases at once). */
int homo3d_field_copy(homo3d_handle* h, int level, int field, float* host, int to_device);
/* FP64 solution of the six load cases ([6][3][n^3] doubles) kept by the refined MG-PCG solver. */
int homo3d_solution_copy(homo3d_handle* h, double* host, int to_device);
int homo3d_macro_rhs(homo3d_handle* h);
int homo3d_residual(homo3d_handle* h, int level);
int homo3d_apply_K(homo3d_handle* h);              /* r = K u on the finest level (pinned) */
int homo3d_smooth(homo3d_handle* h, int level, double omega, int reverse);
/* The same finest-level operations on packed-FP16 storage, the form the PCG preconditioner runs
   (no reference counterpart: the reference smooths FP32 fields).  u16 = fp16(u[0]) with the pin zeroed;
   apply_K16: r[0] = K u16, or f - K u16 when residual; smooth16: one 8-colour sweep on u16 held in
   FP16, then u[0] = u16.  Need n % 8 == 0 (HOMO3D_NO_SUPPORT otherwise). */
int homo3d_apply_K16(homo3d_handle* h, int residual);
int homo3d_smooth16(homo3d_handle* h, double omega, int reverse);
/* The preconditioner's level-0 pre-smoothing: one forward sweep from u16 = 0 (whatever u[0] holds) with
 * f[0], then u[0] = u16 and r[0] = f - K u16, the residual formed from the own-colour residuals the
 * sweep leaves behind plus the later-colour terms.  Needs n % 8 == 0 and fused colour pairs
 * ((n/2) | 256, 256 | (n/2)^3); HOMO3D_NO_SUPPORT otherwise. */
int homo3d_presmooth16(homo3d_handle* h, double omega);
/* The same pre-smoothing on a coarse level (FP32 u): one forward sweep from u[level] = 0 with
 * f[level], and r[level] = f - A u from the own-colour residuals plus the later-colour terms.  Only
 * levels that run colour pairs (n > 32, (n/2) | 256, 256 | (n/2)^3); HOMO3D_NO_SUPPORT otherwise. */
int homo3d_presmooth(homo3d_handle* h, int level, double omega);
int homo3d_restrict(homo3d_handle* h, int level);  /* f[level] = P^T r[level-1] */
int homo3d_prolong(homo3d_handle* h, int level);   /* u[level] += P u[level+1] */
int homo3d_coarse_solve(homo3d_handle* h);
int homo3d_vcycle(homo3d_handle* h, int symmetric);
/* MG::v_cycle(w_SOR, pre, post) (MG.h:26, MG.cpp:46-70) on the FP32 level-0 fields u, f: `pre` forward
 * GS sweeps with relaxation omega, residual and restriction per level, the coarsest solve, then
 * prolongation and `post` forward sweeps.  0 <= pre, post <= 64, 0 < omega < 2. */
int homo3d_vcycle_sweeps(homo3d_handle* h, double omega, int pre, int post);
int homo3d_get_stencil(homo3d_handle* h, int level, float* host); /* [N][27][3][3] */

/* Density-update kernels.  Arrays are lexicographic with row pitch `pitch` (elements). */
int homo3d_symmetrize(homo3d_handle* h, float* field, int pitch, int sym, int on_device);
int homo3d_oc_filter(homo3d_handle* h, float* sens, const float* rho, int pitch, double radius, int periodic,
                     int on_device);
int homo3d_oc_update(homo3d_handle* h, const float* sens, float* rho, int pitch, double vol, double min_rho,
                     double step, double damp, int on_device);

/* Density filter of the reference's expression templates, rho.conv(radial_convker_t<Spline4 | Linear>(radius))
 * (AutoDiff/TensorExpression.h:1007-1046, TensorExpression.cuh:288-368): normalised radial weights on the
 * (2 int(radius + 0.5) + 1)^3 box, renormalised over the in-grid neighbours unless periodic.  kernel 0 =
 * Spline4, 1 = Linear.  backward = 1 applies the adjoint (in = dJ/dout, out = dJ/din).  in may equal out. */
int homo3d_density_filter(homo3d_handle* h, const float* in, float* out, int pitch, double radius, int kernel,
                          int periodic, int backward, int on_device);
/* Tensor::sgm(k, c) (AutoDiff.h:808-823): out = 1 / (1 + exp(-k (in - c))); backward = 1: y = forward output,
 * in = dJ/dy, out = dJ/dx = in k y (1 - y). */
int homo3d_projection(homo3d_handle* h, const float* in, const float* y, float* out, int pitch, double k, double c,
                      int backward, int on_device);

/* Elementwise density maps of the optimisation loop, which the reference forms with its device expression
 * templates (rho.pow(3) and its chain rule, Tensor::sum; Framework.cu:137-356): x = rho^p;
 * drho = p rho^(p-1) dx (drho may alias dx); *sum = FP64 sum over the n^3 cells; a = value.  Pitched
 * lexicographic arrays on the handle's grid; on_device as elsewhere (-1: from the pointer). */
int homo3d_density_pow(homo3d_handle* h, const float* rho, float* x, int pitch, double p, int on_device);
int homo3d_density_pow_grad(homo3d_handle* h, const float* rho, const float* dx, float* drho, int pitch, double p,
                            int on_device);
int homo3d_field_sum(homo3d_handle* h, const float* a, int pitch, double* sum, int on_device);
int homo3d_field_fill(homo3d_handle* h, float* a, int pitch, float value, int on_device);
/* Device memory on the calling thread's current device, for host code without the HIP headers (no
 * reference counterpart: the reference allocates through CUDA directly).  kind: 0 host to device, 1 device
 * to host, 2 device to device; the copy is synchronous. */
int homo3d_device_alloc(void** ptr, size_t bytes);
int homo3d_device_free(void* ptr);
int homo3d_device_copy(void* dst, const void* src, size_t bytes, int kind);

/* homo3d_oc_update without any allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  sens and rho are device pointers (lexicographic, row pitch `pitch`, rho updated in
 * place); stream is a hipStream_t or NULL for the default stream; 1 <= n <= 1024 (HOMO3D_BAD_ARG
 * beyond).  The bisection's reduction scratch is a fixed 20 KB device array inside the library (one per
 * GPU, held until the process ends); calls that overlap on one GPU share it and race.  The result is
 * bit-identical to homo3d_oc_update's. */
int homo3d_oc_update_noalloc(const float* sens, float* rho, int n, int pitch, double vol, double min_rho, double step,
                             double damp, void* stream);

/* MMA (homo::MMAOptimizer, optimization/mmaOptimizer.h:12-34; mmasubDevice, mmaOpt.h:17-19).
 * FP64 state on the GPU; m <= 4 constraints; constants of mmasub_ker (mmaOpt.cu:846-851), move = 1. */
typedef struct homo3d_mma homo3d_mma;
int homo3d_mma_create(int m, long n, double a0, double a, double c, double d, int device, homo3d_mma** out);
int homo3d_mma_destroy(homo3d_mma* h);
const char* homo3d_mma_last_error(const homo3d_mma* h);
int homo3d_mma_set_bound(homo3d_mma* h, double xmin, double xmax);
/* MMAOptimizer::setBound(float* xmin, float* xmax) (mmaOptimizer.h:30): n per-variable bounds each. */
int homo3d_mma_set_bounds(homo3d_mma* h, const float* xmin, const float* xmax, int on_device);
/* x: n floats updated in place; dfdx: n floats; gval: m floats on the host; dgdx: m arrays of n floats.
 * on_device selects device (1) or host (0) for x, dfdx and the dgdx arrays. */
int homo3d_mma_update(homo3d_mma* h, int itn, float* x, const float* dfdx, const float* gval, const float* const* dgdx,
                      int on_device);
/* Read back the FP64 iterate and asymptotes (host arrays of n doubles, any may be NULL). */
int homo3d_mma_state(homo3d_mma* h, double* x, double* low, double* upp, int* newton_iters);

/* Seeded randc density (project generator), written to host memory, n^3 floats, pitch n. */
int homo3d_randc_density(int n, uint64_t seed, double vol, int sym, float* out);

/* Learned warm start (no counterpart in the reference, whose warm start is the previous design's stored
 * solutions, homogenization.cu:471-477 / grid.cpp:459-462).  A periodic 3D U-Net (levels, channels
 * base_channels * 2^l, bf16 weights and activations, FP32 accumulation on the matrix cores) maps the
 * density on a grid^3 grid to the six load cases' displacements, trained on the GPU on (density, solution)
 * pairs the solver generates.  The prediction, upsampled to the solve grid, becomes the FP64 iterate of
 * the next warm refined solve, which stays the solver and the judge of its error.
 * Parameter layout (FP32): per layer W[cout][taps][cin] then b[cout], in the order enc0.c1, enc0.c2, ...,
 * enc{L-1}.c2, dec{L-2}.c1, dec{L-2}.c2, ..., dec0.c2, head (1x1, 18 outputs). */
typedef struct homo3d_net homo3d_net;
typedef struct {
    int grid;            /* network grid g (power of two, g >= 8 * 2^(levels-1)); default 64 */
    int levels;          /* U-Net levels (4) */
    int base_channels;   /* channels of level 0 (16; multiple of 8) */
    int batch;           /* largest batch of a forward or training step (4) */
    double target_scale; /* the network predicts target_scale * u / n, mean-free per channel (8) */
    uint64_t seed;       /* initial weights (He-normal) */
    int device;
} homo3d_net_config;
void homo3d_net_default_config(homo3d_net_config* cfg);
int homo3d_net_create(const homo3d_net_config* cfg, homo3d_net** out);
int homo3d_net_destroy(homo3d_net* net);
const char* homo3d_net_last_error(const homo3d_net* net);
long homo3d_net_num_params(const homo3d_net* net);
/* FP32 master parameters to (to_device = 1; resets the Adam state) or from the network. */
int homo3d_net_params(homo3d_net* net, float* host, int to_device);
/* rho: batch x grid^3 densities (host, x fastest); out: batch x grid^3 x 18 predictions (host). */
int homo3d_net_forward(homo3d_net* net, const float* rho, int batch, float* out);
/* Loss (mean over batch, voxels and channels of (P - mean_v P - target)^2, target mean-free) and its
 * gradient (num_params floats, may be NULL); target: batch x grid^3 x 18 (host). */
int homo3d_net_loss_grad(homo3d_net* net, const float* rho, const float* target, int batch, double* loss, float* grad);
/* One Adam step (lr, betas 0.9 / 0.999) on a host batch. */
int homo3d_net_train_step(homo3d_net* net, const float* rho, const float* target, int batch, double lr, double* loss);
/* Device-resident training set: reserve room for `count` pairs; homo3d_net_dataset_add appends the pair of a
 * solved handle (its density pooled to the network grid and its FP64 solution sampled there; n % grid == 0). */
int homo3d_net_dataset_reserve(homo3d_net* net, int count);
int homo3d_net_dataset_add(homo3d_net* net, homo3d_handle* h);
int homo3d_net_dataset_size(const homo3d_net* net);
/* Adam epochs over the first `count` pairs of the training set (all when count <= 0) in seeded random order;
 * losses[e] = mean batch loss of epoch e. */
int homo3d_net_train(homo3d_net* net, int epochs, int count, double lr, uint64_t seed, double* losses);
/* Mean loss over pairs [first, first + count) of the training set, without updating. */
int homo3d_net_eval(homo3d_net* net, int first, int count, double* loss);
/* Predict the solution of the handle's current density and make it the FP64 iterate that the next warm
 * solve (homo3d_solve with warm = 1) starts from.  n % grid == 0. */
int homo3d_net_warm_start(homo3d_handle* h, homo3d_net* net);

/* Batched uneven designs (no reference counterpart: the reference solves one design per Homogenization,
 * homogenization.cu:464-575).  One handle and stream per design (sizes[i]^3, each as homo3d_create with the
 * base configuration); homo3d_batch_solve drives up to max_inflight designs at once from host threads, so the
 * small grids' launches overlap on the GPU.  x[i]: design i's density (x = rho^p, unpitched, host or device
 * memory); stats[i] and C[36 i .. 36 i + 35] receive its solve statistics and C^H (either may be NULL).
 * autotune = 1 runs homo3d_autotune on each design's first solve.  Returns the worst status. */
typedef struct homo3d_batch homo3d_batch;
int homo3d_batch_create(const homo3d_config* base, int count, const int* sizes, int max_inflight, homo3d_batch** out);
int homo3d_batch_destroy(homo3d_batch* batch);
const char* homo3d_batch_last_error(const homo3d_batch* batch);
homo3d_handle* homo3d_batch_handle(homo3d_batch* batch, int i);
int homo3d_batch_solve(homo3d_batch* batch, const float* const* x, double tol, int autotune, homo3d_solve_stats* stats,
                       double* C);

/* Kernel timing: enable event pairs around the finest-level kernel launches, then read totals per
 * kernel: "fine_gs" (full GS colour passes), "fine_gs_zero" (colour passes of sweeps that start from
 * zero), "fine_apply" (the CG operator pass: q = K z + beta q_old with the fused p, x updates and p.q;
 * also homo3d_apply_K), "fine_apply_first" (the first CG pass of an inner solve: q = K z, p = z) and
 * "fine_residual" (f - K u inside the V-cycle).  Each event pair can delay the next launch by tens of
 * microseconds; homo3d_profile_only restricts the pairs to one kernel kind (NULL or "": all kinds). */
int homo3d_profile_enable(homo3d_handle* h, int enable);
int homo3d_profile_only(homo3d_handle* h, const char* kernel);
int homo3d_profile_get(homo3d_handle* h, const char* kernel, long* launches, double* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* HOMO3D_H */
