// Host side of the homo3d C ABI: hierarchy set-up, V-cycle, stand-alone MG and MG-PCG drivers,
// C^H / sensitivity and density updates.  Every device operation is a kernel of kernels.hip
// launched on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/homo3d.h"
#include "kernels.h"
#include "net.h"

using namespace homo3d;

namespace {

constexpr int kMaxLevels = 16;
constexpr int kRhs = 6;
constexpr int kDotBlocks = 1024;

struct Level {
    int n = 0;
    size_t N = 0;
    int lex = 0;  // 1: x + n(y + n z) (coarsest level); 0: parity-blocked (see store_index)
    float* u = nullptr;
    float* f = nullptr;
    float* r = nullptr;
    void* st = nullptr;   // levels >= 1: FP32 [243][N], or packed FP16 [122][N] words when st_half
    int st_half = 0;
};

struct ProfEvent {
    int name = -1;
    int a = -1, b = -1;  // indices into the handle's event pool
};

// Timed finest-level kernels: full GS colour passes, zero-start GS colour passes, the CG operator pass
// (K z + beta q with the fused p, x updates and p.q; the first step of an inner solve separately), and
// the V-cycle residual f - K u.
const char* const kProfNames[] = {"fine_gs", "fine_apply", "fine_residual", "fine_gs_zero", "fine_apply_first"};
constexpr int kProfKinds = 5;
enum { kProfGs = 0, kProfApply = 1, kProfResidual = 2, kProfGsZero = 3, kProfApplyFirst = 4 };
constexpr size_t kEventPool = 8192;
// Levels of kSmallLevelMinN < n <= kSmallLevelN (the largest that run their smoother one load case per
// launch) take kPrecondSmallSweeps sweeps per side in the PCG preconditioner.  Measured at 256^3,
// solves/s: 2 sweeps at 32^3 only 7.79; at 32^3 and 16^3 7.71; at 32^3, 16^3 and 8^3 7.69; 3 at
// 32^3 7.75; 2 at 16^3 and 8^3 7.75; 2 at 8^3 only 6.79; 1 everywhere 6.39 (an older build).
constexpr int kSmallLevelN = 32;
constexpr int kSmallLevelMinN = 16;
constexpr int kPrecondSmallSweeps = 2;

}  // namespace

struct homo3d_handle {
    homo3d_config cfg{};
    int nl = 0;
    Level lv[kMaxLevels];
    hipStream_t stream = nullptr;
    float* d_ke = nullptr;
    void* d_ke16 = nullptr;     // element matrix as matrix-core operands (ke_mfma_table), or null
    FineMfma mfma{};            // its launch parameters
    bool use_mfma = false;      // run the finest-level operator in element form (config or homo3d_autotune)
    const FineMfma* mf() const { return d_ke16 && use_mfma ? &mfma : nullptr; }
    bool stage_tiles = true;    // CG pass: z node boxes staged in LDS where fine_cg_staged_ok (homo3d_stage_tiles)
    float* d_kchi = nullptr;
    float* d_G = nullptr;
    float* d_w = nullptr;
    float* d_x = nullptr;       // raw cell values x (FP32 input), for FP64 weights
    float* d_stage = nullptr;   // pitched staging buffer for host densities / sensitivities
    size_t stage_elems = 0;
    double* d_Ainv = nullptr;       // the coarsest level's dense FP64 inverse (mc x mc)
    int* d_cstat = nullptr;         // factorisation status on the device
    double* d_cpiv = nullptr;       // its pivot rows and columns (4 mc)
    int mc = 0;
    // the coarsest factorisation runs on the GPU at the density update, on a stream of its own behind
    // the Galerkin kernels (ev_gal), so it overlaps the solve's finest-level work; its status is read,
    // and the main stream made to wait for it, at the first coarse solve after it
    hipStream_t cstream = nullptr;
    hipEvent_t ev_gal = nullptr;
    int* h_cstat = nullptr;          // pinned landing place of the status
    hipEvent_t ev_coarse = nullptr;  // ... and the event that says it has landed
    bool coarse_pending = false;
    double* h_rr = nullptr;          // pinned landing place of the CG residual norms
    hipEvent_t ev_rr = nullptr;      // ... and the event that says they have landed
    // solution fields (warm start) and PCG work vectors, level-0 sized, 6 load cases
    float* X = nullptr;
    float* F = nullptr;
    double* U64 = nullptr;      // FP64 solution of the refinement loop
    double* d_ke64 = nullptr;
    double* d_kchi64 = nullptr;
    size_t partial_elems = 0;
    bool x_dirty = false;       // X was written by the caller: widen it before a warm refined solve
    float* P = nullptr;
    float* Q = nullptr;
    float* Z = nullptr;
    float* Rr = nullptr;
    double* d_partial = nullptr;
    double* d_sc = nullptr;
    double* d_tmp = nullptr;
    float* d_ftmp = nullptr;
    int* d_gperm = nullptr;
    int* d_gflip = nullptr;
    double ke[576];
    bool density_set = false;
    std::string err;
    bool prof = false;
    int prof_only = -1;  // the one profiled kind, or -1 for all
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    std::vector<ProfEvent> pending;
    long prof_count[kProfKinds] = {};
    double prof_ms[kProfKinds] = {};
};

namespace {

int set_err(homo3d_handle* h, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, expr)                                                                               \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return set_err(h, HOMO3D_HIP_ERROR, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------------------------
// Element matrix (templateMatrix.cpp:63-166): 2x2x2 Gauss quadrature of B^T C B on the unit hex.
// ---------------------------------------------------------------------------------------------
void element_matrix_lame(double lam, double mu, double* ke) {
    double C[6][6] = {};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[i][j] = lam;
    for (int i = 0; i < 3; i++) C[i][i] += 2 * mu;
    for (int i = 3; i < 6; i++) C[i][i] = mu;
    std::vector<double> K(576, 0.0);
    const double g = std::sqrt(1.0 / 3.0);
    for (int q = 0; q < 8; q++) {
        const double pt[3] = {(1 + (2 * (q & 1) - 1) * g) / 2, (1 + (2 * ((q >> 1) & 1) - 1) * g) / 2,
                              (1 + (2 * (q >> 2) - 1) * g) / 2};
        double B[6][24] = {};
        for (int b = 0; b < 8; b++) {
            const int id[3] = {b & 1, (b >> 1) & 1, b >> 2};
            double lin[3], dn[3];
            for (int k = 0; k < 3; k++) lin[k] = id[k] ? pt[k] : 1 - pt[k];
            for (int k = 0; k < 3; k++) dn[k] = (id[k] ? 1.0 : -1.0) * lin[(k + 1) % 3] * lin[(k + 2) % 3];
            for (int k = 0; k < 3; k++) B[k][3 * b + k] = dn[k];
            B[3][3 * b + 1] = dn[2];
            B[3][3 * b + 2] = dn[1];
            B[4][3 * b + 0] = dn[2];
            B[4][3 * b + 2] = dn[0];
            B[5][3 * b + 0] = dn[1];
            B[5][3 * b + 1] = dn[0];
        }
        for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) {
                double s = 0;
                for (int a = 0; a < 6; a++)
                    for (int c = 0; c < 6; c++) s += B[a][i] * C[a][c] * B[c][j];
                K[i * 24 + j] += s * 0.125;
            }
    }
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) ke[i * 24 + j] = 0.5 * (K[i * 24 + j] + K[j * 24 + i]);
}

void lame_params(double E, double nu, double& lam, double& mu) {
    lam = nu * E / ((1 + nu) * (1 - 2 * nu));
    mu = E / (2 * (1 + nu));
}

void element_matrix(double E, double nu, double* ke) {
    double lam, mu;
    lame_params(E, nu, lam, mu);
    element_matrix_lame(lam, mu, ke);
}

// The Lame tables 72 KE(lambda = 1, mu = 0) and 72 KE(0, 1) (templateMatrix.cpp:8-61): small integers,
// recovered from the quadrature above.  False if an entry is not within 1e-9 of an integer.
bool lame_tables(double* kl, double* km) {
    element_matrix_lame(72.0, 0.0, kl);
    element_matrix_lame(0.0, 72.0, km);
    for (int i = 0; i < 576; i++) {
        const double a = std::nearbyint(kl[i]), b = std::nearbyint(km[i]);
        if (std::fabs(a - kl[i]) > 1e-9 || std::fabs(b - km[i]) > 1e-9) return false;
        kl[i] = a;
        km[i] = b;
    }
    return true;
}

double macro_disp(int node, int s, int c) {
    const double px = node & 1, py = (node >> 1) & 1, pz = node >> 2;
    switch (s) {
        case 0: return c == 0 ? px : 0;
        case 1: return c == 1 ? py : 0;
        case 2: return c == 2 ? pz : 0;
        case 3: return c == 1 ? pz / 2 : (c == 2 ? py / 2 : 0);
        case 4: return c == 0 ? pz / 2 : (c == 2 ? px / 2 : 0);
        default: return c == 0 ? py / 2 : (c == 1 ? px / 2 : 0);
    }
}

double hat(int o) { return o == 0 ? 1.0 : ((o == 1 || o == -1) ? 0.5 : 0.0); }

// Level sizes (Grid::buildRoot / coarse2, grid.cpp:21-151) in integer arithmetic.
// Upper limit: the finest-level gathers address each field through one buffer resource with 32-bit
// byte offsets, and the largest field, the FP64 iterate [6][3][n^3], must stay below 2^32 bytes:
// 144 n^3 < 2^32 holds up to n = 309, and the next solvable size after 256 is 320 (c * 2^6).
constexpr int kMaxReso = 256;
static_assert(144ull * kMaxReso * kMaxReso * kMaxReso < (1ull << 32), "FP64 field exceeds 32-bit offsets");
int level_sizes(int n, int* sizes) {
    if (n < 8 || n > kMaxReso) return -1;
    int fl = 0;
    while ((1 << (fl + 1)) <= n) fl++;
    const int k = std::max(fl - 2, 0);
    if (n % (1 << k) != 0) return -1;  // only sizes c * 2^k are solved as requested
    const int c = n >> k;
    for (int l = 0; l <= k; l++) sizes[l] = c << (k - l);
    return k + 1;
}

// Storage index of vertex (x, y, z) on a level.  All levels but the coarsest are parity-blocked: the
// eight classes (x&1, y&1, z&1) are contiguous (n/2)^3 blocks, block b = (x&1) | (y&1)<<1 | (z&1)<<2,
// lexicographic inside; a GS colour is one block.  The coarsest level is lexicographic.
size_t store_index(const Level& L, int x, int y, int z) {
    if (L.lex) return x + (size_t)L.n * (y + (size_t)L.n * z);
    const size_t h = L.n / 2;
    const size_t b = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
    return b * h * h * h + (x >> 1) + h * ((y >> 1) + h * (z >> 1));
}

// Copy `planes` planes of L.N values between lexicographic order (the C ABI's) and storage order.
template <typename T>
void permute_planes(const Level& L, const T* in, T* out, int planes, bool to_storage) {
    for (int z = 0; z < L.n; z++)
        for (int y = 0; y < L.n; y++)
            for (int x = 0; x < L.n; x++) {
                const size_t v = x + (size_t)L.n * (y + (size_t)L.n * z), sv = store_index(L, x, y, z);
                for (int p = 0; p < planes; p++) {
                    if (to_storage)
                        out[(size_t)p * L.N + sv] = in[(size_t)p * L.N + v];
                    else
                        out[(size_t)p * L.N + v] = in[(size_t)p * L.N + sv];
                }
            }
}

// Symmetry groups (TensorExpression.cuh:72-262) in the canonical order shared with the checker:
// element = (perm, flips), out[k] = flip_k ? n-1-p[perm[k]] : p[perm[k]].
const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
const int kPermSign[6] = {1, -1, -1, 1, 1, -1};

void symmetry_group(int sym, std::vector<int>& perm, std::vector<int>& flip) {
    perm.clear();
    flip.clear();
    auto add = [&](int p, int f) {
        perm.push_back(kPerm[p][0] | (kPerm[p][1] << 2) | (kPerm[p][2] << 4));
        flip.push_back(f);
    };
    if (sym == HOMO3D_SYM_REFLECT3) {
        for (int f = 0; f < 8; f++) add(0, f);
    } else if (sym == HOMO3D_SYM_REFLECT6) {
        for (int p = 0; p < 6; p++)
            for (int f = 0; f < 8; f++) add(p, f);
    } else if (sym == HOMO3D_SYM_ROTATE3) {
        for (int p = 0; p < 6; p++)
            for (int f = 0; f < 8; f++) {
                const int nflip = (f & 1) + ((f >> 1) & 1) + (f >> 2);
                if (kPermSign[p] * ((nflip & 1) ? -1 : 1) == 1) add(p, f);
            }
    }
}

void prof_flush(homo3d_handle* h) {
    if (h->pending.empty()) return;
    (void)hipStreamSynchronize(h->stream);
    for (auto& ev : h->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, h->pool[ev.a], h->pool[ev.b]) != hipSuccess) continue;
        h->prof_count[ev.name] += 1;
        h->prof_ms[ev.name] += ms;
    }
    h->pending.clear();
    h->pool_used = 0;
}

// Event pair around one launch of a profiled kernel kind (kProfNames).
void prof_begin(homo3d_handle* h, int kind, ProfEvent& ev) {
    if (!h->prof || (h->prof_only >= 0 && kind != h->prof_only)) return;
    if (h->pool_used + 2 > h->pool.size()) prof_flush(h);
    while (h->pool.size() < h->pool_used + 2 && h->pool.size() < kEventPool) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        h->pool.push_back(e);
    }
    if (h->pool_used + 2 > h->pool.size()) return;
    ev.name = kind;
    ev.a = (int)h->pool_used++;
    ev.b = (int)h->pool_used++;
    (void)hipEventRecord(h->pool[ev.a], h->stream);
}

void prof_end(homo3d_handle* h, ProfEvent& ev) {
    if (!h->prof || ev.name < 0) return;
    (void)hipEventRecord(h->pool[ev.b], h->stream);
    h->pending.push_back(ev);
}

// ---------------------------------------------------------------------------------------------
// Level operations
// ---------------------------------------------------------------------------------------------
// One 8-colour sweep.  zero_start: u is taken as zero on entry whatever it holds (forward sweeps
// only); every colour pass then writes its block without reading later colours.
// u_half: the level-0 u is packed FP16 (coarse levels are always FP32).
// dot_out (level 0, reverse sweeps of a packed-FP16 u with fused colour pairs; returns false otherwise):
// also the per-load-case f . u after the sweep, accumulated by the pair kernels as each vertex gets its
// final value.
// rown (zero-start forward sweeps with fused colour pairs, else ignored): the passes also store their
// own-colour residuals there, for op_residual_after_zero_sweep.  [pair_begin, pair_end): the
// colour pairs to launch (level 0 with fused pairs, forward sweeps; a sweep split in two launches the
// same kernels as a whole one).
bool op_smooth(homo3d_handle* h, int l, void* u, const float* f, float omega, int reverse, bool zero_start = false,
               int u_half = 0, double* dot_out = nullptr, float* rown = nullptr, int pair_begin = 0,
               int pair_end = 4) {
    const Level& L = h->lv[l];
    if (l == 0 && !zero_start) launch_zero_pin(h->stream, L.N, u, u_half);
    const int zs = zero_start && !reverse ? 1 : 0;
    if (l == 0 && fine_gs_pairs_ok(L.n)) {  // colours two at a time (same result as one by one)
        const bool dot = dot_out && reverse && u_half;
        const int G = (int)(L.N / 8 / 256);  // workgroups per pair launch
        for (int cc = reverse ? 0 : 2 * pair_begin; cc < (reverse ? 8 : 2 * pair_end); cc += 2) {
            ProfEvent ev;
            prof_begin(h, zs ? kProfGsZero : kProfGs, ev);
            launch_fine_gs2(h->stream, L.n, reverse ? 7 - cc : cc, omega, reverse, zs, h->d_ke, h->d_w, u, u_half, f,
                            dot ? h->d_partial + (size_t)(cc / 2) * G : nullptr, 4 * G, zs ? rown : nullptr);
            prof_end(h, ev);
        }
        if (dot) launch_sum_partials(h->stream, kRhs, 4 * G, h->d_partial, dot_out);
        return dot;
    }
    if (l > 0 && stencil_gs_pairs_ok(L.n, L.lex)) {  // colours two at a time (same result as one by one)
        for (int cc = 0; cc < 8; cc += 2)
            launch_stencil_gs2(h->stream, L.n, reverse ? 7 - cc : cc, omega, reverse, zs, L.st, L.st_half, u, u_half, f,
                               zs ? rown : nullptr);
        return false;
    }
    for (int cc = 0; cc < 8; cc++) {
        const int c = reverse ? 7 - cc : cc;
        if (l == 0) {
            ProfEvent ev;
            prof_begin(h, zs ? kProfGsZero : kProfGs, ev);
            launch_fine_gs(h->stream, L.n, c, omega, reverse, zs, h->d_ke, h->d_w, u, u_half, f);
            prof_end(h, ev);
        } else {
            launch_stencil_gs(h->stream, L.n, c, omega, reverse, zs, L.st, L.st_half, u, u_half, f);
        }
    }
    return false;
}

// After a zero-start forward sweep that stored its own-colour residuals in r (op_smooth rown):
// r = f - K u from the later-colour terms alone.  Level 0 needs fused colour pairs; coarse levels
// their colour pairs (six load cases per launch).
bool residual_after_zero_sweep_ok(const homo3d_handle* h, int l = 0) {
    const Level& L = h->lv[l];
    return l == 0 ? fine_gs_pairs_ok(L.n) != 0 : stencil_gs_pairs_ok(L.n, L.lex) != 0;
}

void op_residual_after_zero_sweep(homo3d_handle* h, int l, const void* u, int u_half, float* r) {
    const Level& L = h->lv[l];
    if (l > 0) {
        launch_stencil_residual_later(h->stream, L.n, L.st, L.st_half, u, u_half, r);
        return;
    }
    ProfEvent ev;
    prof_begin(h, kProfResidual, ev);
    launch_fine_residual_later(h->stream, L.n, h->d_ke, h->d_w, u, u_half, r);
    prof_end(h, ev);
}

void op_residual(homo3d_handle* h, int l, const void* u, const float* f, float* r, int u_half = 0) {
    const Level& L = h->lv[l];
    if (l == 0) {
        ProfEvent ev;
        prof_begin(h, kProfResidual, ev);
        launch_fine_apply(h->stream, L.n, h->d_ke, h->mf(), h->d_w, u, u_half, f, r);
        prof_end(h, ev);
    } else {
        launch_stencil_apply(h->stream, L.n, L.lex, L.st, L.st_half, u, u_half, f, r);
    }
}

// The coarsest level's FP64 inverse is formed on the GPU at the density update (launch_coarse_factor,
// on cstream); the first coarse solve after it reads the factorisation's status and makes the main
// stream wait for it (by then the solve's finest-level work, queued behind the Galerkin kernels, has
// kept the GPU busy meanwhile).  A system that is not positive definite
// leaves the handle without a density.
int coarse_factor(homo3d_handle* h) {
    if (!h->coarse_pending) return HOMO3D_OK;
    h->coarse_pending = false;
    HIPCHK(h, hipEventSynchronize(h->ev_coarse));
    HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_coarse, 0));
    if (*h->h_cstat != 0) {
        h->density_set = false;
        return set_err(h, HOMO3D_DIVERGED, "coarsest system is not positive definite (pivot %d of %d)", *h->h_cstat - 1,
                       h->mc);
    }
    return HOMO3D_OK;
}

// One V(1,1) cycle on the level-0 pair (u0, f0) (MG::v_cycle, MG.cpp:46-70).  When rn2 is given it
// receives, per load case, the squared norm of the level-0 residual after pre-smoothing.  u0_zero:
// the cycle starts from u0 = 0 (u0's contents are ignored), as the coarse levels always do.
// u0_half: u0 is packed FP16 (level 0 only; every other field is FP32).  dot_out: per-load-case
// f0 . u0 of the result (the r . z of a CG step).
// small_sweeps: smoothing sweeps per side on the levels with kSmallLevelMinN < n <= kSmallLevelN;
// every other level takes one.
// head_done: vcycle_head has already launched the first colour pair of the level-0 pre-smoothing.
int vcycle(homo3d_handle* h, void* u0, const float* f0, int symmetric, double* rn2, bool u0_zero = false,
           int u0_half = 0, double* dot_out = nullptr, int small_sweeps = 1, bool head_done = false) {
    const float omega = (float)h->cfg.sor_omega;
    // packed FP16 u: level 0 when the caller's iterate is, and then every coarser level (but the
    // coarsest) whose smoother runs six load cases per launch; the small levels stay FP32
    auto u_half_of = [&](int l) {
        if (l == 0) return u0_half;
        return u0_half && l < h->nl - 1 && stencil_u_half_ok(h->lv[l].n, h->lv[l].lex) ? 1 : 0;
    };
    for (int l = 0; l < h->nl - 1; l++) {
        Level& L = h->lv[l];
        void* u = l == 0 ? u0 : L.u;
        const float* f = l == 0 ? f0 : L.f;
        const int uh = u_half_of(l);
        const int sweeps = l > 0 && L.n <= kSmallLevelN && L.n > kSmallLevelMinN ? small_sweeps : 1;
        // a single zero-start sweep leaves its own-colour residuals behind; the residual then needs
        // only the later-colour terms
        const bool rown = (l > 0 || u0_zero) && sweeps == 1 && residual_after_zero_sweep_ok(h, l);
        op_smooth(h, l, u, f, omega, 0, l > 0 || u0_zero, uh, nullptr, rown ? L.r : nullptr,
                  l == 0 && head_done ? 1 : 0);
        for (int k = 1; k < sweeps; k++) op_smooth(h, l, u, f, omega, 0, false, uh);
        if (rown)
            op_residual_after_zero_sweep(h, l, u, uh, L.r);
        else
            op_residual(h, l, u, f, L.r, uh);
        if (l == 0 && rn2) launch_dot6(h->stream, 3 * L.N, L.r, L.r, h->d_partial, kDotBlocks, rn2);
        launch_restrict(h->stream, L.n, h->lv[l + 1].lex, L.r, h->lv[l + 1].f);
    }
    const Level& Lc = h->lv[h->nl - 1];
    int rc = coarse_factor(h);
    if (rc) return rc;
    launch_coarse_solve(h->stream, Lc.n, h->d_Ainv, Lc.f, Lc.u);
    for (int l = h->nl - 2; l >= 0; l--) {
        Level& L = h->lv[l];
        void* u = l == 0 ? u0 : L.u;
        const float* f = l == 0 ? f0 : L.f;
        const int uh = u_half_of(l);
        launch_prolong(h->stream, L.n, h->lv[l + 1].lex, h->lv[l + 1].u, u_half_of(l + 1), u, uh, l == 0 ? 1 : 0);
        const int sweeps = l > 0 && L.n <= kSmallLevelN && L.n > kSmallLevelMinN ? small_sweeps : 1;
        for (int k = 1; k < sweeps; k++) op_smooth(h, l, u, f, omega, symmetric ? 1 : 0, false, uh);
        const bool fused = op_smooth(h, l, u, f, omega, symmetric ? 1 : 0, false, uh, l == 0 ? dot_out : nullptr);
        if (l == 0 && dot_out && !fused) launch_dot_rz(h->stream, 3 * L.N, f0, u0, uh, h->d_partial, kDotBlocks, dot_out);
    }
    return HOMO3D_OK;
}

// MG::v_cycle(w_SOR, pre, post) (MG.cpp:46-70) with arbitrary sweep counts and relaxation factor, on
// the FP32 level-0 fields: `pre` forward sweeps, residual, restriction on the way down; the coarsest
// solve; prolongation and `post` forward sweeps on the way up.  A coarse level's correction starts
// at zero (the first sweep skips the not yet visited colours; with pre = 0 it is cleared).
int vcycle_sweeps(homo3d_handle* h, float omega, int pre, int post) {
    for (int l = 0; l < h->nl - 1; l++) {
        Level& L = h->lv[l];
        float* u = l == 0 ? h->X : L.u;
        const float* f = l == 0 ? h->F : L.f;
        if (l > 0 && pre == 0) HIPCHK(h, hipMemsetAsync(L.u, 0, sizeof(float) * 3 * kRhs * L.N, h->stream));
        for (int k = 0; k < pre; k++) op_smooth(h, l, u, f, omega, 0, l > 0 && k == 0, 0);
        op_residual(h, l, u, f, L.r, 0);
        launch_restrict(h->stream, L.n, h->lv[l + 1].lex, L.r, h->lv[l + 1].f);
    }
    const Level& Lc = h->lv[h->nl - 1];
    int rc = coarse_factor(h);
    if (rc) return rc;
    launch_coarse_solve(h->stream, Lc.n, h->d_Ainv, Lc.f, Lc.u);
    for (int l = h->nl - 2; l >= 0; l--) {
        Level& L = h->lv[l];
        float* u = l == 0 ? h->X : L.u;
        const float* f = l == 0 ? h->F : L.f;
        launch_prolong(h->stream, L.n, h->lv[l + 1].lex, h->lv[l + 1].u, 0, u, 0, l == 0 ? 1 : 0);
        for (int k = 0; k < post; k++) op_smooth(h, l, u, f, omega, 0, false, 0);
    }
    return HOMO3D_OK;
}

// The first launch of a preconditioning V-cycle (u0_zero, level 0 with fused colour pairs): colour
// pair 0 of the pre-smoothing, queued ahead so that the GPU has work while the host reads the CG
// norms; vcycle(..., head_done = true) then continues from pair 1.
bool vcycle_head_ok(const homo3d_handle* h) { return fine_gs_pairs_ok(h->lv[0].n) != 0; }

void vcycle_head(homo3d_handle* h, void* u0, const float* f0, int u0_half) {
    Level& L = h->lv[0];
    op_smooth(h, 0, u0, f0, (float)h->cfg.sor_omega, 0, true, u0_half, nullptr,
              residual_after_zero_sweep_ok(h) ? L.r : nullptr, 0, 1);
}

int ensure_stage(homo3d_handle* h, size_t elems) {
    if (h->stage_elems >= elems) return HOMO3D_OK;
    if (h->d_stage) (void)hipFree(h->d_stage);
    h->d_stage = nullptr;
    HIPCHK(h, hipMalloc(&h->d_stage, sizeof(float) * elems));
    h->stage_elems = elems;
    return HOMO3D_OK;
}

// Bring a pitched lexicographic array to the device: returns a device pointer (the caller's when
// on_device, else the staging buffer).
int stage_in(homo3d_handle* h, const float* a, int pitch, int on_device, const float** dptr) {
    const int n = h->lv[0].n;
    if (on_device) {
        *dptr = a;
        return HOMO3D_OK;
    }
    const size_t elems = (size_t)pitch * n * n;
    int rc = ensure_stage(h, elems);
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(h->d_stage, a, sizeof(float) * elems, hipMemcpyHostToDevice, h->stream));
    *dptr = h->d_stage;
    return HOMO3D_OK;
}

// Return a pitched lexicographic result: only the n live columns of each row are written, so the
// caller's padding (pitch > n) keeps its values, as the reference's pitched tensor views do.
int stage_out(homo3d_handle* h, float* dst, const float* src, int pitch, int on_device) {
    const int n = h->lv[0].n;
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (pitch == n)
        HIPCHK(h, hipMemcpyAsync(dst, src, sizeof(float) * (size_t)n * n * n, kind, h->stream));
    else
        HIPCHK(h, hipMemcpy2DAsync(dst, sizeof(float) * pitch, src, sizeof(float) * pitch, sizeof(float) * n,
                                   (size_t)n * n, kind, h->stream));
    return HOMO3D_OK;
}

constexpr int kMaxOuter = 12;
// Inner solve target per refinement round (measured at 256^3: 1e-4 and 1e-6 cost more CG steps than
// they save in FP64 residual rounds; 3e-3 and 1e-2 need a third round).
constexpr double kInnerTol = 1e-3;

// MG-preconditioned CG on K x = r for the load cases not marked done: x (h->X) starts at zero, r
// (h->Rr) holds the right-hand side and is updated in place to the recursive residual.  A load case
// stops when ||r|| <= tol[r] * sqrt(r0n2[r]).  Symmetric V(1,1) preconditioner, FP64 inner products.
// iters[] accumulates the iterations spent per load case.
// On return *x_pending tells whether x still lacks the last step's alpha p (launch_refine_accumulate
// adds it on the way into the FP64 iterate).
int pcg_inner(homo3d_handle* h, const double* r0n2, const double* tol, const bool* done, int maxit, int* iters,
              bool* x_pending_out) {
    *x_pending_out = false;
    const Level& L0 = h->lv[0];
    const size_t M = 3 * L0.N;
    double sc[6 * kRhs];
    bool act[kRhs];
    bool any = false;
    for (int r = 0; r < kRhs; r++) {
        act[r] = !done[r] && r0n2[r] > 0;
        sc[5 * kRhs + r] = act[r] ? 1.0 : 0.0;
        any = any || act[r];
    }
    if (!any) return HOMO3D_OK;
    if (M % 4) return set_err(h, HOMO3D_NO_SUPPORT, "CG vectors need 4 | 3n^3");
    HIPCHK(h, hipMemcpyAsync(h->d_sc + 5 * kRhs, sc + 5 * kRhs, sizeof(double) * kRhs, hipMemcpyHostToDevice,
                             h->stream));
    // the preconditioner's level-0 iterate z in packed FP16 when enabled and the layout allows it (p, q, r
    // and x stay FP32: an FP16 search direction costs more CG steps than it saves)
    const int zh = (h->cfg.precond_half && L0.n % 8 == 0) ? 1 : 0;
    int rc = vcycle(h, h->Z, h->Rr, 1, nullptr, true, zh, h->d_sc, kPrecondSmallSweeps);  // with rz
    if (rc) return rc;
    // One CG step: q = K p with p.q, formed as K z + beta q_old since p = z + beta p_old (the operator
    // gathers the packed z instead of the FP32 p).  The same pass advances x by the previous step's
    // alpha p_old and forms the new p, so p, x, r and q are each streamed once per step; the caller
    // adds the last step's alpha p.  While the host waits for the residual norms, the next V-cycle's
    // first colour pair is already queued (if the step turns out to be the last, it only rewrites z
    // and the level-0 residual, which nothing reads afterwards).
    if (!h->h_rr) {
        HIPCHK(h, hipHostMalloc((void**)&h->h_rr, sizeof(double) * kRhs, hipHostMallocDefault));
        HIPCHK(h, hipEventCreateWithFlags(&h->ev_rr, hipEventDisableTiming));
    }
    const bool ahead = vcycle_head_ok(h);
    double rr[kRhs];
    bool x_pending = false;
    for (int it = 0; it < maxit; it++) {
        {
            ProfEvent ev;
            prof_begin(h, it ? kProfApply : kProfApplyFirst, ev);
            launch_fine_apply_cg(h->stream, L0.n, h->d_ke, h->mf(), h->d_w, h->Z, zh, h->stage_tiles ? 1 : 0, h->P,
                                 h->X, h->d_sc, it == 0, h->Q, h->d_partial, kDotBlocks, h->d_sc + kRhs);  // pq
            prof_end(h, ev);
        }
        launch_pcg_alpha(h->stream, h->d_sc);
        launch_pcg_update_r(h->stream, M, h->d_sc, h->Q, h->Rr, h->d_partial, kDotBlocks, h->d_tmp);
        x_pending = true;  // alpha p is added by the next operator pass, or by the pass after the loop
        HIPCHK(h, hipMemcpyAsync(h->h_rr, h->d_tmp, sizeof(rr), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipEventRecord(h->ev_rr, h->stream));
        const bool head = ahead && it + 1 < maxit;
        if (head) vcycle_head(h, h->Z, h->Rr, zh);
        HIPCHK(h, hipEventSynchronize(h->ev_rr));
        std::copy(h->h_rr, h->h_rr + kRhs, rr);
        bool changed = false, all = true;
        for (int r = 0; r < kRhs; r++) {
            if (!act[r]) continue;
            iters[r]++;
            if (!std::isfinite(rr[r])) {
                *x_pending_out = true;
                return set_err(h, HOMO3D_DIVERGED, "non-finite CG residual norm (load case %d)", r);
            }
            if (rr[r] <= tol[r] * tol[r] * r0n2[r]) {
                act[r] = false;
                sc[5 * kRhs + r] = 0.0;
                changed = true;
            } else {
                all = false;
            }
        }
        if (all) break;
        if (changed)
            HIPCHK(h, hipMemcpyAsync(h->d_sc + 5 * kRhs, sc + 5 * kRhs, sizeof(double) * kRhs, hipMemcpyHostToDevice,
                                     h->stream));
        rc = vcycle(h, h->Z, h->Rr, 1, nullptr, true, zh, h->d_sc + 3 * kRhs, kPrecondSmallSweeps,
                    head);  // with rz_new
        if (rc) return rc;
        launch_pcg_beta(h->stream, h->d_sc);
    }
    *x_pending_out = x_pending;
    return HOMO3D_OK;
}

// on_device < 0: classify the pointer (device or managed memory -> 1, anything else -> 0)
int resolve_on_device(const void* p, int on_device) {
    if (on_device >= 0) return on_device;
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) ? 1 : 0;
}

int require_density(homo3d_handle* h) {
    if (h->cfg.solver == HOMO3D_SOLVER_NONE)
        return set_err(h, HOMO3D_BAD_ARG, "density-only handle: no solver hierarchy");
    if (!h->density_set) return set_err(h, HOMO3D_BAD_ARG, "density not set: call homo3d_update_density first");
    return HOMO3D_OK;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

void homo3d_default_config(homo3d_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->reso[0] = c->reso[1] = c->reso[2] = 64;
    c->youngs_modulus = 10.0;
    c->poisson_ratio = 0.3;
    c->emin = 1e-9;
    c->fem_rel_thres = 1e-2;
    c->max_iter = 200;
    c->solver = HOMO3D_SOLVER_MG;
    c->device = 0;
    c->sor_omega = 1.0;
    c->diag_clamp = 1e-4;
    c->precond_half = 1;
    c->coarse_half = 1;
    c->fine_mfma = 0;
}

int homo3d_create(const homo3d_config* cfg, homo3d_handle** out) {
    if (!cfg || !out) return HOMO3D_BAD_ARG;
    *out = nullptr;
    const int n = cfg->reso[0];
    if (cfg->reso[1] != n || cfg->reso[2] != n) return HOMO3D_NO_SUPPORT;
    int sizes[kMaxLevels];
    const int nl = cfg->solver == HOMO3D_SOLVER_NONE ? (n >= 1 && n <= 1024 ? 1 : -1) : level_sizes(n, sizes);
    if (cfg->solver == HOMO3D_SOLVER_NONE && nl == 1) sizes[0] = n;
    if (nl < 1 || (cfg->solver != HOMO3D_SOLVER_NONE && nl < 2)) return HOMO3D_NO_SUPPORT;
    if (!(cfg->poisson_ratio > -1 && cfg->poisson_ratio < 0.5) || !(cfg->youngs_modulus > 0)) return HOMO3D_BAD_ARG;
    homo3d_handle* h = new homo3d_handle();
    h->cfg = *cfg;
    h->nl = nl;
    const bool dens_only = cfg->solver == HOMO3D_SOLVER_NONE;
    auto fail = [&](int code) {
        homo3d_destroy(h);
        return code;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) return fail(HOMO3D_HIP_ERROR);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(HOMO3D_HIP_ERROR);
    if (dens_only) {
        h->lv[0].n = n;
        h->lv[0].N = (size_t)n * n * n;
        h->partial_elems = 21 * kDotBlocks;
        if (hipMalloc(&h->d_partial, sizeof(double) * h->partial_elems) || hipMalloc(&h->d_tmp, sizeof(double) * 64) ||
            hipMalloc(&h->d_ftmp, sizeof(float) * 4096) || hipMalloc(&h->d_gperm, sizeof(int) * 48) ||
            hipMalloc(&h->d_gflip, sizeof(int) * 48))
            return fail(HOMO3D_HIP_ERROR);
        *out = h;
        return HOMO3D_OK;
    }
    for (int l = 0; l < nl; l++) {
        Level& L = h->lv[l];
        L.n = sizes[l];
        L.N = (size_t)L.n * L.n * L.n;
        L.lex = l == nl - 1;
        const size_t bytes = sizeof(float) * 3 * kRhs * L.N;
        if (hipMalloc(&L.u, bytes) || hipMalloc(&L.f, bytes) || hipMalloc(&L.r, bytes) || hipMemset(L.u, 0, bytes) ||
            hipMemset(L.f, 0, bytes) || hipMemset(L.r, 0, bytes))
            return fail(HOMO3D_HIP_ERROR);
        if (l > 0) {
            L.st_half = cfg->coarse_half && l < nl - 1 ? 1 : 0;  // the coarsest stays FP32 for the dense solve
            if (hipMalloc(&L.st, sizeof(uint32_t) * (L.st_half ? 122 : 243) * L.N)) return fail(HOMO3D_HIP_ERROR);
        }
    }
    const size_t N0 = h->lv[0].N, vb = sizeof(float) * 3 * kRhs * N0;
    if (hipMalloc(&h->X, vb) || hipMalloc(&h->F, vb) || hipMalloc(&h->P, vb) || hipMalloc(&h->Q, vb) ||
        hipMalloc(&h->Z, vb) || hipMalloc(&h->Rr, vb) || hipMalloc(&h->d_w, sizeof(float) * N0) ||
        hipMalloc(&h->d_x, sizeof(float) * N0))
        return fail(HOMO3D_HIP_ERROR);
    if (hipMemset(h->X, 0, vb) || hipMemset(h->F, 0, vb)) return fail(HOMO3D_HIP_ERROR);
    // per-block partial sums (<= 6 per 256 vertices) and per-wave ones (6 per 64 vertices)
    h->partial_elems = std::max((size_t)21 * kDotBlocks, 6 * ((N0 + 63) / 64));
    if (hipMalloc(&h->U64, sizeof(double) * 3 * kRhs * N0) || hipMemset(h->U64, 0, sizeof(double) * 3 * kRhs * N0))
        return fail(HOMO3D_HIP_ERROR);
    if (hipMalloc(&h->d_partial, sizeof(double) * h->partial_elems) || hipMalloc(&h->d_sc, sizeof(double) * 64) ||
        hipMalloc(&h->d_tmp, sizeof(double) * 64) || hipMalloc(&h->d_ftmp, sizeof(float) * 4096))
        return fail(HOMO3D_HIP_ERROR);
    // element matrix, macro loads and the level-1 Galerkin table
    element_matrix(cfg->youngs_modulus, cfg->poisson_ratio, h->ke);
    std::vector<float> kef(576), kchi(6 * 24), G(64 * 27 * 9, 0.f);
    for (int i = 0; i < 576; i++) kef[i] = (float)h->ke[i];
    for (int s = 0; s < 6; s++)
        for (int i = 0; i < 24; i++) {
            double t = 0;
            for (int b = 0; b < 8; b++)
                for (int c = 0; c < 3; c++) t += h->ke[i * 24 + 3 * b + c] * macro_disp(b, s, c);
            kchi[s * 24 + i] = (float)t;
        }
    for (int c = 0; c < 64; c++) {
        const int ox = (c & 3) - 2, oy = ((c >> 2) & 3) - 2, oz = (c >> 4) - 2;
        for (int D = 0; D < 27; D++) {
            const int Dx = D % 3 - 1, Dy = D / 3 % 3 - 1, Dz = D / 9 - 1;
            double S[9] = {};
            for (int a = 0; a < 8; a++) {
                const double wa = hat(ox + (a & 1)) * hat(oy + ((a >> 1) & 1)) * hat(oz + (a >> 2));
                if (wa == 0) continue;
                for (int b = 0; b < 8; b++) {
                    const double wb = hat(ox + (b & 1) - 2 * Dx) * hat(oy + ((b >> 1) & 1) - 2 * Dy) *
                                      hat(oz + (b >> 2) - 2 * Dz);
                    if (wb == 0) continue;
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) S[i * 3 + j] += wa * wb * h->ke[(3 * a + i) * 24 + 3 * b + j];
                }
            }
            for (int k = 0; k < 9; k++) G[(c * 27 + D) * 9 + k] = (float)S[k];
        }
    }
    if (hipMalloc(&h->d_ke64, sizeof(double) * 576) || hipMalloc(&h->d_kchi64, sizeof(double) * 144))
        return fail(HOMO3D_HIP_ERROR);
    {
        std::vector<double> kchi64(144);
        for (int i = 0; i < 144; i++) {
            const int s_ = i / 24, r_ = i % 24;
            double t = 0;
            for (int b = 0; b < 8; b++)
                for (int c = 0; c < 3; c++) t += h->ke[r_ * 24 + 3 * b + c] * macro_disp(b, s_, c);
            kchi64[i] = t;
        }
        if (hipMemcpy(h->d_ke64, h->ke, sizeof(double) * 576, hipMemcpyHostToDevice) ||
            hipMemcpy(h->d_kchi64, kchi64.data(), sizeof(double) * 144, hipMemcpyHostToDevice))
            return fail(HOMO3D_HIP_ERROR);
    }
    if (hipMalloc(&h->d_ke, sizeof(float) * 576) || hipMalloc(&h->d_kchi, sizeof(float) * 144) ||
        hipMalloc(&h->d_G, sizeof(float) * G.size()) || hipMalloc(&h->d_gperm, sizeof(int) * 48) ||
        hipMalloc(&h->d_gflip, sizeof(int) * 48))
        return fail(HOMO3D_HIP_ERROR);
    double kl[576], km[576];
    // the matrix-core operand table is built whenever the element form can run, so that homo3d_autotune
    // can choose it; cfg->fine_mfma selects it from the start
    h->use_mfma = cfg->fine_mfma != 0;
    if (fine_apply_mfma_ok(n) && lame_tables(kl, km)) {
        std::vector<unsigned char> t16(kKeMfmaBytes);
        double lam, mu;
        lame_params(cfg->youngs_modulus, cfg->poisson_ratio, lam, mu);
        h->mfma.nt = ke_mfma_table(kl, km, lam, mu, t16.data(), &h->mfma.c0, &h->mfma.c1);
        if (hipMalloc(&h->d_ke16, kKeMfmaBytes) ||
            hipMemcpy(h->d_ke16, t16.data(), kKeMfmaBytes, hipMemcpyHostToDevice))
            return fail(HOMO3D_HIP_ERROR);
        h->mfma.table = h->d_ke16;
    }
    if (hipMemcpy(h->d_ke, kef.data(), sizeof(float) * 576, hipMemcpyHostToDevice) ||
        hipMemcpy(h->d_kchi, kchi.data(), sizeof(float) * 144, hipMemcpyHostToDevice) ||
        hipMemcpy(h->d_G, G.data(), sizeof(float) * G.size(), hipMemcpyHostToDevice))
        return fail(HOMO3D_HIP_ERROR);
    const int nc = sizes[nl - 1];
    h->mc = 3 * nc * nc * nc;
    if (hipMalloc(&h->d_Ainv, sizeof(double) * h->mc * h->mc) || hipMalloc(&h->d_cstat, sizeof(int)) ||
        hipMalloc(&h->d_cpiv, sizeof(double) * 4 * h->mc))
        return fail(HOMO3D_HIP_ERROR);
    if (hipDeviceSynchronize() != hipSuccess) return fail(HOMO3D_HIP_ERROR);
    *out = h;
    return HOMO3D_OK;
}

int homo3d_destroy(homo3d_handle* h) {
    if (!h) return HOMO3D_OK;
    // teardown: errors are not actionable here, so they are ignored
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->cstream) (void)hipStreamSynchronize(h->cstream);
    for (hipEvent_t e : h->pool) (void)hipEventDestroy(e);
    for (int l = 0; l < kMaxLevels; l++) {
        (void)hipFree(h->lv[l].u);
        (void)hipFree(h->lv[l].f);
        (void)hipFree(h->lv[l].r);
        (void)hipFree(h->lv[l].st);
    }
    void* ptrs[] = {h->d_x, h->U64, h->d_ke64, h->d_kchi64, h->d_ke, h->d_ke16, h->d_kchi, h->d_G, h->d_w, h->d_stage, h->d_Ainv, h->d_cstat, h->d_cpiv, h->X, h->F, h->P, h->Q, h->Z,
                    h->Rr, h->d_partial, h->d_sc, h->d_tmp, h->d_ftmp, h->d_gperm, h->d_gflip};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->cstream) (void)hipStreamDestroy(h->cstream);
    if (h->ev_gal) (void)hipEventDestroy(h->ev_gal);
    if (h->h_cstat) (void)hipHostFree(h->h_cstat);
    if (h->ev_coarse) (void)hipEventDestroy(h->ev_coarse);
    if (h->h_rr) (void)hipHostFree(h->h_rr);
    if (h->ev_rr) (void)hipEventDestroy(h->ev_rr);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return HOMO3D_OK;
}

const char* homo3d_last_error(const homo3d_handle* h) { return h ? h->err.c_str() : "null handle"; }

int homo3d_levels(const homo3d_handle* h, int* sizes, int max_levels) {
    if (!h) return -1;
    for (int l = 0; l < h->nl && l < max_levels; l++) sizes[l] = h->lv[l].n;
    return h->nl;
}

void* homo3d_stream(homo3d_handle* h) { return h ? (void*)h->stream : nullptr; }

// Homogenization::update (homogenization.cpp:79-85): cell weights w = Emin + x, Galerkin stencils of
// every coarse level (MG::updateStencils, MG.cpp:341-354) and the coarsest FP64 factorisation.
int homo3d_update_density(homo3d_handle* h, const float* x, int pitch, int on_device) {
    if (!h || !x) return HOMO3D_BAD_ARG;
    if (h->cfg.solver == HOMO3D_SOLVER_NONE) return set_err(h, HOMO3D_BAD_ARG, "density-only handle");
    on_device = resolve_on_device(x, on_device);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    const float* dx = nullptr;
    int rc = stage_in(h, x, pitch, on_device, &dx);
    if (rc) return rc;
    launch_set_weights(h->stream, n, pitch, h->cfg.emin, dx, h->d_w, h->d_x);
    // a factorisation still running from the previous update reads the coarsest stencil rewritten below
    if (h->ev_coarse) HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_coarse, 0));
    const float clamp = (float)h->cfg.diag_clamp;
    launch_galerkin_fine(h->stream, n, h->lv[1].lex, h->d_G, h->d_ke, h->d_w, h->lv[1].st, h->lv[1].st_half, clamp);
    for (int l = 2; l < h->nl; l++)
        launch_galerkin_coarse(h->stream, h->lv[l - 1].n, h->lv[l].lex, h->lv[l - 1].st, h->lv[l - 1].st_half,
                               h->lv[l].st, h->lv[l].st_half, clamp);
    // coarsest dense system: assembled and inverted on the GPU; its status is read at the
    // first coarse solve (coarse_factor)
    const Level& Lc = h->lv[h->nl - 1];
    if (!h->h_cstat) {
        HIPCHK(h, hipHostMalloc((void**)&h->h_cstat, sizeof(int), hipHostMallocDefault));
        HIPCHK(h, hipEventCreateWithFlags(&h->ev_coarse, hipEventDisableTiming));
        HIPCHK(h, hipEventCreateWithFlags(&h->ev_gal, hipEventDisableTiming));
        HIPCHK(h, hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking));
    }
    // behind everything queued so far on the main stream (the Galerkin kernels, and earlier coarse
    // solves that read the inverse)
    HIPCHK(h, hipEventRecord(h->ev_gal, h->stream));
    HIPCHK(h, hipStreamWaitEvent(h->cstream, h->ev_gal, 0));
    launch_coarse_factor(h->cstream, Lc.n, static_cast<const float*>(Lc.st), h->d_Ainv, h->d_cpiv, h->d_cstat);
    HIPCHK(h, hipMemcpyAsync(h->h_cstat, h->d_cstat, sizeof(int), hipMemcpyDeviceToHost, h->cstream));
    HIPCHK(h, hipEventRecord(h->ev_coarse, h->cstream));
    h->coarse_pending = true;
    h->density_set = true;
    return HOMO3D_OK;
}

// The six macro-strain solves of Homogenization::elasticMatrix (homogenization.cu:471-477), run in
// lockstep.  HOMO3D_SOLVER_MG: stand-alone V(1,1) cycles until the level-0 residual computed after
// pre-smoothing satisfies ||r||/(||f||+1e-10) <= tol for every load case (MG::solveEquation,
// MG.cpp:113-203, incl. one reset of the guess on explosion).  HOMO3D_SOLVER_PCG: MG-preconditioned
// CG with a symmetric V-cycle and FP64 inner products.  Solutions stay in the handle (warm start).
int homo3d_solve(homo3d_handle* h, double tol, int warm, homo3d_solve_stats* stats) {
    if (!h) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    auto t0 = std::chrono::steady_clock::now();
    const Level& L0 = h->lv[0];
    const size_t M = 3 * L0.N;  // floats per load case
    const size_t vb = sizeof(float) * kRhs * M;
    if (tol <= 0) tol = h->cfg.fem_rel_thres;
    const int maxit = h->cfg.max_iter;
    launch_macro_rhs(h->stream, L0.n, h->d_kchi, h->d_w, h->F);
    double fn2[kRhs];
    int iters[kRhs] = {0, 0, 0, 0, 0, 0};
    bool done[kRhs] = {false, false, false, false, false, false};
    int status = HOMO3D_OK;
    if (h->cfg.solver == HOMO3D_SOLVER_MG) {
        if (!warm) HIPCHK(h, hipMemsetAsync(h->X, 0, vb, h->stream));
        launch_dot6(h->stream, M, h->F, h->F, h->d_partial, kDotBlocks, h->d_tmp);
        HIPCHK(h, hipMemcpyAsync(fn2, h->d_tmp, sizeof(fn2), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        int overflow[kRhs] = {2, 2, 2, 2, 2, 2};
        int it = 0;
        bool all = false;
        while (!all && it++ < maxit) {
            rc = vcycle(h, h->X, h->F, 0, h->d_sc);
            if (rc) return rc;
            double rn2[kRhs];
            HIPCHK(h, hipMemcpyAsync(rn2, h->d_sc, sizeof(rn2), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            all = true;
            for (int r = 0; r < kRhs; r++) {
                if (done[r]) continue;
                const double rel = std::sqrt(rn2[r]) / (std::sqrt(fn2[r]) + 1e-10);
                iters[r] = it;
                if (rel <= tol) {
                    done[r] = true;
                    continue;
                }
                if (rel > 10 || it >= maxit - 1) {
                    if (--overflow[r] <= 0) {
                        status = HOMO3D_DIVERGED;
                        done[r] = true;
                        continue;
                    }
                    HIPCHK(h, hipMemsetAsync(h->X + r * M, 0, sizeof(float) * M, h->stream));
                }
                all = false;
            }
        }
        for (int r = 0; r < kRhs; r++)
            if (!done[r]) status = std::max(status, (int)HOMO3D_NOT_CONVERGED);
    } else {
        // FP64 iterative refinement: r = f(w) - K U64 in FP64, inner FP32 MG-PCG on K d = r, U64 += d.
        // A cold start keeps U64 = 0 implicit: the first residual is f itself (which also gives |f|^2)
        // and the first correction is copied rather than added.
        const size_t M6 = kRhs * M;
        bool u_zero = !warm;
        if (warm) {
            launch_rhs_norm64(h->stream, L0.n, h->d_kchi64, h->d_x, h->cfg.emin, h->d_partial, h->d_tmp);
            HIPCHK(h, hipMemcpyAsync(fn2, h->d_tmp, sizeof(fn2), hipMemcpyDeviceToHost, h->stream));
            if (h->x_dirty) launch_widen_f64(h->stream, L0.N, h->X, h->U64);
            launch_zero_pin64(h->stream, L0.N, h->U64);
        }
        h->x_dirty = false;
        double rel[kRhs];
        bool warm_checked = false;
        for (int outer = 0;; outer++) {
            double rn2[kRhs];
            launch_fine_residual64(h->stream, L0.n, h->d_ke64, h->d_kchi64, h->d_x, h->cfg.emin, h->U64, h->Rr,
                                   h->d_partial, h->d_tmp, u_zero ? 0 : 1);
            HIPCHK(h, hipMemcpyAsync(rn2, h->d_tmp, sizeof(rn2), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            if (outer == 0 && u_zero) std::copy(rn2, rn2 + kRhs, fn2);
            // A warm start whose FP64 residual is not below ||f|| is worse than none: that load case
            // restarts from zero (the residual is formed once more with its slice of U64 cleared).
            if (outer == 0 && warm && !warm_checked) {
                warm_checked = true;
                int reset = 0, nonzero = 0;
                for (int r = 0; r < kRhs; r++)
                    if (fn2[r] > 0) {
                        nonzero++;
                        if (!(rn2[r] < fn2[r])) reset++;
                    }
                if (reset > 0 && reset == nonzero) {
                    // every load case falls back (those with f = 0 have u = 0): the cold start, whose
                    // first residual is f itself and whose first correction is copied into U64
                    u_zero = true;
                    outer = -1;
                    continue;
                }
                if (reset > 0) {
                    for (int r = 0; r < kRhs; r++)
                        if (fn2[r] > 0 && !(rn2[r] < fn2[r])) launch_zero_case64(h->stream, L0.N, r, h->U64);
                    outer = -1;
                    continue;
                }
            }
            bool all = true, finite = true;
            for (int r = 0; r < kRhs; r++) {
                rel[r] = fn2[r] > 0 ? std::sqrt(rn2[r] / fn2[r]) : 0.0;
                done[r] = rel[r] <= tol;
                all = all && done[r];
                finite = finite && std::isfinite(rn2[r]) && std::isfinite(fn2[r]);
            }
            if (!finite) {
                status = HOMO3D_DIVERGED;
                u_zero = false;
                break;
            }
            if (all || outer >= kMaxOuter) break;
            int total = 0;
            for (int r = 0; r < kRhs; r++) total += iters[r];
            if (total >= maxit * kRhs) break;
            double tol_in[kRhs];
            for (int r = 0; r < kRhs; r++) tol_in[r] = std::max(kInnerTol, 0.5 * tol / std::max(rel[r], 1e-300));
            HIPCHK(h, hipMemsetAsync(h->X, 0, sizeof(float) * M6, h->stream));
            bool x_pending = false;
            rc = pcg_inner(h, rn2, tol_in, done, maxit, iters, &x_pending);
            if (rc == HOMO3D_DIVERGED) {
                status = rc;
                for (int r = 0; r < kRhs; r++) rel[r] = std::numeric_limits<double>::quiet_NaN();
                u_zero = false;
                break;
            }
            if (rc) return rc;
            // U64 = or += the correction x + alpha p (the last CG step's x update, fused)
            launch_refine_accumulate(h->stream, L0.N, h->d_sc, x_pending ? 1 : 0, h->X, h->P, h->U64, u_zero ? 0 : 1);
            u_zero = false;
        }
        if (u_zero) HIPCHK(h, hipMemsetAsync(h->U64, 0, sizeof(double) * M6, h->stream));
        for (int r = 0; r < kRhs; r++)
            if (!done[r]) status = std::max(status, (int)HOMO3D_NOT_CONVERGED);
        launch_round_f32(h->stream, L0.N, h->U64, h->X);
        HIPCHK(h, hipStreamSynchronize(h->stream));
        prof_flush(h);
        if (status == HOMO3D_DIVERGED) set_err(h, status, "non-finite residual in the refined MG-PCG solve");
        if (stats) {
            stats->status = status;
            for (int r = 0; r < kRhs; r++) {
                stats->iters[r] = iters[r];
                stats->relres[r] = rel[r];
            }
            stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return status;
    }
    // true final residual per load case
    op_residual(h, 0, h->X, h->F, h->Rr);
    double rr2[kRhs];
    launch_dot6(h->stream, M, h->Rr, h->Rr, h->d_partial, kDotBlocks, h->d_tmp);
    HIPCHK(h, hipMemcpyAsync(rr2, h->d_tmp, sizeof(rr2), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_flush(h);
    if (stats) {
        stats->status = status;
        for (int r = 0; r < kRhs; r++) {
            stats->iters[r] = iters[r];
            stats->relres[r] = fn2[r] > 0 ? std::sqrt(rr2[r] / fn2[r]) : 0.0;
        }
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (status == HOMO3D_DIVERGED) set_err(h, status, "MG numerical explode");
    return status;
}

// Autotuner of the finest-level operator's tile form: the row form (512-thread workgroups over
// 128 x 2 x 2 parity tiles, neighbour gathers) against the element form on the matrix cores
// (32 x 4 x 4 vertex tiles, one 16x16x32 MMA per cell group), timed with HIP events on this handle's
// grid and density (r = K z on the packed-FP16 iterate, the form every CG pass and V-cycle residual
// runs); the faster one is used from then on.
int homo3d_autotune(homo3d_handle* h, homo3d_tune_report* rep) {
    if (!h) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    const Level& L0 = h->lv[0];
    homo3d_tune_report r{};
    r.row_ms = r.mfma_ms = -1;
    const bool half = L0.n % 8 == 0;
    const bool can_mfma = h->d_ke16 != nullptr && half;
    if (!half) {
        h->use_mfma = false;
        if (rep) *rep = r;
        return HOMO3D_OK;
    }
    launch_pack_half(h->stream, L0.N, h->X, h->Z);  // a defined operand (the current iterate)
    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    auto time_form = [&](bool mfma_form, double* ms) -> int {
        h->use_mfma = mfma_form;
        for (int i = 0; i < 3; i++) launch_fine_apply(h->stream, L0.n, h->d_ke, h->mf(), h->d_w, h->Z, 1, nullptr, L0.r);
        HIPCHK(h, hipEventRecord(e0, h->stream));
        const int reps = 10;
        for (int i = 0; i < reps; i++)
            launch_fine_apply(h->stream, L0.n, h->d_ke, h->mf(), h->d_w, h->Z, 1, nullptr, L0.r);
        HIPCHK(h, hipEventRecord(e1, h->stream));
        HIPCHK(h, hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(h, hipEventElapsedTime(&t, e0, e1));
        *ms = t / reps;
        return HOMO3D_OK;
    };
    rc = time_form(false, &r.row_ms);
    if (!rc && can_mfma) rc = time_form(true, &r.mfma_ms);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    h->use_mfma = can_mfma && r.mfma_ms < r.row_ms;
    r.chosen_mfma = h->use_mfma ? 1 : 0;
    if (rep) *rep = r;
    return HOMO3D_OK;
}

// FP64 relres ||f - K u|| / ||f|| per load case of the current FP64 iterate (the caller's level-0 u when
// that was written last), without solving: the measure a warm start is judged by.
int homo3d_stage_tiles(homo3d_handle* h, int on) {
    if (!h) return HOMO3D_BAD_ARG;
    h->stage_tiles = on != 0;
    return HOMO3D_OK;
}

int homo3d_residual_norms(homo3d_handle* h, double relres[6]) {
    if (!h || !relres) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    if (h->cfg.solver != HOMO3D_SOLVER_PCG) return set_err(h, HOMO3D_BAD_ARG, "needs the refined solver's FP64 iterate");
    const Level& L0 = h->lv[0];
    double fn2[kRhs], rn2[kRhs];
    launch_rhs_norm64(h->stream, L0.n, h->d_kchi64, h->d_x, h->cfg.emin, h->d_partial, h->d_tmp);
    HIPCHK(h, hipMemcpyAsync(fn2, h->d_tmp, sizeof(fn2), hipMemcpyDeviceToHost, h->stream));
    if (h->x_dirty) launch_widen_f64(h->stream, L0.N, h->X, h->U64);
    launch_zero_pin64(h->stream, L0.N, h->U64);
    h->x_dirty = false;
    launch_fine_residual64(h->stream, L0.n, h->d_ke64, h->d_kchi64, h->d_x, h->cfg.emin, h->U64, h->Rr, h->d_partial,
                           h->d_tmp, 1);
    HIPCHK(h, hipMemcpyAsync(rn2, h->d_tmp, sizeof(rn2), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int r = 0; r < kRhs; r++) relres[r] = fn2[r] > 0 ? std::sqrt(rn2[r] / fn2[r]) : 0.0;
    return HOMO3D_OK;
}

// C^H from the current solutions (Homogenization::elasticMatrix, homogenization.cu:478-575).
int homo3d_elastic_matrix(homo3d_handle* h, double C[36]) {
    if (!h || !C) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    const Level& L0 = h->lv[0];
    launch_elastic_matrix(h->stream, L0.n, h->d_ke, h->d_w, h->X, h->d_partial, kDotBlocks, h->d_tmp);
    double acc[21];
    HIPCHK(h, hipMemcpyAsync(acc, h->d_tmp, sizeof(acc), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int k = 0;
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++) {
            C[i * 6 + j] = C[j * 6 + i] = acc[k] / (double)L0.N;
            k++;
        }
    return HOMO3D_OK;
}

// dJ/dx_e for dJ/dC = dC (Homogenization::Sensitivity, homogenization.cu:860-948).  lexi_order: the
// lexicographic pitched layout; otherwise the level's own cell order (parity-blocked, n^3 unpitched),
// the counterpart of the reference's GS-order output for lexiOrder = false.
static int sensitivity(homo3d_handle* h, const float dC[36], float* sens, int pitch, int lexi_order, int on_device) {
    if (!h || !dC || !sens) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(sens, on_device);
    int rc = require_density(h);
    if (rc) return rc;
    const int n = h->lv[0].n;
    if (lexi_order && pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    if (!lexi_order) pitch = 0;
    float wdc[21];
    int k = 0;
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++) wdc[k++] = dC[i * 6 + j] + (i != j ? dC[j * 6 + i] : 0.f);
    HIPCHK(h, hipMemcpyAsync(h->d_ftmp, wdc, sizeof(wdc), hipMemcpyHostToDevice, h->stream));
    float* dst = sens;
    const size_t elems = lexi_order ? (size_t)pitch * n * n : h->lv[0].N;
    if (!on_device) {
        rc = ensure_stage(h, elems);
        if (rc) return rc;
        dst = h->d_stage;
    }
    launch_sensitivity(h->stream, n, pitch, h->d_ke, h->X, h->d_ftmp, dst);
    if (!on_device) {
        if (lexi_order)
            rc = stage_out(h, sens, dst, pitch, 0);
        else
            HIPCHK(h, hipMemcpyAsync(sens, dst, sizeof(float) * elems, hipMemcpyDeviceToHost, h->stream));
        if (rc) return rc;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_sensitivity(homo3d_handle* h, const float dC[36], float* sens, int pitch, int on_device) {
    return sensitivity(h, dC, sens, pitch, 1, on_device);
}

int homo3d_sensitivity_order(homo3d_handle* h, const float dC[36], float* sens, int pitch, int lexi_order,
                             int on_device) {
    return sensitivity(h, dC, sens, pitch, lexi_order, on_device);
}

static float* field_ptr(homo3d_handle* h, int level, int field) {
    Level& L = h->lv[level];
    if (level == 0) {
        if (field == HOMO3D_FIELD_U) return h->X;
        if (field == HOMO3D_FIELD_F) return h->F;
        return L.r;
    }
    return field == HOMO3D_FIELD_U ? L.u : (field == HOMO3D_FIELD_F ? L.f : L.r);
}

int homo3d_field_copy(homo3d_handle* h, int level, int field, float* host, int to_device) {
    if (!h || !host || level < 0 || level >= h->nl || field < 0 || field > 2) return HOMO3D_BAD_ARG;
    if (h->cfg.solver == HOMO3D_SOLVER_NONE) return set_err(h, HOMO3D_BAD_ARG, "density-only handle");
    float* d = field_ptr(h, level, field);
    const Level& L = h->lv[level];
    const size_t bytes = sizeof(float) * 3 * kRhs * L.N;
    std::vector<float> tmp(3 * kRhs * L.N);
    if (to_device) {
        permute_planes(L, host, tmp.data(), 3 * kRhs, true);
        HIPCHK(h, hipMemcpyAsync(d, tmp.data(), bytes, hipMemcpyHostToDevice, h->stream));
        if (level == 0 && field == HOMO3D_FIELD_U) h->x_dirty = true;
        HIPCHK(h, hipStreamSynchronize(h->stream));
    } else {
        HIPCHK(h, hipMemcpyAsync(tmp.data(), d, bytes, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        permute_planes(L, tmp.data(), host, 3 * kRhs, false);
    }
    return HOMO3D_OK;
}

int homo3d_solution_copy(homo3d_handle* h, double* host, int to_device) {
    if (!h || !host) return HOMO3D_BAD_ARG;
    if (h->cfg.solver == HOMO3D_SOLVER_NONE) return set_err(h, HOMO3D_BAD_ARG, "density-only handle");
    const Level& L = h->lv[0];
    const size_t bytes = sizeof(double) * 3 * kRhs * L.N;
    // host [18][n^3] planes <-> the iterate's pair planes in level storage order
    std::vector<double> tmp(3 * kRhs * L.N), pairs(3 * kRhs * L.N);
    if (to_device) {
        permute_planes(L, host, tmp.data(), 3 * kRhs, true);
        for (int c = 0; c < 3 * kRhs; c++)
            for (size_t v = 0; v < L.N; v++) pairs[u64_index(c, v, L.N)] = tmp[c * L.N + v];
        HIPCHK(h, hipMemcpyAsync(h->U64, pairs.data(), bytes, hipMemcpyHostToDevice, h->stream));
        launch_zero_pin64(h->stream, L.N, h->U64);
        launch_round_f32(h->stream, L.N, h->U64, h->X);
        h->x_dirty = false;
        HIPCHK(h, hipStreamSynchronize(h->stream));
    } else {
        HIPCHK(h, hipMemcpyAsync(pairs.data(), h->U64, bytes, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (int c = 0; c < 3 * kRhs; c++)
            for (size_t v = 0; v < L.N; v++) tmp[c * L.N + v] = pairs[u64_index(c, v, L.N)];
        permute_planes(L, tmp.data(), host, 3 * kRhs, false);
    }
    return HOMO3D_OK;
}

int homo3d_macro_rhs(homo3d_handle* h) {
    if (!h) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    launch_macro_rhs(h->stream, h->lv[0].n, h->d_kchi, h->d_w, h->F);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_residual(homo3d_handle* h, int level) {
    if (!h || level < 0 || level >= h->nl) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    if (level == 0) launch_zero_pin(h->stream, h->lv[0].N, h->X, 0);
    op_residual(h, level, field_ptr(h, level, 0), field_ptr(h, level, 1), h->lv[level].r);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_flush(h);
    return HOMO3D_OK;
}

int homo3d_apply_K(homo3d_handle* h) {
    if (!h) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    launch_zero_pin(h->stream, h->lv[0].N, h->X, 0);
    ProfEvent ev;
    prof_begin(h, kProfApply, ev);
    launch_fine_apply(h->stream, h->lv[0].n, h->d_ke, nullptr, h->d_w, h->X, 0, nullptr, h->lv[0].r);
    prof_end(h, ev);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_flush(h);
    return HOMO3D_OK;
}

// Finest-level operator and smoother on packed-FP16 storage (the form the PCG preconditioner runs),
// on an FP16-rounded copy of the level-0 u held in the handle's z buffer.
static int fp16_ready(homo3d_handle* h) {
    if (!h) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    if (h->lv[0].n % 8) return set_err(h, HOMO3D_NO_SUPPORT, "packed FP16 fields need n %% 8 == 0 (n = %d)", h->lv[0].n);
    return HOMO3D_OK;
}

int homo3d_apply_K16(homo3d_handle* h, int residual) {
    int rc = fp16_ready(h);
    if (rc) return rc;
    const Level& L = h->lv[0];
    launch_pack_half(h->stream, L.N, h->X, h->Z);
    launch_zero_pin(h->stream, L.N, h->Z, 1);
    ProfEvent ev;
    prof_begin(h, residual ? kProfResidual : kProfApply, ev);
    launch_fine_apply(h->stream, L.n, h->d_ke, h->mf(), h->d_w, h->Z, 1, residual ? h->F : nullptr,
                      L.r);
    prof_end(h, ev);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_flush(h);
    return HOMO3D_OK;
}

int homo3d_smooth16(homo3d_handle* h, double omega, int reverse) {
    int rc = fp16_ready(h);
    if (rc) return rc;
    const Level& L = h->lv[0];
    launch_pack_half(h->stream, L.N, h->X, h->Z);
    op_smooth(h, 0, h->Z, h->F, (float)omega, reverse, false, 1);
    launch_unpack_half(h->stream, L.N, h->Z, h->X);
    h->x_dirty = true;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_flush(h);
    return HOMO3D_OK;
}

int homo3d_presmooth16(homo3d_handle* h, double omega) {
    int rc = fp16_ready(h);
    if (rc) return rc;
    if (!residual_after_zero_sweep_ok(h)) return set_err(h, HOMO3D_NO_SUPPORT, "needs fused colour pairs at level 0");
    const Level& L = h->lv[0];
    op_smooth(h, 0, h->Z, h->F, (float)omega, 0, true, 1, nullptr, L.r);
    op_residual_after_zero_sweep(h, 0, h->Z, 1, L.r);
    launch_unpack_half(h->stream, L.N, h->Z, h->X);
    h->x_dirty = true;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_flush(h);
    return HOMO3D_OK;
}

int homo3d_presmooth(homo3d_handle* h, int level, double omega) {
    if (!h || level < 1 || level >= h->nl - 1) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    if (!residual_after_zero_sweep_ok(h, level))
        return set_err(h, HOMO3D_NO_SUPPORT, "level %d has no colour pairs", level);
    Level& L = h->lv[level];
    op_smooth(h, level, L.u, L.f, (float)omega, 0, true, 0, nullptr, L.r);
    op_residual_after_zero_sweep(h, level, L.u, 0, L.r);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_smooth(homo3d_handle* h, int level, double omega, int reverse) {
    if (!h || level < 0 || level >= h->nl - 1) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    op_smooth(h, level, field_ptr(h, level, 0), field_ptr(h, level, 1), (float)omega, reverse);
    if (level == 0) h->x_dirty = true;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_flush(h);
    return HOMO3D_OK;
}

int homo3d_restrict(homo3d_handle* h, int level) {
    if (!h || level < 1 || level >= h->nl) return HOMO3D_BAD_ARG;
    if (h->cfg.solver == HOMO3D_SOLVER_NONE) return set_err(h, HOMO3D_BAD_ARG, "density-only handle");
    launch_restrict(h->stream, h->lv[level - 1].n, h->lv[level].lex, h->lv[level - 1].r, field_ptr(h, level, 1));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_prolong(homo3d_handle* h, int level) {
    if (!h || level < 0 || level >= h->nl - 1) return HOMO3D_BAD_ARG;
    if (h->cfg.solver == HOMO3D_SOLVER_NONE) return set_err(h, HOMO3D_BAD_ARG, "density-only handle");
    launch_prolong(h->stream, h->lv[level].n, h->lv[level + 1].lex, field_ptr(h, level + 1, 0), 0, field_ptr(h, level, 0), 0,
                   level == 0 ? 1 : 0);
    if (level == 0) h->x_dirty = true;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_coarse_solve(homo3d_handle* h) {
    if (!h) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    const Level& Lc = h->lv[h->nl - 1];
    rc = coarse_factor(h);
    if (rc) return rc;
    launch_coarse_solve(h->stream, Lc.n, h->d_Ainv, Lc.f, Lc.u);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_vcycle(homo3d_handle* h, int symmetric) {
    if (!h) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    rc = vcycle(h, h->X, h->F, symmetric, nullptr);
    h->x_dirty = true;
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_flush(h);
    return HOMO3D_OK;
}

int homo3d_vcycle_sweeps(homo3d_handle* h, double omega, int pre, int post) {
    if (!h || pre < 0 || post < 0 || pre > 64 || post > 64 || !(omega > 0 && omega < 2)) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    rc = vcycle_sweeps(h, (float)omega, pre, post);
    h->x_dirty = true;
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_flush(h);
    return HOMO3D_OK;
}

int homo3d_get_stencil(homo3d_handle* h, int level, float* host) {
    if (!h || !host || level < 1 || level >= h->nl) return HOMO3D_BAD_ARG;
    if (h->cfg.solver == HOMO3D_SOLVER_NONE) return set_err(h, HOMO3D_BAD_ARG, "density-only handle");
    const Level& L = h->lv[level];
    const int words = L.st_half ? 122 : 243;
    std::vector<uint32_t> soa((size_t)words * L.N);
    HIPCHK(h, hipMemcpyAsync(soa.data(), L.st, sizeof(uint32_t) * soa.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    auto entry = [&](int k, size_t sv) -> float {
        const uint32_t wv = soa[(size_t)(L.st_half ? k >> 1 : k) * L.N + sv];
        if (!L.st_half) {
            float f;
            std::memcpy(&f, &wv, sizeof(f));
            return f;
        }
        const uint16_t bits = (uint16_t)((k & 1) ? wv >> 16 : wv & 0xFFFFu);
        _Float16 hv;
        std::memcpy(&hv, &bits, sizeof(hv));
        return (float)hv;
    };
    for (int z = 0; z < L.n; z++)
        for (int y = 0; y < L.n; y++)
            for (int x = 0; x < L.n; x++) {
                const size_t v = x + (size_t)L.n * (y + (size_t)L.n * z), sv = store_index(L, x, y, z);
                for (int k = 0; k < 243; k++) host[v * 243 + k] = entry(k, sv);
            }
    return HOMO3D_OK;
}

int homo3d_symmetrize(homo3d_handle* h, float* field, int pitch, int sym, int on_device) {
    if (!h || !field) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(field, on_device);
    if (sym == HOMO3D_SYM_NONE) return HOMO3D_OK;
    if (sym < 0 || sym > 3) return set_err(h, HOMO3D_BAD_ARG, "unknown symmetry %d", sym);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    std::vector<int> perm, flip;
    symmetry_group(sym, perm, flip);
    HIPCHK(h, hipMemcpyAsync(h->d_gperm, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_gflip, flip.data(), sizeof(int) * flip.size(), hipMemcpyHostToDevice, h->stream));
    const size_t elems = (size_t)pitch * n * n;
    int rc = ensure_stage(h, 2 * elems);
    if (rc) return rc;
    float* in = h->d_stage;
    float* out = h->d_stage + elems;
    if (on_device)
        HIPCHK(h, hipMemcpyAsync(in, field, sizeof(float) * elems, hipMemcpyDeviceToDevice, h->stream));
    else
        HIPCHK(h, hipMemcpyAsync(in, field, sizeof(float) * elems, hipMemcpyHostToDevice, h->stream));
    launch_symmetrize(h->stream, n, pitch, (int)perm.size(), h->d_gperm, h->d_gflip, in, out);
    launch_copy_pitched(h->stream, n, pitch, pitch, out, in);
    HIPCHK(h, hipMemcpyAsync(field, in, sizeof(float) * elems, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

// Sigmund filter in place on sens (OCOptimizer::filterSens, OCOptimizer.cu:185-195).
int homo3d_oc_filter(homo3d_handle* h, float* sens, const float* rho, int pitch, double radius, int periodic,
                     int on_device) {
    if (!h || !sens || !rho) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(sens, on_device);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    const int hw = (int)(radius + 0.5);
    const int side = 2 * hw + 1;
    if (side * side * side > 4096) return set_err(h, HOMO3D_NO_SUPPORT, "filter radius too large");
    std::vector<float> wk(side * side * side);
    double W = 0;
    std::vector<double> wd(wk.size());
    for (int i = 0; i < side * side * side; i++) {
        const int o[3] = {i % side - hw, (i / side) % side - hw, i / (side * side) - hw};
        const double r2 = (o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) / (radius * radius);
        wd[i] = r2 < 1 ? 1 - std::sqrt(r2) : 0.0;
        W += wd[i];
    }
    for (size_t i = 0; i < wk.size(); i++) wk[i] = (float)(wd[i] / W);
    HIPCHK(h, hipMemcpyAsync(h->d_ftmp, wk.data(), sizeof(float) * wk.size(), hipMemcpyHostToDevice, h->stream));
    const size_t elems = (size_t)pitch * n * n;
    int rc = ensure_stage(h, 3 * elems);
    if (rc) return rc;
    float* ds = h->d_stage;
    float* dr = h->d_stage + elems;
    float* dout = h->d_stage + 2 * elems;
    const hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIPCHK(h, hipMemcpyAsync(ds, sens, sizeof(float) * elems, kin, h->stream));
    HIPCHK(h, hipMemcpyAsync(dr, rho, sizeof(float) * elems, kin, h->stream));
    launch_oc_filter(h->stream, n, pitch, hw, periodic, h->d_ftmp, ds, dr, dout);
    launch_copy_pitched(h->stream, n, pitch, pitch, dout, ds);
    HIPCHK(h, hipMemcpyAsync(sens, ds, sizeof(float) * elems, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

// Elementwise density maps of the optimisation drivers: the reference forms rho^3, its chain rule and the
// volume sum with its expression templates on the device (Framework.cu:137-200 rho.pow(3), Tensor::sum).
int homo3d_density_pow(homo3d_handle* h, const float* rho, float* x, int pitch, double p, int on_device) {
    if (!h || !rho || !x) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(x, on_device);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    const size_t elems = (size_t)pitch * n * n;
    const float* dr = rho;
    float* dx = x;
    if (!on_device) {
        int rc = ensure_stage(h, 2 * elems);
        if (rc) return rc;
        HIPCHK(h, hipMemcpyAsync(h->d_stage, rho, sizeof(float) * elems, hipMemcpyHostToDevice, h->stream));
        dr = h->d_stage;
        dx = h->d_stage + elems;
    }
    launch_density_pow(h->stream, n, pitch, (float)p, dr, dx);
    if (!on_device) {
        int rc = stage_out(h, x, dx, pitch, 0);
        if (rc) return rc;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_density_pow_grad(homo3d_handle* h, const float* rho, const float* dx, float* drho, int pitch, double p,
                            int on_device) {
    if (!h || !rho || !dx || !drho) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(drho, on_device);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    const size_t elems = (size_t)pitch * n * n;
    const float *dr = rho, *dd = dx;
    float* dg = drho;
    if (!on_device) {
        int rc = ensure_stage(h, 3 * elems);
        if (rc) return rc;
        HIPCHK(h, hipMemcpyAsync(h->d_stage, rho, sizeof(float) * elems, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->d_stage + elems, dx, sizeof(float) * elems, hipMemcpyHostToDevice, h->stream));
        dr = h->d_stage;
        dd = h->d_stage + elems;
        dg = h->d_stage + 2 * elems;
    }
    launch_density_pow_grad(h->stream, n, pitch, (float)p, dr, dd, dg);
    if (!on_device) {
        int rc = stage_out(h, drho, dg, pitch, 0);
        if (rc) return rc;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_field_sum(homo3d_handle* h, const float* a, int pitch, double* sum, int on_device) {
    if (!h || !a || !sum) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(a, on_device);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    const float* da = a;
    if (!on_device) {
        int rc = stage_in(h, a, pitch, 0, &da);
        if (rc) return rc;
    }
    constexpr int nb = 1024;
    launch_field_sum(h->stream, n, pitch, da, h->d_partial, nb);
    std::vector<double> part(nb);
    HIPCHK(h, hipMemcpyAsync(part.data(), h->d_partial, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double t = 0;
    for (double v : part) t += v;
    *sum = t;
    return HOMO3D_OK;
}

int homo3d_field_fill(homo3d_handle* h, float* a, int pitch, float value, int on_device) {
    if (!h || !a) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(a, on_device);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    if (!on_device) {
        for (int z = 0; z < n; z++)
            for (int y = 0; y < n; y++)
                for (int x = 0; x < n; x++) a[x + ((size_t)y + (size_t)z * n) * pitch] = value;
        return HOMO3D_OK;
    }
    launch_field_fill(h->stream, n, pitch, value, a);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

// Device memory for host code that does not include the HIP headers (the optimisation drivers keep their
// density fields resident), on the calling thread's current device.
int homo3d_device_alloc(void** ptr, size_t bytes) {
    if (!ptr) return HOMO3D_BAD_ARG;
    *ptr = nullptr;
    if (bytes == 0) return HOMO3D_OK;
    return hipMalloc(ptr, bytes) == hipSuccess ? HOMO3D_OK : HOMO3D_HIP_ERROR;
}

int homo3d_device_free(void* ptr) { return !ptr || hipFree(ptr) == hipSuccess ? HOMO3D_OK : HOMO3D_HIP_ERROR; }

int homo3d_device_copy(void* dst, const void* src, size_t bytes, int kind) {
    if (kind < 0 || kind > 2 || (bytes && (!dst || !src))) return HOMO3D_BAD_ARG;
    const hipMemcpyKind k[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    return bytes == 0 || hipMemcpy(dst, src, bytes, k[kind]) == hipSuccess ? HOMO3D_OK : HOMO3D_HIP_ERROR;
}

// OC update of rho in place (OCOptimizer::update, OCOptimizer.cu:218-247).
// Density filter conv(rho, radial_convker_t<Spline4 | Linear>(radius)) (TensorExpression.h:1007-1046,
// TensorExpression.cuh:288-368), forward (backward = 0) or its adjoint (backward = 1: in holds dJ/dout,
// out receives dJ/din).  kernel: 0 = Spline4, 1 = Linear.  in and out may be the same array.
int homo3d_density_filter(homo3d_handle* h, const float* in, float* out, int pitch, double radius, int kernel,
                          int periodic, int backward, int on_device) {
    if (!h || !in || !out || !(radius > 0) || kernel < 0 || kernel > 1) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(out, on_device);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    const int hw = (int)(radius + 0.5);
    const int side = 2 * hw + 1;
    if (side * side * side > 4096) return set_err(h, HOMO3D_NO_SUPPORT, "filter radius too large");
    if (2 * hw + 1 > n && periodic) return set_err(h, HOMO3D_NO_SUPPORT, "periodic filter wider than the grid");
    std::vector<double> wd(side * side * side);
    double W = 0;
    for (int i = 0; i < side * side * side; i++) {
        const int o[3] = {i % side - hw, (i / side) % side - hw, i / (side * side) - hw};
        const double q = (o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) / (radius * radius);
        wd[i] = q < 1 ? (kernel == 0 ? 1 - 6 * q + 8 * q * std::sqrt(q) - 3 * q * q : 1 - std::sqrt(q)) : 0.0;
        W += wd[i];
    }
    std::vector<float> wk(wd.size());
    for (size_t i = 0; i < wk.size(); i++) wk[i] = (float)(wd[i] / W);
    HIPCHK(h, hipMemcpyAsync(h->d_ftmp, wk.data(), sizeof(float) * wk.size(), hipMemcpyHostToDevice, h->stream));
    const size_t elems = (size_t)pitch * n * n;
    int rc = ensure_stage(h, 3 * elems);
    if (rc) return rc;
    float* din = h->d_stage;
    float* dws = h->d_stage + elems;
    float* dres = h->d_stage + 2 * elems;
    HIPCHK(h, hipMemcpyAsync(din, in, sizeof(float) * elems, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                             h->stream));
    if (!backward) {
        launch_conv_filter(h->stream, n, pitch, hw, periodic, h->d_ftmp, din, dres, nullptr);
    } else {
        launch_conv_filter(h->stream, n, pitch, hw, periodic, h->d_ftmp, din, dres, dws);  // weight sums W
        launch_conv_filter_back(h->stream, n, pitch, hw, periodic, h->d_ftmp, din, dws, dres);
    }
    rc = stage_out(h, out, dres, pitch, on_device);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

// sgm(k, c) projection (sgm_exp_t, AutoDiff.h:808-823): out = 1 / (1 + exp(-k (in - c))); with
// backward = 1, y holds the forward output and out = in * k y (1 - y) (in = dJ/dy).
int homo3d_projection(homo3d_handle* h, const float* in, const float* y, float* out, int pitch, double k, double c,
                      int backward, int on_device) {
    if (!h || !in || !out || (backward && !y)) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(out, on_device);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    const size_t elems = (size_t)pitch * n * n;
    int rc = ensure_stage(h, 3 * elems);
    if (rc) return rc;
    float *din = h->d_stage, *dy = h->d_stage + elems, *dres = h->d_stage + 2 * elems;
    const hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIPCHK(h, hipMemcpyAsync(din, in, sizeof(float) * elems, kin, h->stream));
    if (backward) {
        HIPCHK(h, hipMemcpyAsync(dy, y, sizeof(float) * elems, kin, h->stream));
        launch_sgm_back(h->stream, n, pitch, k, dy, din, dres);
    } else {
        launch_sgm(h->stream, n, pitch, k, c, din, dres);
    }
    rc = stage_out(h, out, dres, pitch, on_device);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_oc_update(homo3d_handle* h, const float* sens, float* rho, int pitch, double vol, double min_rho,
                     double step, double damp, int on_device) {
    if (!h || !sens || !rho) return HOMO3D_BAD_ARG;
    on_device = resolve_on_device(rho, on_device);
    const int n = h->lv[0].n;
    if (pitch < n) return set_err(h, HOMO3D_BAD_ARG, "pitch %d < n %d", pitch, n);
    const size_t elems = (size_t)pitch * n * n;
    int rc = ensure_stage(h, 3 * elems);
    if (rc) return rc;
    float* ds = h->d_stage;
    float* dr = h->d_stage + elems;
    float* dn = h->d_stage + 2 * elems;
    const hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIPCHK(h, hipMemcpyAsync(ds, sens, sizeof(float) * elems, kin, h->stream));
    HIPCHK(h, hipMemcpyAsync(dr, rho, sizeof(float) * elems, kin, h->stream));
    const int nb = 1024;
    launch_absmax(h->stream, n, pitch, ds, h->d_ftmp, nb);
    std::vector<float> pm(nb);
    HIPCHK(h, hipMemcpyAsync(pm.data(), h->d_ftmp, sizeof(float) * nb, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double hi = 0, lo = 0;
    for (float v : pm) hi = std::max(hi, (double)v);
    const double Ne = (double)n * n * n;
    for (int it = 0; it < 20; it++) {
        const double g = (hi + lo) / 2;
        double sum = 0;
        launch_oc_trial(h->stream, n, pitch, g, min_rho, step, damp, ds, dr, dn, h->d_partial, nb, h->d_tmp);
        HIPCHK(h, hipMemcpyAsync(&sum, h->d_tmp, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        const double cur = sum / Ne;
        if (cur < vol - 1e-4)
            hi = g;
        else if (cur > vol + 1e-4)
            lo = g;
        else
            break;
    }
    rc = stage_out(h, rho, dn, pitch, on_device);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

// Learned warm start: the network's side is net.hip; these two entry points read or write the handle.
int homo3d_net_dataset_add(homo3d_net* net, homo3d_handle* h) {
    if (!net || !h) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return net_set_error(net, rc, h->err.c_str());
    if (h->cfg.solver != HOMO3D_SOLVER_PCG)
        return net_set_error(net, HOMO3D_BAD_ARG, "training pairs come from the refined solver (HOMO3D_SOLVER_PCG)");
    const int n = h->lv[0].n, g = net_grid(net);
    if (n % g) return net_set_error(net, HOMO3D_NO_SUPPORT, "solve grid is not a multiple of the network grid");
    uint16_t *in = nullptr, *tg = nullptr;
    rc = net_dataset_push(net, &in, &tg);
    if (rc) return rc;
    // the pair is taken on the handle's stream (after its solve), then both streams are in step
    launch_net_input(h->stream, n, g, h->d_x, in);
    launch_net_target(h->stream, n, g, h->U64, in, net_target_scale(net), h->d_tmp, tg);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

int homo3d_net_warm_start(homo3d_handle* h, homo3d_net* net) {
    if (!net || !h) return HOMO3D_BAD_ARG;
    int rc = require_density(h);
    if (rc) return rc;
    const int n = h->lv[0].n, g = net_grid(net);
    if (n % g) return set_err(h, HOMO3D_NO_SUPPORT, "solve grid %d is not a multiple of the network grid %d", n, g);
    hipStream_t ns = net_stream(net);
    HIPCHK(h, hipStreamSynchronize(h->stream));  // the density update has landed
    launch_net_input(ns, n, g, h->d_x, net_input_slot(net));
    const float* P = nullptr;
    rc = net_forward_one(net, &P);
    if (rc) return set_err(h, rc, "network forward failed: %s", homo3d_net_last_error(net));
    launch_net_warm_start(ns, n, g, P, net_pred_ld(net), net_target_scale(net), h->U64, h->X);
    HIPCHK(h, hipStreamSynchronize(ns));
    h->x_dirty = false;  // U64 holds the start; X is its FP32 copy
    return HOMO3D_OK;
}

int homo3d_profile_enable(homo3d_handle* h, int enable) {
    if (!h) return HOMO3D_BAD_ARG;
    prof_flush(h);
    h->prof = enable != 0;
    if (enable)
        for (int k = 0; k < kProfKinds; k++) {
            h->prof_count[k] = 0;
            h->prof_ms[k] = 0;
        }
    return HOMO3D_OK;
}

int homo3d_profile_only(homo3d_handle* h, const char* kernel) {
    if (!h) return HOMO3D_BAD_ARG;
    prof_flush(h);
    if (!kernel || !*kernel) {
        h->prof_only = -1;
        return HOMO3D_OK;
    }
    for (int k = 0; k < kProfKinds; k++)
        if (std::strcmp(kernel, kProfNames[k]) == 0) {
            h->prof_only = k;
            return HOMO3D_OK;
        }
    return set_err(h, HOMO3D_BAD_ARG, "unknown profiled kernel %s", kernel);
}

int homo3d_profile_get(homo3d_handle* h, const char* kernel, long* launches, double* total_ms) {
    if (!h || !kernel) return HOMO3D_BAD_ARG;
    prof_flush(h);
    for (int k = 0; k < kProfKinds; k++)
        if (std::strcmp(kernel, kProfNames[k]) == 0) {
            if (launches) *launches = h->prof_count[k];
            if (total_ms) *total_ms = h->prof_ms[k];
            return HOMO3D_OK;
        }
    return set_err(h, HOMO3D_BAD_ARG, "unknown profiled kernel %s", kernel);
}

// ---------------------------------------------------------------------------------------------
// Batched uneven designs: one handle (and stream) per design; up to max_inflight designs are driven
// at once by host threads, so the small grids' launches overlap on the GPU.
// ---------------------------------------------------------------------------------------------
struct homo3d_batch {
    std::vector<homo3d_handle*> h;
    std::vector<int> tuned;
    int device = 0;
    int inflight = 8;
    std::string err;
};

int homo3d_batch_create(const homo3d_config* base, int count, const int* sizes, int max_inflight,
                        homo3d_batch** out) {
    if (!base || !sizes || !out || count < 1 || max_inflight < 1) return HOMO3D_BAD_ARG;
    *out = nullptr;
    homo3d_batch* b = new homo3d_batch();
    b->device = base->device;
    b->inflight = std::min(max_inflight, 64);
    for (int i = 0; i < count; i++) {
        homo3d_config c = *base;
        c.reso[0] = c.reso[1] = c.reso[2] = sizes[i];
        homo3d_handle* h = nullptr;
        const int rc = homo3d_create(&c, &h);
        if (rc) {
            homo3d_batch_destroy(b);
            return rc;
        }
        b->h.push_back(h);
    }
    b->tuned.assign(count, 0);
    *out = b;
    return HOMO3D_OK;
}

int homo3d_batch_destroy(homo3d_batch* b) {
    if (!b) return HOMO3D_OK;
    for (homo3d_handle* h : b->h) homo3d_destroy(h);
    delete b;
    return HOMO3D_OK;
}

const char* homo3d_batch_last_error(const homo3d_batch* b) { return b ? b->err.c_str() : "null batch"; }

homo3d_handle* homo3d_batch_handle(homo3d_batch* b, int i) {
    return b && i >= 0 && i < (int)b->h.size() ? b->h[i] : nullptr;
}

int homo3d_batch_solve(homo3d_batch* b, const float* const* x, double tol, int autotune, homo3d_solve_stats* stats,
                       double* C) {
    if (!b || !x) return HOMO3D_BAD_ARG;
    const int count = (int)b->h.size();
    std::atomic<int> next{0};
    std::vector<int> status(count, HOMO3D_OK);
    auto worker = [&]() {
        if (hipSetDevice(b->device) != hipSuccess) return;
        for (int i = next++; i < count; i = next++) {
            homo3d_handle* h = b->h[i];
            const int n = h->lv[0].n;
            int rc = homo3d_update_density(h, x[i], n, -1);
            if (!rc && autotune && !b->tuned[i]) {
                rc = homo3d_autotune(h, nullptr);
                b->tuned[i] = 1;
            }
            homo3d_solve_stats st{};
            if (!rc) rc = homo3d_solve(h, tol, 0, &st);
            if (rc == HOMO3D_OK || rc == HOMO3D_NOT_CONVERGED) {
                const int rc2 = C ? homo3d_elastic_matrix(h, C + 36 * (size_t)i) : HOMO3D_OK;
                if (rc2) rc = rc2;
            }
            if (stats) stats[i] = st;
            status[i] = rc;
        }
    };
    const int nt = std::min(b->inflight, count);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back(worker);
    for (auto& t : th) t.join();
    int worst = HOMO3D_OK;
    for (int i = 0; i < count; i++)
        if (status[i] > worst) {
            worst = status[i];
            b->err = "design " + std::to_string(i) + ": " + b->h[i]->err;
        }
    return worst;
}

}  // extern "C"
