// FP64 CPU oracle: multigrid hierarchy, V-cycle, stand-alone MG iteration and MG-PCG.
//
// TEST INFRASTRUCTURE ONLY (see homo_oracle.cpp).  Restates homogenization/MG.cpp,
// homogenization/grid.cpp and homogenization/homogenization.cu on the lexicographic
// periodic layout, using the stateless kernels of homo_oracle.cpp.

#include <cmath>
#include <cstring>
#include <vector>

#include "homo_oracle.h"

namespace {

struct OracleMG {
    int nl = 0;
    int n[16][3];
    long N[16];
    double ke[576];
    double emin = 1e-9;
    double clamp = 1e-4;
    std::vector<double> w;                  // fine cell weights Emin + x_e
    std::vector<std::vector<double>> st;    // stencils of levels >= 1
    std::vector<double> L;                  // Cholesky factor of the coarsest level
    int mc = 0;
    std::vector<std::vector<double>> u, f, r;
    bool ready = false;
};

double dot(const double* a, const double* b, long m) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < m; i++) s += a[i] * b[i];
    return s;
}

void smooth(OracleMG& M, int l, double omega, int reverse) {
    if (l == 0)
        orc_gs_fine(M.n[0], M.ke, M.w.data(), M.u[0].data(), M.f[0].data(), omega, reverse);
    else
        orc_stencil_gs(M.n[l], M.st[l].data(), M.u[l].data(), M.f[l].data(), omega, reverse);
}

void residual(OracleMG& M, int l) {
    if (l == 0)
        orc_residual(M.n[0], M.ke, M.w.data(), M.u[0].data(), M.f[0].data(), M.r[0].data());
    else
        orc_stencil_residual(M.n[l], M.st[l].data(), M.u[l].data(), M.f[l].data(), M.r[l].data());
}

void coarse_solve(OracleMG& M) {
    int l = M.nl - 1;
    long Nc = M.N[l];
    std::vector<double> b(3 * Nc);
    for (long v = 0; v < Nc; v++)
        for (int i = 0; i < 3; i++) b[3 * v + i] = M.f[l][i * Nc + v];
    orc_chol_solve(M.mc, M.L.data(), b.data());
    for (long v = 0; v < Nc; v++)
        for (int i = 0; i < 3; i++) M.u[l][i * Nc + v] = b[3 * v + i];
}

// One V(1,1) cycle (MG::v_cycle, MG.cpp:46-70) on u[0], f[0].  Returns the norm of the
// level-0 residual computed after pre-smoothing (the quantity the reference's stop test reads,
// MG.cpp:139).  symmetric != 0 post-smooths with the reverse sweep.
double vcycle(OracleMG& M, double omega, int symmetric) {
    double rnorm = 0;
    for (int l = 0; l < M.nl - 1; l++) {
        if (l > 0) std::fill(M.u[l].begin(), M.u[l].end(), 0.0);
        smooth(M, l, omega, 0);
        residual(M, l);
        if (l == 0) rnorm = std::sqrt(dot(M.r[0].data(), M.r[0].data(), 3 * M.N[0]));
        orc_restrict(M.n[l], M.r[l].data(), M.f[l + 1].data());
    }
    coarse_solve(M);
    for (int l = M.nl - 2; l >= 0; l--) {
        orc_prolong_add(M.n[l], M.u[l + 1].data(), M.u[l].data(), l == 0 ? 1 : 0);
        smooth(M, l, omega, symmetric ? 1 : 0);
    }
    return rnorm;
}

}  // namespace

extern "C" {

void* orc_mg_create(const int n[3], double E, double nu, double emin) {
    if (n[0] != n[1] || n[1] != n[2]) return nullptr;
    int sizes[16];
    int nl = orc_level_sizes(n[0], sizes, 16);
    if (nl < 2 || sizes[0] != n[0]) return nullptr;
    OracleMG* M = new OracleMG();
    M->nl = nl;
    M->emin = emin;
    orc_ke(E, nu, M->ke);
    M->u.resize(nl);
    M->f.resize(nl);
    M->r.resize(nl);
    M->st.resize(nl);
    for (int l = 0; l < nl; l++) {
        for (int k = 0; k < 3; k++) M->n[l][k] = sizes[l];
        M->N[l] = (long)sizes[l] * sizes[l] * sizes[l];
        M->u[l].assign(3 * M->N[l], 0.0);
        M->f[l].assign(3 * M->N[l], 0.0);
        M->r[l].assign(3 * M->N[l], 0.0);
        if (l > 0) M->st[l].assign(243 * M->N[l], 0.0);
    }
    return M;
}

void orc_mg_destroy(void* h) { delete static_cast<OracleMG*>(h); }

int orc_mg_levels(void* h, int* sizes) {
    OracleMG* M = static_cast<OracleMG*>(h);
    for (int l = 0; l < M->nl; l++) sizes[l] = M->n[l][0];
    return M->nl;
}

const double* orc_mg_ke(void* h) { return static_cast<OracleMG*>(h)->ke; }

// x: cell values x_e = rho_e^p (lexicographic).  Builds w, Galerkin stencils and the coarsest
// Cholesky factor (Homogenization::update + MG::updateStencils, homogenization.cpp:79-85,
// MG.cpp:341-354; coarsest solve replaced by a dense FP64 Cholesky of the pinned system).
int orc_mg_set_density(void* h, const double* x) {
    OracleMG* M = static_cast<OracleMG*>(h);
    long N0 = M->N[0];
    M->w.resize(N0);
    for (long e = 0; e < N0; e++) M->w[e] = M->emin + x[e];
    orc_galerkin_fine(M->n[0], M->ke, M->w.data(), M->st[1].data(), M->clamp);
    for (int l = 2; l < M->nl; l++) orc_galerkin_coarse(M->n[l - 1], M->st[l - 1].data(), M->st[l].data(), M->clamp);
    int lc = M->nl - 1;
    M->mc = (int)(3 * M->N[lc]);
    M->L.assign((size_t)M->mc * M->mc, 0.0);
    orc_dense_from_stencil(M->n[lc], M->st[lc].data(), M->L.data());
    int rc = orc_cholesky(M->mc, M->L.data());
    M->ready = (rc == 0);
    return rc;
}

const double* orc_mg_weights(void* h) { return static_cast<OracleMG*>(h)->w.data(); }

void orc_mg_get_stencil(void* h, int level, double* out) {
    OracleMG* M = static_cast<OracleMG*>(h);
    std::memcpy(out, M->st[level].data(), sizeof(double) * 243 * M->N[level]);
}

// u (in/out), f: [3][N0] for one load case.  Returns the stale residual norm.
double orc_mg_vcycle(void* h, double* u, const double* f, double omega, int symmetric) {
    OracleMG* M = static_cast<OracleMG*>(h);
    long m = 3 * M->N[0];
    std::memcpy(M->u[0].data(), u, sizeof(double) * m);
    std::memcpy(M->f[0].data(), f, sizeof(double) * m);
    double rn = vcycle(*M, omega, symmetric);
    std::memcpy(u, M->u[0].data(), sizeof(double) * m);
    return rn;
}

// Stand-alone MG iteration (MG::solveEquation, MG.cpp:113-203): V-cycles until the stale
// residual ||r||/(||f||+1e-10) <= tol.  A relative residual above 10, or reaching iteration
// maxit-1, resets the guess once; a second time returns HOMO3D-style code 2 (diverged).
// Returns 0 on convergence, 1 when maxit is reached, 2 on divergence.
int orc_mg_solve_mg(void* h, double* u, const double* f, double tol, int maxit, double omega, double* relres,
                    int* iters) {
    OracleMG* M = static_cast<OracleMG*>(h);
    long m = 3 * M->N[0];
    std::memcpy(M->u[0].data(), u, sizeof(double) * m);
    std::memcpy(M->f[0].data(), f, sizeof(double) * m);
    double fnorm = std::sqrt(dot(f, f, m));
    double rel = 1;
    int it = 0, overflow = 2, code = 0;
    while (rel > tol && it++ < maxit) {
        double rn = vcycle(*M, omega, 0);
        rel = rn / (fnorm + 1e-10);
        if (rel > 10 || it >= maxit - 1) {
            overflow--;
            if (overflow <= 0) {
                code = 2;
                break;
            }
            std::fill(M->u[0].begin(), M->u[0].end(), 0.0);
        }
    }
    if (code == 0 && rel > tol) code = 1;
    std::memcpy(u, M->u[0].data(), sizeof(double) * m);
    if (relres) *relres = rel;
    if (iters) *iters = it;
    return code;
}

// MG-preconditioned CG on the pinned system, symmetric V(1,1) preconditioner.  u holds the
// initial guess (pinned entries must be zero) and receives the solution.  Stops when the
// recursive residual satisfies ||r||/||f|| <= tol.  *relres receives the true final relative
// residual.  Returns 0 on convergence, 1 when maxit is reached.
int orc_mg_solve_pcg(void* h, double* u, const double* f, double tol, int maxit, double omega, double* relres,
                     int* iters) {
    OracleMG* M = static_cast<OracleMG*>(h);
    long m = 3 * M->N[0];
    long N0 = M->N[0];
    std::vector<double> x(u, u + m), r(m), z(m), p(m), q(m);
    for (int i = 0; i < 3; i++) x[i * N0] = 0;
    orc_residual(M->n[0], M->ke, M->w.data(), x.data(), f, r.data());
    double fnorm = std::sqrt(dot(f, f, m));
    if (fnorm == 0) {
        std::memcpy(u, x.data(), sizeof(double) * m);
        if (relres) *relres = 0;
        if (iters) *iters = 0;
        return 0;
    }
    auto precond = [&](const double* rr, double* zz) {
        std::fill(M->u[0].begin(), M->u[0].end(), 0.0);
        std::memcpy(M->f[0].data(), rr, sizeof(double) * m);
        vcycle(*M, omega, 1);
        std::memcpy(zz, M->u[0].data(), sizeof(double) * m);
    };
    double rel = std::sqrt(dot(r.data(), r.data(), m)) / fnorm;
    int it = 0, code = 1;
    if (rel <= tol) code = 0;
    else {
        precond(r.data(), z.data());
        p = z;
        double rz = dot(r.data(), z.data(), m);
        while (it < maxit) {
            it++;
            orc_apply_K(M->n[0], M->ke, M->w.data(), p.data(), q.data(), 1);
            double alpha = rz / dot(p.data(), q.data(), m);
            for (long i = 0; i < m; i++) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
            }
            rel = std::sqrt(dot(r.data(), r.data(), m)) / fnorm;
            if (rel <= tol) {
                code = 0;
                break;
            }
            precond(r.data(), z.data());
            double rz2 = dot(r.data(), z.data(), m);
            double beta = rz2 / rz;
            rz = rz2;
            for (long i = 0; i < m; i++) p[i] = z[i] + beta * p[i];
        }
    }
    orc_residual(M->n[0], M->ke, M->w.data(), x.data(), f, r.data());
    if (relres) *relres = std::sqrt(dot(r.data(), r.data(), m)) / fnorm;
    if (iters) *iters = it;
    std::memcpy(u, x.data(), sizeof(double) * m);
    return code;
}

// Full homogenization of one design: 6 macro-strain load cases solved from zero (or from chi
// when warm != 0), then C^H (Homogenization::elasticMatrix, homogenization.cu:464-575).
// mode 0: stand-alone MG (reference), mode 1: MG-PCG.  chi: [6][3][N] in/out.
int orc_mg_homogenize(void* h, double tol, int maxit, int mode, int warm, double* chi, double* C, int* iters) {
    OracleMG* M = static_cast<OracleMG*>(h);
    long N0 = M->N[0];
    std::vector<double> f(3 * N0);
    int worst = 0;
    for (int s = 0; s < 6; s++) {
        orc_macro_rhs(M->n[0], M->ke, M->w.data(), s, f.data());
        double* u = chi + (long)s * 3 * N0;
        if (!warm) std::fill(u, u + 3 * N0, 0.0);
        double rel;
        int it;
        int code = mode == 0 ? orc_mg_solve_mg(h, u, f.data(), tol, maxit, 1.0, &rel, &it)
                             : orc_mg_solve_pcg(h, u, f.data(), tol, maxit, 1.0, &rel, &it);
        if (iters) iters[s] = it;
        if (code > worst) worst = code;
    }
    orc_elastic_matrix(M->n[0], M->ke, M->w.data(), chi, C);
    return worst;
}

}  // extern "C"
