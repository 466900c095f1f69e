// FP64 CPU oracle for the periodic homogenization solve path of lavenklau/homo3d.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke() entry
// and bench.py's cpu_baseline leg load it.  The product library (homo3d_amd/)
// never links, loads or calls it.
//
// Every function restates the reference algorithm on a lexicographic periodic
// index space (vertex/cell id = x + nx*(y + ny*z); cell (x,y,z) has its minimum
// corner at vertex (x,y,z)).  Reference files are cited per function.
//
// Vector layout: one load case is a [3][N] array (component-major); R load cases
// are stored back to back ([R][3][N]).  Coarse stencils are [N][27][3][3] with
// neighbour index d = (dx+1) + 3*(dy+1) + 9*(dz+1).
//
// Parity status per function is listed in the header comment of homo_oracle.h.

#include "homo_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Grid3 {
    int nx, ny, nz;
    long N() const { return (long)nx * ny * nz; }
    static int wrap(int a, int n) {
        a %= n;
        return a < 0 ? a + n : a;
    }
    long id(int x, int y, int z) const {
        return wrap(x, nx) + (long)nx * (wrap(y, ny) + (long)ny * wrap(z, nz));
    }
    void pos(long v, int& x, int& y, int& z) const {
        x = (int)(v % nx);
        y = (int)((v / nx) % ny);
        z = (int)(v / ((long)nx * ny));
    }
};

// ---------------------------------------------------------------------------
// Element stiffness (templateMatrix.cpp:63-166): trilinear 8-node hex of unit
// edge, 2x2x2 Gauss quadrature, isotropic C written with Lame constants.
// Node i sits at (i%2, i/2%2, i/4); DOF 3*i+c.
// ---------------------------------------------------------------------------
void ke_from_lame(double lam, double mu, double* ke) {
    double C[6][6] = {};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[i][j] = lam;
    for (int i = 0; i < 3; i++) C[i][i] += 2 * mu;
    for (int i = 3; i < 6; i++) C[i][i] = mu;
    const double p = std::sqrt(1.0 / 3.0);
    std::vector<double> K(576, 0.0);
    for (int q = 0; q < 8; q++) {
        double g[3] = {((2 * (q % 2) - 1) * p + 1) / 2, ((2 * (q / 2 % 2) - 1) * p + 1) / 2,
                       ((2 * (q / 4) - 1) * p + 1) / 2};
        double dN[8][3];
        for (int i = 0; i < 8; i++) {
            int id[3] = {i % 2, i / 2 % 2, i / 4};
            double r[3];
            for (int k = 0; k < 3; k++) r[k] = id[k] ? g[k] : 1 - g[k];
            for (int k = 0; k < 3; k++) dN[i][k] = (id[k] ? 1.0 : -1.0) * r[(k + 1) % 3] * r[(k + 2) % 3];
        }
        double B[6][24] = {};
        for (int b = 0; b < 8; b++) {
            for (int a = 0; a < 3; a++) B[a][3 * b + a] = dN[b][a];
            B[3][3 * b + 1] = dN[b][2];
            B[3][3 * b + 2] = dN[b][1];
            B[4][3 * b + 0] = dN[b][2];
            B[4][3 * b + 2] = dN[b][0];
            B[5][3 * b + 0] = dN[b][1];
            B[5][3 * b + 1] = dN[b][0];
        }
        double CB[6][24];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 24; j++) {
                double s = 0;
                for (int k = 0; k < 6; k++) s += C[i][k] * B[k][j];
                CB[i][j] = s;
            }
        for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) {
                double s = 0;
                for (int k = 0; k < 6; k++) s += B[k][i] * CB[k][j];
                K[i * 24 + j] += s / 8.0;  // Gauss weight 1, Jacobian (1/2)^3
            }
    }
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) ke[i * 24 + j] = 0.5 * (K[i * 24 + j] + K[j * 24 + i]);
}

void lame_constants(double E, double nu, double& lam, double& mu) {
    lam = nu * E / ((1 + nu) * (1 - 2 * nu));
    mu = E / (2 * (1 + nu));
}

// Unit macro displacement of node `node` for strain `s` (homoCommon.cuh:454-519).
// Voigt order xx, yy, zz, yz, xz, xy; shear strains are engineering strains split 1/2, 1/2.
void macro_disp(int node, int s, double u[3]) {
    double px = node % 2, py = node / 2 % 2, pz = node / 4;
    u[0] = u[1] = u[2] = 0;
    switch (s) {
        case 0: u[0] = px; break;
        case 1: u[1] = py; break;
        case 2: u[2] = pz; break;
        case 3: u[1] = pz / 2; u[2] = py / 2; break;
        case 4: u[0] = pz / 2; u[2] = px / 2; break;
        case 5: u[0] = py / 2; u[1] = px / 2; break;
    }
}

// vertex of node b of the cell with minimum corner (cx,cy,cz)
inline long node_vertex(const Grid3& g, int cx, int cy, int cz, int b) {
    return g.id(cx + b % 2, cy + b / 2 % 2, cz + b / 4);
}

// ---------------------------------------------------------------------------
// Fine-level matrix-free operator.  K = sum_e w_e KE_e with w_e = Emin + x_e
// (grid.cu:415, homogenization.h:49).  The vertex at the origin is pinned:
// its displacement is read as zero and its residual row is zero
// (grid.cu:186-191, 1359, 1429, 545).
// ---------------------------------------------------------------------------
void fine_row(const Grid3& g, const double* ke, const double* w, const double* u, long v, int pin,
              double KeU[3], double Ks[3][3], bool exclude_self) {
    int x, y, z;
    g.pos(v, x, y, z);
    const long N = g.N();
    for (int i = 0; i < 3; i++) {
        KeU[i] = 0;
        for (int j = 0; j < 3; j++) Ks[i][j] = 0;
    }
    for (int e = 0; e < 8; e++) {
        int ex = x + e % 2 - 1, ey = y + e / 2 % 2 - 1, ez = z + e / 4 - 1;
        double we = w[g.id(ex, ey, ez)];
        int a = 7 - e;  // node index of v inside this cell
        double acc[3] = {0, 0, 0};
        for (int b = 0; b < 8; b++) {
            if (b == a && exclude_self) continue;
            long vb = node_vertex(g, ex, ey, ez, b);
            if (pin && vb == 0) continue;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) acc[i] += ke[(3 * a + i) * 24 + 3 * b + j] * u[j * N + vb];
        }
        for (int i = 0; i < 3; i++) KeU[i] += we * acc[i];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Ks[i][j] += we * ke[(3 * a + i) * 24 + 3 * a + j];
    }
}

void block_update(double u[3], const double f[3], const double KeU[3], const double Ks[3][3], double omega,
                  int reverse) {
    if (!reverse) {
        for (int i = 0; i < 3; i++) {
            double s = f[i] - KeU[i];
            for (int j = 0; j < 3; j++)
                if (j != i) s -= Ks[i][j] * u[j];
            u[i] = omega * s / Ks[i][i] + (1 - omega) * u[i];
        }
    } else {
        for (int i = 2; i >= 0; i--) {
            double s = f[i] - KeU[i];
            for (int j = 0; j < 3; j++)
                if (j != i) s -= Ks[i][j] * u[j];
            u[i] = omega * s / Ks[i][i] + (1 - omega) * u[i];
        }
    }
}

// GS colour c (0..7) covers vertices whose coordinate parity is (1-c%2, 1-c/2%2, 1-c/4):
// the reference colours by the parity of the ghost-padded position (pos+1)
// (grid.cpp:155-198, homoCommon.cuh:141-160) and sweeps colours 0..7 (grid.cu:1206-1231).
inline int colour_parity(int c, int k) { return 1 - ((k == 0) ? c % 2 : (k == 1) ? (c / 2) % 2 : c / 4); }

// Stored-stencil row (coarse levels): KeU = sum_{d!=13} st[d] u_{v+d}, Ks = st[13]
void stencil_row(const Grid3& g, const double* st, const double* u, long v, double KeU[3], double Ks[3][3],
                 bool exclude_self) {
    int x, y, z;
    g.pos(v, x, y, z);
    const long N = g.N();
    const double* sv = st + v * 243;
    for (int i = 0; i < 3; i++) KeU[i] = 0;
    for (int d = 0; d < 27; d++) {
        if (d == 13 && exclude_self) continue;
        long nb = g.id(x + d % 3 - 1, y + d / 3 % 3 - 1, z + d / 9 - 1);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) KeU[i] += sv[d * 9 + i * 3 + j] * u[j * N + nb];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Ks[i][j] = sv[13 * 9 + i * 3 + j];
}

// Trilinear prolongation weights of coarse vertices for fine coordinate `a` along one axis.
inline int parents_1d(int a, int nc, int idx[2], double wt[2]) {
    if (a % 2 == 0) {
        idx[0] = Grid3::wrap(a / 2, nc);
        wt[0] = 1.0;
        return 1;
    }
    idx[0] = Grid3::wrap((a - 1) / 2, nc);
    idx[1] = Grid3::wrap((a + 1) / 2, nc);
    wt[0] = wt[1] = 0.5;
    return 2;
}

struct Parents {
    int n = 0;
    long id[8];
    double w[8];
};

Parents parents(const Grid3& gf, const Grid3& gc, long v) {
    int x, y, z;
    gf.pos(v, x, y, z);
    int ix[2], iy[2], iz[2];
    double wx[2], wy[2], wz[2];
    int nx = parents_1d(x, gc.nx, ix, wx), ny = parents_1d(y, gc.ny, iy, wy), nz = parents_1d(z, gc.nz, iz, wz);
    Parents p;
    for (int c = 0; c < nz; c++)
        for (int b = 0; b < ny; b++)
            for (int a = 0; a < nx; a++) {
                p.id[p.n] = ix[a] + (long)gc.nx * (iy[b] + (long)gc.ny * iz[c]);
                p.w[p.n] = wx[a] * wy[b] * wz[c];
                p.n++;
            }
    return p;
}

// neighbour index (0..26) of coarse vertex `to` seen from `from` (periodic, offsets in {-1,0,1})
inline int coarse_offset(const Grid3& g, long from, long to) {
    int fx, fy, fz, tx, ty, tz;
    g.pos(from, fx, fy, fz);
    g.pos(to, tx, ty, tz);
    auto off = [](int a, int b, int n) {
        int d = Grid3::wrap(b - a, n);
        if (d == 0) return 0;
        if (d == 1) return 1;
        if (d == n - 1) return -1;
        return 99;
    };
    int dx = off(fx, tx, g.nx), dy = off(fy, ty, g.ny), dz = off(fz, tz, g.nz);
    if (dx == 99 || dy == 99 || dz == 99) return -1;
    return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
}

// Dense Cholesky (lower), in place on a row-major m x m matrix.  Returns 0 on success.
int cholesky(int m, double* A) {
    for (int j = 0; j < m; j++) {
        double s = A[j * m + j];
        for (int k = 0; k < j; k++) s -= A[j * m + k] * A[j * m + k];
        if (!(s > 0)) return -1;
        double ljj = std::sqrt(s);
        A[j * m + j] = ljj;
        for (int i = j + 1; i < m; i++) {
            double t = A[i * m + j];
            for (int k = 0; k < j; k++) t -= A[i * m + k] * A[j * m + k];
            A[i * m + j] = t / ljj;
        }
        for (int i = 0; i < j; i++) A[i * m + j] = 0;
    }
    return 0;
}

void chol_solve(int m, const double* L, double* b) {
    for (int i = 0; i < m; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= L[i * m + k] * b[k];
        b[i] = s / L[i * m + i];
    }
    for (int i = m - 1; i >= 0; i--) {
        double s = b[i];
        for (int k = i + 1; k < m; k++) s -= L[k * m + i] * b[k];
        b[i] = s / L[i * m + i];
    }
}

}  // namespace

// ===========================================================================
// Stateless kernels
// ===========================================================================
extern "C" {

void orc_ke(double E, double nu, double* ke) {
    double lam, mu;
    lame_constants(E, nu, lam, mu);
    ke_from_lame(lam, mu, ke);
}

void orc_ke_lame(double lam, double mu, double* ke) { ke_from_lame(lam, mu, ke); }

// Resolution rule of Grid::buildRoot (grid.cpp:21-85) in exact integer arithmetic:
// k = max(floor(log2 n) - 2, 0), c = ceil(n / 2^k), solved size c * 2^k, coarsest c.
// The reference evaluates c with floating point pow and, for n = 80, obtains
// 5.0000000000000009 and rounds up to 6; this restatement keeps n = 80.
int orc_level_sizes(int n, int* sizes, int max_levels) {
    if (n < 1 || n > 1024) return -1;
    int fl = 0;
    while ((1 << (fl + 1)) <= n) fl++;
    int k = std::max(fl - 2, 0);
    int c = (n + (1 << k) - 1) >> k;
    int nl = k + 1;
    if (nl > max_levels) return -1;
    for (int l = 0; l < nl; l++) sizes[l] = c << (k - l);
    return nl;
}

// f_s = sum_e w_e KE chi0_s (grid.cu:1818-1919).  Pinned row zeroed.
void orc_macro_rhs(const int n[3], const double* ke, const double* w, int s, double* f) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
    double kchi[24];
    for (int i = 0; i < 24; i++) {
        double t = 0;
        for (int b = 0; b < 8; b++) {
            double u[3];
            macro_disp(b, s, u);
            for (int j = 0; j < 3; j++) t += ke[i * 24 + 3 * b + j] * u[j];
        }
        kchi[i] = t;
    }
#pragma omp parallel for schedule(static)
    for (long v = 0; v < N; v++) {
        int x, y, z;
        g.pos(v, x, y, z);
        double acc[3] = {0, 0, 0};
        for (int e = 0; e < 8; e++) {
            double we = w[g.id(x + e % 2 - 1, y + e / 2 % 2 - 1, z + e / 4 - 1)];
            int a = 7 - e;
            for (int i = 0; i < 3; i++) acc[i] += we * kchi[3 * a + i];
        }
        for (int i = 0; i < 3; i++) f[i * N + v] = (v == 0) ? 0.0 : acc[i];
    }
}

// y = K u (pin != 0: pinned operator, u(0) read as 0 and y(0) = 0).
// Reference: update_residual_otf_kernel_1 / v3_stencilOnLeft_kernel (grid.cu:1277-1461, 3686-3857).
void orc_apply_K(const int n[3], const double* ke, const double* w, const double* u, double* y, int pin) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
#pragma omp parallel for schedule(static)
    for (long v = 0; v < N; v++) {
        double KeU[3], Ks[3][3];
        fine_row(g, ke, w, u, v, pin, KeU, Ks, false);
        for (int i = 0; i < 3; i++) y[i * N + v] = (pin && v == 0) ? 0.0 : KeU[i];
    }
}

// r = f - K u with the pinned operator (grid.cu:1277-1461).
void orc_residual(const int n[3], const double* ke, const double* w, const double* u, const double* f,
                  double* r) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
#pragma omp parallel for schedule(static)
    for (long v = 0; v < N; v++) {
        double KeU[3], Ks[3][3];
        fine_row(g, ke, w, u, v, 1, KeU, Ks, false);
        for (int i = 0; i < 3; i++) r[i * N + v] = (v == 0) ? 0.0 : f[i * N + v] - KeU[i];
    }
}

// One 8-colour block Gauss-Seidel/SOR sweep on the fine level, in place
// (gs_relaxation_otf_kernel, grid.cu:323-551; colour loop grid.cu:1206-1231).
// reverse = 1 runs colours 7..0 and, inside a vertex, components z, y, x
// (the adjoint sweep used as post-smoother by the symmetric V-cycle).
void orc_gs_fine(const int n[3], const double* ke, const double* w, double* u, const double* f, double omega,
                 int reverse) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
    for (int cc = 0; cc < 8; cc++) {
        int c = reverse ? 7 - cc : cc;
        int px = colour_parity(c, 0), py = colour_parity(c, 1), pz = colour_parity(c, 2);
#pragma omp parallel for schedule(static) collapse(2)
        for (int z = pz; z < g.nz; z += 2)
            for (int y = py; y < g.ny; y += 2)
                for (int x = px; x < g.nx; x += 2) {
                    long v = g.id(x, y, z);
                    if (v == 0) {
                        for (int i = 0; i < 3; i++) u[i * N] = 0;
                        continue;
                    }
                    double KeU[3], Ks[3][3];
                    fine_row(g, ke, w, u, v, 1, KeU, Ks, true);
                    double uu[3] = {u[v], u[N + v], u[2 * N + v]};
                    double ff[3] = {f[v], f[N + v], f[2 * N + v]};
                    block_update(uu, ff, KeU, Ks, omega, reverse);
                    for (int i = 0; i < 3; i++) u[i * N + v] = uu[i];
                }
    }
}

// Stored 27-point stencil operator on a coarse level (update_residual_kernel_1, grid.cu:1465-1564).
void orc_stencil_apply(const int n[3], const double* st, const double* u, double* y) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
#pragma omp parallel for schedule(static)
    for (long v = 0; v < N; v++) {
        double KeU[3], Ks[3][3];
        stencil_row(g, st, u, v, KeU, Ks, false);
        for (int i = 0; i < 3; i++) y[i * N + v] = KeU[i];
    }
}

void orc_stencil_residual(const int n[3], const double* st, const double* u, const double* f, double* r) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
#pragma omp parallel for schedule(static)
    for (long v = 0; v < N; v++) {
        double KeU[3], Ks[3][3];
        stencil_row(g, st, u, v, KeU, Ks, false);
        for (int i = 0; i < 3; i++) r[i * N + v] = f[i * N + v] - KeU[i];
    }
}

// Coarse-level colour sweep with the stored stencil (gs_relaxation_kernel, grid.cu:556-689).
// No pin on coarse levels: the reference's check there is disabled (grid.cu:654).
void orc_stencil_gs(const int n[3], const double* st, double* u, const double* f, double omega, int reverse) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
    for (int cc = 0; cc < 8; cc++) {
        int c = reverse ? 7 - cc : cc;
        int px = colour_parity(c, 0), py = colour_parity(c, 1), pz = colour_parity(c, 2);
#pragma omp parallel for schedule(static) collapse(2)
        for (int z = pz; z < g.nz; z += 2)
            for (int y = py; y < g.ny; y += 2)
                for (int x = px; x < g.nx; x += 2) {
                    long v = g.id(x, y, z);
                    double KeU[3], Ks[3][3];
                    stencil_row(g, st, u, v, KeU, Ks, true);
                    double uu[3] = {u[v], u[N + v], u[2 * N + v]};
                    double ff[3] = {f[v], f[N + v], f[2 * N + v]};
                    block_update(uu, ff, KeU, Ks, omega, reverse);
                    for (int i = 0; i < 3; i++) u[i * N + v] = uu[i];
                }
    }
}

// f_c = P^T r_f, weights prod_k (2-|o_k|)/2 over the 27 fine points around 2V
// (restrict_residual_kernel_1, grid.cu:917-994).  nf is the fine size; coarse = nf/2.
void orc_restrict(const int nf[3], const double* rf, double* fc) {
    Grid3 gf{nf[0], nf[1], nf[2]};
    Grid3 gc{nf[0] / 2, nf[1] / 2, nf[2] / 2};
    const long Nf = gf.N(), Nc = gc.N();
#pragma omp parallel for schedule(static)
    for (long V = 0; V < Nc; V++) {
        int X, Y, Z;
        gc.pos(V, X, Y, Z);
        double acc[3] = {0, 0, 0};
        for (int o = 0; o < 27; o++) {
            int ox = o % 3 - 1, oy = o / 3 % 3 - 1, oz = o / 9 - 1;
            double wgt = (2 - std::abs(ox)) * (2 - std::abs(oy)) * (2 - std::abs(oz)) / 8.0;
            long vf = gf.id(2 * X + ox, 2 * Y + oy, 2 * Z + oz);
            for (int i = 0; i < 3; i++) acc[i] += wgt * rf[i * Nf + vf];
        }
        for (int i = 0; i < 3; i++) fc[i * Nc + V] = acc[i];
    }
}

// u_f += P u_c, trilinear (prolongate_correction_kernel_1, grid.cu:1120-1190).
// zero_pin != 0 leaves the pinned origin untouched (root level).
void orc_prolong_add(const int nf[3], const double* uc, double* uf, int zero_pin) {
    Grid3 gf{nf[0], nf[1], nf[2]};
    Grid3 gc{nf[0] / 2, nf[1] / 2, nf[2] / 2};
    const long Nf = gf.N(), Nc = gc.N();
#pragma omp parallel for schedule(static)
    for (long v = 0; v < Nf; v++) {
        if (zero_pin && v == 0) continue;
        Parents p = parents(gf, gc, v);
        for (int i = 0; i < 3; i++) {
            double s = 0;
            for (int k = 0; k < p.n; k++) s += p.w[k] * uc[i * Nc + p.id[k]];
            uf[i * Nf + v] += s;
        }
    }
}

// Level-1 Galerkin stencil P^T Kp P from the fine cells (restrict_stencil_otf_aos_kernel_1,
// grid_ex.cu:936-1061, with the pinned-origin rows of restrict_stencil_arround_dirichelt_boundary,
// grid.cpp:1184-1325): node pairs touching the pinned origin contribute identity when they are the
// same node and nothing otherwise.  Diagonal entries below `clamp` are raised to `clamp`
// (grid_ex.cu:1048-1054).  Cell weights are w_e = Emin + x_e throughout.
void orc_galerkin_fine(const int nf[3], const double* ke, const double* w, double* stc, double clamp) {
    Grid3 gf{nf[0], nf[1], nf[2]};
    Grid3 gc{nf[0] / 2, nf[1] / 2, nf[2] / 2};
    const long Nc = gc.N();
#pragma omp parallel for schedule(dynamic, 16)
    for (long V = 0; V < Nc; V++) {
        int X, Y, Z;
        gc.pos(V, X, Y, Z);
        double st[243] = {};
        // fine cells whose nodes can carry weight of V: cells at 2V + (-2..1)
        for (int c = 0; c < 64; c++) {
            int cx = 2 * X + c % 4 - 2, cy = 2 * Y + c / 4 % 4 - 2, cz = 2 * Z + c / 16 - 2;
            double we = w[gf.id(cx, cy, cz)];
            for (int a = 0; a < 8; a++) {
                long va = node_vertex(gf, cx, cy, cz, a);
                Parents pa = parents(gf, gc, va);
                double wa = 0;
                for (int k = 0; k < pa.n; k++)
                    if (pa.id[k] == V) wa += pa.w[k];
                if (wa == 0) continue;
                for (int b = 0; b < 8; b++) {
                    long vb = node_vertex(gf, cx, cy, cz, b);
                    bool pinned = (va == 0 || vb == 0);
                    if (pinned && a != b) continue;
                    Parents pb = parents(gf, gc, vb);
                    for (int k = 0; k < pb.n; k++) {
                        int d = coarse_offset(gc, V, pb.id[k]);
                        if (d < 0) continue;
                        double f = wa * pb.w[k];
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++) {
                                double kij = pinned ? (i == j ? 1.0 : 0.0) : we * ke[(3 * a + i) * 24 + 3 * b + j];
                                st[d * 9 + i * 3 + j] += f * kij;
                            }
                    }
                }
            }
        }
        for (int k = 0; k < 3; k++)
            if (std::fabs(st[13 * 9 + k * 4]) < clamp) st[13 * 9 + k * 4] = clamp;
        std::memcpy(stc + V * 243, st, sizeof(st));
    }
}

// Coarser Galerkin stencil P^T A_f P from a stored stencil (restrict_stencil_aos_kernel_1,
// grid_ex.cu:1065-1160), with the same diagonal clamp.
void orc_galerkin_coarse(const int nf[3], const double* stf, double* stc, double clamp) {
    Grid3 gf{nf[0], nf[1], nf[2]};
    Grid3 gc{nf[0] / 2, nf[1] / 2, nf[2] / 2};
    const long Nc = gc.N();
#pragma omp parallel for schedule(dynamic, 16)
    for (long V = 0; V < Nc; V++) {
        int X, Y, Z;
        gc.pos(V, X, Y, Z);
        double st[243] = {};
        for (int o = 0; o < 27; o++) {
            int ox = o % 3 - 1, oy = o / 3 % 3 - 1, oz = o / 9 - 1;
            double wa = (2 - std::abs(ox)) * (2 - std::abs(oy)) * (2 - std::abs(oz)) / 8.0;
            int fx = 2 * X + ox, fy = 2 * Y + oy, fz = 2 * Z + oz;
            long v = gf.id(fx, fy, fz);
            const double* sv = stf + v * 243;
            for (int d = 0; d < 27; d++) {
                long vb = gf.id(fx + d % 3 - 1, fy + d / 3 % 3 - 1, fz + d / 9 - 1);
                Parents pb = parents(gf, gc, vb);
                for (int k = 0; k < pb.n; k++) {
                    int dc = coarse_offset(gc, V, pb.id[k]);
                    if (dc < 0) continue;
                    double f = wa * pb.w[k];
                    for (int t = 0; t < 9; t++) st[dc * 9 + t] += f * sv[d * 9 + t];
                }
            }
        }
        for (int k = 0; k < 3; k++)
            if (std::fabs(st[13 * 9 + k * 4]) < clamp) st[13 * 9 + k * 4] = clamp;
        std::memcpy(stc + V * 243, st, sizeof(st));
    }
}

// Dense 3N x 3N matrix (row = 3*v + i) of a stored-stencil level, symmetrised.
void orc_dense_from_stencil(const int n[3], const double* st, double* A) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
    const long m = 3 * N;
    std::fill(A, A + m * m, 0.0);
    for (long v = 0; v < N; v++) {
        int x, y, z;
        g.pos(v, x, y, z);
        for (int d = 0; d < 27; d++) {
            long nb = g.id(x + d % 3 - 1, y + d / 3 % 3 - 1, z + d / 9 - 1);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) A[(3 * v + i) * m + 3 * nb + j] += st[v * 243 + d * 9 + i * 3 + j];
        }
    }
    for (long i = 0; i < m; i++)
        for (long j = i + 1; j < m; j++) {
            double s = 0.5 * (A[i * m + j] + A[j * m + i]);
            A[i * m + j] = A[j * m + i] = s;
        }
}

// Dense 3N x 3N pinned fine matrix (row = 3*v + i): rows and columns of the origin replaced by
// identity.  For brute-force checks at tiny n only.
void orc_dense_fine(const int n[3], const double* ke, const double* w, double* A) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
    const long m = 3 * N;
    std::fill(A, A + m * m, 0.0);
    for (long e = 0; e < N; e++) {
        int cx, cy, cz;
        g.pos(e, cx, cy, cz);
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++) {
                long va = node_vertex(g, cx, cy, cz, a), vb = node_vertex(g, cx, cy, cz, b);
                if (va == 0 || vb == 0) continue;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        A[(3 * va + i) * m + 3 * vb + j] += w[e] * ke[(3 * a + i) * 24 + 3 * b + j];
            }
    }
    for (int i = 0; i < 3; i++) A[i * m + i] = 1.0;
}

int orc_cholesky(int m, double* A) { return cholesky(m, A); }
void orc_chol_solve(int m, const double* L, double* b) { chol_solve(m, L, b); }

// C^H_ij = (1/|Y|) sum_e w_e (chi0_i - chi_i)_e^T KE (chi0_j - chi_j)_e, |Y| = number of cells
// (elasticMatrix_kernel_opt + Homogenization::elasticMatrix, homogenization.cu:208-346, 464-575).
// chi: [6][3][N] solution fields.  C: 36 doubles, row-major, symmetric.
void orc_elastic_matrix(const int n[3], const double* ke, const double* w, const double* chi, double* C) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
    double acc[21] = {};
#pragma omp parallel
    {
        double loc[21] = {};
#pragma omp for schedule(static)
        for (long e = 0; e < N; e++) {
            int cx, cy, cz;
            g.pos(e, cx, cy, cz);
            long vn[8];
            for (int b = 0; b < 8; b++) vn[b] = node_vertex(g, cx, cy, cz, b);
            double U[6][24];
            for (int s = 0; s < 6; s++)
                for (int b = 0; b < 8; b++) {
                    double m0[3];
                    macro_disp(b, s, m0);
                    for (int c = 0; c < 3; c++) U[s][3 * b + c] = m0[c] - chi[(s * 3 + c) * N + vn[b]];
                }
            double KU[6][24];
            for (int s = 0; s < 6; s++)
                for (int i = 0; i < 24; i++) {
                    double t = 0;
                    for (int j = 0; j < 24; j++) t += ke[i * 24 + j] * U[s][j];
                    KU[s][i] = t;
                }
            int k = 0;
            for (int i = 0; i < 6; i++)
                for (int j = i; j < 6; j++) {
                    double t = 0;
                    for (int q = 0; q < 24; q++) t += U[i][q] * KU[j][q];
                    loc[k++] += w[e] * t;
                }
        }
#pragma omp critical
        for (int k = 0; k < 21; k++) acc[k] += loc[k];
    }
    int k = 0;
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++) {
            C[i * 6 + j] = C[j * 6 + i] = acc[k] / (double)N;
            k++;
        }
}

// dJ/dx_e = (1/|Y|) sum_{i<=j} E_ij,e (dC_ij + [i!=j] dC_ji), E_ij,e the cell energy without w_e
// (Sensitivity_kernel_opt_2, homogenization.cu:697-858; exp_penal = 1 so p x^(p-1) = 1).
// sens: N doubles, lexicographic.
void orc_sensitivity(const int n[3], const double* ke, const double* chi, const double* dC, double* sens) {
    Grid3 g{n[0], n[1], n[2]};
    const long N = g.N();
#pragma omp parallel for schedule(static)
    for (long e = 0; e < N; e++) {
        int cx, cy, cz;
        g.pos(e, cx, cy, cz);
        long vn[8];
        for (int b = 0; b < 8; b++) vn[b] = node_vertex(g, cx, cy, cz, b);
        double U[6][24];
        for (int s = 0; s < 6; s++)
            for (int b = 0; b < 8; b++) {
                double m0[3];
                macro_disp(b, s, m0);
                for (int c = 0; c < 3; c++) U[s][3 * b + c] = m0[c] - chi[(s * 3 + c) * N + vn[b]];
            }
        double KU[6][24];
        for (int s = 0; s < 6; s++)
            for (int i = 0; i < 24; i++) {
                double t = 0;
                for (int j = 0; j < 24; j++) t += ke[i * 24 + j] * U[s][j];
                KU[s][i] = t;
            }
        double sv = 0;
        for (int i = 0; i < 6; i++)
            for (int j = i; j < 6; j++) {
                double t = 0;
                for (int q = 0; q < 24; q++) t += U[i][q] * KU[j][q];
                double wgt = dC[i * 6 + j] + (i != j ? dC[j * 6 + i] : 0.0);
                sv += t * wgt;
            }
        sens[e] = sv / (double)N;
    }
}

}  // extern "C"
