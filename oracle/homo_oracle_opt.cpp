// FP64 CPU oracle: density updates and density-field utilities.
//
// TEST INFRASTRUCTURE ONLY (see homo_oracle.cpp).  Restates optimization/OCOptimizer.cu,
// optimization/mmaOpt.cu (Svanberg's MMA), AutoDiff/TensorExpression.cuh (symmetry
// projection) and the seeded "randc" density generator.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "homo_oracle.h"

namespace {

// ---------------------------------------------------------------------------
// Symmetry groups acting on cell indices of a cubic n^3 grid; "negating" a
// coordinate maps p -> n-1-p (TensorExpression.cuh:72-262).
// Element g = (perm, flips): out[k] = flip_k ? n-1-p[perm[k]] : p[perm[k]].
// ---------------------------------------------------------------------------
const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
const int kPermSign[6] = {1, -1, -1, 1, 1, -1};

struct GroupElem {
    int perm, flips;
};

std::vector<GroupElem> group(int sym) {
    std::vector<GroupElem> g;
    if (sym == 0) {  // reflect3: 8 reflections
        for (int f = 0; f < 8; f++) g.push_back({0, f});
    } else if (sym == 1) {  // reflect6: 6 permutations x 8 reflections
        for (int p = 0; p < 6; p++)
            for (int f = 0; f < 8; f++) g.push_back({p, f});
    } else if (sym == 2) {  // rotate3: proper rotations of the cube (det = +1)
        for (int p = 0; p < 6; p++)
            for (int f = 0; f < 8; f++) {
                int nflip = (f & 1) + (f >> 1 & 1) + (f >> 2 & 1);
                int det = kPermSign[p] * ((nflip % 2) ? -1 : 1);
                if (det == 1) g.push_back({p, f});
            }
    }
    return g;
}

inline long apply_g(const GroupElem& g, int n, const int p[3]) {
    int q[3];
    for (int k = 0; k < 3; k++) {
        int a = p[kPerm[g.perm][k]];
        q[k] = ((g.flips >> k) & 1) ? n - 1 - a : a;
    }
    return q[0] + (long)n * (q[1] + (long)n * q[2]);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), Box-Muller normals.
// ---------------------------------------------------------------------------
void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
    uint32_t k[2] = {key_in[0], key_in[1]};
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k[0];
        uint32_t n2 = hi0 ^ c[3] ^ k[1];
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
    for (int i = 0; i < 4; i++) out[i] = c[i];
}

double filter_weight(const int off[3], double r) {
    double r2 = (off[0] * off[0] + off[1] * off[1] + off[2] * off[2]) / (r * r);
    if (r2 < 1) return 1 - std::sqrt(r2);
    return 0;
}

}  // namespace

extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox4x32_10(ctr, key, out); }

// Symmetry projection: every cell receives the average of the field over its orbit under the
// group (reflect3 = 0, reflect6 = 1, rotate3 = 2; other values copy).  The sum runs in FP64 over
// the group elements applied to the orbit's smallest-index member, so every member of an orbit
// receives the same bits.  Cubic grids only.
void orc_symmetrize_f32(int n, int sym, const float* in, float* out) {
    auto g = group(sym);
    long N = (long)n * n * n;
    if (g.empty()) {
        std::memcpy(out, in, sizeof(float) * N);
        return;
    }
#pragma omp parallel for schedule(static)
    for (long e = 0; e < N; e++) {
        int p[3] = {(int)(e % n), (int)(e / n % n), (int)(e / ((long)n * n))};
        long best = e;
        for (auto& ge : g) best = std::min(best, apply_g(ge, n, p));
        int q[3] = {(int)(best % n), (int)(best / n % n), (int)(best / ((long)n * n))};
        double s = 0;
        for (auto& ge : g) s += (double)in[apply_g(ge, n, q)];
        out[e] = (float)(s / (double)g.size());
    }
}

void orc_symmetrize_f64(int n, int sym, const double* in, double* out) {
    auto g = group(sym);
    long N = (long)n * n * n;
    if (g.empty()) {
        std::memcpy(out, in, sizeof(double) * N);
        return;
    }
#pragma omp parallel for schedule(static)
    for (long e = 0; e < N; e++) {
        int p[3] = {(int)(e % n), (int)(e / n % n), (int)(e / ((long)n * n))};
        double s = 0;
        for (auto& ge : g) s += in[apply_g(ge, n, p)];
        out[e] = s / (double)g.size();
    }
}

// Sigmund sensitivity filter (filterSens_Tensor_kernel, OCOptimizer.cu:155-195):
// half-width int(r+0.5), weight max(0, 1-|off|/r) normalised by the full-kernel sum,
// s'_e = sum_nb w s rho / (sum_nb w * rho_e); neighbours outside the box are skipped
// unless periodic.
void orc_oc_filter(const int n[3], const double* sens, const double* rho, double r, int periodic, double* out) {
    int h = (int)(r + 0.5);
    int side = 2 * h + 1;
    std::vector<double> wk(side * side * side);
    double W = 0;
    for (int i = 0; i < side * side * side; i++) {
        int off[3] = {i % side - h, i / side % side - h, i / (side * side) - h};
        wk[i] = filter_weight(off, r);
        W += wk[i];
    }
    for (auto& x : wk) x /= W;
    long N = (long)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (long e = 0; e < N; e++) {
        int p[3] = {(int)(e % n[0]), (int)(e / n[0] % n[1]), (int)(e / ((long)n[0] * n[1]))};
        double sum = 0, ws = 0;
        for (int i = 0; i < side * side * side; i++) {
            int q[3] = {p[0] + i % side - h, p[1] + i / side % side - h, p[2] + i / (side * side) - h};
            bool in = true;
            for (int k = 0; k < 3; k++) {
                if (periodic) q[k] = ((q[k] % n[k]) + n[k]) % n[k];
                if (q[k] < 0 || q[k] >= n[k]) in = false;
            }
            if (!in) continue;
            long id = q[0] + (long)n[0] * (q[1] + (long)n[1] * q[2]);
            sum += sens[id] * rho[id] * wk[i];
            ws += wk[i];
        }
        out[e] = sum / (ws * rho[e]);
    }
}

// OC density update (OCOptimizer::update, OCOptimizer.cu:218-247): bisection (<= 20 steps) on the
// multiplier g in [0, max|s|]; per cell B = -s/g (0.01 when negative),
// rho' = clamp(B^damp * rho, rho +- step, [minRho, 1]); stop when |mean(rho') - vol| <= 1e-4.
// Per-cell arithmetic in FP64 rounded to FP32; volume sums in FP64.  Returns bisection steps.
int orc_oc_update(long ne, const float* sens, float* rho, double vol, double minRho, double step, double damp) {
    double maxs = 0;
    for (long i = 0; i < ne; i++) maxs = std::max(maxs, std::fabs((double)sens[i]));
    double lo = 0, hi = maxs;
    std::vector<float> nr(ne);
    int it = 0;
    for (it = 0; it < 20; it++) {
        double g = (hi + lo) / 2;
        double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (long i = 0; i < ne; i++) {
            double r0 = rho[i];
            double B = -(double)sens[i] / g;
            if (B < 0) B = 0.01;
            double x = (damp == 0.5 ? std::sqrt(B) : std::pow(B, damp)) * r0;
            if (x - r0 < -step) x = r0 - step;
            if (x - r0 > step) x = r0 + step;
            if (x < minRho) x = minRho;
            if (x > 1) x = 1;
            nr[i] = (float)x;
            s += (double)nr[i];
        }
        double cur = s / (double)ne;
        if (cur < vol - 1e-4)
            hi = g;
        else if (cur > vol + 1e-4)
            lo = g;
        else {
            it++;
            break;
        }
    }
    std::memcpy(rho, nr.data(), sizeof(float) * ne);
    return it;
}

// Svanberg's MMA step (mmasub_ker + subsolv_g, mmaOpt.cu:429-1019; constants mmaOpt.cu:846-851;
// the (m+1)x(m+1) Newton system of solveLinearHost, mmaOpt.cpp:48-73, solved here by Gaussian
// elimination with partial pivoting).  x, xold1, xold2, low, upp: n doubles, updated in place.
// dfdx: m x n row-major.  Returns the number of Newton iterations.
int orc_mma_update(int m, long n, int iter, double* x, const double* xmin, const double* xmax, double* xold1,
                   double* xold2, const double* df0dx, const double* fval, const double* dfdx, double* low,
                   double* upp, double a0, const double* a, const double* c, const double* d, double move) {
    const double epsimin = 1e-7, raa0 = 1e-5, albefa = 0.1, asyinit = 0.5, asyincr = 1.2, asydecr = 0.7;
    std::vector<double> alfa(n), beta(n), p0(n), q0(n), P((size_t)m * n), Q((size_t)m * n), b(m);
    for (long j = 0; j < n; j++) {
        double dx = xmax[j] - xmin[j];
        if (iter <= 2) {
            low[j] = x[j] - asyinit * dx;
            upp[j] = x[j] + asyinit * dx;
        } else {
            double zzz = (x[j] - xold1[j]) * (xold1[j] - xold2[j]);
            double fac = zzz > 0 ? asyincr : (zzz < 0 ? asydecr : 1.0);
            low[j] = x[j] - fac * (xold1[j] - low[j]);
            upp[j] = x[j] + fac * (upp[j] - xold1[j]);
            low[j] = std::max(low[j], x[j] - 10 * dx);
            low[j] = std::min(low[j], x[j] - 0.01 * dx);
            upp[j] = std::min(upp[j], x[j] + 10 * dx);
            upp[j] = std::max(upp[j], x[j] + 0.01 * dx);
        }
        alfa[j] = std::max(std::max(low[j] + albefa * (x[j] - low[j]), x[j] - move * dx), xmin[j]);
        beta[j] = std::min(std::min(upp[j] - albefa * (upp[j] - x[j]), x[j] + move * dx), xmax[j]);
        double xmamiinv = 1.0 / std::max(dx, 1e-5);
        double ux1 = upp[j] - x[j], xl1 = x[j] - low[j];
        double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
        double dp = std::max(df0dx[j], 0.0), dm = std::max(-df0dx[j], 0.0);
        double pq0 = 1e-3 * (dp + dm) + raa0 * xmamiinv;
        p0[j] = (dp + pq0) * ux2;
        q0[j] = (dm + pq0) * xl2;
        for (int i = 0; i < m; i++) {
            double g = dfdx[(size_t)i * n + j];
            double pp = std::max(g, 0.0), qq = std::max(-g, 0.0);
            double pq = 1e-3 * (pp + qq) + raa0 * xmamiinv;
            P[(size_t)i * n + j] = (pp + pq) * ux2;
            Q[(size_t)i * n + j] = (qq + pq) * xl2;
        }
    }
    for (int i = 0; i < m; i++) {
        double s = 0;
        for (long j = 0; j < n; j++)
            s += P[(size_t)i * n + j] / (upp[j] - x[j]) + Q[(size_t)i * n + j] / (x[j] - low[j]);
        b[i] = s - fval[i];
    }

    // ---- subsolv: primal-dual interior point ----
    std::vector<double> xs(n), xsi(n), eta(n), y(m, 1.0), lam(m, 1.0), mu(m), s(m, 1.0);
    double z = 1, zet = 1;
    for (long j = 0; j < n; j++) {
        xs[j] = 0.5 * (alfa[j] + beta[j]);
        xsi[j] = std::max(1.0 / (xs[j] - alfa[j]), 1.0);
        eta[j] = std::max(1.0 / (beta[j] - xs[j]), 1.0);
    }
    for (int i = 0; i < m; i++) mu[i] = std::max(1.0, 0.5 * c[i]);
    std::vector<double> plam(n), qlam(n), gvec(m), GG((size_t)m * n), delx(n), diagx(n), dx(n), dxsi(n), deta(n);
    std::vector<double> dely(m), dellam(m), diagy(m), dlam(m), dy(m), dmu(m), ds(m);
    int itera = 0;
    double epsi = 1;

    auto pq_lambda = [&](const std::vector<double>& xx, const std::vector<double>& ll) {
        for (long j = 0; j < n; j++) {
            double pl = p0[j], ql = q0[j];
            for (int i = 0; i < m; i++) {
                pl += P[(size_t)i * n + j] * ll[i];
                ql += Q[(size_t)i * n + j] * ll[i];
            }
            plam[j] = pl;
            qlam[j] = ql;
        }
        for (int i = 0; i < m; i++) {
            double g = 0;
            for (long j = 0; j < n; j++)
                g += P[(size_t)i * n + j] / (upp[j] - xx[j]) + Q[(size_t)i * n + j] / (xx[j] - low[j]);
            gvec[i] = g;
        }
    };
    auto residual = [&](double& nrm, double& mx) {
        pq_lambda(xs, lam);
        double s2 = 0, m2 = 0;
        auto acc = [&](double v) {
            s2 += v * v;
            m2 = std::max(m2, std::fabs(v));
        };
        for (long j = 0; j < n; j++) {
            double ux1 = upp[j] - xs[j], xl1 = xs[j] - low[j];
            double dpsidx = plam[j] / (ux1 * ux1) - qlam[j] / (xl1 * xl1);
            acc(dpsidx - xsi[j] + eta[j]);
        }
        for (int i = 0; i < m; i++) acc(c[i] + d[i] * y[i] - mu[i] - lam[i]);
        double alam = 0;
        for (int i = 0; i < m; i++) alam += a[i] * lam[i];
        acc(a0 - zet - alam);
        for (int i = 0; i < m; i++) acc(gvec[i] - a[i] * z - y[i] + s[i] - b[i]);
        for (long j = 0; j < n; j++) acc(xsi[j] * (xs[j] - alfa[j]) - epsi);
        for (long j = 0; j < n; j++) acc(eta[j] * (beta[j] - xs[j]) - epsi);
        for (int i = 0; i < m; i++) acc(mu[i] * y[i] - epsi);
        acc(zet * z - epsi);
        for (int i = 0; i < m; i++) acc(lam[i] * s[i] - epsi);
        nrm = std::sqrt(s2);
        mx = m2;
    };

    while (epsi > epsimin) {
        double residunorm, residumax;
        residual(residunorm, residumax);
        int ittt = 0;
        while (residumax > 0.9 * epsi && ittt < 200) {
            ittt++;
            itera++;
            pq_lambda(xs, lam);
            for (long j = 0; j < n; j++) {
                double ux1 = upp[j] - xs[j], xl1 = xs[j] - low[j];
                double ux2 = ux1 * ux1, xl2 = xl1 * xl1, ux3 = ux1 * ux2, xl3 = xl1 * xl2;
                for (int i = 0; i < m; i++) GG[(size_t)i * n + j] = P[(size_t)i * n + j] / ux2 - Q[(size_t)i * n + j] / xl2;
                double dpsidx = plam[j] / ux2 - qlam[j] / xl2;
                delx[j] = dpsidx - epsi / (xs[j] - alfa[j]) + epsi / (beta[j] - xs[j]);
                diagx[j] = 2 * (plam[j] / ux3 + qlam[j] / xl3) + xsi[j] / (xs[j] - alfa[j]) + eta[j] / (beta[j] - xs[j]);
            }
            double alam = 0;
            for (int i = 0; i < m; i++) alam += a[i] * lam[i];
            double delz = a0 - alam - epsi / z;
            std::vector<double> diaglamyi(m);
            for (int i = 0; i < m; i++) {
                dely[i] = c[i] + d[i] * y[i] - lam[i] - epsi / y[i];
                dellam[i] = gvec[i] - a[i] * z - y[i] - b[i] + epsi / lam[i];
                diagy[i] = d[i] + mu[i] / y[i];
                diaglamyi[i] = s[i] / lam[i] + 1.0 / diagy[i];
            }
            // (m+1) x (m+1) system [Alam a; a^T -zet/z] [dlam; dz] = [blam; delz]
            int M1 = m + 1;
            std::vector<double> A((size_t)M1 * M1, 0.0), rhs(M1, 0.0);
            for (int i = 0; i < m; i++) {
                double gd = 0;
                for (long j = 0; j < n; j++) gd += GG[(size_t)i * n + j] * delx[j] / diagx[j];
                rhs[i] = dellam[i] + dely[i] / diagy[i] - gd;
                for (int k = 0; k < m; k++) {
                    double t = 0;
                    for (long j = 0; j < n; j++) t += GG[(size_t)i * n + j] * GG[(size_t)k * n + j] / diagx[j];
                    A[i * M1 + k] = t;
                }
                A[i * M1 + i] += diaglamyi[i];
                A[i * M1 + m] = a[i];
                A[m * M1 + i] = a[i];
            }
            A[m * M1 + m] = -zet / z;
            rhs[m] = delz;
            // Gaussian elimination with partial pivoting
            std::vector<int> piv(M1);
            for (int k = 0; k < M1; k++) piv[k] = k;
            for (int k = 0; k < M1; k++) {
                int pr = k;
                for (int i = k + 1; i < M1; i++)
                    if (std::fabs(A[i * M1 + k]) > std::fabs(A[pr * M1 + k])) pr = i;
                if (pr != k) {
                    for (int t = 0; t < M1; t++) std::swap(A[k * M1 + t], A[pr * M1 + t]);
                    std::swap(rhs[k], rhs[pr]);
                }
                for (int i = k + 1; i < M1; i++) {
                    double f = A[i * M1 + k] / A[k * M1 + k];
                    for (int t = k; t < M1; t++) A[i * M1 + t] -= f * A[k * M1 + t];
                    rhs[i] -= f * rhs[k];
                }
            }
            std::vector<double> sol(M1);
            for (int k = M1 - 1; k >= 0; k--) {
                double t = rhs[k];
                for (int j2 = k + 1; j2 < M1; j2++) t -= A[k * M1 + j2] * sol[j2];
                sol[k] = t / A[k * M1 + k];
            }
            for (int i = 0; i < m; i++) dlam[i] = sol[i];
            double dz = sol[m];
            for (long j = 0; j < n; j++) {
                double gt = 0;
                for (int i = 0; i < m; i++) gt += GG[(size_t)i * n + j] * dlam[i];
                dx[j] = -delx[j] / diagx[j] - gt / diagx[j];
                dxsi[j] = -xsi[j] + epsi / (xs[j] - alfa[j]) - (xsi[j] * dx[j]) / (xs[j] - alfa[j]);
                deta[j] = -eta[j] + epsi / (beta[j] - xs[j]) + (eta[j] * dx[j]) / (beta[j] - xs[j]);
            }
            for (int i = 0; i < m; i++) {
                dy[i] = -dely[i] / diagy[i] + dlam[i] / diagy[i];
                dmu[i] = -mu[i] + epsi / y[i] - (mu[i] * dy[i]) / y[i];
                ds[i] = -s[i] + epsi / lam[i] - (s[i] * dlam[i]) / lam[i];
            }
            double dzet = -zet + epsi / z - zet * dz / z;
            double stm = 1.0;
            auto upd = [&](double v, double dv, double sign) { stm = std::max(stm, sign * 1.01 * dv / v); };
            for (int i = 0; i < m; i++) {
                upd(y[i], dy[i], -1);
                upd(lam[i], dlam[i], -1);
                upd(mu[i], dmu[i], -1);
                upd(s[i], ds[i], -1);
            }
            upd(z, dz, -1);
            upd(zet, dzet, -1);
            for (long j = 0; j < n; j++) {
                upd(xsi[j], dxsi[j], -1);
                upd(eta[j], deta[j], -1);
                upd(xs[j] - alfa[j], dx[j], -1);
                upd(beta[j] - xs[j], dx[j], 1);
            }
            double steg = 1.0 / stm;
            auto xold = xs, xsiold = xsi, etaold = eta, yold = y, lamold = lam, muold = mu, sold = s;
            double zold = z, zetold = zet;
            int itto = 0;
            double resinew = 2 * residunorm, resimax = residumax;
            while (resinew > residunorm && itto < 50) {
                itto++;
                for (long j = 0; j < n; j++) {
                    xs[j] = xold[j] + steg * dx[j];
                    xsi[j] = xsiold[j] + steg * dxsi[j];
                    eta[j] = etaold[j] + steg * deta[j];
                }
                for (int i = 0; i < m; i++) {
                    y[i] = yold[i] + steg * dy[i];
                    lam[i] = lamold[i] + steg * dlam[i];
                    mu[i] = muold[i] + steg * dmu[i];
                    s[i] = sold[i] + steg * ds[i];
                }
                z = zold + steg * dz;
                zet = zetold + steg * dzet;
                residual(resinew, resimax);
                steg /= 2;
            }
            residunorm = resinew;
            residumax = resimax;
        }
        epsi *= 0.1;
    }
    for (long j = 0; j < n; j++) {
        xold2[j] = xold1[j];
        xold1[j] = x[j];
        x[j] = xs[j];
    }
    return itera;
}

// Seeded "randc" density (the project generator; differs from the reference's time-seeded
// randTri, Framework.cuh:237-310):
//   rho~(p) = (1/8) sum_{a,b,c in 1..4} sum_{t in {cos,sin}^3} alpha * Tx(2 pi a px) Ty(2 pi b py) Tz(2 pi c pz),
//   p = (cell index + 0.5)/n, alpha ~ N(0,1) from Philox4x32-10 (key = seed) with Box-Muller;
//   rho1 = 1/(1+exp(-3 rho~)), rho2 = symmetry projection of rho1,
//   rho = clamp(s * rho2, 1e-3, 1) with s found by 100 bisection steps so that mean(rho) = vol.
// Coefficient k = ((((a-1)*4+(b-1))*4+(c-1))*8 + t), t bit j selects sin on axis j.
void orc_randc_coeffs(uint64_t seed, double* alpha) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t j = 0; j < 128; j++) {
        uint32_t ctr[4] = {j, 0x72616e64u, 0, 0}, r[4];
        philox4x32_10(ctr, key, r);
        for (int h = 0; h < 2; h++) {
            double u1 = ((double)r[2 * h] + 0.5) / 4294967296.0;
            double u2 = ((double)r[2 * h + 1] + 0.5) / 4294967296.0;
            double rad = std::sqrt(-2.0 * std::log(u1));
            alpha[4 * j + 2 * h] = rad * std::cos(2 * M_PI * u2);
            alpha[4 * j + 2 * h + 1] = rad * std::sin(2 * M_PI * u2);
        }
    }
}

void orc_randc_density(int n, uint64_t seed, double vol, int sym, float* out) {
    double alpha[512];
    orc_randc_coeffs(seed, alpha);
    orc_randc_field(n, alpha, vol, sym, out);
}

// The density of a given coefficient set alpha[512] (index as above).
void orc_randc_field(int n, const double* alpha, double vol, int sym, float* out) {
    // per-axis tables T[f][trig][i]
    std::vector<double> T(4 * 2 * n);
    for (int f = 0; f < 4; f++)
        for (int i = 0; i < n; i++) {
            double arg = 2 * M_PI * (f + 1) * ((i + 0.5) / n);
            T[(f * 2 + 0) * n + i] = std::cos(arg);
            T[(f * 2 + 1) * n + i] = std::sin(arg);
        }
    long N = (long)n * n * n;
    std::vector<float> rho1(N);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < n; z++) {
        // Cz[a][b][tx][ty] = sum_{c,tz} alpha * Tz
        double Cz[4][4][2][2];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++)
                for (int tx = 0; tx < 2; tx++)
                    for (int ty = 0; ty < 2; ty++) {
                        double s = 0;
                        for (int c = 0; c < 4; c++)
                            for (int tz = 0; tz < 2; tz++) {
                                int t = tx | (ty << 1) | (tz << 2);
                                s += alpha[(((a * 4 + b) * 4 + c) * 8) + t] * T[(c * 2 + tz) * n + z];
                            }
                        Cz[a][b][tx][ty] = s;
                    }
        for (int y = 0; y < n; y++) {
            double Byz[4][2];
            for (int a = 0; a < 4; a++)
                for (int tx = 0; tx < 2; tx++) {
                    double s = 0;
                    for (int b = 0; b < 4; b++)
                        for (int ty = 0; ty < 2; ty++) s += Cz[a][b][tx][ty] * T[(b * 2 + ty) * n + y];
                    Byz[a][tx] = s;
                }
            for (int x = 0; x < n; x++) {
                double s = 0;
                for (int a = 0; a < 4; a++)
                    for (int tx = 0; tx < 2; tx++) s += Byz[a][tx] * T[(a * 2 + tx) * n + x];
                s /= 8.0;
                rho1[x + (long)n * (y + (long)n * z)] = (float)(1.0 / (1.0 + std::exp(-3.0 * s)));
            }
        }
    }
    std::vector<float> rho2(N);
    orc_symmetrize_f32(n, sym, rho1.data(), rho2.data());
    double lo = 0, hi = 1e3;
    for (int it = 0; it < 100; it++) {
        double sc = 0.5 * (lo + hi), s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (long e = 0; e < N; e++) s += std::min(1.0, std::max(1e-3, sc * (double)rho2[e]));
        if (s / (double)N > vol)
            hi = sc;
        else
            lo = sc;
    }
    double sc = 0.5 * (lo + hi);
#pragma omp parallel for schedule(static)
    for (long e = 0; e < N; e++) out[e] = (float)std::min(1.0, std::max(1e-3, sc * (double)rho2[e]));
}

}  // extern "C"
