// Seeded "randc" density generator (host).  Replaces the reference's time-seeded randTri
// (homogenization/Framework.cuh:237-310, culib/lib.cuh:1141-1146) with a reproducible definition:
//
//   rho~(p) = (1/8) sum_{a,b,c in 1..4} sum_{t in {cos,sin}^3} alpha_k Tx(2 pi a px) Ty(2 pi b py) Tz(2 pi c pz)
//   p = (cell index + 0.5) / n; alpha_k ~ N(0,1): Philox4x32-10 (Salmon et al., SC'11; key = seed,
//   counter = (j, 0x72616e64, 0, 0), j = 0..127) and the Box-Muller transform;
//   rho1 = 1/(1+exp(-3 rho~)); rho2 = symmetry projection of rho1 (group average in FP64 over the
//   orbit's smallest-index member); rho = clamp(s rho2, 1e-3, 1) with s from 100 bisection steps
//   so that mean(rho) = vol.  Coefficient k = ((((a-1)*4+(b-1))*4+(c-1))*8 + t), bit j of t selects sin
//   on axis j.
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/homo3d.h"

namespace {

void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t m0 = (uint64_t)0xD2511F53u * c0, m1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(m1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(m0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)m1;
        c3 = (uint32_t)m0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
const int kPermSign[6] = {1, -1, -1, 1, 1, -1};

void group(int sym, std::vector<std::pair<int, int>>& g) {
    g.clear();
    if (sym == HOMO3D_SYM_REFLECT3)
        for (int f = 0; f < 8; f++) g.push_back({0, f});
    else if (sym == HOMO3D_SYM_REFLECT6)
        for (int p = 0; p < 6; p++)
            for (int f = 0; f < 8; f++) g.push_back({p, f});
    else if (sym == HOMO3D_SYM_ROTATE3)
        for (int p = 0; p < 6; p++)
            for (int f = 0; f < 8; f++) {
                const int nf = (f & 1) + ((f >> 1) & 1) + (f >> 2);
                if (kPermSign[p] * ((nf & 1) ? -1 : 1) == 1) g.push_back({p, f});
            }
}

inline long gapply(const std::pair<int, int>& g, int n, const int p[3]) {
    int q[3];
    for (int k = 0; k < 3; k++) {
        const int a = p[kPerm[g.first][k]];
        q[k] = ((g.second >> k) & 1) ? n - 1 - a : a;
    }
    return q[0] + (long)n * (q[1] + (long)n * q[2]);
}

// Run body(begin, end) over [0, count) split into contiguous ranges on up to 16 host threads.
template <typename F>
void parallel_ranges(long count, const F& body) {
    const long nt = std::max(1L, std::min<long>(16, std::min<long>((long)std::thread::hardware_concurrency(), count / 4096)));
    if (nt <= 1) {
        body(0L, count);
        return;
    }
    std::vector<std::thread> th;
    for (long t = 0; t < nt; t++) th.emplace_back([&, t] { body(count * t / nt, count * (t + 1) / nt); });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" int homo3d_randc_density(int n, uint64_t seed, double vol, int sym, float* out) {
    if (n < 1 || !out || !(vol > 0 && vol <= 1)) return HOMO3D_BAD_ARG;
    double alpha[512];
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t j = 0; j < 128; j++) {
        const uint32_t ctr[4] = {j, 0x72616e64u, 0, 0};
        uint32_t r[4];
        philox(ctr, key, r);
        for (int h = 0; h < 2; h++) {
            const double u1 = ((double)r[2 * h] + 0.5) / 4294967296.0;
            const double u2 = ((double)r[2 * h + 1] + 0.5) / 4294967296.0;
            const double rad = std::sqrt(-2.0 * std::log(u1));
            alpha[4 * j + 2 * h] = rad * std::cos(2 * M_PI * u2);
            alpha[4 * j + 2 * h + 1] = rad * std::sin(2 * M_PI * u2);
        }
    }
    std::vector<double> T(8 * (size_t)n);
    for (int f = 0; f < 4; f++)
        for (int i = 0; i < n; i++) {
            const double arg = 2 * M_PI * (f + 1) * ((i + 0.5) / n);
            T[(f * 2 + 0) * (size_t)n + i] = std::cos(arg);
            T[(f * 2 + 1) * (size_t)n + i] = std::sin(arg);
        }
    const long N = (long)n * n * n;
    std::vector<float> rho1(N), rho2(N);
    parallel_ranges((long)n * 4096, [&](long zb, long ze) {
    for (int z = (int)(zb / 4096); z < (int)(ze / 4096); z++) {
        double Cz[4][4][2][2];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++)
                for (int tx = 0; tx < 2; tx++)
                    for (int ty = 0; ty < 2; ty++) {
                        double s = 0;
                        for (int c = 0; c < 4; c++)
                            for (int tz = 0; tz < 2; tz++)
                                s += alpha[(((a * 4 + b) * 4 + c) * 8) + (tx | (ty << 1) | (tz << 2))] *
                                     T[(c * 2 + tz) * (size_t)n + z];
                        Cz[a][b][tx][ty] = s;
                    }
        for (int y = 0; y < n; y++) {
            double Byz[4][2];
            for (int a = 0; a < 4; a++)
                for (int tx = 0; tx < 2; tx++) {
                    double s = 0;
                    for (int b = 0; b < 4; b++)
                        for (int ty = 0; ty < 2; ty++) s += Cz[a][b][tx][ty] * T[(b * 2 + ty) * (size_t)n + y];
                    Byz[a][tx] = s;
                }
            for (int x = 0; x < n; x++) {
                double s = 0;
                for (int a = 0; a < 4; a++)
                    for (int tx = 0; tx < 2; tx++) s += Byz[a][tx] * T[(a * 2 + tx) * (size_t)n + x];
                s /= 8.0;
                rho1[x + (long)n * (y + (long)n * z)] = (float)(1.0 / (1.0 + std::exp(-3.0 * s)));
            }
        }
    }
    });
    std::vector<std::pair<int, int>> g;
    group(sym, g);
    if (g.empty()) {
        rho2 = rho1;
    } else {
        parallel_ranges(N, [&](long eb, long ee) {
            for (long e = eb; e < ee; e++) {
                const int p[3] = {(int)(e % n), (int)(e / n % n), (int)(e / ((long)n * n))};
                long best = e;
                for (auto& ge : g) best = std::min(best, gapply(ge, n, p));
                const int q[3] = {(int)(best % n), (int)(best / n % n), (int)(best / ((long)n * n))};
                double s = 0;
                for (auto& ge : g) s += (double)rho1[gapply(ge, n, q)];
                rho2[e] = (float)(s / (double)g.size());
            }
        });
    }
    // mean(clamp(s rho2)) over the cells for each bisection step: 64 fixed ranges summed in order (the
    // same result for any thread count)
    constexpr int kParts = 64;
    double lo = 0, hi = 1e3;
    for (int it = 0; it < 100; it++) {
        const double sc = 0.5 * (lo + hi);
        double part[kParts];
        parallel_ranges(kParts * 4096L, [&](long pb, long pe) {
            for (long k = pb / 4096; k < pe / 4096; k++) {
                double s = 0;
                for (long e = N * k / kParts; e < N * (k + 1) / kParts; e++)
                    s += std::min(1.0, std::max(1e-3, sc * (double)rho2[e]));
                part[k] = s;
            }
        });
        double s = 0;
        for (int k = 0; k < kParts; k++) s += part[k];
        if (s / (double)N > vol)
            hi = sc;
        else
            lo = sc;
    }
    const double sc = 0.5 * (lo + hi);
    parallel_ranges(N, [&](long eb, long ee) {
        for (long e = eb; e < ee; e++) out[e] = (float)std::min(1.0, std::max(1e-3, sc * (double)rho2[e]));
    });
    return HOMO3D_OK;
}
