// Host driver of the GPU MMA step (homo::MMAOptimizer, optimization/mmaOptimizer.cu:18-77, and
// mmasub_g / subsolv_g, optimization/mmaOpt.cu:429-1062).  Per-variable work and every reduction
// run on the GPU in FP64; the m-sized dual variables and the (m+1) x (m+1) Newton system
// (solveLinearHost, mmaOpt.cpp:48-73; here Gaussian elimination with partial pivoting) stay on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/homo3d.h"
#include "kernels.h"

using namespace homo3d;

struct homo3d_mma {
    int m = 0;
    long n = 0;
    double a0 = 1, a[kMmaMaxM], c[kMmaMaxM], d[kMmaMaxM];
    hipStream_t stream = nullptr;
    int nb = 1024;
    double *x = nullptr, *xmin = nullptr, *xmax = nullptr, *xold1 = nullptr, *xold2 = nullptr, *low = nullptr,
           *upp = nullptr, *df0 = nullptr, *dfdx = nullptr, *alfa = nullptr, *beta = nullptr, *p0 = nullptr,
           *q0 = nullptr, *P = nullptr, *Q = nullptr, *xs = nullptr, *xsi = nullptr, *eta = nullptr, *xso = nullptr,
           *xsio = nullptr, *etao = nullptr, *delx = nullptr, *diagx = nullptr, *dx = nullptr, *dxsi = nullptr,
           *deta = nullptr, *part = nullptr, *partmax = nullptr, *out = nullptr, *dlam = nullptr;
    float* stage = nullptr;
    bool initialised = false;
    int last_newton = 0;
    std::string err;
};

namespace {

int mma_err(homo3d_mma* h, int code, const char* msg) {
    if (h) h->err = msg;
    return code;
}

#define MMACHK(h, expr)                                                          \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) return mma_err(h, HOMO3D_HIP_ERROR, hipGetErrorString(e_)); \
    } while (0)

bool solve_small(int M1, std::vector<double>& A, std::vector<double>& rhs, std::vector<double>& sol) {
    for (int k = 0; k < M1; k++) {
        int pr = k;
        for (int i = k + 1; i < M1; i++)
            if (std::fabs(A[i * M1 + k]) > std::fabs(A[pr * M1 + k])) pr = i;
        if (A[pr * M1 + k] == 0) return false;
        if (pr != k) {
            for (int t = 0; t < M1; t++) std::swap(A[k * M1 + t], A[pr * M1 + t]);
            std::swap(rhs[k], rhs[pr]);
        }
        for (int i = k + 1; i < M1; i++) {
            const double f = A[i * M1 + k] / A[k * M1 + k];
            for (int t = k; t < M1; t++) A[i * M1 + t] -= f * A[k * M1 + t];
            rhs[i] -= f * rhs[k];
        }
    }
    sol.assign(M1, 0.0);
    for (int k = M1 - 1; k >= 0; k--) {
        double t = rhs[k];
        for (int j = k + 1; j < M1; j++) t -= A[k * M1 + j] * sol[j];
        sol[k] = t / A[k * M1 + k];
    }
    return true;
}

}  // namespace

extern "C" {

int homo3d_mma_create(int m, long n, double a0, double a, double c, double d, int device, homo3d_mma** out) {
    if (!out || m < 1 || m > kMmaMaxM || n < 1) return HOMO3D_BAD_ARG;
    *out = nullptr;
    homo3d_mma* h = new homo3d_mma();
    h->m = m;
    h->n = n;
    h->a0 = a0;
    for (int i = 0; i < kMmaMaxM; i++) {
        h->a[i] = a;
        h->c[i] = c;
        h->d[i] = d;
    }
    auto fail = [&](int code) {
        homo3d_mma_destroy(h);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(HOMO3D_HIP_ERROR);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(HOMO3D_HIP_ERROR);
    double** vecs[] = {&h->x, &h->xmin, &h->xmax, &h->xold1, &h->xold2, &h->low, &h->upp, &h->df0, &h->alfa,
                       &h->beta, &h->p0, &h->q0, &h->xs, &h->xsi, &h->eta, &h->xso, &h->xsio, &h->etao,
                       &h->delx, &h->diagx, &h->dx, &h->dxsi, &h->deta};
    for (double** v : vecs)
        if (hipMalloc(v, sizeof(double) * n) != hipSuccess) return fail(HOMO3D_HIP_ERROR);
    if (hipMalloc(&h->dfdx, sizeof(double) * n * m) || hipMalloc(&h->P, sizeof(double) * n * m) ||
        hipMalloc(&h->Q, sizeof(double) * n * m) || hipMalloc(&h->part, sizeof(double) * h->nb * kMmaNewtonSums) ||
        hipMalloc(&h->partmax, sizeof(double) * h->nb) || hipMalloc(&h->out, sizeof(double) * 64) ||
        hipMalloc(&h->dlam, sizeof(double) * kMmaMaxM) || hipMalloc(&h->stage, sizeof(float) * n))
        return fail(HOMO3D_HIP_ERROR);
    std::vector<double> ones(n, 1.0);
    if (hipMemset(h->xmin, 0, sizeof(double) * n) || hipMemset(h->xold1, 0, sizeof(double) * n) ||
        hipMemset(h->xold2, 0, sizeof(double) * n) || hipMemset(h->low, 0, sizeof(double) * n) ||
        hipMemset(h->upp, 0, sizeof(double) * n) ||
        hipMemcpy(h->xmax, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice))
        return fail(HOMO3D_HIP_ERROR);
    *out = h;
    return HOMO3D_OK;
}

int homo3d_mma_destroy(homo3d_mma* h) {
    if (!h) return HOMO3D_OK;
    // teardown: errors are not actionable here, so they are ignored
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* ptrs[] = {h->x, h->xmin, h->xmax, h->xold1, h->xold2, h->low, h->upp, h->df0, h->dfdx, h->alfa, h->beta,
                    h->p0, h->q0, h->P, h->Q, h->xs, h->xsi, h->eta, h->xso, h->xsio, h->etao, h->delx, h->diagx,
                    h->dx, h->dxsi, h->deta, h->part, h->partmax, h->out, h->dlam, h->stage};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return HOMO3D_OK;
}

const char* homo3d_mma_last_error(const homo3d_mma* h) { return h ? h->err.c_str() : "null handle"; }

// MMAOptimizer::setBound(float xmin, float xmax) (mmaOptimizer.cu:53-57)
int homo3d_mma_set_bound(homo3d_mma* h, double xmin, double xmax) {
    if (!h || !(xmax > xmin)) return HOMO3D_BAD_ARG;
    std::vector<double> lo(h->n, xmin), hi(h->n, xmax);
    MMACHK(h, hipMemcpyAsync(h->xmin, lo.data(), sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
    MMACHK(h, hipMemcpyAsync(h->xmax, hi.data(), sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
    MMACHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

// MMAOptimizer::setBound(float* xmin, float* xmax) (mmaOptimizer.h:30): per-variable bounds, n floats
// each, in host (on_device = 0) or device memory (1); -1 classifies the pointers.
int homo3d_mma_set_bounds(homo3d_mma* h, const float* xmin, const float* xmax, int on_device) {
    if (!h || !xmin || !xmax) return HOMO3D_BAD_ARG;
    if (on_device < 0) {
        hipPointerAttribute_t attr;
        if (hipPointerGetAttributes(&attr, xmin) != hipSuccess) {
            (void)hipGetLastError();
            on_device = 0;
        } else {
            on_device = (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) ? 1 : 0;
        }
    }
    if (on_device) {
        launch_f32_to_f64(h->stream, h->n, xmin, h->xmin);
        launch_f32_to_f64(h->stream, h->n, xmax, h->xmax);
        MMACHK(h, hipStreamSynchronize(h->stream));
        return HOMO3D_OK;
    }
    std::vector<double> lo(h->n), hi(h->n);
    for (long i = 0; i < h->n; i++) {
        if (!(xmax[i] >= xmin[i])) return mma_err(h, HOMO3D_BAD_ARG, "setBound: xmax < xmin");
        lo[i] = xmin[i];
        hi[i] = xmax[i];
    }
    MMACHK(h, hipMemcpyAsync(h->xmin, lo.data(), sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
    MMACHK(h, hipMemcpyAsync(h->xmax, hi.data(), sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
    MMACHK(h, hipStreamSynchronize(h->stream));
    return HOMO3D_OK;
}

// One MMA step (MMAOptimizer::update, mmaOptimizer.cu:61-77): x (n floats) is updated in place.
// dfdx: n floats; gval: m floats (host); dgdx: m pointers to n floats.  Arrays are device
// pointers when on_device != 0, host pointers otherwise.
int homo3d_mma_update(homo3d_mma* h, int itn, float* x, const float* dfdx, const float* gval, const float* const* dgdx,
                      int on_device) {
    if (!h || !x || !dfdx || !gval || !dgdx) return HOMO3D_BAD_ARG;
    if (on_device < 0) {
        hipPointerAttribute_t attr;
        if (hipPointerGetAttributes(&attr, x) != hipSuccess) {
            (void)hipGetLastError();
            on_device = 0;
        } else {
            on_device = (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) ? 1 : 0;
        }
    }
    const int m = h->m;
    const long n = h->n;
    const int nb = h->nb;
    hipStream_t s = h->stream;
    const hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    auto widen = [&](const float* src, double* dst) -> int {
        MMACHK(h, hipMemcpyAsync(h->stage, src, sizeof(float) * n, kin, s));
        launch_f32_to_f64(s, n, h->stage, dst);
        return HOMO3D_OK;
    };
    int rc;
    if ((rc = widen(x, h->x))) return rc;
    if (!h->initialised) {
        MMACHK(h, hipMemcpyAsync(h->xold1, h->x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        MMACHK(h, hipMemcpyAsync(h->xold2, h->x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        h->initialised = true;
    }
    if ((rc = widen(dfdx, h->df0))) return rc;
    for (int i = 0; i < m; i++)
        if ((rc = widen(dgdx[i], h->dfdx + (size_t)i * n))) return rc;
    double hout[64];
    launch_mma_setup(s, m, n, itn, 1.0, h->x, h->xmin, h->xmax, h->xold1, h->xold2, h->df0, h->dfdx, h->low, h->upp,
                     h->alfa, h->beta, h->p0, h->q0, h->P, h->Q, h->part, nb, h->out);
    MMACHK(h, hipMemcpyAsync(hout, h->out, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    MMACHK(h, hipStreamSynchronize(s));
    double b[kMmaMaxM];
    for (int i = 0; i < m; i++) b[i] = hout[i] - (double)gval[i];

    // subsolv
    std::vector<double> y(m, 1.0), lam(m, 1.0), mu(m), sv(m, 1.0), gvec(m);
    double z = 1, zet = 1, epsi = 1;
    for (int i = 0; i < m; i++) mu[i] = std::max(1.0, 0.5 * h->c[i]);
    launch_mma_init(s, n, h->alfa, h->beta, h->xs, h->xsi, h->eta);
    auto residual = [&](double& nrm, double& mx) -> int {
        launch_mma_resid(s, m, n, lam.data(), epsi, h->xs, h->xsi, h->eta, h->low, h->upp, h->alfa, h->beta, h->p0,
                         h->q0, h->P, h->Q, h->part, h->partmax, nb, h->out);
        MMACHK(h, hipMemcpyAsync(hout, h->out, sizeof(double) * (2 + m), hipMemcpyDeviceToHost, s));
        MMACHK(h, hipStreamSynchronize(s));
        double s2 = hout[0], m2 = hout[1 + m];
        for (int i = 0; i < m; i++) gvec[i] = hout[1 + i];
        auto acc = [&](double v) {
            s2 += v * v;
            m2 = std::max(m2, std::fabs(v));
        };
        double alam = 0;
        for (int i = 0; i < m; i++) {
            acc(h->c[i] + h->d[i] * y[i] - mu[i] - lam[i]);
            alam += h->a[i] * lam[i];
        }
        acc(h->a0 - zet - alam);
        for (int i = 0; i < m; i++) acc(gvec[i] - h->a[i] * z - y[i] + sv[i] - b[i]);
        for (int i = 0; i < m; i++) acc(mu[i] * y[i] - epsi);
        acc(zet * z - epsi);
        for (int i = 0; i < m; i++) acc(lam[i] * sv[i] - epsi);
        nrm = std::sqrt(s2);
        mx = m2;
        return HOMO3D_OK;
    };
    int itera = 0;
    while (epsi > 1e-7) {
        double residunorm, residumax;
        if ((rc = residual(residunorm, residumax))) return rc;
        int ittt = 0;
        while (residumax > 0.9 * epsi && ittt < 200) {
            ittt++;
            itera++;
            launch_mma_newton(s, m, n, lam.data(), epsi, h->xs, h->xsi, h->eta, h->low, h->upp, h->alfa, h->beta,
                              h->p0, h->q0, h->P, h->Q, h->delx, h->diagx, h->part, nb, h->out);
            MMACHK(h, hipMemcpyAsync(hout, h->out, sizeof(double) * kMmaNewtonSums, hipMemcpyDeviceToHost, s));
            MMACHK(h, hipStreamSynchronize(s));
            double alam = 0;
            for (int i = 0; i < m; i++) alam += h->a[i] * lam[i];
            const double delz = h->a0 - alam - epsi / z;
            std::vector<double> dely(m), dellam(m), diagy(m);
            const int M1 = m + 1;
            std::vector<double> A((size_t)M1 * M1, 0.0), rhs(M1, 0.0), sol;
            for (int i = 0; i < m; i++) {
                dely[i] = h->c[i] + h->d[i] * y[i] - lam[i] - epsi / y[i];
                dellam[i] = gvec[i] - h->a[i] * z - y[i] - b[i] + epsi / lam[i];
                diagy[i] = h->d[i] + mu[i] / y[i];
                rhs[i] = dellam[i] + dely[i] / diagy[i] - hout[i];
                for (int q = i; q < m; q++) {
                    const double v = hout[kMmaMaxM + i * kMmaMaxM - i * (i - 1) / 2 + (q - i)];
                    A[i * M1 + q] = A[q * M1 + i] = v;
                }
                A[i * M1 + i] += sv[i] / lam[i] + 1.0 / diagy[i];
                A[i * M1 + m] = h->a[i];
                A[m * M1 + i] = h->a[i];
            }
            A[m * M1 + m] = -zet / z;
            rhs[m] = delz;
            if (!solve_small(M1, A, rhs, sol)) return mma_err(h, HOMO3D_DIVERGED, "singular MMA Newton system");
            std::vector<double> dlam(sol.begin(), sol.begin() + m), dy(m), dmu(m), ds(m);
            const double dz = sol[m];
            MMACHK(h, hipMemcpyAsync(h->dlam, dlam.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
            launch_mma_dir(s, m, n, epsi, h->dlam, h->xs, h->xsi, h->eta, h->low, h->upp, h->alfa, h->beta, h->P,
                           h->Q, h->delx, h->diagx, h->dx, h->dxsi, h->deta, h->partmax, nb, h->out);
            double stmx;
            MMACHK(h, hipMemcpyAsync(&stmx, h->out, sizeof(double), hipMemcpyDeviceToHost, s));
            MMACHK(h, hipStreamSynchronize(s));
            for (int i = 0; i < m; i++) {
                dy[i] = -dely[i] / diagy[i] + dlam[i] / diagy[i];
                dmu[i] = -mu[i] + epsi / y[i] - (mu[i] * dy[i]) / y[i];
                ds[i] = -sv[i] + epsi / lam[i] - (sv[i] * dlam[i]) / lam[i];
            }
            const double dzet = -zet + epsi / z - zet * dz / z;
            double stm = std::max(1.0, stmx);
            for (int i = 0; i < m; i++) {
                stm = std::max(stm, -1.01 * dy[i] / y[i]);
                stm = std::max(stm, -1.01 * dlam[i] / lam[i]);
                stm = std::max(stm, -1.01 * dmu[i] / mu[i]);
                stm = std::max(stm, -1.01 * ds[i] / sv[i]);
            }
            stm = std::max(stm, -1.01 * dz / z);
            stm = std::max(stm, -1.01 * dzet / zet);
            double steg = 1.0 / stm;
            MMACHK(h, hipMemcpyAsync(h->xso, h->xs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            MMACHK(h, hipMemcpyAsync(h->xsio, h->xsi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            MMACHK(h, hipMemcpyAsync(h->etao, h->eta, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            const std::vector<double> yold = y, lamold = lam, muold = mu, sold = sv;
            const double zold = z, zetold = zet;
            int itto = 0;
            double resinew = 2 * residunorm, resimax = residumax;
            while (resinew > residunorm && itto < 50) {
                itto++;
                launch_mma_step(s, n, steg, h->xso, h->xsio, h->etao, h->dx, h->dxsi, h->deta, h->xs, h->xsi, h->eta);
                for (int i = 0; i < m; i++) {
                    y[i] = yold[i] + steg * dy[i];
                    lam[i] = lamold[i] + steg * dlam[i];
                    mu[i] = muold[i] + steg * dmu[i];
                    sv[i] = sold[i] + steg * ds[i];
                }
                z = zold + steg * dz;
                zet = zetold + steg * dzet;
                if ((rc = residual(resinew, resimax))) return rc;
                steg /= 2;
            }
            residunorm = resinew;
            residumax = resimax;
        }
        epsi *= 0.1;
    }
    h->last_newton = itera;
    // xold2 <- xold1 <- x <- solution (mmasub_g, mmaOpt.cu:1057-1059)
    MMACHK(h, hipMemcpyAsync(h->xold2, h->xold1, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    MMACHK(h, hipMemcpyAsync(h->xold1, h->x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    MMACHK(h, hipMemcpyAsync(h->x, h->xs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    launch_f64_to_f32(s, n, h->x, h->stage);
    MMACHK(h, hipMemcpyAsync(x, h->stage, sizeof(float) * n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                             s));
    MMACHK(h, hipStreamSynchronize(s));
    return HOMO3D_OK;
}

int homo3d_mma_state(homo3d_mma* h, double* x, double* low, double* upp, int* newton_iters) {
    if (!h) return HOMO3D_BAD_ARG;
    const size_t b = sizeof(double) * h->n;
    if (x) MMACHK(h, hipMemcpyAsync(x, h->x, b, hipMemcpyDeviceToHost, h->stream));
    if (low) MMACHK(h, hipMemcpyAsync(low, h->low, b, hipMemcpyDeviceToHost, h->stream));
    if (upp) MMACHK(h, hipMemcpyAsync(upp, h->upp, b, hipMemcpyDeviceToHost, h->stream));
    MMACHK(h, hipStreamSynchronize(h->stream));
    if (newton_iters) *newton_iters = h->last_newton;
    return HOMO3D_OK;
}

}  // extern "C"
