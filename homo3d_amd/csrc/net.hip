// Learned warm start: a small periodic 3D U-Net that maps a density field to the six load cases'
// displacement fields, trained and run on the GPU by the kernels in this file.
//
// The reference has no learned component; its warm start is the previous design's stored solutions
// (homogenization.cu:471-477, grid.cpp:459-462).  Here the network's prediction, upsampled to the
// solve grid, takes that place (homo3d_net_warm_start in api.cpp), and the FP64-refined MG-PCG remains
// the solver and the judge of the prediction's error.
//
// Architecture (levels L, channels c_l = C0 2^l, grid g / 2^l at level l; g >= 8 * 2^(L-1)):
//   encoder: E_l = relu(conv3(in_l)), S_l = relu(conv3(E_l)), in_0 = density, in_{l+1} = avgpool2(S_l)
//   decoder: CAT_l = [up2(D_{l+1}) | S_l], D_l = relu(conv3(relu(conv3(CAT_l)))), D_{L-1} = S_{L-1}
//   head:    P = conv1(D_0) (18 channels: u of load case s, component k at channel 3 s + k)
// conv3 is a periodic 3x3x3 correlation, tap t = (dx+1) + 3 (dy+1) + 9 (dz+1) reading X[v + (dx,dy,dz)].
//
// Storage: activations are bf16, channels-last [B][g^3][ld] (voxel v = x + g (y + g z) + g^3 b), so a
// lane's 8 consecutive channels are one 16-byte load.  S_l lives inside CAT_l (channel offset
// c_{l+1}), which makes the concatenation free.  Master weights, Adam state and gradients are FP32.
//
// Matrix cores: every convolution pass is an implicit GEMM on v_mfma_f32_16x16x32_bf16 (bf16
// operands, FP32 accumulation):
//   forward / input gradient (k_conv):  out[v][co] = sum_{t,ci} W[co][t][ci] X[v + d_t][ci]
//       A = weights (16 output channels x 32 of (t, ci)), B = the 16 voxels' shifted inputs.  The input
//       gradient is the same pass over dZ with the flipped, transposed weights Wb[ci][t][co] = W[co][26-t][ci].
//   weight gradient (k_wgrad):  dW[co][t][ci] = sum_v dZ[v][co] X[v + d_t][ci]
//       the reduction runs over voxels, so dZ and X are first transposed to channel-major rows with
//       X in three x-shifted copies (k_to_cm); a lane's 8 voxels are then one aligned 16-byte load for
//       every tap (y and z shifts move whole rows).  Split over voxel ranges, partial sums reduced
//       after (k_wgrad_reduce): deterministic, no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/homo3d.h"
#include "net.h"

using namespace homo3d;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kOut = 18;      // output channels
constexpr int kInCh = 8;      // stored input channels (density + 7 zeros): the 8-channel loads stay aligned
constexpr int kPld = 24;      // row stride of the FP32 prediction
constexpr int kDpLd = 32;     // row stride of the head's output gradient (18 used, zeros after)
constexpr int kDpC = 24;      // channels of it read by the head's input-gradient pass (multiple of 8)
constexpr int kWgChunks = 32; // 32-voxel chunks per wave in the weight gradient

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

// ---------------------------------------------------------------------------------------------------
// Convolution (forward and input gradient): implicit GEMM.  A wave computes 16 MB output channels for
// 16 NB voxels; a block holds 4 waves (64 NB consecutive voxels); blockIdx.y selects the channel block.
// cin % 8 == 0; the weight rows are padded to gridDim.y * 16 MB and to kpad (multiple of 32) with zeros.
// ---------------------------------------------------------------------------------------------------
template <int MB, int NB, bool TAP27, bool RELU, bool F32OUT>
__global__ __launch_bounds__(256) void k_conv(const uint16_t* __restrict__ in, int ldi, int cin,
                                              const uint16_t* __restrict__ w, int kpad,
                                              const float* __restrict__ bias, void* __restrict__ out, int ldo,
                                              int cout, int lg, long nvox) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, grp = lane >> 4;
    const long vb = ((long)blockIdx.x * 4 + wave) * (16 * NB);
    if (vb >= nvox) return;
    const int cb = blockIdx.y * 16 * MB;
    const int m = (1 << lg) - 1;
    int vx[NB], vy[NB], vz[NB];
    long vs[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        const long v = vb + 16 * j + col;
        vx[j] = (int)(v & m);
        vy[j] = (int)((v >> lg) & m);
        vz[j] = (int)((v >> (2 * lg)) & m);
        vs[j] = (v >> (3 * lg)) << (3 * lg);
    }
    f32x4 acc[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int ntaps = TAP27 ? 27 : 1;
    int tap = 0, ci = 8 * grp;  // (tap, ci) of k0 = 32 kc + 8 grp, advanced incrementally
    while (ci >= cin) {
        ci -= cin;
        tap++;
    }
    const int kch = kpad >> 5;
    for (int kc = 0; kc < kch; kc++) {
        const int k0 = kc * 32 + 8 * grp;
        bf16x8 a[MB], b[NB];
#pragma unroll
        for (int i = 0; i < MB; i++) a[i] = load8(w + (long)(cb + 16 * i + col) * kpad + k0);
        if (tap < ntaps) {
            const int dx = TAP27 ? tap % 3 - 1 : 0, dy = TAP27 ? (tap / 3) % 3 - 1 : 0, dz = TAP27 ? tap / 9 - 1 : 0;
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const long nb = vs[j] + (((long)((vz[j] + dz) & m) << (2 * lg)) | ((long)((vy[j] + dy) & m) << lg) |
                                         (long)((vx[j] + dx) & m));
                b[j] = load8(in + nb * ldi + ci);
            }
        } else {
#pragma unroll
            for (int j = 0; j < NB; j++) b[j] = bf16x8{};
        }
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        ci += 32;
        while (ci >= cin) {
            ci -= cin;
            tap++;
        }
    }
    // D layout: lane (col, grp) holds voxel vb + 16 j + col, channels cb + 16 i + 4 grp + r
#pragma unroll
    for (int i = 0; i < MB; i++) {
        const int co0 = cb + 16 * i + 4 * grp;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const long v = vb + 16 * j + col;
            float o[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                o[r] = acc[i][j][r] + (bias ? bias[co0 + r] : 0.f);
                if (RELU) o[r] = fmaxf(o[r], 0.f);
            }
            if (F32OUT) {
                float* dst = reinterpret_cast<float*>(out) + v * ldo + co0;
                if (co0 + 3 < cout) {
                    *reinterpret_cast<float4*>(dst) = make_float4(o[0], o[1], o[2], o[3]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (co0 + r < cout) dst[r] = o[r];
                }
            } else {
                uint16_t* dst = reinterpret_cast<uint16_t*>(out) + v * ldo + co0;
                if (co0 + 3 < cout) {
                    const uint32_t lo = (uint32_t)f2bf(o[0]) | ((uint32_t)f2bf(o[1]) << 16);
                    const uint32_t hi = (uint32_t)f2bf(o[2]) | ((uint32_t)f2bf(o[3]) << 16);
                    *reinterpret_cast<uint2*>(dst) = make_uint2(lo, hi);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (co0 + r < cout) dst[r] = f2bf(o[r]);
                }
            }
        }
    }
}

// Channel-major copy for the weight gradient: out[s][c][v] = in[v with x -> x + s + s0 - 1][c] for
// slots s < ns, c < cp (zeros for c >= C).  One thread per voxel.
__global__ __launch_bounds__(256) void k_to_cm(const uint16_t* __restrict__ in, int ldi, int C, int cp, int lg,
                                               long nvox, int s0, int ns, uint16_t* __restrict__ out) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int m = (1 << lg) - 1;
    const int x = (int)(v & m);
    for (int s = 0; s < ns; s++) {
        const int dx = s + s0 - 1;
        const long src = (v - x) + ((x + dx) & m);
        for (int c = 0; c < cp; c++) out[((long)s * cp + c) * nvox + v] = c < C ? in[src * ldi + c] : (uint16_t)0;
    }
}

// Weight gradient partial sums: one wave per (voxel range, 16 output channels, 16 input channels).
// dzc: [czp][nvox] channel-major dZ; xs: [ns][cxp][nvox] (ns = 3 x-shifted copies for 3x3x3, 1 for 1x1).
// partial[split][czp][taps][cxp].
template <bool TAP27>
__global__ __launch_bounds__(64) void k_wgrad(const uint16_t* __restrict__ dzc, int czp, const uint16_t* __restrict__ xs,
                                              int cxp, int lg, long nvox, float* __restrict__ partial) {
    constexpr int NT = TAP27 ? 27 : 1;
    const int lane = threadIdx.x, col = lane & 15, grp = lane >> 4;
    const int split = blockIdx.x, cb = blockIdx.y * 16, nb = blockIdx.z * 16;
    const int m = (1 << lg) - 1;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const long vbeg = (long)split * kWgChunks * 32;
    for (int c = 0; c < kWgChunks; c++) {
        const long v0 = vbeg + 32L * c;
        if (v0 >= nvox) break;
        const long vl = v0 + 8 * grp;  // the lane's 8 voxels: one x-row segment (g >= 8)
        const bf16x8 a = load8(dzc + (long)(cb + col) * nvox + vl);
        const long row = vl >> lg;
        const int xseg = (int)(vl & m);
        const int y = (int)(row & m), z = (int)((row >> lg) & m);
        const long sb = (row >> (2 * lg)) << (2 * lg);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int dx = TAP27 ? t % 3 - 1 : 0, dy = TAP27 ? (t / 3) % 3 - 1 : 0, dz = TAP27 ? t / 9 - 1 : 0;
            const int s = TAP27 ? dx + 1 : 0;
            const long r2 = sb + (((long)((z + dz) & m)) << lg) + ((y + dy) & m);
            const bf16x8 b = load8(xs + ((long)s * cxp + nb + col) * nvox + (r2 << lg) + xseg);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
    // lane (col, grp) holds input channel nb + col, output channels cb + 4 grp + r
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++)
            partial[(((long)split * czp + cb + 4 * grp + r) * NT + t) * cxp + nb + col] = acc[t][r];
}

// dW[co][t][ci] (co < cout, ci < cin; the flat parameter layout) = sum over splits.
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float* __restrict__ partial, int splits, int czp, int taps,
                                                      int cxp, int cout, int cin, float* __restrict__ dW) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)cout * taps * cin;
    if (i >= total) return;
    const int ci = (int)(i % cin), t = (int)((i / cin) % taps), co = (int)(i / ((long)cin * taps));
    float s = 0.f;
    for (int k = 0; k < splits; k++) s += partial[(((long)k * czp + co) * taps + t) * cxp + ci];
    dW[i] = s;
}

// db[co] = sum_v dZ[v][co], from the channel-major rows.
__global__ __launch_bounds__(1024) void k_rowsum(const uint16_t* __restrict__ dzc, long nvox, float* __restrict__ db) {
    __shared__ float sh[1024];
    const int co = blockIdx.x;
    float s = 0.f;
    for (long v = threadIdx.x; v < nvox; v += blockDim.x) s += bf2f(dzc[(long)co * nvox + v]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 512; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = sh[0];
}

// dZ = dY where Y > 0 else 0 (dY dense [v][C], overwritten; Y a strided view).
__global__ __launch_bounds__(256) void k_relu_mask(uint16_t* __restrict__ dy, int C, const uint16_t* __restrict__ y,
                                                   int ldy, long nvox) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvox * C) return;
    const long v = i / C;
    const int c = (int)(i - v * C);
    if (!(bf2f(y[v * ldy + c]) > 0.f)) dy[i] = 0;
}

// avgpool2: out[vc][c] (dense, grid g/2) = mean of the 8 children in the view (in, ldi), grid g.
__global__ __launch_bounds__(256) void k_pool(const uint16_t* __restrict__ in, int ldi, int C, int lgc, long nvc,
                                              uint16_t* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvc * C) return;
    const long vc = i / C;
    const int c = (int)(i - vc * C);
    const int mc = (1 << lgc) - 1, lg = lgc + 1;
    const int x = (int)(vc & mc), y = (int)((vc >> lgc) & mc), z = (int)((vc >> (2 * lgc)) & mc);
    const long b = vc >> (3 * lgc);
    float s = 0.f;
    for (int k = 0; k < 8; k++) {
        const long vf = (b << (3 * lg)) + ((long)(2 * z + (k >> 2)) << (2 * lg)) + ((long)(2 * y + ((k >> 1) & 1)) << lg) +
                        (2 * x + (k & 1));
        s += bf2f(in[vf * ldi + c]);
    }
    out[i] = f2bf(s * 0.125f);
}

// up2 into the first C channels of the view (out, ldo) at grid g from the dense [vc][C] at grid g/2.
__global__ __launch_bounds__(256) void k_up(const uint16_t* __restrict__ in, int C, int lg, long nvox,
                                            uint16_t* __restrict__ out, int ldo) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvox * C) return;
    const long v = i / C;
    const int c = (int)(i - v * C);
    const int m = (1 << lg) - 1, lgc = lg - 1;
    const int x = (int)(v & m), y = (int)((v >> lg) & m), z = (int)((v >> (2 * lg)) & m);
    const long b = v >> (3 * lg);
    const long vc = (b << (3 * lgc)) + ((long)(z >> 1) << (2 * lgc)) + ((long)(y >> 1) << lgc) + (x >> 1);
    out[v * ldo + c] = in[vc * C + c];
}

// gradient of up2: out[vc][c] (dense) = sum of the children's (in, ldi) entries c.
__global__ __launch_bounds__(256) void k_up_back(const uint16_t* __restrict__ in, int ldi, int C, int lgc, long nvc,
                                                 uint16_t* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvc * C) return;
    const long vc = i / C;
    const int c = (int)(i - vc * C);
    const int mc = (1 << lgc) - 1, lg = lgc + 1;
    const int x = (int)(vc & mc), y = (int)((vc >> lgc) & mc), z = (int)((vc >> (2 * lgc)) & mc);
    const long b = vc >> (3 * lgc);
    float s = 0.f;
    for (int k = 0; k < 8; k++) {
        const long vf = (b << (3 * lg)) + ((long)(2 * z + (k >> 2)) << (2 * lg)) + ((long)(2 * y + ((k >> 1) & 1)) << lg) +
                        (2 * x + (k & 1));
        s += bf2f(in[vf * ldi + c]);
    }
    out[i] = f2bf(s);
}

// Gradient of the skip tensor S_l (relu applied): out[v][c] = S > 0 ? dcat[v][cu + c] + dIn[parent][c] / 8 : 0,
// where dIn is the gradient of in_{l+1} = avgpool2(S_l) (dense [vc][C], grid g/2).
__global__ __launch_bounds__(256) void k_skip_grad(const uint16_t* __restrict__ s, const uint16_t* __restrict__ dcat,
                                                   int ld, int cu, int C, const uint16_t* __restrict__ din, int lg,
                                                   long nvox, uint16_t* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvox * C) return;
    const long v = i / C;
    const int c = (int)(i - v * C);
    if (!(bf2f(s[v * ld + cu + c]) > 0.f)) {
        out[i] = 0;
        return;
    }
    const int m = (1 << lg) - 1, lgc = lg - 1;
    const int x = (int)(v & m), y = (int)((v >> lg) & m), z = (int)((v >> (2 * lg)) & m);
    const long b = v >> (3 * lg);
    const long vc = (b << (3 * lgc)) + ((long)(z >> 1) << (2 * lgc)) + ((long)(y >> 1) << lgc) + (x >> 1);
    out[i] = f2bf(bf2f(dcat[v * ld + cu + c]) + 0.125f * bf2f(din[vc * C + c]));
}

// Stiffness weight of voxel v: w = rho^3 of the network input (the cell stiffness of the SIMP law).
__device__ __forceinline__ float net_weight(const uint16_t* in, long v) {
    const float r = bf2f(in[v * kInCh]);
    return r * r * r;
}

// Per (sample, channel) stiffness-weighted mean of the prediction P [v][kPld] over the sample's g^3
// voxels, and (channel 0 blocks) the sample's mean weight.
__global__ __launch_bounds__(256) void k_chan_mean(const float* __restrict__ P, const uint16_t* __restrict__ in, long nsv,
                                                   float* __restrict__ mean, float* __restrict__ wbar) {
    __shared__ float sh[2][256];
    const int b = blockIdx.x, c = blockIdx.y;
    float s = 0.f, ws = 0.f;
    for (long v = threadIdx.x; v < nsv; v += blockDim.x) {
        const long vv = (long)b * nsv + v;
        const float w = net_weight(in, vv);
        s += w * P[vv * kPld + c];
        ws += w;
    }
    sh[0][threadIdx.x] = s;
    sh[1][threadIdx.x] = ws;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + k];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        mean[b * kOut + c] = sh[1][0] > 0.f ? sh[0][0] / sh[1][0] : 0.f;
        if (c == 0) wbar[b] = sh[1][0] / (float)nsv;
    }
}

// Loss terms with stiffness weights (w / mean w per sample): d = P - mean_w P - T;
// dP = 2 (w / wbar) d / count (bf16, [v][kDpLd], channels >= 18 stay 0); per-block FP64 partial sums of
// (w / wbar) d^2.  T is mean-free under the same weights, so the weighted mean needs no gradient term.
__global__ __launch_bounds__(256) void k_head_loss(const float* __restrict__ P, const float* __restrict__ mean,
                                                   const float* __restrict__ wbar, const uint16_t* __restrict__ in,
                                                   const uint16_t* __restrict__ T, long nvox, int lgs, double inv_count,
                                                   uint16_t* __restrict__ dP, double* __restrict__ partial) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nvox * kOut; i += (long)gridDim.x * blockDim.x) {
        const long v = i / kOut;
        const int c = (int)(i - v * kOut);
        const long b = v >> (3 * lgs);
        const float wb = wbar[b];
        const float wt = wb > 0.f ? net_weight(in, v) / wb : 0.f;
        const float d = P[v * kPld + c] - mean[b * kOut + c] - bf2f(T[i]);
        dP[v * kDpLd + c] = f2bf((float)(2.0 * inv_count) * wt * d);
        s += (double)wt * d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// Adam on the FP32 master parameters (c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)).
__global__ __launch_bounds__(256) void k_adam(long n, const float* __restrict__ g, float* __restrict__ p,
                                              float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2,
                                              float c1, float c2, float eps) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// bf16 weight layouts of one layer from the master W[co][t][ci] (cin real input channels):
// forward Wf[r][k], k = t cinp + ci (rows r < rows_f, zero outside co < cout, ci < cin, k < taps cinp);
// input-gradient Wb[r][k], r = ci, k = t' coutp + co, Wb = W[co][flip t'][ci] (flip t = 26 - t for 3x3x3).
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ W, const float* __restrict__ bias, int cout,
                                              int cin, int cinp, int taps, int rows_f, int kpad_f,
                                              uint16_t* __restrict__ wf, float* __restrict__ bias_pad, int coutp,
                                              int rows_b, int kpad_b, uint16_t* __restrict__ wb) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nf = (long)rows_f * kpad_f, nbk = wb ? (long)rows_b * kpad_b : 0;
    if (i < nf) {
        const int r = (int)(i / kpad_f), k = (int)(i % kpad_f);
        const int t = k / cinp, ci = k % cinp;
        float v = 0.f;
        if (r < cout && t < taps && ci < cin) v = W[((long)r * taps + t) * cin + ci];
        wf[i] = f2bf(v);
        if (k == 0) bias_pad[r] = r < cout ? bias[r] : 0.f;
    } else if (i < nf + nbk) {
        const long j = i - nf;
        const int r = (int)(j / kpad_b), k = (int)(j % kpad_b);
        const int t = k / coutp, co = k % coutp;
        float v = 0.f;
        if (r < cin && t < taps && co < cout) v = W[((long)co * taps + (taps == 27 ? 26 - t : t)) * cin + r];
        wb[j] = f2bf(v);
    }
}

// density (FP32, [B][g^3]) -> input slots (bf16 [v][8], channel 0)
__global__ __launch_bounds__(256) void k_set_input(const float* __restrict__ rho, long nvox, uint16_t* __restrict__ in) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    uint4 z = make_uint4(f2bf(rho[v]), 0, 0, 0);
    *reinterpret_cast<uint4*>(in + v * kInCh) = z;
}

__global__ __launch_bounds__(256) void k_f32_to_bf16(const float* __restrict__ a, long n, uint16_t* __restrict__ b) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

inline unsigned nblk(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

int round_up(int a, int b) { return (a + b - 1) / b * b; }

}  // namespace

// =====================================================================================================
// Host side
// =====================================================================================================
struct NetLayer {
    int cin, cinp, cout, taps;
    bool relu;
    long off_w, off_b;      // into the flat parameter vector
    int mb_f, rows_f, kpad_f;
    int coutp, mb_b, rows_b, kpad_b;
    uint16_t *wf = nullptr, *wb = nullptr;
    float* bias_pad = nullptr;
};

struct NetLevel {
    int g, lg, c;
    long nvox;                   // batch_max g^3
    uint16_t* in = nullptr;      // level 0: input [v][8]; else avgpool2(S_{l-1}) [v][c_{l-1}]
    uint16_t* e = nullptr;       // E_l [v][c]
    uint16_t* cat = nullptr;     // CAT_l [v][c_{l+1} + c] (S_l at offset c_{l+1}); l < L-1
    uint16_t* s = nullptr;       // S_{L-1} [v][c] at the bottom level
    uint16_t* d1 = nullptr;      // decoder first conv [v][c]
    uint16_t* d = nullptr;       // D_l [v][c]
    uint16_t* ga = nullptr;      // gradient scratch [v][c] x 2
    uint16_t* gb = nullptr;
    uint16_t* dcat = nullptr;    // gradient of CAT_l
    uint16_t* din = nullptr;     // gradient of in_l (l > 0) [v][c_{l-1}]
};

struct homo3d_net {
    homo3d_net_config cfg{};
    hipStream_t stream = nullptr;
    int L = 0;
    std::vector<NetLayer> layers;
    std::vector<NetLevel> lv;
    long nparam = 0;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    long adam_t = 0;
    float* P = nullptr;          // prediction [v][kPld]
    uint16_t* T = nullptr;       // targets [v][18]
    uint16_t* dP = nullptr;      // [v][kDpLd]
    float* mean = nullptr;       // [batch][18]
    float* wbar = nullptr;       // [batch] mean stiffness weight
    double* lpart = nullptr;     // loss partial sums
    int lblocks = 1024;
    float* rho = nullptr;        // host-staged densities [batch][g^3]
    float* tstage = nullptr;     // host-staged targets [batch][g^3][18]
    uint16_t *zc = nullptr, *xc = nullptr;  // channel-major scratch for the weight gradient
    size_t zc_elems = 0, xc_elems = 0;
    float* partial = nullptr;
    size_t partial_elems = 0;
    // dataset (device resident): inputs [count][g^3][8], targets [count][g^3][18]
    uint16_t *ds_in = nullptr, *ds_tg = nullptr;
    int ds_cap = 0, ds_count = 0;
    std::string err;
};

namespace {

int nerr(homo3d_net* n, int code, const char* fmt, ...) {
    if (n) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        n->err = buf;
    }
    return code;
}

#define NETCHK(n, expr)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return nerr(n, HOMO3D_HIP_ERROR, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int pick_mb(int cout) { return cout >= 128 ? 8 : cout >= 64 ? 4 : cout >= 32 ? 2 : 1; }

template <int MB, bool TAP27, bool RELU, bool F32OUT>
void conv_mb(hipStream_t s, const uint16_t* in, int ldi, int cin, const uint16_t* w, int kpad, const float* bias,
             void* out, int ldo, int cout, int lg, long nvox) {
    constexpr int NB = 8 / MB;
    const dim3 grid((unsigned)((nvox + 64 * NB - 1) / (64 * NB)), (unsigned)((cout + 16 * MB - 1) / (16 * MB)));
    hipLaunchKernelGGL((k_conv<MB, NB, TAP27, RELU, F32OUT>), grid, dim3(256), 0, s, in, ldi, cin, w, kpad, bias, out, ldo,
                       cout, lg, nvox);
}

template <bool TAP27, bool RELU, bool F32OUT>
void conv_any(hipStream_t s, int mb, const uint16_t* in, int ldi, int cin, const uint16_t* w, int kpad, const float* bias,
              void* out, int ldo, int cout, int lg, long nvox) {
    switch (mb) {
        case 1: conv_mb<1, TAP27, RELU, F32OUT>(s, in, ldi, cin, w, kpad, bias, out, ldo, cout, lg, nvox); break;
        case 2: conv_mb<2, TAP27, RELU, F32OUT>(s, in, ldi, cin, w, kpad, bias, out, ldo, cout, lg, nvox); break;
        case 4: conv_mb<4, TAP27, RELU, F32OUT>(s, in, ldi, cin, w, kpad, bias, out, ldo, cout, lg, nvox); break;
        default: conv_mb<8, TAP27, RELU, F32OUT>(s, in, ldi, cin, w, kpad, bias, out, ldo, cout, lg, nvox); break;
    }
}

// forward pass of one layer: out = act(conv(in) + b)
void layer_fwd(homo3d_net* n, const NetLayer& Ly, const uint16_t* in, int ldi, void* out, int ldo, int lg, long nvox,
               bool f32out) {
    if (Ly.taps == 27)
        conv_any<true, true, false>(n->stream, Ly.mb_f, in, ldi, Ly.cinp, Ly.wf, Ly.kpad_f, Ly.bias_pad, out, ldo, Ly.cout,
                                    lg, nvox);
    else if (f32out)
        conv_any<false, false, true>(n->stream, Ly.mb_f, in, ldi, Ly.cinp, Ly.wf, Ly.kpad_f, Ly.bias_pad, out, ldo,
                                     Ly.cout, lg, nvox);
}

// input gradient of one layer: dX (dense or strided [v][ldo]) = conv(dZ, Wb)
void layer_dx(homo3d_net* n, const NetLayer& Ly, const uint16_t* dz, int ldz, int cz, uint16_t* dx, int ldo, int lg,
              long nvox) {
    if (Ly.taps == 27)
        conv_any<true, false, false>(n->stream, Ly.mb_b, dz, ldz, cz, Ly.wb, Ly.kpad_b, nullptr, dx, ldo, Ly.cin, lg, nvox);
    else
        conv_any<false, false, false>(n->stream, Ly.mb_b, dz, ldz, cz, Ly.wb, Ly.kpad_b, nullptr, dx, ldo, Ly.cin, lg,
                                      nvox);
}

// weight and bias gradient of one layer into n->g: X the layer input view (ldx, cx stored channels),
// dZ the output gradient view (ldz, cz channels with cz >= cout).
int layer_dw(homo3d_net* n, const NetLayer& Ly, const uint16_t* x, int ldx, int cx, const uint16_t* dz, int ldz, int cz,
             int lg, long nvox) {
    const int czp = round_up(cz, 16), cxp = round_up(cx, 16);
    const int ns = Ly.taps == 27 ? 3 : 1;
    if ((size_t)czp * nvox > n->zc_elems || (size_t)ns * cxp * nvox > n->xc_elems)
        return nerr(n, HOMO3D_BAD_ARG, "weight-gradient scratch too small");
    hipLaunchKernelGGL(k_to_cm, dim3(nblk(nvox)), dim3(256), 0, n->stream, dz, ldz, cz, czp, lg, nvox, 1, 1, n->zc);
    hipLaunchKernelGGL(k_to_cm, dim3(nblk(nvox)), dim3(256), 0, n->stream, x, ldx, cx, cxp, lg, nvox, ns == 3 ? 0 : 1, ns,
                       n->xc);
    const int splits = (int)((nvox + 32L * kWgChunks - 1) / (32L * kWgChunks));
    if ((size_t)splits * czp * Ly.taps * cxp > n->partial_elems) return nerr(n, HOMO3D_BAD_ARG, "partial buffer too small");
    const dim3 grid((unsigned)splits, (unsigned)(czp / 16), (unsigned)(cxp / 16));
    if (Ly.taps == 27)
        hipLaunchKernelGGL(k_wgrad<true>, grid, dim3(64), 0, n->stream, n->zc, czp, n->xc, cxp, lg, nvox, n->partial);
    else
        hipLaunchKernelGGL(k_wgrad<false>, grid, dim3(64), 0, n->stream, n->zc, czp, n->xc, cxp, lg, nvox, n->partial);
    const long tot = (long)Ly.cout * Ly.taps * Ly.cin;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(nblk(tot)), dim3(256), 0, n->stream, n->partial, splits, czp, Ly.taps, cxp,
                       Ly.cout, Ly.cin, n->g + Ly.off_w);
    hipLaunchKernelGGL(k_rowsum, dim3(Ly.cout), dim3(1024), 0, n->stream, n->zc, nvox, n->g + Ly.off_b);
    return HOMO3D_OK;
}

void relu_mask(homo3d_net* n, uint16_t* dy, int C, const uint16_t* y, int ldy, long nvox) {
    hipLaunchKernelGGL(k_relu_mask, dim3(nblk(nvox * C)), dim3(256), 0, n->stream, dy, C, y, ldy, nvox);
}

int pack_weights(homo3d_net* n) {
    for (const NetLayer& Ly : n->layers) {
        const long tot = (long)Ly.rows_f * Ly.kpad_f + (Ly.wb ? (long)Ly.rows_b * Ly.kpad_b : 0);
        hipLaunchKernelGGL(k_pack, dim3(nblk(tot)), dim3(256), 0, n->stream, n->p + Ly.off_w, n->p + Ly.off_b, Ly.cout,
                           Ly.cin, Ly.cinp, Ly.taps, Ly.rows_f, Ly.kpad_f, Ly.wf, Ly.bias_pad, Ly.coutp, Ly.rows_b,
                           Ly.kpad_b, Ly.wb);
    }
    NETCHK(n, hipGetLastError());
    return HOMO3D_OK;
}

// Forward pass over `batch` samples already in lv[0].in; prediction into n->P.
void forward(homo3d_net* n, int batch) {
    const int L = n->L;
    int li = 0;
    for (int l = 0; l < L; l++) {
        NetLevel& V = n->lv[l];
        const long nv = (long)batch << (3 * V.lg);
        const NetLayer& c1 = n->layers[li++];
        const NetLayer& c2 = n->layers[li++];
        const int ldi = l == 0 ? kInCh : n->lv[l - 1].c;
        layer_fwd(n, c1, V.in, ldi, V.e, V.c, V.lg, nv, false);
        if (l < L - 1) {
            const int cu = n->lv[l + 1].c, ld = cu + V.c;
            layer_fwd(n, c2, V.e, V.c, V.cat + cu, ld, V.lg, nv, false);
            const long nvc = nv / 8;
            hipLaunchKernelGGL(k_pool, dim3(nblk(nvc * V.c)), dim3(256), 0, n->stream, V.cat + cu, ld, V.c, V.lg - 1, nvc,
                               n->lv[l + 1].in);
        } else {
            layer_fwd(n, c2, V.e, V.c, V.s, V.c, V.lg, nv, false);
        }
    }
    // decoder
    for (int l = L - 2; l >= 0; l--) {
        NetLevel& V = n->lv[l];
        const NetLevel& W = n->lv[l + 1];
        const long nv = (long)batch << (3 * V.lg);
        const uint16_t* dn = l == L - 2 ? W.s : W.d;
        const int cu = W.c, ld = cu + V.c;
        hipLaunchKernelGGL(k_up, dim3(nblk(nv * cu)), dim3(256), 0, n->stream, dn, cu, V.lg, nv, V.cat, ld);
        const NetLayer& c1 = n->layers[li++];
        const NetLayer& c2 = n->layers[li++];
        layer_fwd(n, c1, V.cat, ld, V.d1, V.c, V.lg, nv, false);
        layer_fwd(n, c2, V.d1, V.c, V.d, V.c, V.lg, nv, false);
    }
    const NetLayer& hd = n->layers[li];
    const NetLevel& V0 = n->lv[0];
    const uint16_t* top = L > 1 ? V0.d : V0.s;
    layer_fwd(n, hd, top, V0.c, n->P, kPld, V0.lg, (long)batch << (3 * V0.lg), true);
}

// Loss of the prediction against n->T and the full parameter gradient into n->g.
int loss_backward(homo3d_net* n, int batch, double* loss) {
    const int L = n->L;
    const NetLevel& V0 = n->lv[0];
    const long nv0 = (long)batch << (3 * V0.lg);
    const long nsv = 1L << (3 * V0.lg);
    hipLaunchKernelGGL(k_chan_mean, dim3(batch, kOut), dim3(256), 0, n->stream, n->P, V0.in, nsv, n->mean, n->wbar);
    const double inv = 1.0 / ((double)nv0 * kOut);
    hipLaunchKernelGGL(k_head_loss, dim3(n->lblocks), dim3(256), 0, n->stream, n->P, n->mean, n->wbar, V0.in, n->T, nv0,
                       V0.lg, inv, n->dP, n->lpart);
    NETCHK(n, hipMemsetAsync(n->g, 0, sizeof(float) * n->nparam, n->stream));
    int li = (int)n->layers.size() - 1;
    int rc;
    // head
    {
        const NetLayer& hd = n->layers[li--];
        const uint16_t* top = L > 1 ? V0.d : V0.s;
        if ((rc = layer_dw(n, hd, top, V0.c, V0.c, n->dP, kDpLd, kOut, V0.lg, nv0))) return rc;
        layer_dx(n, hd, n->dP, kDpLd, kDpC, V0.ga, V0.c, V0.lg, nv0);  // dD_0
    }
    // decoder levels, top down in the parameter list (level 0 first)
    for (int l = 0; l <= L - 2; l++) {
        NetLevel& V = n->lv[l];
        const long nv = (long)batch << (3 * V.lg);
        const NetLayer& c2 = n->layers[li--];
        const NetLayer& c1 = n->layers[li--];
        const int cu = n->lv[l + 1].c, ld = cu + V.c;
        relu_mask(n, V.ga, V.c, V.d, V.c, nv);  // dZ of dec l.c2
        if ((rc = layer_dw(n, c2, V.d1, V.c, V.c, V.ga, V.c, V.c, V.lg, nv))) return rc;
        layer_dx(n, c2, V.ga, V.c, V.c, V.gb, V.c, V.lg, nv);
        relu_mask(n, V.gb, V.c, V.d1, V.c, nv);  // dZ of dec l.c1
        if ((rc = layer_dw(n, c1, V.cat, ld, ld, V.gb, V.c, V.c, V.lg, nv))) return rc;
        layer_dx(n, c1, V.gb, V.c, V.c, V.dcat, ld, V.lg, nv);
        // gradient of D_{l+1} (or S_{L-1}) = up2^T of the first cu channels
        NetLevel& W = n->lv[l + 1];
        const long nvc = nv / 8;
        hipLaunchKernelGGL(k_up_back, dim3(nblk(nvc * cu)), dim3(256), 0, n->stream, V.dcat, ld, cu, W.lg, nvc, W.ga);
    }
    // encoder, bottom up; lv[L-1].ga holds the gradient of S_{L-1}
    for (int l = L - 1; l >= 0; l--) {
        NetLevel& V = n->lv[l];
        const long nv = (long)batch << (3 * V.lg);
        const NetLayer& c2 = n->layers[2 * l + 1];
        const NetLayer& c1 = n->layers[2 * l];
        const uint16_t* S;
        int lds;
        if (l == L - 1) {
            S = V.s;
            lds = V.c;
            relu_mask(n, V.ga, V.c, S, lds, nv);
        } else {
            const int cu = n->lv[l + 1].c;
            S = V.cat + cu;
            lds = cu + V.c;
            hipLaunchKernelGGL(k_skip_grad, dim3(nblk(nv * V.c)), dim3(256), 0, n->stream, V.cat, V.dcat, lds, cu, V.c,
                               n->lv[l + 1].din, V.lg, nv, V.ga);
        }
        if ((rc = layer_dw(n, c2, V.e, V.c, V.c, V.ga, V.c, V.c, V.lg, nv))) return rc;
        layer_dx(n, c2, V.ga, V.c, V.c, V.gb, V.c, V.lg, nv);
        relu_mask(n, V.gb, V.c, V.e, V.c, nv);  // dZ of enc l.c1
        const int ldi = l == 0 ? kInCh : n->lv[l - 1].c;
        if ((rc = layer_dw(n, c1, V.in, ldi, l == 0 ? kInCh : ldi, V.gb, V.c, V.c, V.lg, nv))) return rc;
        if (l > 0) layer_dx(n, c1, V.gb, V.c, V.c, V.din, ldi, V.lg, nv);
    }
    NETCHK(n, hipGetLastError());
    if (loss) {
        std::vector<double> part(n->lblocks);
        NETCHK(n, hipMemcpyAsync(part.data(), n->lpart, sizeof(double) * n->lblocks, hipMemcpyDeviceToHost, n->stream));
        NETCHK(n, hipStreamSynchronize(n->stream));
        double s = 0;
        for (double v : part) s += v;
        *loss = s * inv;
    }
    return HOMO3D_OK;
}

int adam_step(homo3d_net* n, double lr) {
    n->adam_t++;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, (double)n->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, (double)n->adam_t)));
    hipLaunchKernelGGL(k_adam, dim3(nblk(n->nparam)), dim3(256), 0, n->stream, n->nparam, n->g, n->p, n->m1, n->m2,
                       (float)lr, (float)b1, (float)b2, c1, c2, 1e-8f);
    return pack_weights(n);
}

int stage_inputs(homo3d_net* n, const float* rho, int batch) {
    const long nsv = 1L << (3 * n->lv[0].lg);
    NETCHK(n, hipMemcpyAsync(n->rho, rho, sizeof(float) * nsv * batch, hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_set_input, dim3(nblk(nsv * batch)), dim3(256), 0, n->stream, n->rho, nsv * batch, n->lv[0].in);
    return HOMO3D_OK;
}

int stage_targets(homo3d_net* n, const float* t, int batch) {
    const long cnt = (1L << (3 * n->lv[0].lg)) * batch * kOut;
    NETCHK(n, hipMemcpyAsync(n->tstage, t, sizeof(float) * cnt, hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(cnt)), dim3(256), 0, n->stream, n->tstage, cnt, n->T);
    return HOMO3D_OK;
}

void destroy(homo3d_net* n) {
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (NetLayer& Ly : n->layers) {
        (void)hipFree(Ly.wf);
        (void)hipFree(Ly.wb);
        (void)hipFree(Ly.bias_pad);
    }
    for (NetLevel& V : n->lv) {
        void* ptrs[] = {V.in, V.e, V.cat, V.s, V.d1, V.d, V.ga, V.gb, V.dcat, V.din};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
    }
    void* ptrs[] = {n->p, n->g, n->m1, n->m2, n->P, n->T, n->dP, n->mean, n->wbar, n->lpart, n->rho, n->tstage, n->zc, n->xc,
                    n->partial, n->ds_in, n->ds_tg};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

}  // namespace

// ---------------------------------------------------------------------------------------------------
// Internal interface for api.cpp
// ---------------------------------------------------------------------------------------------------
namespace homo3d {
int net_grid(const homo3d_net* n) { return n->cfg.grid; }
int net_pred_ld(const homo3d_net*) { return kPld; }
float net_target_scale(const homo3d_net* n) { return (float)n->cfg.target_scale; }
hipStream_t net_stream(const homo3d_net* n) { return n->stream; }
uint16_t* net_input_slot(homo3d_net* n) { return n->lv[0].in; }
int net_forward_one(homo3d_net* n, const float** pred) {
    forward(n, 1);
    NETCHK(n, hipGetLastError());
    *pred = n->P;
    return HOMO3D_OK;
}
int net_dataset_push(homo3d_net* n, uint16_t** input, uint16_t** target) {
    if (n->ds_count >= n->ds_cap) return nerr(n, HOMO3D_BAD_ARG, "dataset full (%d pairs)", n->ds_cap);
    const long nsv = (long)n->cfg.grid * n->cfg.grid * n->cfg.grid;
    *input = n->ds_in + (long)n->ds_count * nsv * kInCh;
    *target = n->ds_tg + (long)n->ds_count * nsv * kOut;
    n->ds_count++;
    return HOMO3D_OK;
}
int net_set_error(homo3d_net* n, int code, const char* msg) { return nerr(n, code, "%s", msg); }
}  // namespace homo3d

// ---------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------
extern "C" {

void homo3d_net_default_config(homo3d_net_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->grid = 64;
    c->levels = 4;
    c->base_channels = 16;
    c->batch = 4;
    c->target_scale = 8.0;
    c->seed = 0;
    c->device = 0;
}

int homo3d_net_create(const homo3d_net_config* cfg, homo3d_net** out) {
    if (!cfg || !out) return HOMO3D_BAD_ARG;
    *out = nullptr;
    const int g = cfg->grid, L = cfg->levels, c0 = cfg->base_channels, B = cfg->batch;
    int lg = 0;
    while ((1 << lg) < g) lg++;
    if ((1 << lg) != g || L < 1 || L > 6 || (g >> (L - 1)) < 8 || c0 < 8 || c0 % 8 || (c0 << (L - 1)) > 256 || B < 1 ||
        B > 64 || !(cfg->target_scale > 0))
        return HOMO3D_NO_SUPPORT;
    homo3d_net* n = new homo3d_net();
    n->cfg = *cfg;
    n->L = L;
    auto fail = [&](int code) {
        destroy(n);
        return code;
    };
    if (hipSetDevice(cfg->device) != hipSuccess) return fail(HOMO3D_HIP_ERROR);
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) return fail(HOMO3D_HIP_ERROR);
    // layers (parameter order: enc0.c1, enc0.c2, ..., enc{L-1}.c2, dec{L-2}.c1, dec{L-2}.c2, ..., dec0.c2, head)
    auto add = [&](int cin, int cinp, int cout, int taps, bool relu, bool need_dx) {
        NetLayer Ly;
        Ly.cin = cin;
        Ly.cinp = cinp;
        Ly.cout = cout;
        Ly.taps = taps;
        Ly.relu = relu;
        Ly.off_w = n->nparam;
        n->nparam += (long)cout * taps * cin;
        Ly.off_b = n->nparam;
        n->nparam += cout;
        Ly.mb_f = pick_mb(cout);
        Ly.rows_f = round_up(cout, 16 * Ly.mb_f);
        Ly.kpad_f = round_up(taps * cinp, 32);
        Ly.coutp = round_up(cout, 8);
        Ly.mb_b = pick_mb(cin);
        Ly.rows_b = round_up(cin, 16 * Ly.mb_b);
        Ly.kpad_b = round_up(taps * Ly.coutp, 32);
        n->layers.push_back(Ly);
        NetLayer& Q = n->layers.back();
        return !(hipMalloc(&Q.wf, sizeof(uint16_t) * Q.rows_f * Q.kpad_f) ||
                 hipMalloc(&Q.bias_pad, sizeof(float) * Q.rows_f) ||
                 (need_dx && hipMalloc(&Q.wb, sizeof(uint16_t) * Q.rows_b * Q.kpad_b)));
    };
    std::vector<int> ch(L);
    for (int l = 0; l < L; l++) ch[l] = c0 << l;
    bool ok = true;
    for (int l = 0; l < L && ok; l++) {
        ok = add(l == 0 ? 1 : ch[l - 1], l == 0 ? kInCh : ch[l - 1], ch[l], 27, true, l > 0) &&
             add(ch[l], ch[l], ch[l], 27, true, true);
    }
    for (int l = L - 2; l >= 0 && ok; l--)
        ok = add(ch[l + 1] + ch[l], ch[l + 1] + ch[l], ch[l], 27, true, true) && add(ch[l], ch[l], ch[l], 27, true, true);
    ok = ok && add(ch[0], ch[0], kOut, 1, false, true);
    if (!ok) return fail(HOMO3D_HIP_ERROR);
    // activations
    n->lv.resize(L);
    size_t max_zc = 0, max_xc = 0, max_part = 0;
    for (int l = 0; l < L; l++) {
        NetLevel& V = n->lv[l];
        V.g = g >> l;
        V.lg = lg - l;
        V.c = ch[l];
        V.nvox = (long)B << (3 * V.lg);
        const long nv = V.nvox;
        const int cin = l == 0 ? kInCh : ch[l - 1];
        const int cu = l < L - 1 ? ch[l + 1] : 0;
        auto mal = [&](uint16_t** q, long elems) { return hipMalloc((void**)q, sizeof(uint16_t) * elems) == hipSuccess; };
        ok = mal(&V.in, nv * cin) && mal(&V.e, nv * V.c) && mal(&V.ga, nv * std::max(V.c, 1)) && mal(&V.gb, nv * V.c);
        if (l > 0) ok = ok && mal(&V.din, nv * cin);
        if (l < L - 1)
            ok = ok && mal(&V.cat, nv * (cu + V.c)) && mal(&V.dcat, nv * (cu + V.c)) && mal(&V.d1, nv * V.c) &&
                 mal(&V.d, nv * V.c);
        else
            ok = ok && mal(&V.s, nv * V.c);
        if (!ok) return fail(HOMO3D_HIP_ERROR);
        if (l == 0 && hipMemset(V.in, 0, sizeof(uint16_t) * nv * kInCh)) return fail(HOMO3D_HIP_ERROR);
        // weight-gradient scratch: the widest layer input (CAT) and output of this level
        const int cx = round_up(std::max(cin, cu + V.c), 16);
        max_xc = std::max(max_xc, (size_t)3 * cx * nv);
        max_zc = std::max(max_zc, (size_t)round_up(std::max(V.c, kOut), 16) * nv);
        const long splits = (nv + 32L * kWgChunks - 1) / (32L * kWgChunks);
        max_part = std::max(max_part, (size_t)splits * round_up(std::max(V.c, kOut), 16) * 27 * cx);
    }
    n->zc_elems = max_zc;
    n->xc_elems = max_xc;
    n->partial_elems = max_part;
    const long nv0 = n->lv[0].nvox;
    if (hipMalloc(&n->p, sizeof(float) * n->nparam) || hipMalloc(&n->g, sizeof(float) * n->nparam) ||
        hipMalloc(&n->m1, sizeof(float) * n->nparam) || hipMalloc(&n->m2, sizeof(float) * n->nparam) ||
        hipMalloc(&n->P, sizeof(float) * nv0 * kPld) || hipMalloc(&n->T, sizeof(uint16_t) * nv0 * kOut) ||
        hipMalloc(&n->dP, sizeof(uint16_t) * nv0 * kDpLd) || hipMalloc(&n->mean, sizeof(float) * B * kOut) || hipMalloc(&n->wbar, sizeof(float) * B) ||
        hipMalloc(&n->lpart, sizeof(double) * n->lblocks) || hipMalloc(&n->rho, sizeof(float) * nv0) ||
        hipMalloc(&n->tstage, sizeof(float) * nv0 * kOut) ||
        hipMalloc(&n->zc, sizeof(uint16_t) * max_zc) || hipMalloc(&n->xc, sizeof(uint16_t) * max_xc) ||
        hipMalloc(&n->partial, sizeof(float) * max_part))
        return fail(HOMO3D_HIP_ERROR);
    if (hipMemset(n->dP, 0, sizeof(uint16_t) * nv0 * kDpLd) || hipMemset(n->m1, 0, sizeof(float) * n->nparam) ||
        hipMemset(n->m2, 0, sizeof(float) * n->nparam) || hipMemset(n->P, 0, sizeof(float) * nv0 * kPld))
        return fail(HOMO3D_HIP_ERROR);
    // He-normal initial weights (head: sqrt(1 / cin) / 4), zero biases
    std::vector<float> hp(n->nparam, 0.f);
    std::mt19937_64 rng(cfg->seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (const NetLayer& Ly : n->layers) {
        const double sd = Ly.relu ? std::sqrt(2.0 / (Ly.taps * Ly.cin)) : std::sqrt(1.0 / Ly.cin) / 4;
        for (long i = 0; i < (long)Ly.cout * Ly.taps * Ly.cin; i++) hp[Ly.off_w + i] = (float)(nd(rng) * sd);
    }
    if (hipMemcpy(n->p, hp.data(), sizeof(float) * n->nparam, hipMemcpyHostToDevice)) return fail(HOMO3D_HIP_ERROR);
    if (pack_weights(n)) return fail(HOMO3D_HIP_ERROR);
    if (hipStreamSynchronize(n->stream)) return fail(HOMO3D_HIP_ERROR);
    *out = n;
    return HOMO3D_OK;
}

int homo3d_net_destroy(homo3d_net* n) {
    if (n) destroy(n);
    return HOMO3D_OK;
}

const char* homo3d_net_last_error(const homo3d_net* n) { return n ? n->err.c_str() : "null network"; }

long homo3d_net_num_params(const homo3d_net* n) { return n ? n->nparam : -1; }

int homo3d_net_params(homo3d_net* n, float* host, int to_device) {
    if (!n || !host) return HOMO3D_BAD_ARG;
    if (to_device) {
        NETCHK(n, hipMemcpyAsync(n->p, host, sizeof(float) * n->nparam, hipMemcpyHostToDevice, n->stream));
        NETCHK(n, hipMemsetAsync(n->m1, 0, sizeof(float) * n->nparam, n->stream));
        NETCHK(n, hipMemsetAsync(n->m2, 0, sizeof(float) * n->nparam, n->stream));
        n->adam_t = 0;
        int rc = pack_weights(n);
        if (rc) return rc;
    } else {
        NETCHK(n, hipMemcpyAsync(host, n->p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->stream));
    }
    NETCHK(n, hipStreamSynchronize(n->stream));
    return HOMO3D_OK;
}

int homo3d_net_forward(homo3d_net* n, const float* rho, int batch, float* out) {
    if (!n || !rho || !out) return HOMO3D_BAD_ARG;
    if (batch < 1 || batch > n->cfg.batch) return nerr(n, HOMO3D_BAD_ARG, "batch %d outside [1, %d]", batch, n->cfg.batch);
    int rc = stage_inputs(n, rho, batch);
    if (rc) return rc;
    forward(n, batch);
    NETCHK(n, hipGetLastError());
    const long nsv = 1L << (3 * n->lv[0].lg);
    NETCHK(n, hipMemcpy2DAsync(out, sizeof(float) * kOut, n->P, sizeof(float) * kPld, sizeof(float) * kOut,
                               nsv * batch, hipMemcpyDeviceToHost, n->stream));
    NETCHK(n, hipStreamSynchronize(n->stream));
    return HOMO3D_OK;
}

int homo3d_net_loss_grad(homo3d_net* n, const float* rho, const float* target, int batch, double* loss, float* grad) {
    if (!n || !rho || !target || !loss) return HOMO3D_BAD_ARG;
    if (batch < 1 || batch > n->cfg.batch) return nerr(n, HOMO3D_BAD_ARG, "batch %d outside [1, %d]", batch, n->cfg.batch);
    int rc = stage_inputs(n, rho, batch);
    if (!rc) rc = stage_targets(n, target, batch);
    if (rc) return rc;
    forward(n, batch);
    rc = loss_backward(n, batch, loss);
    if (rc) return rc;
    if (grad) NETCHK(n, hipMemcpyAsync(grad, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->stream));
    NETCHK(n, hipStreamSynchronize(n->stream));
    return HOMO3D_OK;
}

int homo3d_net_train_step(homo3d_net* n, const float* rho, const float* target, int batch, double lr, double* loss) {
    double l = 0;
    int rc = homo3d_net_loss_grad(n, rho, target, batch, &l, nullptr);
    if (rc) return rc;
    rc = adam_step(n, lr);
    if (rc) return rc;
    NETCHK(n, hipStreamSynchronize(n->stream));
    if (loss) *loss = l;
    return HOMO3D_OK;
}

int homo3d_net_dataset_reserve(homo3d_net* n, int count) {
    if (!n || count < 1) return HOMO3D_BAD_ARG;
    NETCHK(n, hipStreamSynchronize(n->stream));
    if (n->ds_in) (void)hipFree(n->ds_in);
    if (n->ds_tg) (void)hipFree(n->ds_tg);
    n->ds_in = n->ds_tg = nullptr;
    n->ds_cap = n->ds_count = 0;
    const long nsv = (long)n->cfg.grid * n->cfg.grid * n->cfg.grid;
    NETCHK(n, hipMalloc(&n->ds_in, sizeof(uint16_t) * nsv * kInCh * count));
    NETCHK(n, hipMalloc(&n->ds_tg, sizeof(uint16_t) * nsv * kOut * count));
    n->ds_cap = count;
    return HOMO3D_OK;
}

int homo3d_net_dataset_size(const homo3d_net* n) { return n ? n->ds_count : -1; }

// Adam epochs over the first `count` pairs of the device-resident dataset (all when count <= 0): each epoch
// visits them in a seeded random order in batches of the configured size (the last partial batch is
// dropped).  losses[e] = mean batch loss of epoch e.
int homo3d_net_train(homo3d_net* n, int epochs, int count, double lr, uint64_t seed, double* losses) {
    if (!n || epochs < 0) return HOMO3D_BAD_ARG;
    const int B = n->cfg.batch;
    if (count <= 0 || count > n->ds_count) count = n->ds_count;
    if (count < B) return nerr(n, HOMO3D_BAD_ARG, "training on %d pairs, fewer than one batch", count);
    const long nsv = (long)n->cfg.grid * n->cfg.grid * n->cfg.grid;
    std::mt19937_64 rng(seed);
    std::vector<int> order(count);
    for (int i = 0; i < count; i++) order[i] = i;
    for (int e = 0; e < epochs; e++) {
        std::shuffle(order.begin(), order.end(), rng);
        double sum = 0;
        int steps = 0;
        for (int s = 0; s + B <= count; s += B) {
            for (int b = 0; b < B; b++) {
                const long k = order[s + b];
                NETCHK(n, hipMemcpyAsync(n->lv[0].in + b * nsv * kInCh, n->ds_in + k * nsv * kInCh,
                                         sizeof(uint16_t) * nsv * kInCh, hipMemcpyDeviceToDevice, n->stream));
                NETCHK(n, hipMemcpyAsync(n->T + b * nsv * kOut, n->ds_tg + k * nsv * kOut, sizeof(uint16_t) * nsv * kOut,
                                         hipMemcpyDeviceToDevice, n->stream));
            }
            forward(n, B);
            double l = 0;
            int rc = loss_backward(n, B, &l);
            if (rc) return rc;
            rc = adam_step(n, lr);
            if (rc) return rc;
            sum += l;
            steps++;
        }
        if (losses) losses[e] = steps ? sum / steps : 0.0;
    }
    NETCHK(n, hipStreamSynchronize(n->stream));
    return HOMO3D_OK;
}

// Mean loss of the network over dataset pairs [first, first + count) (no update).
int homo3d_net_eval(homo3d_net* n, int first, int count, double* loss) {
    if (!n || !loss || first < 0 || count < 1 || first + count > n->ds_count) return HOMO3D_BAD_ARG;
    const long nsv = (long)n->cfg.grid * n->cfg.grid * n->cfg.grid;
    double sum = 0;
    for (int k = first; k < first + count; k++) {
        NETCHK(n, hipMemcpyAsync(n->lv[0].in, n->ds_in + (long)k * nsv * kInCh, sizeof(uint16_t) * nsv * kInCh,
                                 hipMemcpyDeviceToDevice, n->stream));
        NETCHK(n, hipMemcpyAsync(n->T, n->ds_tg + (long)k * nsv * kOut, sizeof(uint16_t) * nsv * kOut,
                                 hipMemcpyDeviceToDevice, n->stream));
        forward(n, 1);
        double l = 0;
        int rc = loss_backward(n, 1, &l);
        if (rc) return rc;
        sum += l;
    }
    *loss = sum / count;
    return HOMO3D_OK;
}

}  // extern "C"
