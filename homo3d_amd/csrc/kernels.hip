// gfx950 kernels of the periodic homogenization solve path.
//
// Grid convention: a cubic periodic level of n^3 vertices; vertex/cell id v = x + n*(y + n*z); cell
// (x,y,z) has its minimum corner at vertex (x,y,z); node b of a cell sits at corner +
// (b&1, b>>1&1, b>>2) (homogenization/templateMatrix.cpp:79-93).  Fields hold R = 6 load cases:
// component c of load case r of vertex v lives at (r*3 + c)*N + v.  Coarse stencils are stored
// structure-of-arrays: entry k = d*9 + i*3 + j (d = (dx+1) + 3(dy+1) + 9(dz+1)) of vertex v at k*N + v.
//
// The finest level is matrix-free: the 27 3x3 blocks of a vertex row are rebuilt from the 8 cell
// weights w_e = Emin + x_e and the element matrix KE, which is read with wave-uniform indices so
// the compiler keeps it in scalar registers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/homo3d.h"
#include "kernels.h"
#include "net.h"

namespace homo3d {

namespace {

__device__ __forceinline__ int wrap_dec(int a, int n) { return a == 0 ? n - 1 : a - 1; }
__device__ __forceinline__ int wrap_inc(int a, int n) { return a == n - 1 ? 0 : a + 1; }

// Storage order of a level's vertices (and of the cells, which share the vertex index).  Even levels
// are parity-blocked: the 8 coordinate-parity classes are contiguous (n/2)^3 blocks, block
// b = (x&1) | (y&1)<<1 | (z&1)<<2, and inside a block (x>>1) + h((y>>1) + h(z>>1)).  A GS colour is
// one block, so colour passes read their own vertices and every neighbour class contiguously.
// The coarsest level (possibly odd) is lexicographic.  Vertex (0,0,0) has index 0 in both orders.
struct Lay {
    int n;        // points per axis
    int h;        // n / 2
    uint32_t h3;  // h^3
    int lex;      // 1: x + n(y + n z)
    // index contribution of (already wrapped) coordinate a on axis k
    __device__ __forceinline__ uint32_t part(int a, int k) const {
        if (lex) return k == 0 ? (uint32_t)a : (k == 1 ? (uint32_t)a * n : (uint32_t)a * n * n);
        const uint32_t stride = k == 0 ? 1u : (k == 1 ? (uint32_t)h : (uint32_t)h * h);
        return ((uint32_t)(a & 1) << k) * h3 + (uint32_t)(a >> 1) * stride;
    }
    __device__ __forceinline__ uint32_t idx(int x, int y, int z) const { return part(x, 0) + part(y, 1) + part(z, 2); }
    __device__ __forceinline__ void pos(uint32_t t, int& x, int& y, int& z) const {
        if (lex) {
            x = (int)(t % n);
            y = (int)((t / n) % n);
            z = (int)(t / ((uint32_t)n * n));
            return;
        }
        const uint32_t b = t / h3, l = t - b * h3;
        x = 2 * (int)(l % h) + (int)(b & 1);
        y = 2 * (int)((l / h) % h) + (int)((b >> 1) & 1);
        z = 2 * (int)(l / ((uint32_t)h * h)) + (int)(b >> 2);
    }
    __device__ __forceinline__ int wrap(int a) const { return a < 0 ? a + n : (a >= n ? a - n : a); }
};

// Workgroup id -> work chunk.  The hardware deals workgroup g to XCD g % 8; this gives each XCD a
// contiguous run of chunks, so the neighbour rows that adjacent chunks share stay in one L2.
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t g, uint32_t G) {
    if (G & 7u) return g;
    return (g & 7u) * (G >> 3) + (g >> 3);
}

// Full-level kernels run 512-thread workgroups.  On a parity-blocked level the 8 waves of a
// workgroup take the same 64 local positions in the 8 parity blocks, so a workgroup covers a
// 128 x 2 x 2 vertex tile and the neighbour reads of its waves overlap in L1/L2; on a
// lexicographic level the 512 threads take consecutive vertices.  False past the end of the level.
constexpr int kFullThreads = 512;
__device__ __forceinline__ bool full_vertex(const Lay& L, uint32_t chunk, uint32_t& v) {
    if (L.lex) {
        v = chunk * (uint32_t)kFullThreads + threadIdx.x;
        return v < (uint32_t)L.n * L.n * L.n;
    }
    const uint32_t t = chunk * 64u + (threadIdx.x & 63u);
    v = (threadIdx.x >> 6) * L.h3 + t;
    return t < L.h3;
}

// Buffer-resource loads: one SGPR descriptor per array, a 32-bit VGPR byte offset per element and an
// SGPR byte offset per (load case, component) plane.  Offsets past the array read as zero.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc make_rsrc(const void* p, size_t bytes) {
    const uint32_t nrec = bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)nrec, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(Rsrc r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float bload<float>(Rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

// Packed FP16 storage of a 6-load-case field: word plane (p * 3 + i) holds components i of load cases
// 2p (.x) and 2p + 1 (.y), so a neighbour of all six load cases is 9 32-bit words.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
template <>
__device__ __forceinline__ half2_t bload<half2_t>(Rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(half2_t, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

// The FP64 iterate of the refinement loop stores its 18 components as nine planes of 16-byte pairs
// (u64_index), so a neighbour of all six load cases is nine 16-byte loads rather than eighteen
// 8-byte ones: the gather's cost is per load instruction.
typedef double double2_t __attribute__((ext_vector_type(2)));
template <>
__device__ __forceinline__ double2_t bload<double2_t>(Rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double2_t, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}

__device__ __forceinline__ void bstore(float v, Rsrc r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
}

// load cases per stored word of a field type
template <typename A>
struct CasesPerWord {
    static constexpr int value = 1;
};
template <>
struct CasesPerWord<half2_t> {
    static constexpr int value = 2;
};

// Packed FP16 six-load-case fields (half2_t words): word q = p * 3 + i of vertex v (component i of load
// cases 2p (.x) and 2p + 1 (.y)) is word hw_index(q, v, N) of the field.  Words 0-3 and 4-7 form two
// planes of 16-byte quads and word 8 a plane of single words, so the nine words of a vertex are two
// 16-byte loads and one 4-byte load: a 27-neighbour gather costs the texture addresser per load
// instruction, not per byte (nine 4-byte loads per neighbour take about four times as long).
__host__ __device__ __forceinline__ size_t hw_index(int q, size_t v, size_t N) {
    return q < 8 ? (size_t)(q >> 2) * 4 * N + 4 * v + (size_t)(q & 3) : 8 * N + v;
}
// word q of vertex v in a field of A words: packed FP16 as above, FP32 planes q * N + v
template <typename A>
__device__ __forceinline__ size_t fw_index(int q, size_t v, size_t N) {
    if constexpr (CasesPerWord<A>::value == 2)
        return hw_index(q, v, N);
    else
        return (size_t)q * N + v;
}
// the nine words of vertex idx of a packed FP16 field through a buffer resource over it (plane = 4 N
// bytes; the byte offsets stay below 36 N)
__device__ __forceinline__ void hload9(uint32_t (&wv)[9], Rsrc r, uint32_t idx, uint32_t plane) {
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(idx * 16u), 0, 0);
    const auto b = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(idx * 16u), (int)(4u * plane), 0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        wv[k] = (uint32_t)a[k];
        wv[4 + k] = (uint32_t)b[k];
    }
    wv[8] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)(idx * 4u), (int)(8u * plane), 0);
}

// The element matrix staged in LDS: KE (576 entries) followed by 243 block constants (kBlockConsts),
// padded so that what follows it stays 16-byte aligned.
constexpr int kBlockConsts = 576;
constexpr int kKeStaged = 832;

// Sum over the cells that contain both v and v + d of the cell weights w_e, each with the sign
// prod over the axes m in SM of (e_m ? -1 : +1).  Cell e (0..7) sits at offset (e&1, e>>1&1, e>>2) - 1
// from v; it holds v + d iff e_m = (d_m + 1) / 2 on every axis with d_m != 0.
template <int D, int SM, typename T, typename W>
__device__ __forceinline__ T cell_form(const W wc[8]) {
    constexpr int dm[3] = {D % 3 - 1, D / 3 % 3 - 1, D / 9 - 1};
    T s = T(0);
    bool first = true;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        bool in = true;
#pragma unroll
        for (int m = 0; m < 3; m++)
            if (dm[m] != 0 && ((e >> m) & 1) != (dm[m] + 1) / 2) in = false;
        if (!in) continue;
        const bool neg = (__builtin_popcount(e & SM) & 1) != 0;
        const T we = (T)wc[e];
        s = first ? (neg ? -we : we) : (neg ? s - we : s + we);
        first = false;
    }
    return s;
}

// 3x3 block of the assembled fine operator between vertex v and its neighbour v+d:
// S = sum over the cells e containing both of w_e * KE[node(v)][node(v+d)], v being node 7-e of cell e.
// For the isotropic cube element, entry (i, j) of KE[a][b] over those cells is one constant times
// s_i s_j (i != j; s_m = +1 at the upper node of axis m, the sign of every axis along which d is
// non-zero being the same for all the cells) and one constant (i = j), so
// S_ij = C_d,ij * (sum of w_e, with the sign s_m(e) for each axis m in {i, j} along which d is zero).
// C_d,ij (ke[kBlockConsts + 9 d + 3 i + j], set by stage_ke) is the entry of the cell with e_m = 1
// exactly where d_m = +1.  Four signed sums of two to eight weights and nine products per block,
// instead of nine FMAs per cell.
template <int D, typename T, typename W>
__device__ __forceinline__ void fine_block(const W wc[8], const T* ke, T S[9]) {
    constexpr int fm = (D % 3 == 1 ? 1 : 0) | (D / 3 % 3 == 1 ? 2 : 0) | (D / 9 == 1 ? 4 : 0);  // axes with d_m = 0
    const T* cb = ke + kBlockConsts + 9 * D;
    const T w0 = cell_form<D, 0, T>(wc);
    const T fxy = cell_form<D, fm & 3, T>(wc), fxz = cell_form<D, fm & 5, T>(wc), fyz = cell_form<D, fm & 6, T>(wc);
    S[0] = cb[0] * w0;
    S[4] = cb[4] * w0;
    S[8] = cb[8] * w0;
    S[1] = cb[1] * fxy;
    S[3] = cb[3] * fxy;
    S[2] = cb[2] * fxz;
    S[6] = cb[6] * fxz;
    S[5] = cb[5] * fyz;
    S[7] = cb[7] * fyz;
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < BS / 64; i++) s += sh[i];
    }
    __syncthreads();
    return s;
}

struct FineCtx {
    uint32_t X[3], Y[3], Z[3];  // index parts of the coordinates -1, 0, +1 per axis
    float wc[8];
    double wd[8];  // FP64 weights Emin + x (refinement residual only)
    __device__ __forceinline__ uint32_t nb(int d) const { return X[d % 3] + Y[d / 3 % 3] + Z[d / 9]; }
    // cell at offset (ex-1, ey-1, ez-1) from the vertex
    __device__ __forceinline__ uint32_t cell(int e) const { return X[e & 1] + Y[(e >> 1) & 1] + Z[e >> 2]; }
};

__device__ __forceinline__ void fine_parts(FineCtx& c, int x, int y, int z, const Lay& L) {
    c.X[0] = L.part(wrap_dec(x, L.n), 0); c.X[1] = L.part(x, 0); c.X[2] = L.part(wrap_inc(x, L.n), 0);
    c.Y[0] = L.part(wrap_dec(y, L.n), 1); c.Y[1] = L.part(y, 1); c.Y[2] = L.part(wrap_inc(y, L.n), 1);
    c.Z[0] = L.part(wrap_dec(z, L.n), 2); c.Z[1] = L.part(z, 2); c.Z[2] = L.part(wrap_inc(z, L.n), 2);
}

__device__ __forceinline__ void fine_ctx(FineCtx& c, int x, int y, int z, const Lay& L, const float* __restrict__ w) {
    fine_parts(c, x, y, z, L);
    const Rsrc rw = make_rsrc(w, (size_t)L.n * L.n * L.n * sizeof(float));
#pragma unroll
    for (int e = 0; e < 8; e++) c.wc[e] = bload<float>(rw, c.cell(e) * 4u, 0);
}

// FP64 cell weights w = Emin + x from the raw cell values x (FP32 input data, exact in FP64).
__device__ __forceinline__ void fine_ctx64(FineCtx& c, int x, int y, int z, const Lay& L, const float* __restrict__ xc,
                                           double emin) {
    fine_parts(c, x, y, z, L);
    const Rsrc rw = make_rsrc(xc, (size_t)L.n * L.n * L.n * sizeof(float));
#pragma unroll
    for (int e = 0; e < 8; e++) {
        c.wd[e] = emin + (double)bload<float>(rw, c.cell(e) * 4u, 0);
        c.wc[e] = (float)c.wd[e];
    }
}

// Neighbour masks: bit d (d = 0..26, offset (d%3-1, d/3%3-1, d/9-1)) selects the neighbours a row
// gathers.  kRowAll is the whole row, kRowOffDiag the row without the vertex itself (d = 13).
constexpr uint32_t kRowAll = (1u << 27) - 1u;
constexpr int kSumThreads = 1024;       // workgroup of k_sum_partials
constexpr uint32_t kRowOffDiag = kRowAll & ~(1u << 13);

// Colour c of a sweep that starts from u = 0: the neighbours of colours above c still hold zero, so
// only those of colours below c are gathered.  A neighbour at offset d has colour c ^ flip(d), where
// flip sets bit k when the offset is odd along axis k.
__host__ __device__ constexpr uint32_t zero_start_mask(int c) {
    uint32_t m = 0;
    for (int d = 0; d < 27; d++) {
        const int f = (d % 3 != 1 ? 1 : 0) | (d / 3 % 3 != 1 ? 2 : 0) | (d / 9 != 1 ? 4 : 0);
        if ((c ^ f) < c) m |= 1u << d;
    }
    return m;
}

// first selected neighbour after d (27 when none)
constexpr int next_nb(uint32_t mask, int d) {
    for (int k = d + 1; k < 27; k++)
        if ((mask >> k) & 1u) return k;
    return 27;
}

// zero-start colour of a GS pass (-1: ordinary pass) -> gathered neighbours
constexpr uint32_t gs_mask(int zc) { return zc < 0 ? kRowOffDiag : zero_start_mask(zc); }

// neighbours of a colour-c vertex whose colour comes after c in a forward sweep
__host__ __device__ constexpr uint32_t later_mask(int c) {
    uint32_t m = 0;
    for (int d = 0; d < 27; d++) {
        const int f = (d % 3 != 1 ? 1 : 0) | (d / 3 % 3 != 1 ? 2 : 0) | (d / 9 != 1 ? 4 : 0);
        if ((c ^ f) > c) m |= 1u << d;
    }
    return m;
}

// the stored words of one neighbour: W = R / (load cases per word) words per component; r0 counts
// load cases (a multiple of the cases per word)
// (qo = 4 * vertex index; plane = 4 N bytes)
template <int W, typename A>
__device__ __forceinline__ void load_nb(A (&dst)[W][3], Rsrc ru, uint32_t qo, uint32_t plane, int r0) {
    constexpr int cpw = CasesPerWord<A>::value;
    if constexpr (cpw == 2 && W == 3) {  // all six load cases (r0 = 0): the vertex's nine words
        uint32_t wv[9];
        hload9(wv, ru, qo >> 2, plane);
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int i = 0; i < 3; i++) dst[p][i] = __builtin_bit_cast(A, wv[p * 3 + i]);
    } else if constexpr (cpw == 2) {
#pragma unroll
        for (int p = 0; p < W; p++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const int q = (r0 / 2 + p) * 3 + i;
                dst[p][i] = q < 8 ? bload<A>(ru, 4u * qo + 4u * (uint32_t)(q & 3), (uint32_t)(q >> 2) * 4u * plane)
                                  : bload<A>(ru, qo, 8u * plane);
            }
    } else {
#pragma unroll
        for (int p = 0; p < W; p++)
#pragma unroll
            for (int i = 0; i < 3; i++) dst[p][i] = bload<A>(ru, qo, (uint32_t)((r0 / cpw + p) * 3 + i) * plane);
    }
}

// The two x-neighbours (d = 12, 14) of a vertex taken from LDS instead of memory: `base` holds 3W word
// planes of 256 slots (one per thread of the workgroup), slot[0] / slot[1] the x-1 / x+1 neighbour.
template <typename A>
struct XNbLds {
    const A* base = nullptr;
    int slot[2] = {0, 0};
};

template <int W, typename A>
__device__ __forceinline__ void load_nb_lds(A (&dst)[W][3], const XNbLds<A>& xl, int k) {
#pragma unroll
    for (int p = 0; p < W; p++)
#pragma unroll
        for (int i = 0; i < 3; i++) dst[p][i] = xl.base[(p * 3 + i) * 256 + xl.slot[k]];
}

// the stored words of neighbour d: from memory, or from LDS for the x-neighbours when XL
template <int W, typename A, bool XL, int D>
__device__ __forceinline__ void load_nb_at(A (&dst)[W][3], const FineCtx& c, Rsrc ru, uint32_t plane, int r0,
                                           const XNbLds<A>& xl) {
    if constexpr (XL && (D == 12 || D == 14))
        load_nb_lds<W, A>(dst, xl, D == 12 ? 0 : 1);
    else
        load_nb<W, A>(dst, ru, c.nb(D) * (uint32_t)sizeof(A), plane, r0);
}

// acc[r] += S u_r for the load cases of one stored word triple
template <int W, typename T, typename A>
__device__ __forceinline__ void block_fma(const T (&S)[9], const A (&cur)[W][3], T (&acc)[W * CasesPerWord<A>::value][3]) {
#pragma unroll
    for (int p = 0; p < W; p++) {
        if constexpr (CasesPerWord<A>::value == 2) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const float u0 = (float)cur[p][0][h], u1 = (float)cur[p][1][h], u2 = (float)cur[p][2][h];
                T(&a)[3] = acc[2 * p + h];
                a[0] = fmaf(S[0], u0, fmaf(S[1], u1, fmaf(S[2], u2, a[0])));
                a[1] = fmaf(S[3], u0, fmaf(S[4], u1, fmaf(S[5], u2, a[1])));
                a[2] = fmaf(S[6], u0, fmaf(S[7], u1, fmaf(S[8], u2, a[2])));
            }
        } else {
            const T u0 = (T)cur[p][0], u1 = (T)cur[p][1], u2 = (T)cur[p][2];
            acc[p][0] = fma(S[0], u0, fma(S[1], u1, fma(S[2], u2, acc[p][0])));
            acc[p][1] = fma(S[3], u0, fma(S[4], u1, fma(S[5], u2, acc[p][1])));
            acc[p][2] = fma(S[6], u0, fma(S[7], u1, fma(S[8], u2, acc[p][2])));
        }
    }
}

// Keep a neighbour's products at its step: the products read values loaded one step earlier, and without
// this the instruction selector may sink all of them below the last scheduling barrier, keeping every
// neighbour's operands live (hundreds of spilled registers).  The empty volatile statement emits nothing.
template <int R, typename T>
__device__ __forceinline__ void pin_acc(T (&acc)[R][3]) {
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 3; i++) __asm__ volatile("" : "+v"(acc[r][i]));
}

// One neighbour of the row: issue the loads of the next neighbour, build the 3x3 block of this one
// from the cell weights and KE (in LDS), accumulate, recurse.
template <int R, uint32_t MASK, int D, typename T, typename A, bool XL>
__device__ __forceinline__ void gather_step(const FineCtx& c, int n, const T* ke, Rsrc ru, uint32_t plane, int r0,
                                            const A (&cur)[R / CasesPerWord<A>::value][3], T (&acc)[R][3],
                                            const XNbLds<A>& xl) {
    constexpr int W = R / CasesPerWord<A>::value;
    constexpr int DN = next_nb(MASK, D);
    // keep each neighbour's block build next to its use (bounds register pressure)
    __builtin_amdgcn_sched_barrier(0);
    pin_acc<R, T>(acc);
    A nxt[W][3];
    if constexpr (DN < 27) load_nb_at<W, A, XL, DN>(nxt, c, ru, plane, r0, xl);
    T S[9];
    if constexpr (sizeof(T) == 8)
        fine_block<D>(c.wd, ke, S);
    else
        fine_block<D>(c.wc, ke, S);
    block_fma<W, T, A>(S, cur, acc);
    if constexpr (DN < 27) gather_step<R, MASK, DN, T, A, XL>(c, n, ke, ru, plane, r0, nxt, acc, xl);
}

// The same walk with the loads issued two neighbours ahead (cur: this neighbour, nx1: the next one,
// already in flight): more loads in flight per wave for the latency-bound gather.
template <int R, uint32_t MASK, int D, typename T, typename A>
__device__ __forceinline__ void gather_step2(const FineCtx& c, int n, const T* ke, Rsrc ru, uint32_t plane, int r0,
                                             const A (&cur)[R / CasesPerWord<A>::value][3],
                                             const A (&nx1)[R / CasesPerWord<A>::value][3], T (&acc)[R][3]) {
    constexpr int W = R / CasesPerWord<A>::value;
    constexpr int DN = next_nb(MASK, D);
    constexpr int DN2 = DN < 27 ? next_nb(MASK, DN) : 27;
    __builtin_amdgcn_sched_barrier(0);
    pin_acc<R, T>(acc);
    A nx2[W][3];
    if constexpr (DN2 < 27) load_nb<W, A>(nx2, ru, c.nb(DN2) * (uint32_t)sizeof(A), plane, r0);
    T S[9];
    if constexpr (sizeof(T) == 8)
        fine_block<D>(c.wd, ke, S);
    else
        fine_block<D>(c.wc, ke, S);
    block_fma<W, T, A>(S, cur, acc);
    if constexpr (DN < 27) {
        if constexpr (DN2 < 27)
            gather_step2<R, MASK, DN, T, A>(c, n, ke, ru, plane, r0, nx1, nx2, acc);
        else
            gather_step2<R, MASK, DN, T, A>(c, n, ke, ru, plane, r0, nx1, nx1, acc);
    }
}

// acc[r][i] = sum_d S_d u_{v+d} over the neighbours d selected by MASK, load cases r0..r0+R-1.
// ke points to the element matrix staged in LDS.  The pinned origin's entries of u are zero by
// construction (every entry point zeroes them before the fine operator reads u).  XL: the two
// x-neighbours come from LDS (xl).
template <int R, uint32_t MASK, typename T, typename A, bool XL = false, int PF = 1>
__device__ __forceinline__ void fine_gather(const FineCtx& c, int n, size_t N, const T* ke, const A* __restrict__ u,
                                            T (&acc)[R][3], int r0 = 0, const XNbLds<A>& xl = XNbLds<A>()) {
    constexpr int W = R / CasesPerWord<A>::value;
    const Rsrc ru = make_rsrc(u, (size_t)(18 / CasesPerWord<A>::value) * N * sizeof(A));
    const uint32_t plane = (uint32_t)(N * sizeof(A));
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 3; i++) acc[r][i] = T(0);
    constexpr int D0 = next_nb(MASK, -1);
    if constexpr (D0 < 27) {
        A first[W][3];
        load_nb_at<W, A, XL, D0>(first, c, ru, plane, r0, xl);
        if constexpr (PF == 2 && !XL && next_nb(MASK, D0) < 27) {
            A second[W][3];
            load_nb<W, A>(second, ru, c.nb(next_nb(MASK, D0)) * (uint32_t)sizeof(A), plane, r0);
            gather_step2<R, MASK, D0, T, A>(c, n, ke, ru, plane, r0, first, second, acc);
        } else {
            gather_step<R, MASK, D0, T, A, XL>(c, n, ke, ru, plane, r0, first, acc, xl);
        }
    }
}

// Stage the element matrix in LDS (all threads of the block take part), followed, when CONSTS (ke_s
// then holds kKeStaged entries), by the block constants of fine_block: C_d,ij = KE[7 - e0][b0](i, j)
// for the cell e0 with e0_m = 1 exactly where d_m = +1 and b0_m = d_m + 1 - e0_m.
template <typename T, bool CONSTS = true>
__device__ __forceinline__ void stage_ke(T* ke_s, const T* __restrict__ ke) {
    for (int i = threadIdx.x; i < 576; i += blockDim.x) ke_s[i] = ke[i];
    __syncthreads();
    if constexpr (!CONSTS) return;
    for (int t = threadIdx.x; t < 243; t += blockDim.x) {
        const int d = t / 9, i = t / 3 % 3, j = t % 3;
        int e0 = 0, b0 = 0;
        for (int m = 0; m < 3; m++) {
            const int dm = (m == 0 ? d % 3 : (m == 1 ? d / 3 % 3 : d / 9)) - 1;
            const int em = dm > 0 ? 1 : 0;
            e0 |= em << m;
            b0 |= (dm + 1 - em) << m;
        }
        ke_s[kBlockConsts + t] = ke_s[(3 * (7 - e0) + i) * 24 + 3 * b0 + j];
    }
    __syncthreads();
}

// Block Gauss-Seidel / SOR update of one vertex: x, y, z in turn (z, y, x when REVERSE)
// (gs_relaxation_otf_kernel update formula, grid.cu:530-541).
// inv holds 1 / Ks[0], 1 / Ks[4], 1 / Ks[8] (shared by all load cases of the vertex).
template <bool REVERSE>
__device__ __forceinline__ void block_update(float u[3], const float f[3], const float KeU[3], const float Ks[9],
                                             const float inv[3], float omega) {
    if (!REVERSE) {
        u[0] = omega * (f[0] - KeU[0] - Ks[1] * u[1] - Ks[2] * u[2]) * inv[0] + (1.f - omega) * u[0];
        u[1] = omega * (f[1] - KeU[1] - Ks[3] * u[0] - Ks[5] * u[2]) * inv[1] + (1.f - omega) * u[1];
        u[2] = omega * (f[2] - KeU[2] - Ks[6] * u[0] - Ks[7] * u[1]) * inv[2] + (1.f - omega) * u[2];
    } else {
        u[2] = omega * (f[2] - KeU[2] - Ks[6] * u[0] - Ks[7] * u[1]) * inv[2] + (1.f - omega) * u[2];
        u[1] = omega * (f[1] - KeU[1] - Ks[3] * u[0] - Ks[5] * u[2]) * inv[1] + (1.f - omega) * u[1];
        u[0] = omega * (f[0] - KeU[0] - Ks[1] * u[1] - Ks[2] * u[2]) * inv[0] + (1.f - omega) * u[0];
    }
}

// GS colour c covers the vertices with coordinate parity (1-c&1, 1-c>>1&1, 1-c>>2): the reference
// colours by the parity of the ghost-padded position (grid.cpp:155-198) and sweeps colours 0..7.
// In the parity-blocked order that is block 7 - c.
__device__ __forceinline__ uint32_t colour_block(int colour) { return (uint32_t)(7 - colour); }

// ---------------------------------------------------------------------------------------------
// Finest level (matrix-free)
// ---------------------------------------------------------------------------------------------

// MODE 0: out = K u; MODE 1: out = f - K u.  Pinned origin row is zero.
// (update_residual_otf_kernel_1 grid.cu:1277-1461; v3_stencilOnLeft_kernel grid.cu:3686-3857)
// MODE 0: out = K u, or with beta (per load case) out = K u + beta out_old, the CG recurrence
// K p = K z + beta K p_old that applies K to the (cheaper to gather) preconditioned vector z.
// DOT (MODE 0): also the per-workgroup FP64 partial sums of dotv . out per load case (the p . Kp of a
// CG step) in partial[r * gridDim.x + block].  The launcher uses it only when every lane of the grid
// holds a vertex (64 | (n/2)^3), so no thread leaves before the workgroup reduction.
// A: storage of u (float, or half2_t pairs of load cases).  BETA selects the recurrence form at compile
// time (a run-time test in the epilogue makes the compiler spill the gather pipeline).
// PX (with DOT, u = z): the CG vector updates of the step, fused: p = z (PX 1, first step) or
// x += alpha p_old, p = z + beta p_old (PX 2), per load case with alpha, beta from the device; the dot
// is then p . out with the new p (dotv unused).  Load cases that have converged carry alpha = beta = 0.
template <int R, int MODE, bool DOT, typename A, bool BETA = false, int PX = 0, int PF = 1>
__global__ __launch_bounds__(kFullThreads, 4) void k_fine_apply(Lay L, size_t N, const float* __restrict__ ke,
                                                                const float* __restrict__ w,
                                                                const A* __restrict__ u,
                                                                const float* __restrict__ f, float* __restrict__ out,
                                                                double* __restrict__ partial,
                                                                const float* __restrict__ dotv,
                                                                const double* __restrict__ beta,
                                                                float* __restrict__ pp = nullptr,
                                                                float* __restrict__ xx = nullptr,
                                                                const double* __restrict__ alpha = nullptr) {
    __shared__ float ke_s[kKeStaged];
    stage_ke(ke_s, ke);
    uint32_t v;
    if (!full_vertex(L, xcd_chunk(blockIdx.x, gridDim.x), v)) return;
    int x, y, z;
    L.pos(v, x, y, z);
    FineCtx c;
    fine_ctx(c, x, y, z, L, w);
    float acc[R][3];
    fine_gather<R, kRowAll, float, A, false, PF>(c, L.n, N, ke_s, u, acc);
    double dot[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        float b = 0.f, a = 0.f;
        if constexpr (BETA) b = (float)beta[r];
        if constexpr (PX == 2) a = (float)alpha[r];
        dot[r] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const size_t o = (size_t)(r * 3 + i) * N + v;
            float val = (MODE == 1) ? f[o] - acc[r][i] : acc[r][i];
            if constexpr (BETA) val = fmaf(b, out[o], val);
            val = (v == 0) ? 0.f : val;
            out[o] = val;
            float dv = 0.f;
            if constexpr (PX != 0) {
                float zc;  // z of the vertex itself
                if constexpr (CasesPerWord<A>::value == 2)
                    zc = (float)u[hw_index((r >> 1) * 3 + i, v, N)][r & 1];
                else
                    zc = (float)u[o];
                dv = zc;
                if constexpr (PX == 2) {
                    const float po = pp[o];
                    xx[o] = fmaf(a, po, xx[o]);
                    dv = fmaf(b, po, zc);
                }
                pp[o] = dv;
            } else if constexpr (DOT) {
                dv = dotv[o];
            }
            if constexpr (DOT) dot[r] = fma((double)dv, (double)val, dot[r]);
        }
    }
    if constexpr (DOT) {
        __shared__ double sh[kFullThreads / 64];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double d = block_sum<kFullThreads>(dot[r], sh);
            if (threadIdx.x == 0) partial[r * gridDim.x + blockIdx.x] = d;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// CG pass with each workgroup's node box of z staged in LDS
// ---------------------------------------------------------------------------------------------
// The step of k_fine_apply<6, 0, true, half2_t, BETA, PX> on the same 128 x 2 x 2 vertex tiles, with the
// same arithmetic in the same order (so the same results bit for bit), for a parity-blocked level with
// 64 | n/2.  The row form gathers 243 single words per vertex from memory, and on gfx950 it is the texture
// addresser that those gathers keep busy, not HBM; here the workgroup loads the 130 x 4 x 4 node box of
// the packed-FP16 z (9 word planes) once, row by row with coalesced loads (about 37 per thread), and
// reads the 27 neighbours of each vertex from LDS.
// Box layout: word plane q, then z row lz and y row ly (coordinate - (2 by - 1), bz alike), then x
// parity s, then the half-x position k = (x >> 1) - (bx0 - s).  A row of one parity is contiguous in
// memory and in LDS, so the 64 lanes of a wave (consecutive x >> 1) read 64 consecutive words.
constexpr int kStK = 65;                    // half-x positions per staged row: 64 owned + 1 halo
constexpr int kStRow = 2 * kStK;            // both x parities of one (y, z) row
constexpr int kStRows = 16;                 // 4 z rows x 4 y rows
constexpr int kStPlane = 2112;              // 16 x 130 words padded to a multiple of 64 (256-byte steps)
constexpr int kStPos = kStRows * kStRow;   // staged positions per word plane
constexpr int kStIter = (kStPos + kFullThreads - 1) / kFullThreads;
static_assert(kStPlane >= kStRows * kStRow && kStPlane % 64 == 0, "staged plane stride");

// neighbour d of the vertex: its 9 words from the staged box; xb[dx + 1] holds the thread's word index of
// the x-neighbour dx in row (py, pz) of every plane
template <int D>
__device__ __forceinline__ void staged_nb(half2_t (&dst)[3][3], const uint32_t* zs, const int (&xb)[3]) {
    constexpr int dx = D % 3, dy = D / 3 % 3, dz = D / 9;
    const uint32_t* a = zs + xb[dx] + (dz * 4 + dy) * kStRow;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int i = 0; i < 3; i++) dst[p][i] = __builtin_bit_cast(half2_t, a[(p * 3 + i) * kStPlane]);
}

template <int D>
__device__ __forceinline__ void staged_step(const FineCtx& c, const float* ke, const uint32_t* zs, const int (&xb)[3],
                                            const half2_t (&cur)[3][3], float (&acc)[6][3]) {
    __builtin_amdgcn_sched_barrier(0);
    half2_t nxt[3][3];
    if constexpr (D < 26) staged_nb<D + 1>(nxt, zs, xb);
    float S[9];
    fine_block<D>(c.wc, ke, S);
    block_fma<3, float, half2_t>(S, cur, acc);
    if constexpr (D < 26) staged_step<D + 1>(c, ke, zs, xb, nxt, acc);
}

// Load the workgroup's node box of the packed-FP16 z (tile of chunk t0: half-x bx0.. bx0 + 63, half-y by,
// half-z bz) into zs: position pos = ((lz * 4 + ly) * 2 + s) * 65 + k of the box, one address and its
// nine words per position.  The caller publishes it with a workgroup barrier.
__device__ __forceinline__ void stage_z_box(uint32_t* zs, const uint32_t* __restrict__ z, const Lay& L, size_t N,
                                            int bx0, int by, int bz) {
    const int h = L.h;
    const Rsrc rz = make_rsrc(z, 9 * N * sizeof(uint32_t));
    const uint32_t plane = (uint32_t)(N * sizeof(uint32_t));
    uint32_t tmp[kStIter][9];
#pragma unroll
    for (int it = 0; it < kStIter; it++) {
        const int pos = it * kFullThreads + (int)threadIdx.x;
        if (kStPos % kFullThreads == 0 || pos < kStPos) {
            const int row = pos / kStK, k = pos - row * kStK;
            const int s = row & 1, ly = (row >> 1) & 3, lz = row >> 3;
            int hx = bx0 - s + k;
            hx = hx < 0 ? hx + h : (hx >= h ? hx - h : hx);
            const int yy = L.wrap(2 * by - 1 + ly), zz = L.wrap(2 * bz - 1 + lz);
            const uint32_t blk = (uint32_t)(s | (yy & 1) << 1 | (zz & 1) << 2);
            const uint32_t id = blk * L.h3 + (uint32_t)hx +
                                (uint32_t)h * ((uint32_t)(yy >> 1) + (uint32_t)h * (uint32_t)(zz >> 1));
            hload9(tmp[it], rz, id, plane);
        }
    }
#pragma unroll
    for (int it = 0; it < kStIter; it++) {
        const int pos = it * kFullThreads + (int)threadIdx.x;
        if (kStPos % kFullThreads == 0 || pos < kStPos) {
#pragma unroll
            for (int q = 0; q < 9; q++) zs[q * kStPlane + pos] = tmp[it][q];
        }
    }
}

// word index of the x-neighbours dx = -1, 0, 1 of a thread's vertex in the staged box (row (py, pz) of
// every plane): parity s' = (px + dx) & 1, half-x position k = lane + (px + dx - s') / 2 + s'
__device__ __forceinline__ void staged_xb(int (&xb)[3], int lane, int px, int py, int pz) {
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
        const int sp = (px + dx) & 1;
        xb[dx + 1] = lane + (pz * 4 + py) * kStRow + sp * kStK + (px + dx - sp) / 2 + sp;
    }
}

template <bool BETA, int PX>
__global__ __launch_bounds__(kFullThreads, 4) void k_fine_cg_staged(Lay L, size_t N, const float* __restrict__ ke,
                                                                    const float* __restrict__ w,
                                                                    const uint32_t* __restrict__ z,
                                                                    float* __restrict__ out,
                                                                    double* __restrict__ partial,
                                                                    const double* __restrict__ beta,
                                                                    float* __restrict__ pp, float* __restrict__ xx,
                                                                    const double* __restrict__ alpha) {
    static_assert(BETA == (PX == 2), "the first pass (PX 1) has no beta; later passes (PX 2) have both");
    // one LDS array with KE first: its wave-uniform reads then take immediate offsets below 64 KB
    __shared__ uint32_t lds[kKeStaged + 9 * kStPlane];
    float* ke_s = reinterpret_cast<float*>(lds);
    uint32_t* zs = lds + kKeStaged;
    const int h = L.h;
    const uint32_t t0 = xcd_chunk(blockIdx.x, gridDim.x) * 64u;
    const int bx0 = (int)(t0 % (uint32_t)h), by = (int)((t0 / (uint32_t)h) % (uint32_t)h);
    const int bz = (int)(t0 / ((uint32_t)h * h));
    const int b = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t v = (uint32_t)b * L.h3 + t0 + (uint32_t)lane;
    // PX 2: the vertex's p_old and x are loaded with the box and q_old right after it, so their latency
    // overlaps the staging and the neighbour loop; x += alpha p_old is stored before the loop, q_old and
    // p_old stay in registers
    // (buffer accesses: one 32-bit offset v * 4 per thread, the (load case, component) plane in an SGPR)
    const uint32_t vo = v * 4u, fplane = (uint32_t)(N * sizeof(float));
    const size_t fbytes = 18 * N * sizeof(float);
    const Rsrc rq = make_rsrc(out, fbytes), rp = make_rsrc(pp, fbytes), rx = make_rsrc(xx, fbytes);
    float qo[6][3], po[6][3];
    {
        float xo[6][3];
        if constexpr (PX == 2) {
#pragma unroll
            for (int r = 0; r < 6; r++)
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    po[r][i] = bload<float>(rp, vo, (uint32_t)(r * 3 + i) * fplane);
                    xo[r][i] = bload<float>(rx, vo, (uint32_t)(r * 3 + i) * fplane);
                }
        }
        stage_z_box(zs, z, L, N, bx0, by, bz);
        if constexpr (PX == 2) {
#pragma unroll
            for (int r = 0; r < 6; r++) {
                const float a = (float)alpha[r];
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const uint32_t so = (uint32_t)(r * 3 + i) * fplane;
                    bstore(fmaf(a, po[r][i], xo[r][i]), rx, vo, so);
                    qo[r][i] = bload<float>(rq, vo, so);
                }
            }
        }
    }
    stage_ke(ke_s, ke);  // ends with the workgroup barrier that also publishes the staged box
    const int px = b & 1, py = (b >> 1) & 1, pz = b >> 2;
    const int x = 2 * (bx0 + lane) + px, y = 2 * by + py, zc = 2 * bz + pz;
    FineCtx c;
    fine_ctx(c, x, y, zc, L, w);
    int xb[3];
    staged_xb(xb, lane, px, py, pz);
    float acc[6][3];
#pragma unroll
    for (int r = 0; r < 6; r++)
#pragma unroll
        for (int i = 0; i < 3; i++) acc[r][i] = 0.f;
    {
        half2_t first[3][3];
        staged_nb<0>(first, zs, xb);
        staged_step<0>(c, ke_s, zs, xb, first, acc);
    }
    half2_t zv[3][3];  // z of the vertex itself (d = 13)
    staged_nb<13>(zv, zs, xb);
    double dot[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
        float bt = 0.f;
        if constexpr (BETA) bt = (float)beta[r];
        dot[r] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const uint32_t so = (uint32_t)(r * 3 + i) * fplane;
            float val = acc[r][i];
            if constexpr (BETA) val = fmaf(bt, qo[r][i], val);
            val = (v == 0) ? 0.f : val;
            bstore(val, rq, vo, so);
            const float zr = (float)zv[r >> 1][i][r & 1];
            float dv = zr;
            if constexpr (PX == 2) dv = fmaf(bt, po[r][i], zr);
            bstore(dv, rp, vo, so);
            dot[r] = fma((double)dv, (double)val, dot[r]);
        }
    }
    __shared__ double sh[kFullThreads / 64];
#pragma unroll
    for (int r = 0; r < 6; r++) {
        const double d = block_sum<kFullThreads>(dot[r], sh);
        if (threadIdx.x == 0) partial[r * gridDim.x + blockIdx.x] = d;
    }
}

// ---------------------------------------------------------------------------------------------
// Finest-level operator in element form on the matrix cores (packed-FP16 u, n % 32 == 0)
// ---------------------------------------------------------------------------------------------
// K u = sum over cells e of w_e KE u_e (the per-cell form of grid.cu:1277-1461).  A workgroup owns a
// tile of 32 x 4 x 4 vertices.  For one load-case pair at a time it stages the 34 x 6 x 6 node box
// of u in LDS, forms KE u_e for the 33 x 5 x 5 cells that touch the tile with
// v_mfma_f32_16x16x32_f16 and adds w_e KE u_e into the tile's vertices in LDS.  KE enters through its
// Lame tables, KE = lambda/72 Kl + mu/72 Km with small-integer Kl, Km (templateMatrix.cpp:8-61),
// which FP16 holds exactly: every product is exact, the sums are FP32, and a constant (rigid
// translation) field gives exactly zero, as it must; CG on this operator relies on that.
//
// Lanes: column c = lane % 16 of a wave is one cell; lane group g = lane / 16 carries the cell's two
// nodes (ox, g & 1, g >> 1), ox = 0, 1.  B holds their 6 words (3 components x 2 nodes) as K entries
// 8g .. 8g + 5, and the three 16-row output tiles return to the lane the same 6 rows of Kl u_e and of
// Km u_e (table layout: ke_mfma_table).  The 16 cells of a group are consecutive cells of one grid
// row, so the face shared by cells c and c + 1 is summed across lanes (DPP row shift).
//
// Sums without atomics: a vertex row (y, z) receives from the four cell rows (ey, ez) in
// {y-1, y} x {z-1, z}, one of each (ey, ez) parity class.  Each class writes its own LDS plane, and
// every vertex of a plane is written exactly once per load case: plain stores, no clearing, no
// ordering between waves.  The few contributions that meet on a plane at the x-seams of the row
// groups (x = 0, 16: first lane of a group; x = 0 also from the cells at x = -1, which form
// separate groups) go to small per-class arrays of their own.  The epilogue adds it all up.
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kEfTX = 16, kEfTY = 2, kEfTZ = 2;                                     // parity positions
constexpr int kEfVX = 2 * kEfTX, kEfVY = 2 * kEfTY, kEfVZ = 2 * kEfTZ;              // owned vertices
constexpr int kEfVerts = kEfVX * kEfVY * kEfVZ;
constexpr int kEfBX = kEfVX + 2, kEfBY = kEfVY + 2, kEfBZ = kEfVZ + 2;              // node box
constexpr int kEfNodes = kEfBX * kEfBY * kEfBZ;
constexpr int kEfCX = kEfVX + 1, kEfCY = kEfVY + 1, kEfCZ = kEfVZ + 1;              // cells, from -1
constexpr int kEfCells = kEfCX * kEfCY * kEfCZ;
constexpr int kEfRows = kEfCY * kEfCZ;
constexpr int kEfRowGroups = kEfVX / 16;                 // groups of cells x = 0 .. VX-1 per row
constexpr int kEfFull = kEfRows * kEfRowGroups;
constexpr int kEfGroups = kEfFull + (kEfRows + 15) / 16;  // + groups of the x = -1 cells
constexpr int kEfYZ = kEfVY * kEfVZ;                      // vertex rows of the tile
// one LDS row per (cell-row class, case of the pair, component): the tile's vertices, the x-seam and
// x = 0 contributions, and a slot per lane that absorbs the stores of lanes with nothing for the tile
// (stores stay unconditional: no exec-mask branches in the cell loop)
constexpr int kEfSeam = kEfVerts, kEfHalo = kEfSeam + 2 * kEfYZ, kEfDummy = kEfHalo + kEfYZ;
constexpr int kEfSlots = kEfDummy + 64;
constexpr int kEfThreads = 512;
constexpr int kEfStage = (3 * kEfNodes + kEfThreads - 1) / kEfThreads;  // staged words per thread and pair
static_assert(kEfVerts == kEfThreads, "epilogue maps one vertex per thread and plane");
static_assert(kEfRowGroups == 2, "x-seams at x = 0 and x = 16");

// row shift of a 16-lane row: lane c reads lane c + 1, lane 15 reads 0
__device__ __forceinline__ float dpp_next(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x101, 0xf, 0xf, true));
}

// One group of 16 cells and the two load cases of a packed word pair: the B operands (the lane's 6 node
// words, split into the two load cases) from the node box ...
__device__ __forceinline__ void ef_load_b(const uint32_t* nodes, int nb0, half8_t (&b)[2]) {
    half2_t wv[6];
#pragma unroll
    for (int k = 0; k < 6; k++) wv[k] = __builtin_bit_cast(half2_t, nodes[(k % 3) * kEfNodes + nb0 + k / 3]);
#pragma unroll
    for (int k = 0; k < 6; k++) {
        b[0][k] = wv[k].x;
        b[1][k] = wv[k].y;
    }
    b[0][6] = b[0][7] = b[1][6] = b[1][7] = (_Float16)0.f;
}

// ... and NT + 1 output tiles per case (NT = 1: one exact table T; NT = 2: Kl and Km), giving the lane's
// 6 rows of w_e KE u_e per case in val[case][k] (k = 3 ox + component); w0, w1 = w_e c0, w_e c1.
template <int NT>
__device__ __forceinline__ void ef_mma(const half8_t (&b)[2], half8_t a0, half8_t a1, half8_t a2, float w0, float w1,
                                       float (&val)[2][6]) {
    const float4_t zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int hh = 0; hh < 2; hh++) {
        const float4_t d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b[hh], zero, 0, 0, 0);
        const float4_t d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, b[hh], zero, 0, 0, 0);
        if constexpr (NT == 1) {
            // rows k = 0..3 in tile 0, k = 4, 5 in tile 1
            val[hh][0] = w0 * d0[0];
            val[hh][1] = w0 * d0[1];
            val[hh][2] = w0 * d0[2];
            val[hh][3] = w0 * d0[3];
            val[hh][4] = w0 * d1[0];
            val[hh][5] = w0 * d1[1];
        } else {
            // Kl rows in tiles 0/1, Km rows in tiles 1/2
            const float4_t d2 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2, b[hh], zero, 0, 0, 0);
            val[hh][0] = fmaf(w0, d0[0], w1 * d1[2]);
            val[hh][1] = fmaf(w0, d0[1], w1 * d1[3]);
            val[hh][2] = fmaf(w0, d0[2], w1 * d2[0]);
            val[hh][3] = fmaf(w0, d0[3], w1 * d2[1]);
            val[hh][4] = fmaf(w0, d1[0], w1 * d2[2]);
            val[hh][5] = fmaf(w0, d1[1], w1 * d2[3]);
        }
    }
}

// A group of cells x = 16 half .. 16 half + 15 of cell row gi / 2 and where its lanes' results go
struct EfRowGroup {
    int nb0;   // node box index of the lane's node (0, gy, gz)
    int wi;    // cell weight index
    int v1;    // tile vertex of the lane's node (1, gy, gz): x = 16 half + col + 1
    int yz;    // vertex row of the lane's nodes
    int cls;   // parity class of the cell row
    int half;
    bool own;  // the lane's vertex row lies in the tile
};

__device__ __forceinline__ EfRowGroup ef_row_group(int gi, int col, int gy, int gz, int lane_nb, int lane_v) {
    EfRowGroup G;
    const int row = gi >> 1;
    G.half = gi & 1;
    const int ey = row % kEfCY - 1, ez = row / kEfCY - 1;
    G.cls = ((ey + 1) & 1) | (((ez + 1) & 1) << 1);
    G.wi = ((ez + 1) * kEfCY + (ey + 1)) * kEfCX + 16 * G.half + 1 + col;
    G.nb0 = ((ez + 1) * kEfBY + (ey + 1)) * kEfBX + 16 * G.half + 1 + lane_nb;
    const int vy = ey + gy, vz = ez + gz;
    G.own = vy >= 0 && vy < kEfVY && vz >= 0 && vz < kEfVZ;
    G.v1 = (ez * kEfVY + ey) * kEfVX + 16 * G.half + 1 + lane_v;
    G.yz = vz * kEfVY + vy;
    return G;
}

// MODE 0: out = K u (BETA: + beta out_old); MODE 1: out = f - K u.  DOT (MODE 0): per-workgroup FP64
// partial sums of dotv . out per load case in partial[r * gridDim.x + block].  Pinned origin row zero.
// PX: the CG vector updates fused as in k_fine_apply (z of the vertex read from the staged node box).
// KE = c0 T (NT = 1) or c0 Kl + c1 Km (NT = 2).  Needs 72 n^3 < 2^32 (32-bit buffer offsets).
template <int MODE, bool DOT, bool BETA, int NT, int PX = 0>
__global__ __launch_bounds__(kEfThreads, 2) void k_fine_apply_ef(Lay L, size_t N, const half8_t* __restrict__ amat,
                                                              float c0, float c1, const float* __restrict__ w,
                                                              const uint32_t* __restrict__ u,
                                                              const float* __restrict__ f, float* __restrict__ out,
                                                              double* __restrict__ partial,
                                                              const float* __restrict__ dotv,
                                                              const double* __restrict__ beta,
                                                              float* __restrict__ pp, float* __restrict__ xx,
                                                              const double* __restrict__ alpha) {
    __shared__ uint32_t nodes[3 * kEfNodes];       // packed words of the node box, one load-case pair
    __shared__ float acc[4][6][kEfSlots];          // per cell-row class, (case of the pair, component)
    __shared__ float wcell[kEfCells];              // w_e of the cells from -1
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, col = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: group loops stay scalar
    const int gy = g & 1, gz = g >> 1;
    const uint32_t wg = xcd_chunk(blockIdx.x, gridDim.x);
    const uint32_t ntx = (uint32_t)(L.h / kEfTX), nty = (uint32_t)(L.h / kEfTY);
    const int x0 = kEfVX * (int)(wg % ntx), y0 = kEfVY * (int)((wg / ntx) % nty), z0 = kEfVZ * (int)(wg / (ntx * nty));
    const half8_t a0 = amat[lane], a1 = amat[64 + lane], a2 = NT == 2 ? amat[128 + lane] : a1;
    const Rsrc ru = make_rsrc(u, 9 * N * sizeof(uint32_t));
    const Rsrc ro = make_rsrc(out, 18 * N * sizeof(float));
    const Rsrc rf = make_rsrc(f, f ? 18 * N * sizeof(float) : 0);
    const Rsrc rd = make_rsrc(PX ? pp : dotv, (PX ? pp != nullptr : dotv != nullptr) ? 18 * N * sizeof(float) : 0);
    const Rsrc rx = make_rsrc(xx, xx ? 18 * N * sizeof(float) : 0);
    for (int c = tid; c < kEfCells; c += kEfThreads) {
        const int ex = c % kEfCX - 1, ey = (c / kEfCX) % kEfCY - 1, ez = c / (kEfCX * kEfCY) - 1;
        wcell[c] = w[L.idx(L.wrap(x0 + ex), L.wrap(y0 + ey), L.wrap(z0 + ez))];
    }
    // node-box words of this thread: component i and vertex index of each (~0u: none).  The next pair's
    // words (q = 3 p + i, hw_index) are fetched into registers before the current pair is computed.
    uint32_t nidx[kEfStage], pre[kEfStage];
    int ncomp[kEfStage];
#pragma unroll
    for (int k = 0; k < kEfStage; k++) {
        const int s = tid + k * kEfThreads, i = s / kEfNodes, b = s - i * kEfNodes;
        const int bx = b % kEfBX, by = (b / kEfBX) % kEfBY, bz = b / (kEfBX * kEfBY);
        ncomp[k] = i;
        nidx[k] = s < 3 * kEfNodes ? L.idx(L.wrap(x0 - 1 + bx), L.wrap(y0 - 1 + by), L.wrap(z0 - 1 + bz)) : ~0u;
    }
    auto fetch = [&](int p) {
#pragma unroll
        for (int k = 0; k < kEfStage; k++)
            if (nidx[k] != ~0u)
                pre[k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(
                    ru, (int)(4u * (uint32_t)hw_index(3 * p + ncomp[k], nidx[k], N)), 0, 0);
    };
    auto put = [&]() {
#pragma unroll
        for (int k = 0; k < kEfStage; k++)
            if (nidx[k] != ~0u) nodes[tid + k * kEfThreads] = pre[k];
    };
    fetch(0);
    put();
    __syncthreads();
    // epilogue vertex of this thread
    const int vx = tid % kEfVX, vyz = tid / kEfVX;
    const uint32_t vg = L.idx(x0 + vx, y0 + vyz % kEfVY, z0 + vyz / kEfVY), vb = 4u * vg;
    const int nb_self = ((vyz / kEfVY + 1) * kEfBY + (vyz % kEfVY + 1)) * kEfBX + vx + 1;  // the vertex in the node box
    const uint32_t plane_bytes = (uint32_t)(N * sizeof(float));
    // lane offsets of the full-row groups: node (0, gy, gz) of cell (col, 0, 0) in the node box; vertex
    // (col + 1, gy, gz) in the tile
    const int lane_nb = (gz * kEfBY + gy) * kEfBX + col, lane_v = (gz * kEfVY + gy) * kEfVX + col;
    double dot[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int p = 0; p < 3; p++) {
        if (p < 2) fetch(p + 1);
        // epilogue inputs of this pair, in flight during the cell products
        float fin[6], oin[6], din[6], xin[6];
#pragma unroll
        for (int pl = 0; pl < 6; pl++) {
            const uint32_t so = (uint32_t)(6 * p + pl) * plane_bytes;
            if constexpr (MODE == 1) fin[pl] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rf, (int)vb, (int)so, 0));
            if constexpr (BETA) oin[pl] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(ro, (int)vb, (int)so, 0));
            if constexpr ((DOT && PX == 0) || PX == 2)  // dotv, or p_old
                din[pl] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rd, (int)vb, (int)so, 0));
            if constexpr (PX == 2) xin[pl] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, (int)vb, (int)so, 0));
        }
        // cells x = 0 .. VX-1 of every cell row, 16 per group
        for (int gi = wave; gi < kEfFull; gi += kEfThreads / 64) {
            const EfRowGroup g = ef_row_group(gi, col, gy, gz, lane_nb, lane_v);
            half8_t b[2];
            ef_load_b(nodes, g.nb0, b);
            const float wc = wcell[g.wi];
            float val[2][6];
            ef_mma<NT>(b, a0, a1, a2, wc * c0, wc * c1, val);
            // vertex x = 16 half + col + 1 (all lanes but the last of the second half), and the seam x = 16 half
            // of the first lane
            float* row = &acc[g.cls][0][0];
            const int s1 = g.own && (g.half == 0 || col < 15) ? g.v1 : kEfDummy + lane;
            const int sm = g.own && col == 0 ? kEfSeam + g.half * kEfYZ + g.yz : kEfDummy + lane;
#pragma unroll
            for (int hh = 0; hh < 2; hh++)
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    // node ox = 1 of cell c is node ox = 0 of cell c + 1
                    row[(hh * 3 + i) * kEfSlots + s1] = val[hh][3 + i] + dpp_next(val[hh][i]);
                    row[(hh * 3 + i) * kEfSlots + sm] = val[hh][i];
                }
        }
        // cells x = -1 of the cell rows, 16 rows per group
        for (int gi = kEfFull + wave; gi < kEfGroups; gi += kEfThreads / 64) {
            const int k = 16 * (gi - kEfFull) + col;
            const bool valid = k < kEfRows;
            const int kk = valid ? k : 0, ey = kk % kEfCY - 1, ez = kk / kEfCY - 1;
            const int cls = ((ey + 1) & 1) | (((ez + 1) & 1) << 1);
            const float wc = wcell[((ez + 1) * kEfCY + (ey + 1)) * kEfCX];
            half8_t b[2];
            ef_load_b(nodes, ((ez + 1 + gz) * kEfBY + (ey + 1 + gy)) * kEfBX, b);
            float val[2][6];
            ef_mma<NT>(b, a0, a1, a2, wc * c0, wc * c1, val);
            const int vy = ey + gy, vz = ez + gz;
            const int hs = valid && vy >= 0 && vy < kEfVY && vz >= 0 && vz < kEfVZ ? kEfHalo + vz * kEfVY + vy
                                                                                   : kEfDummy + lane;
            float* row = &acc[cls][0][0];
#pragma unroll
            for (int hh = 0; hh < 2; hh++)
#pragma unroll
                for (int i = 0; i < 3; i++) row[(hh * 3 + i) * kEfSlots + hs] = val[hh][3 + i];  // node x = 0
        }
        // z of this thread's vertex for the fused CG update, read before the barrier: after it, the node
        // box is refilled with the next pair while other threads may still be in the epilogue
        uint32_t zself[3];
        if constexpr (PX != 0) {
#pragma unroll
            for (int i = 0; i < 3; i++) zself[i] = nodes[i * kEfNodes + nb_self];
        }
        __syncthreads();
#pragma unroll
        for (int pl = 0; pl < 6; pl++) {
            const int r = 2 * p + pl / 3;
            float val = 0.f;
            if (vx != 0) val = (acc[0][pl][tid] + acc[1][pl][tid]) + (acc[2][pl][tid] + acc[3][pl][tid]);
            if (vx == 0 || vx == 16) {
                const int sl = kEfSeam + (vx >> 4) * kEfYZ + vyz;
                val += (acc[0][pl][sl] + acc[1][pl][sl]) + (acc[2][pl][sl] + acc[3][pl][sl]);
            }
            if (vx == 0) {
                const int sl = kEfHalo + vyz;
                val += (acc[0][pl][sl] + acc[1][pl][sl]) + (acc[2][pl][sl] + acc[3][pl][sl]);
            }
            const uint32_t so = (uint32_t)(6 * p + pl) * plane_bytes;
            if constexpr (MODE == 1) val = fin[pl] - val;
            if constexpr (BETA) val = fmaf((float)beta[r], oin[pl], val);
            if (vg == 0) val = 0.f;
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, val), ro, (int)vb, (int)so, 0);
            float dv = 0.f;
            if constexpr (PX != 0) {
                const half2_t zw = __builtin_bit_cast(half2_t, zself[pl % 3]);
                const float zc = (float)zw[pl / 3];
                dv = zc;
                if constexpr (PX == 2) {
                    __builtin_amdgcn_raw_buffer_store_b32(
                        __builtin_bit_cast(uint32_t, fmaf((float)alpha[r], din[pl], xin[pl])), rx, (int)vb, (int)so, 0);
                    dv = fmaf((float)beta[r], din[pl], zc);
                }
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, dv), rd, (int)vb, (int)so, 0);
            } else if constexpr (DOT) {
                dv = din[pl];
            }
            if constexpr (DOT) dot[r] = fma((double)dv, (double)val, dot[r]);
        }
        if (p < 2) put();
        __syncthreads();
    }
    if constexpr (DOT) {
        __shared__ double sh[kEfThreads / 64];
#pragma unroll
        for (int r = 0; r < 6; r++) {
            const double t = block_sum<kEfThreads>(dot[r], sh);
            if (threadIdx.x == 0) partial[r * gridDim.x + blockIdx.x] = t;
        }
    }
}

// GS / SOR update of the colour vertex v (load cases r0 .. r0 + R - 1), in place.  A: storage of u
// (float, or half2_t pairs of load cases); f is FP32.  ZC >= 0: colour ZC of a forward sweep that
// starts from u = 0 (colours above ZC are never read, so u needs no clearing beforehand; the old value
// of the vertex itself is zero).  XL: the x-neighbours come from LDS; TO_LDS: the new words are also
// stored to slot threadIdx.x of out_lds.  DOT: dot[r] += f . u_new over the vertex (stored values;
// FP32 over the vertex's 3 components, FP64 across vertices).  ROWN (ZC >= 0): also
// rown = f - sum over the neighbours of colours up to ZC and the vertex itself of K_vj u_j, with the
// stored values; once the sweep is complete, subtracting the later colours' terms gives f - K u.
// The pinned origin keeps u = 0 (and rown = 0).
template <int R, bool REVERSE, int ZC, typename A, bool XL, bool TO_LDS, bool DOT = false, bool ROWN = false>
__device__ __forceinline__ void gs_vertex(const Lay& L, size_t N, uint32_t v, float omega, const float* ke_s,
                                          const float* __restrict__ w, A* __restrict__ u,
                                          const float* __restrict__ f, int r0, const XNbLds<A>& xl, A* out_lds,
                                          float* dot = nullptr, float* __restrict__ rown = nullptr) {
    static_assert(!ROWN || ZC >= 0, "the own-colour residual needs a zero-start forward pass");
    constexpr int cpw = CasesPerWord<A>::value;
    int x, y, z;
    L.pos(v, x, y, z);
    FineCtx c;
    fine_ctx(c, x, y, z, L, w);
    float acc[R][3];
    fine_gather<R, gs_mask(ZC), float, A, XL>(c, L.n, N, ke_s, u, acc, r0, xl);
    float Ks[9];
    fine_block<13, float>(c.wc, ke_s, Ks);
    const float inv[3] = {1.f / Ks[0], 1.f / Ks[4], 1.f / Ks[8]};
    const bool pin = v == 0;
#pragma unroll
    for (int p = 0; p < R / cpw; p++) {
        float uu[cpw][3], ff[cpw][3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const size_t o = fw_index<A>((r0 / cpw + p) * 3 + i, v, N);
            if constexpr (cpw == 2) {
                const half2_t wv = ZC >= 0 ? half2_t{0, 0} : u[o];
                uu[0][i] = (float)wv.x;
                uu[1][i] = (float)wv.y;
            } else {
                uu[0][i] = ZC >= 0 ? 0.f : u[o];
            }
#pragma unroll
            for (int h = 0; h < cpw; h++) ff[h][i] = f[(size_t)((r0 + p * cpw + h) * 3 + i) * N + v];
        }
#pragma unroll
        for (int h = 0; h < cpw; h++) block_update<REVERSE>(uu[h], ff[h], acc[p * cpw + h], Ks, inv, omega);
        float us[cpw][3];  // the stored values
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const size_t o = fw_index<A>((r0 / cpw + p) * 3 + i, v, N);
            A word;
            if constexpr (cpw == 2)
                word = half2_t{(_Float16)uu[0][i], (_Float16)uu[1][i]};
            else
                word = uu[0][i];
            if (pin) word = A(0);
            if constexpr (cpw == 2) {
                us[0][i] = (float)word.x;
                us[1][i] = (float)word.y;
            } else {
                us[0][i] = word;
            }
            u[o] = word;
            if constexpr (TO_LDS) out_lds[(p * 3 + i) * 256 + threadIdx.x] = word;
            if constexpr (DOT) {
                if constexpr (cpw == 2) {
                    dot[p * 2] = fmaf(ff[0][i], (float)word.x, dot[p * 2]);
                    dot[p * 2 + 1] = fmaf(ff[1][i], (float)word.y, dot[p * 2 + 1]);
                } else {
                    dot[p] = fmaf(ff[0][i], word, dot[p]);
                }
            }
        }
        if constexpr (ROWN) {
#pragma unroll
            for (int hh = 0; hh < cpw; hh++)
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const float kv = fmaf(Ks[3 * i], us[hh][0], fmaf(Ks[3 * i + 1], us[hh][1], Ks[3 * i + 2] * us[hh][2]));
                    const float ro = (ff[hh][i] - acc[p * cpw + hh][i]) - kv;
                    rown[(size_t)((r0 + p * cpw + hh) * 3 + i) * N + v] = pin ? 0.f : ro;
                }
        }
    }
}

// One colour of the fine block GS / SOR sweep, in place (gs_relaxation_otf_kernel, grid.cu:323-551).
template <int R, bool REVERSE, int ZC, typename A>
__global__ __launch_bounds__(256, 4) void k_fine_gs(Lay L, size_t N, int colour, float omega,
                                                    const float* __restrict__ ke, const float* __restrict__ w,
                                                    A* __restrict__ u, const float* __restrict__ f) {
    __shared__ float ke_s[kKeStaged];
    stage_ke(ke_s, ke);
    const uint32_t t = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= L.h3) return;
    gs_vertex<R, REVERSE, ZC, A, false, false>(L, N, colour_block(colour) * L.h3 + t, omega, ke_s, w, u, f,
                                              R == 6 ? 0 : (int)blockIdx.y * R, XNbLds<A>(), nullptr);
}

// Two consecutive colours of a sweep in one launch: cfirst, then cfirst + 1 (forward) or cfirst - 1
// (REVERSE), cfirst even resp. odd.  DOT: per-workgroup FP64 sums of f . u_new over the pair's vertices
// per load case to partial[r * pstride + blockIdx.x] (over the four pair launches of a sweep that is
// the last to write u, this is the r . z of the CG step).  The two colours differ only in x-parity, and a vertex of the
// second depends on the first only through its x-neighbours, which lie in the same grid row.  Each
// workgroup owns whole rows (256 % (n/2) == 0, (n/2)^3 % 256 == 0, checked by the launcher): it
// updates the first colour of its rows, keeps the new values in LDS, and updates the second colour
// reading those x-neighbours from LDS.  Other workgroups never read the rows' first-colour values
// before they are final, so the result equals the two colour passes in sequence.
// ROWN (zero-start passes): both colours also store their own-colour residuals to rown (gs_vertex).
template <int R, bool REVERSE, int ZC, typename A, bool DOT = false, bool ROWN = false>
__global__ __launch_bounds__(256, 4) void k_fine_gs2(Lay L, size_t N, int cfirst, float omega,
                                                     const float* __restrict__ ke, const float* __restrict__ w,
                                                     A* __restrict__ u, const float* __restrict__ f,
                                                     double* __restrict__ partial, int pstride,
                                                     float* __restrict__ rown) {
    constexpr int W = R / CasesPerWord<A>::value;
    __shared__ float ke_s[kKeStaged];
    __shared__ A xs[3 * W * 256];
    stage_ke(ke_s, ke);
    const uint32_t t = xcd_chunk(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
    const int c2 = REVERSE ? cfirst - 1 : cfirst + 1;
    float dot[R];
#pragma unroll
    for (int r = 0; r < R; r++) dot[r] = 0.f;
    gs_vertex<R, REVERSE, ZC, A, false, true, DOT, ROWN>(L, N, colour_block(cfirst) * L.h3 + t, omega, ke_s, w, u,
                                                        f, 0, XNbLds<A>(), xs, dot, rown);
    __syncthreads();
    XNbLds<A> xl;
    xl.base = xs;
    const int h = L.h, lx = (int)(t % (uint32_t)h), row0 = (int)threadIdx.x - lx;
    const bool odd = (colour_block(c2) & 1u) != 0;  // x-parity of the second colour
    xl.slot[0] = row0 + (odd ? lx : (lx == 0 ? h - 1 : lx - 1));
    xl.slot[1] = row0 + (odd ? (lx + 1 == h ? 0 : lx + 1) : lx);
    gs_vertex<R, REVERSE, (ZC < 0 ? -1 : ZC + 1), A, true, false, DOT, ROWN>(
        L, N, colour_block(c2) * L.h3 + t, omega, ke_s, w, u, f, 0, xl, nullptr, dot, rown);
    if constexpr (DOT) {
        __shared__ double sh[4];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double sdot = block_sum<256>((double)dot[r], sh);
            if (threadIdx.x == 0) partial[(size_t)r * pstride + blockIdx.x] = sdot;
        }
    }
}

// The residual f - K u after a zero-start forward sweep whose colour passes stored their own-colour
// residuals in r (gs_vertex ROWN): r -= sum over the neighbours of colours after C of K_vj u_j, for
// a vertex v of colour C (C < 7; colour 7 has no later neighbours, so its r is final).  About half
// the neighbour terms of a full residual pass, averaged over the colours.
// zero is 0 at run time: chaining the three components through it keeps the compiler converting the
// gathered FP16 words inside the FMAs (v_fma_mix); with independent components it converts every word
// up front and spills.
template <int R, int C, typename A>
__device__ __forceinline__ void later_vertex(const Lay& L, size_t N, uint32_t v, const float* ke_s,
                                             const float* __restrict__ w, const A* __restrict__ u,
                                             float* __restrict__ r, float zero) {
    int x, y, z;
    L.pos(v, x, y, z);
    FineCtx c;
    fine_ctx(c, x, y, z, L, w);
    float acc[R][3];
    fine_gather<R, later_mask(C), float, A>(c, L.n, N, ke_s, u, acc);
#pragma unroll
    for (int q = 0; q < R; q++) {
        float ro[3], rn[3];
#pragma unroll
        for (int i = 0; i < 3; i++) ro[i] = r[(size_t)(q * 3 + i) * N + v];
        rn[0] = ro[0] - acc[q][0];
        rn[1] = fmaf(-zero, rn[0], ro[1] - acc[q][1]);
        rn[2] = fmaf(-zero, rn[1], ro[2] - acc[q][2]);
#pragma unroll
        for (int i = 0; i < 3; i++) r[(size_t)(q * 3 + i) * N + v] = (v == 0) ? 0.f : rn[i];
    }
}

// Colours 0..6 in one launch, on the 128 x 2 x 2 vertex tiles of full_vertex (64 positions t of the
// eight parity blocks; colour c is block 7 - c) with four waves per tile: wave w takes colours w and
// 7 - w at the tile's positions in turn (a wave-uniform branch).  Colour c has 26, 24, 20, 18, 8, 6, 2,
// 0 later neighbours for c = 0..7, so each wave gathers 26 (one wave per colour left the colour-0
// wave to finish alone).  The tile's blocks share their neighbour rows in L1/L2, so the gathered u is
// fetched from HBM about once.  Parity-blocked levels only (make_lay(n, 0)).
template <int R, typename A>
__global__ __launch_bounds__(256) void k_fine_resid_later(Lay L, size_t N, const float* __restrict__ ke,
                                                          const float* __restrict__ w, const A* __restrict__ u,
                                                          float* __restrict__ r, float zero) {
    __shared__ float ke_s[kKeStaged];
    stage_ke(ke_s, ke);
    const uint32_t t = xcd_chunk(blockIdx.x, gridDim.x) * 64u + (threadIdx.x & 63u);
    if (t >= L.h3) return;
    const uint32_t h3 = L.h3;
    switch (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) {
        case 0: later_vertex<R, 0, A>(L, N, 7 * h3 + t, ke_s, w, u, r, zero); break;  // colour 7: none
        case 1:
            later_vertex<R, 1, A>(L, N, 6 * h3 + t, ke_s, w, u, r, zero);
            later_vertex<R, 6, A>(L, N, 1 * h3 + t, ke_s, w, u, r, zero);
            break;
        case 2:
            later_vertex<R, 2, A>(L, N, 5 * h3 + t, ke_s, w, u, r, zero);
            later_vertex<R, 5, A>(L, N, 2 * h3 + t, ke_s, w, u, r, zero);
            break;
        default:
            later_vertex<R, 3, A>(L, N, 4 * h3 + t, ke_s, w, u, r, zero);
            later_vertex<R, 4, A>(L, N, 3 * h3 + t, ke_s, w, u, r, zero);
            break;
    }
}

// Macro-strain right-hand sides f_s = sum_e w_e KE chi0_s for the 6 strains (grid.cu:1818-1919).
// kchi: 6 x 24 table of KE * chi0_s.  Pinned row zero.
__global__ __launch_bounds__(kFullThreads) void k_macro_rhs(Lay L, size_t N, const float* __restrict__ kchi,
                                                            const float* __restrict__ w, float* __restrict__ f) {
    uint32_t v;
    if (!full_vertex(L, xcd_chunk(blockIdx.x, gridDim.x), v)) return;
    int x, y, z;
    L.pos(v, x, y, z);
    FineCtx c;
    fine_ctx(c, x, y, z, L, w);
#pragma unroll
    for (int s = 0; s < 6; s++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float a = 0.f;
#pragma unroll
            for (int e = 0; e < 8; e++) a = fmaf(c.wc[e], kchi[s * 24 + 3 * (7 - e) + i], a);
            f[(size_t)(s * 3 + i) * N + v] = (v == 0) ? 0.f : a;
        }
}

// Iterative-refinement residual in FP64 (the row of grid.cu:1277-1461 in FP64): r = f(w) - K u with u
// the FP64 solution, f the macro-strain loads rebuilt from the cell weights in FP64 (kchi64 = KE chi0
// in FP64) and every product accumulated in FP64.  Writes r rounded to FP32 (the inner solve's
// right-hand side) and per-block FP64 partial sums of |r|^2 per load case.  DO_K = false gives r = f
// (u = 0, nothing read from u).  All six load cases in one pass:
// the 27 FP64 blocks are built once per vertex and applied to every load case.  The neighbour loop
// stays rolled (d is uniform over the wave): fully unrolled, the FP64 gather spills.  At most 128
// VGPRs, so that two 512-thread workgroups share a CU (the FP64 gather is latency-bound).
template <bool DO_K>
__global__ __launch_bounds__(kFullThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_fine_residual64(
    Lay L, size_t N, const double* __restrict__ ke, const double* __restrict__ kchi, const float* __restrict__ xc,
    double emin, const double* __restrict__ u, float* __restrict__ r32, double* __restrict__ partial) {
    constexpr int R = 6;
    __shared__ double sh[kFullThreads / 64];
    __shared__ double ke_s[kKeStaged];
    stage_ke(ke_s, ke);
    uint32_t v;
    const bool live = full_vertex(L, xcd_chunk(blockIdx.x, gridDim.x), v);
    double nrm[R];
#pragma unroll
    for (int r = 0; r < R; r++) nrm[r] = 0.0;
    if (live) {
        int x, y, z;
        L.pos(v, x, y, z);
        FineCtx c;
        fine_ctx64(c, x, y, z, L, xc, emin);
        double acc[R][3];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int i = 0; i < 3; i++) acc[r][i] = 0.0;
        if (DO_K) {
            const Rsrc ru = make_rsrc(u, 18 * N * sizeof(double));
            const uint32_t plane = (uint32_t)(2 * N * sizeof(double));  // one plane of component pairs
            // cell weights and index parts as plain scalars: the loop below runs over a run-time d
            double wd[8];
#pragma unroll
            for (int e = 0; e < 8; e++) wd[e] = c.wd[e];
            const uint32_t X0 = c.X[0], X1 = c.X[1], X2 = c.X[2], Y0 = c.Y[0], Y1 = c.Y[1], Y2 = c.Y[2], Z0 = c.Z[0],
                           Z1 = c.Z[1], Z2 = c.Z[2];
#pragma unroll 1
            for (int d = 0; d < 27; d++) {
                const int dx = d % 3 - 1, dy = d / 3 % 3 - 1, dz = d / 9 - 1;
                const uint32_t qo = ((dx < 0 ? X0 : (dx == 0 ? X1 : X2)) + (dy < 0 ? Y0 : (dy == 0 ? Y1 : Y2)) +
                                     (dz < 0 ? Z0 : (dz == 0 ? Z1 : Z2))) *
                                    (uint32_t)sizeof(double2_t);
                double un[R][3];
#pragma unroll
                for (int q = 0; q < 9; q++) {
                    const double2_t w = bload<double2_t>(ru, qo, (uint32_t)q * plane);
                    un[(2 * q) / 3][(2 * q) % 3] = w.x;
                    un[(2 * q + 1) / 3][(2 * q + 1) % 3] = w.y;
                }
                // the block from signed sums of the weights of the cells holding v and v + d
                // (fine_block's form, with d known only at run time): cell e at offset
                // (e&1, e>>1&1, e>>2) - 1 holds v + d iff e_m = (d_m + 1) / 2 wherever d_m != 0; the sign
                // of axis m is -1 for e_m = 1 where d_m = 0
                double w0 = 0.0, fxy = 0.0, fxz = 0.0, fyz = 0.0;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
                    if ((dx == 0 || ex == (dx + 1) / 2) && (dy == 0 || ey == (dy + 1) / 2) &&
                        (dz == 0 || ez == (dz + 1) / 2)) {
                        const double sx = (dx == 0 && ex) ? -1.0 : 1.0, sy = (dy == 0 && ey) ? -1.0 : 1.0,
                                     sz = (dz == 0 && ez) ? -1.0 : 1.0;
                        w0 += wd[e];
                        fxy = fma(sx * sy, wd[e], fxy);
                        fxz = fma(sx * sz, wd[e], fxz);
                        fyz = fma(sy * sz, wd[e], fyz);
                    }
                }
                const double* cb = ke_s + kBlockConsts + 9 * d;
                const double S[9] = {cb[0] * w0,  cb[1] * fxy, cb[2] * fxz, cb[3] * fxy, cb[4] * w0,
                                     cb[5] * fyz, cb[6] * fxz, cb[7] * fyz, cb[8] * w0};
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int i = 0; i < 3; i++)
                        acc[r][i] = fma(S[i * 3], un[r][0], fma(S[i * 3 + 1], un[r][1], fma(S[i * 3 + 2], un[r][2], acc[r][i])));
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double f = 0.0;
#pragma unroll
                for (int e = 0; e < 8; e++) f = fma(c.wd[e], kchi[r * 24 + 3 * (7 - e) + i], f);
                const double val = (v == 0) ? 0.0 : f - acc[r][i];
                r32[(size_t)(r * 3 + i) * N + v] = (float)val;
                nrm[r] = fma(val, val, nrm[r]);
            }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const double s = block_sum<kFullThreads>(nrm[r], sh);
        if (threadIdx.x == 0) partial[r * gridDim.x + blockIdx.x] = s;
    }
}

// |f|^2 per load case of the FP64 macro-strain loads (pinned row excluded).
__global__ __launch_bounds__(kFullThreads) void k_rhs_norm64(Lay L, size_t N, const double* __restrict__ kchi,
                                                             const float* __restrict__ xc, double emin,
                                                             double* __restrict__ partial) {
    __shared__ double sh[kFullThreads / 64];
    uint32_t v;
    const bool live = full_vertex(L, xcd_chunk(blockIdx.x, gridDim.x), v);
    double nrm[6] = {0, 0, 0, 0, 0, 0};
    if (live && v != 0) {
        int x, y, z;
        L.pos(v, x, y, z);
        FineCtx c;
        fine_ctx64(c, x, y, z, L, xc, emin);
#pragma unroll
        for (int s = 0; s < 6; s++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double f = 0.0;
#pragma unroll
                for (int e = 0; e < 8; e++) f = fma(c.wd[e], kchi[s * 24 + 3 * (7 - e) + i], f);
                nrm[s] = fma(f, f, nrm[s]);
            }
    }
#pragma unroll
    for (int s = 0; s < 6; s++) {
        const double t = block_sum<kFullThreads>(nrm[s], sh);
        if (threadIdx.x == 0) partial[s * gridDim.x + blockIdx.x] = t;
    }
}

// u (+)= x + alpha p for all six load cases (x, p FP32 planes [18][N], N % 2 == 0; u in pairs,
// u64_index; alpha of load case r = sc[12 + r], taken as 0 unless pending): the last CG step's x
// update, rounded to FP32 as the CG pass rounds it, added to the FP64 iterate (add) or copied into
// it, in one pass.  Grid row q covers the pair plane of components 2q and 2q + 1.
__global__ __launch_bounds__(256) void k_refine_accumulate(size_t N, const double* __restrict__ sc, int pending,
                                                           const float* __restrict__ x, const float* __restrict__ p,
                                                           double* __restrict__ u, int add) {
    const int q = blockIdx.y, c0 = 2 * q, c1 = 2 * q + 1;
    const float a0 = pending ? (float)sc[12 + c0 / 3] : 0.f, a1 = pending ? (float)sc[12 + c1 / 3] : 0.f;
    const float2* x0 = reinterpret_cast<const float2*>(x + (size_t)c0 * N);
    const float2* x1 = reinterpret_cast<const float2*>(x + (size_t)c1 * N);
    const float2* p0 = reinterpret_cast<const float2*>(p + (size_t)c0 * N);
    const float2* p1 = reinterpret_cast<const float2*>(p + (size_t)c1 * N);
    double2* u2 = reinterpret_cast<double2*>(u + (size_t)q * 2 * N);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N / 2; i += (size_t)gridDim.x * blockDim.x) {
        float2 xa = x0[i], xb = x1[i];
        if (a0 != 0.f) {
            const float2 pa = p0[i];
            xa.x = fmaf(a0, pa.x, xa.x);
            xa.y = fmaf(a0, pa.y, xa.y);
        }
        if (a1 != 0.f) {
            const float2 pb = p1[i];
            xb.x = fmaf(a1, pb.x, xb.x);
            xb.y = fmaf(a1, pb.y, xb.y);
        }
        // vertices 2i and 2i + 1: pairs (c0, c1) of each
        double2 lo = make_double2((double)xa.x, (double)xb.x), hi = make_double2((double)xa.y, (double)xb.y);
        if (add) {
            const double2 ulo = u2[2 * i], uhi = u2[2 * i + 1];
            lo.x += ulo.x;
            lo.y += ulo.y;
            hi.x += uhi.x;
            hi.y += uhi.y;
        }
        u2[2 * i] = lo;
        u2[2 * i + 1] = hi;
    }
}

// FP64 iterate (pairs) -> FP32 planes [18][N], and back; one thread per (pair plane, vertex)
__global__ __launch_bounds__(256) void k_round_f32(size_t N, const double* __restrict__ u, float* __restrict__ x) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * N) return;
    const size_t q = t / N, v = t - q * N;
    const double2 w = reinterpret_cast<const double2*>(u)[t];
    x[2 * q * N + v] = (float)w.x;
    x[(2 * q + 1) * N + v] = (float)w.y;
}

__global__ __launch_bounds__(256) void k_widen_f64(size_t N, const float* __restrict__ x, double* __restrict__ u) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * N) return;
    const size_t q = t / N, v = t - q * N;
    reinterpret_cast<double2*>(u)[t] = make_double2((double)x[2 * q * N + v], (double)x[(2 * q + 1) * N + v]);
}

// zero the components of load case r (3 r .. 3 r + 2) of the FP64 iterate, every vertex
__global__ __launch_bounds__(256) void k_zero_case64(size_t N, int r, double* __restrict__ u) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * N) return;
    const size_t k = t / N, v = t - k * N;
    u[u64_index(3 * r + (int)k, v, N)] = 0.0;
}

// zero the pinned origin's components of the FP64 iterate
__global__ void k_zero_pin64(size_t N, double* __restrict__ u) {
    if (threadIdx.x < 18) u[u64_index((int)threadIdx.x, 0, N)] = 0.0;
}

// zero the pinned origin's entries of a 6-load-case field of `planes` word planes
// FP32 planes [6][3][N] <-> packed FP16 words (word q = p * 3 + i: cases 2p, 2p + 1; hw_index)
__global__ __launch_bounds__(256) void k_pack_half(size_t N, const float* __restrict__ u, half2_t* __restrict__ z) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= 9 * N) return;
    const size_t q = t / N, v = t - q * N, p = q / 3, i = q - 3 * p;
    half2_t h;
    h.x = (_Float16)u[((2 * p) * 3 + i) * N + v];
    h.y = (_Float16)u[((2 * p + 1) * 3 + i) * N + v];
    z[hw_index((int)q, v, N)] = h;
}
__global__ __launch_bounds__(256) void k_unpack_half(size_t N, const half2_t* __restrict__ z, float* __restrict__ u) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= 9 * N) return;
    const size_t q = t / N, v = t - q * N, p = q / 3, i = q - 3 * p;
    const half2_t h = z[hw_index((int)q, v, N)];
    u[((2 * p) * 3 + i) * N + v] = (float)h.x;
    u[((2 * p + 1) * 3 + i) * N + v] = (float)h.y;
}

// (half: the 9 words of a packed FP16 field)
template <typename T, bool HALF = false>
__global__ void k_zero_pin(size_t N, int planes, T* __restrict__ u) {
    if ((int)threadIdx.x < planes) u[HALF ? hw_index((int)threadIdx.x, 0, N) : (size_t)threadIdx.x * N] = T(0);
}

// ---------------------------------------------------------------------------------------------
// Coarse levels (stored 27-point stencils)
// ---------------------------------------------------------------------------------------------

// A stored coarse stencil: FP32 entries [243][N] (entry e = 9d + 3i + j), or packed FP16 [122][N]
// words, word w holding entries 2w (.x) and 2w + 1 (.y).  Both have 4-byte words per vertex.
template <typename SA>
struct StencilIO;
template <>
struct StencilIO<float> {
    static constexpr int words = 243;
    __device__ static float get(Rsrc rs, uint32_t vo, uint32_t plane, int e) {
        return bload<float>(rs, vo, (uint32_t)e * plane);
    }
    __device__ static void put(float* st, size_t N, uint32_t v, int e, float val) { st[(size_t)e * N + v] = val; }
};
template <>
struct StencilIO<half2_t> {
    static constexpr int words = 122;
    __device__ static float get(Rsrc rs, uint32_t vo, uint32_t plane, int e) {
        return (float)bload<half2_t>(rs, vo, (uint32_t)(e >> 1) * plane)[e & 1];
    }
    // one 16-bit store per entry: the two halves of a word may come from different threads
    __device__ static void put(half2_t* st, size_t N, uint32_t v, int e, float val) {
        reinterpret_cast<_Float16*>(st + (size_t)(e >> 1) * N + v)[e & 1] = (_Float16)val;
    }
};

// One neighbour of a coarse row: issue the loads of the next selected neighbour (its stencil block and
// words), accumulate this one, recurse.  The barrier keeps the loads one neighbour ahead: with the whole
// row unrolled freely the compiler issues every neighbour's loads at once and needs 256 VGPRs.
struct StencilRow {
    Rsrc rs, ru;
    uint32_t X[3], Y[3], Z[3], vo, plane;
};
template <typename SA, typename A, int W, bool XL, int D>
__device__ __forceinline__ void stencil_nb(float (&S)[9], A (&cur)[W][3], const StencilRow& c, int r0,
                                           const XNbLds<A>& xl) {
#pragma unroll
    for (int k = 0; k < 9; k++) S[k] = StencilIO<SA>::get(c.rs, c.vo, c.plane, D * 9 + k);
    if constexpr (XL && (D == 12 || D == 14)) {
        load_nb_lds<W, A>(cur, xl, D == 12 ? 0 : 1);
    } else {
        const uint32_t qo = (c.X[D % 3] + c.Y[D / 3 % 3] + c.Z[D / 9]) * 4u;
        load_nb<W, A>(cur, c.ru, qo, c.plane, r0);
    }
}
template <int R, uint32_t MASK, int D, typename SA, typename A, bool XL>
__device__ __forceinline__ void stencil_step(const StencilRow& c, int r0, const XNbLds<A>& xl, const float (&S)[9],
                                             const A (&cur)[R / CasesPerWord<A>::value][3], float (&acc)[R][3]) {
    constexpr int W = R / CasesPerWord<A>::value;
    constexpr int DN = next_nb(MASK, D);
    __builtin_amdgcn_sched_barrier(0);
    pin_acc<R, float>(acc);
    float Sn[9];
    A nxt[W][3];
    if constexpr (DN < 27) stencil_nb<SA, A, W, XL, DN>(Sn, nxt, c, r0, xl);
    block_fma<W, float, A>(S, cur, acc);
    if constexpr (DN < 27) stencil_step<R, MASK, DN, SA, A, XL>(c, r0, xl, Sn, nxt, acc);
}

// acc = sum over the neighbours selected by MASK of the stored 3x3 blocks times u, load cases r0 .. r0 + R - 1.
// A: storage of u (float planes, or half2_t words of load-case pairs; r0 then even).  XL: the two
// x-neighbours (d = 12, 14) come from LDS (xl).
template <int R, uint32_t MASK, typename SA, typename A = float, bool XL = false>
__device__ __forceinline__ void stencil_gather(int x, int y, int z, const Lay& L, size_t N, uint32_t v,
                                               const SA* __restrict__ st, const A* __restrict__ u,
                                               float (&acc)[R][3], int r0, const XNbLds<A>& xl = XNbLds<A>()) {
    constexpr int W = R / CasesPerWord<A>::value;
    StencilRow c;
    c.X[0] = L.part(wrap_dec(x, L.n), 0); c.X[1] = L.part(x, 0); c.X[2] = L.part(wrap_inc(x, L.n), 0);
    c.Y[0] = L.part(wrap_dec(y, L.n), 1); c.Y[1] = L.part(y, 1); c.Y[2] = L.part(wrap_inc(y, L.n), 1);
    c.Z[0] = L.part(wrap_dec(z, L.n), 2); c.Z[1] = L.part(z, 2); c.Z[2] = L.part(wrap_inc(z, L.n), 2);
    c.rs = make_rsrc(st, (size_t)StencilIO<SA>::words * N * 4);
    c.ru = make_rsrc(u, (size_t)(18 / CasesPerWord<A>::value) * N * sizeof(A));
    c.plane = (uint32_t)(N * sizeof(float));  // stencil and u words are 4 bytes
    c.vo = v * 4u;
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 3; i++) acc[r][i] = 0.f;
    constexpr int D0 = next_nb(MASK, -1);
    if constexpr (D0 < 27) {
        float S[9];
        A cur[W][3];
        stencil_nb<SA, A, W, XL, D0>(S, cur, c, r0, xl);
        stencil_step<R, MASK, D0, SA, A, XL>(c, r0, xl, S, cur, acc);
    }
}

// MODE 0: out = A u; MODE 1: out = f - A u (update_residual_kernel_1, grid.cu:1465-1564).  A: u storage.
template <int R, int MODE, typename SA, typename A = float>
__global__ __launch_bounds__(kFullThreads) void k_stencil_apply(Lay L, size_t N, const SA* __restrict__ st,
                                                                const A* __restrict__ u,
                                                                const float* __restrict__ f, float* __restrict__ out) {
    uint32_t v;
    if (!full_vertex(L, xcd_chunk(blockIdx.x, gridDim.x), v)) return;
    int x, y, z;
    L.pos(v, x, y, z);
    float acc[R][3];
    const int r0 = blockIdx.y * R;
    stencil_gather<R, kRowAll, SA, A>(x, y, z, L, N, v, st, u, acc, r0);
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const size_t o = (size_t)((r0 + r) * 3 + i) * N + v;
            out[o] = (MODE == 1) ? f[o] - acc[r][i] : acc[r][i];
        }
}

// GS / SOR update of the colour vertex v of a coarse level (load cases r0 .. r0 + R - 1), in place.
// ZC as in k_fine_gs.  A: u storage (float, or half2_t load-case pairs with FP32 arithmetic).  XL: the
// x-neighbours come from LDS; TO_LDS: the new words also go to slot threadIdx.x of out_lds.  ROWN
// (ZC >= 0): also the own-colour residual f - (A u)_{colours <= ZC} to rown, as gs_vertex does.
template <int R, bool REVERSE, int ZC, typename SA, typename A, bool XL, bool TO_LDS, bool ROWN = false>
__device__ __forceinline__ void stencil_gs_vertex(const Lay& L, size_t N, uint32_t v, float omega,
                                                  const SA* __restrict__ st, A* __restrict__ u,
                                                  const float* __restrict__ f, int r0, const XNbLds<A>& xl,
                                                  A* out_lds, float* __restrict__ rown = nullptr) {
    static_assert(!ROWN || ZC >= 0, "the own-colour residual needs a zero-start forward pass");
    int x, y, z;
    L.pos(v, x, y, z);
    float acc[R][3];
    stencil_gather<R, gs_mask(ZC), SA, A, XL>(x, y, z, L, N, v, st, u, acc, r0, xl);
    float Ks[9];
    const Rsrc rs = make_rsrc(st, (size_t)StencilIO<SA>::words * N * 4);
#pragma unroll
    for (int k = 0; k < 9; k++) Ks[k] = StencilIO<SA>::get(rs, v * 4u, (uint32_t)(N * 4), 13 * 9 + k);
    const float inv[3] = {1.f / Ks[0], 1.f / Ks[4], 1.f / Ks[8]};
    constexpr int cpw = CasesPerWord<A>::value;
#pragma unroll
    for (int p = 0; p < R / cpw; p++) {
        A wd[3];
#pragma unroll
        for (int i = 0; i < 3; i++) wd[i] = ZC >= 0 ? A(0) : u[fw_index<A>((r0 / cpw + p) * 3 + i, v, N)];
#pragma unroll
        for (int hh = 0; hh < cpw; hh++) {
            const int r = p * cpw + hh;
            float uu[3], ff[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                if constexpr (cpw == 2)
                    uu[i] = (float)wd[i][hh];
                else
                    uu[i] = (float)wd[i];
                ff[i] = f[(size_t)((r0 + r) * 3 + i) * N + v];
            }
            block_update<REVERSE>(uu, ff, acc[r], Ks, inv, omega);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                if constexpr (cpw == 2)
                    wd[i][hh] = (_Float16)uu[i];
                else
                    wd[i] = uu[i];
            }
            if constexpr (ROWN) {
                float us[3];  // the stored values
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    if constexpr (cpw == 2)
                        us[i] = (float)wd[i][hh];
                    else
                        us[i] = (float)wd[i];
                }
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const float kv = fmaf(Ks[3 * i], us[0], fmaf(Ks[3 * i + 1], us[1], Ks[3 * i + 2] * us[2]));
                    rown[(size_t)((r0 + r) * 3 + i) * N + v] = (ff[i] - acc[r][i]) - kv;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            u[fw_index<A>((r0 / cpw + p) * 3 + i, v, N)] = wd[i];
            if constexpr (TO_LDS) out_lds[(p * 3 + i) * 256 + threadIdx.x] = wd[i];
        }
    }
}

// One colour of the coarse block GS sweep (gs_relaxation_kernel, grid.cu:556-689); no pin on coarse levels.
template <int R, bool REVERSE, int ZC, typename SA, typename A = float>
__global__ __launch_bounds__(256) void k_stencil_gs(Lay L, size_t N, int colour, float omega,
                                                    const SA* __restrict__ st, A* __restrict__ u,
                                                    const float* __restrict__ f) {
    const uint32_t t = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= L.h3) return;
    stencil_gs_vertex<R, REVERSE, ZC, SA, A, false, false>(L, N, colour_block(colour) * L.h3 + t, omega, st, u, f,
                                                           blockIdx.y * R, XNbLds<A>(), nullptr);
}

// Two consecutive colours of a coarse sweep in one launch, as k_fine_gs2 does at the finest level
// (whole grid rows per workgroup; the first colour's new words stay in LDS for the second colour's
// x-neighbours).
template <int R, bool REVERSE, int ZC, typename SA, typename A, bool ROWN = false>
__global__ __launch_bounds__(256) void k_stencil_gs2(Lay L, size_t N, int cfirst, float omega,
                                                     const SA* __restrict__ st, A* __restrict__ u,
                                                     const float* __restrict__ f, float* __restrict__ rown) {
    constexpr int W = R / CasesPerWord<A>::value;
    __shared__ A xs[3 * W * 256];
    const uint32_t t = xcd_chunk(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
    const int c2 = REVERSE ? cfirst - 1 : cfirst + 1;
    const int r0 = blockIdx.y * R;
    stencil_gs_vertex<R, REVERSE, ZC, SA, A, false, true, ROWN>(L, N, colour_block(cfirst) * L.h3 + t, omega, st, u, f,
                                                                r0, XNbLds<A>(), xs, rown);
    __syncthreads();
    XNbLds<A> xl;
    xl.base = xs;
    const int h = L.h, lx = (int)(t % (uint32_t)h), row0 = (int)threadIdx.x - lx;
    const bool odd = (colour_block(c2) & 1u) != 0;  // x-parity of the second colour
    xl.slot[0] = row0 + (odd ? lx : (lx == 0 ? h - 1 : lx - 1));
    xl.slot[1] = row0 + (odd ? (lx + 1 == h ? 0 : lx + 1) : lx);
    stencil_gs_vertex<R, REVERSE, (ZC < 0 ? -1 : ZC + 1), SA, A, true, false, ROWN>(
        L, N, colour_block(c2) * L.h3 + t, omega, st, u, f, r0, xl, nullptr, rown);
}

// The coarse-level counterpart of k_fine_resid_later: r = r_own - the later-colour terms, on the same
// parity tiles with four waves (wave w takes colours w and 7 - w; colour 7 is final).  Load cases
// r0 .. r0 + R - 1 with r0 = blockIdx.y R.  zero as in k_fine_resid_later.
template <int R, int C, typename SA, typename A>
__device__ __forceinline__ void stencil_later_vertex(const Lay& L, size_t N, uint32_t v, const SA* __restrict__ st,
                                                     const A* __restrict__ u, float* __restrict__ r, int r0,
                                                     float zero) {
    int x, y, z;
    L.pos(v, x, y, z);
    float acc[R][3];
    stencil_gather<R, later_mask(C), SA, A>(x, y, z, L, N, v, st, u, acc, r0);
#pragma unroll
    for (int q = 0; q < R; q++) {
        float ro[3], rn[3];
#pragma unroll
        for (int i = 0; i < 3; i++) ro[i] = r[(size_t)((r0 + q) * 3 + i) * N + v];
        rn[0] = ro[0] - acc[q][0];
        rn[1] = fmaf(-zero, rn[0], ro[1] - acc[q][1]);
        rn[2] = fmaf(-zero, rn[1], ro[2] - acc[q][2]);
#pragma unroll
        for (int i = 0; i < 3; i++) r[(size_t)((r0 + q) * 3 + i) * N + v] = rn[i];
    }
}

template <int R, typename SA, typename A>
__global__ __launch_bounds__(256) void k_stencil_resid_later(Lay L, size_t N, const SA* __restrict__ st,
                                                             const A* __restrict__ u, float* __restrict__ r,
                                                             float zero) {
    const uint32_t t = xcd_chunk(blockIdx.x, gridDim.x) * 64u + (threadIdx.x & 63u);
    if (t >= L.h3) return;
    const uint32_t h3 = L.h3;
    const int r0 = blockIdx.y * R;
    switch (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) {
        case 0: stencil_later_vertex<R, 0, SA, A>(L, N, 7 * h3 + t, st, u, r, r0, zero); break;
        case 1:
            stencil_later_vertex<R, 1, SA, A>(L, N, 6 * h3 + t, st, u, r, r0, zero);
            stencil_later_vertex<R, 6, SA, A>(L, N, 1 * h3 + t, st, u, r, r0, zero);
            break;
        case 2:
            stencil_later_vertex<R, 2, SA, A>(L, N, 5 * h3 + t, st, u, r, r0, zero);
            stencil_later_vertex<R, 5, SA, A>(L, N, 2 * h3 + t, st, u, r, r0, zero);
            break;
        default:
            stencil_later_vertex<R, 3, SA, A>(L, N, 4 * h3 + t, st, u, r, r0, zero);
            stencil_later_vertex<R, 4, SA, A>(L, N, 3 * h3 + t, st, u, r, r0, zero);
            break;
    }
}

// Coarse-vertex kernels that read the fine level around 2V take coarse vertices in lexicographic
// order whatever the coarse storage, so that consecutive lanes read consecutive fine local positions.
__device__ __forceinline__ bool coarse_lex_vertex(const Lay& Lc, int& X, int& Y, int& Z, uint32_t& V) {
    const uint32_t i = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const uint32_t n = (uint32_t)Lc.n;
    if (i >= n * n * n) return false;
    X = (int)(i % n);
    Y = (int)((i / n) % n);
    Z = (int)(i / (n * n));
    V = Lc.idx(X, Y, Z);
    return true;
}

// ---------------------------------------------------------------------------------------------
// Grid transfers (restrict_residual_kernel_1 grid.cu:917-994; prolongate_correction_kernel_1
// grid.cu:1120-1190).  Weights prod_k (2-|o_k|)/2; R = P^T.
// ---------------------------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(256) void k_restrict(Lay Lf, Lay Lc, const float* __restrict__ rf,
                                                  float* __restrict__ fc) {
    const size_t Nc = (size_t)Lc.n * Lc.n * Lc.n, Nf = (size_t)Lf.n * Lf.n * Lf.n;
    int X, Y, Z;
    uint32_t V;
    if (!coarse_lex_vertex(Lc, X, Y, Z, V)) return;
    const uint32_t xs[3] = {Lf.part(wrap_dec(2 * X, Lf.n), 0), Lf.part(2 * X, 0), Lf.part(2 * X + 1, 0)};
    const uint32_t ys[3] = {Lf.part(wrap_dec(2 * Y, Lf.n), 1), Lf.part(2 * Y, 1), Lf.part(2 * Y + 1, 1)};
    const uint32_t zs[3] = {Lf.part(wrap_dec(2 * Z, Lf.n), 2), Lf.part(2 * Z, 2), Lf.part(2 * Z + 1, 2)};
    float acc[R][3];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 3; i++) acc[r][i] = 0.f;
    const int r0 = blockIdx.y * R;
#pragma unroll
    for (int o = 0; o < 27; o++) {
        const int ox = o % 3, oy = o / 3 % 3, oz = o / 9;
        const float wgt = (float)((ox == 1 ? 2 : 1) * (oy == 1 ? 2 : 1) * (oz == 1 ? 2 : 1)) * 0.125f;
        const uint32_t q = xs[ox] + ys[oy] + zs[oz];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int i = 0; i < 3; i++)
                acc[r][i] = fmaf(wgt, rf[(size_t)((r0 + r) * 3 + i) * Nf + q], acc[r][i]);
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 3; i++) fc[(size_t)((r0 + r) * 3 + i) * Nc + V] = acc[r][i];
}

template <int R>
__global__ __launch_bounds__(kFullThreads) void k_prolong(Lay Lf, Lay Lc, const float* __restrict__ uc,
                                                          float* __restrict__ uf, int pin) {
    const size_t Nc = (size_t)Lc.n * Lc.n * Lc.n, Nf = (size_t)Lf.n * Lf.n * Lf.n;
    uint32_t v;
    if (!full_vertex(Lf, xcd_chunk(blockIdx.x, gridDim.x), v)) return;
    if (pin && v == 0) return;
    const int nc = Lc.n;
    int x, y, z;
    Lf.pos(v, x, y, z);
    // per axis: even -> one parent weight 1, odd -> two parents weight 1/2
    int ix[2], iy[2], iz[2];
    float wx[2], wy[2], wz[2];
    ix[0] = x >> 1; ix[1] = (x & 1) ? ((x >> 1) + 1 == nc ? 0 : (x >> 1) + 1) : ix[0];
    wx[0] = (x & 1) ? 0.5f : 1.f; wx[1] = (x & 1) ? 0.5f : 0.f;
    iy[0] = y >> 1; iy[1] = (y & 1) ? ((y >> 1) + 1 == nc ? 0 : (y >> 1) + 1) : iy[0];
    wy[0] = (y & 1) ? 0.5f : 1.f; wy[1] = (y & 1) ? 0.5f : 0.f;
    iz[0] = z >> 1; iz[1] = (z & 1) ? ((z >> 1) + 1 == nc ? 0 : (z >> 1) + 1) : iz[0];
    wz[0] = (z & 1) ? 0.5f : 1.f; wz[1] = (z & 1) ? 0.5f : 0.f;
    float acc[R][3];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 3; i++) acc[r][i] = 0.f;
    const int r0 = blockIdx.y * R;
#pragma unroll
    for (int p = 0; p < 8; p++) {
        const float wgt = wx[p & 1] * wy[(p >> 1) & 1] * wz[p >> 2];
        if (wgt == 0.f) continue;
        const uint32_t q = Lc.idx(ix[p & 1], iy[(p >> 1) & 1], iz[p >> 2]);
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int i = 0; i < 3; i++)
                acc[r][i] = fmaf(wgt, uc[(size_t)((r0 + r) * 3 + i) * Nc + q], acc[r][i]);
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 3; i++) uf[(size_t)((r0 + r) * 3 + i) * Nf + v] += acc[r][i];
}

// Vectorised transfers for nf % 8 == 0 (so nc = nf / 2 and the fine local row length are multiples of
// 4): each thread handles four consecutive x of one load case (r = blockIdx.y) with float4 / float2
// accesses.  Same weights and sums as k_restrict / k_prolong.
//
// Restriction: coarse X = 4j..4j+3 of row (Y, Z).  Along x the fine taps are 2X (even block, local X),
// 2X+1 (odd block, local X) and 2X-1 (odd block, local X-1): two float4 rows and one scalar.
__global__ __launch_bounds__(256) void k_restrict4(Lay Lf, Lay Lc, const float* __restrict__ rf,
                                                   float* __restrict__ fc) {
    const int nf = Lf.n, nc = Lc.n, q = nc / 4;
    const uint32_t i = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)q * nc * nc) return;
    const int j = (int)(i % q), Y = (int)((i / q) % nc), Z = (int)(i / ((uint32_t)q * nc));
    const size_t Nf = (size_t)nf * nf * nf, Nc = (size_t)nc * nc * nc;
    const int r = blockIdx.y;
    const uint32_t xe = Lf.part(8 * j, 0), xo = Lf.part(8 * j + 1, 0), xp = Lf.part(wrap_dec(8 * j, nf), 0);
    const uint32_t ys[3] = {Lf.part(wrap_dec(2 * Y, nf), 1), Lf.part(2 * Y, 1), Lf.part(2 * Y + 1, 1)};
    const uint32_t zs[3] = {Lf.part(wrap_dec(2 * Z, nf), 2), Lf.part(2 * Z, 2), Lf.part(2 * Z + 1, 2)};
    const float wt[3] = {0.5f, 1.f, 0.5f};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float* pl = rf + (size_t)(r * 3 + c) * Nf;
        float a[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int oz = 0; oz < 3; oz++)
#pragma unroll
            for (int oy = 0; oy < 3; oy++) {
                const uint32_t base = ys[oy] + zs[oz];
                const float4 e = *reinterpret_cast<const float4*>(pl + base + xe);
                const float4 o = *reinterpret_cast<const float4*>(pl + base + xo);
                const float op = pl[base + xp];
                const float w = wt[oy] * wt[oz];
                a[0] = fmaf(w, e.x + 0.5f * (o.x + op), a[0]);
                a[1] = fmaf(w, e.y + 0.5f * (o.y + o.x), a[1]);
                a[2] = fmaf(w, e.z + 0.5f * (o.z + o.y), a[2]);
                a[3] = fmaf(w, e.w + 0.5f * (o.w + o.z), a[3]);
            }
        float* out = fc + (size_t)(r * 3 + c) * Nc;
        if (Lc.lex) {
            *reinterpret_cast<float4*>(out + Lc.idx(4 * j, Y, Z)) = make_float4(a[0], a[1], a[2], a[3]);
        } else {
            const uint32_t yz = Lc.part(Y, 1) + Lc.part(Z, 2);
            *reinterpret_cast<float2*>(out + yz + Lc.part(4 * j, 0)) = make_float2(a[0], a[2]);
            *reinterpret_cast<float2*>(out + yz + Lc.part(4 * j + 1, 0)) = make_float2(a[1], a[3]);
        }
    }
}

// Prolongation: the 8 waves of a workgroup take the 8 fine parity blocks; a lane adds P u_c to four
// consecutive local x (fine x = 2lx + px) of one row.  Parents along x: lx (weight 1) when px = 0;
// lx and lx + 1 (1/2 each) when px = 1.
__global__ __launch_bounds__(kFullThreads) void k_prolong4(Lay Lf, Lay Lc, const float* __restrict__ uc,
                                                           float* __restrict__ uf, int pin) {
    const uint32_t hf = (uint32_t)Lf.h, q = hf / 4;
    const uint32_t t4 = xcd_chunk(blockIdx.x, gridDim.x) * 64u + (threadIdx.x & 63u);
    if (t4 >= Lf.h3 / 4) return;
    const uint32_t b = threadIdx.x >> 6;
    const int px = b & 1, py = (b >> 1) & 1, pz = (int)(b >> 2);
    const uint32_t j = t4 % q, row = t4 / q;
    const int ly = (int)(row % hf), lz = (int)(row / hf);
    const int nc = Lc.n;
    const int Ys[2] = {ly, ly + 1 == nc ? 0 : ly + 1}, Zs[2] = {lz, lz + 1 == nc ? 0 : lz + 1};
    const float wy[2] = {py ? 0.5f : 1.f, 0.5f}, wz[2] = {pz ? 0.5f : 1.f, 0.5f};
    const int X4 = (int)(4 * j + 4) == nc ? 0 : (int)(4 * j + 4);
    const size_t Nf = (size_t)Lf.n * Lf.n * Lf.n, Nc = (size_t)nc * nc * nc;
    const int r = blockIdx.y;
    const bool skip0 = pin && b == 0 && row == 0 && j == 0;  // pinned origin
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float* cp = uc + (size_t)(r * 3 + c) * Nc;
        float a[4] = {0.f, 0.f, 0.f, 0.f};
        for (int iz = 0; iz <= pz; iz++)
            for (int iy = 0; iy <= py; iy++) {
                float cv[5];
                if (Lc.lex) {
                    const uint32_t base = (uint32_t)nc * (Ys[iy] + (uint32_t)nc * Zs[iz]);
                    const float4 v4 = *reinterpret_cast<const float4*>(cp + base + 4 * j);
                    cv[0] = v4.x; cv[1] = v4.y; cv[2] = v4.z; cv[3] = v4.w;
                    cv[4] = px ? cp[base + X4] : 0.f;
                } else {
                    const uint32_t yz = Lc.part(Ys[iy], 1) + Lc.part(Zs[iz], 2);
                    const float2 ev = *reinterpret_cast<const float2*>(cp + yz + Lc.part(4 * j, 0));
                    const float2 od = *reinterpret_cast<const float2*>(cp + yz + Lc.part(4 * j + 1, 0));
                    cv[0] = ev.x; cv[1] = od.x; cv[2] = ev.y; cv[3] = od.y;
                    cv[4] = px ? cp[yz + Lc.part(X4, 0)] : 0.f;
                }
                const float w = wy[iy] * wz[iz];
#pragma unroll
                for (int k = 0; k < 4; k++) a[k] = fmaf(w, px ? 0.5f * (cv[k] + cv[k + 1]) : cv[k], a[k]);
            }
        float4* dst = reinterpret_cast<float4*>(uf + (size_t)(r * 3 + c) * Nf + b * Lf.h3 + row * hf + 4 * j);
        float4 u4 = *dst;
        u4.x += skip0 ? 0.f : a[0];
        u4.y += a[1];
        u4.z += a[2];
        u4.w += a[3];
        *dst = u4;
    }
}

// Prolongation into a packed FP16 fine field (hw_index words): uf += P uc, one thread per fine vertex
// and all nine of its words (two 16-byte and one 4-byte read-modify-write), on the 512-thread parity
// tiles of full_vertex.  Per axis an even fine coordinate has one coarse parent (weight 1) and an odd one
// two (weight 1/2 each): the x parents are averaged first, then the (y, z) parents accumulate with their
// weights.  CT: coarse storage, FP32 planes (lexicographic or parity-blocked) or packed FP16 words
// (parity-blocked).  The pinned origin (pin) keeps its value.
template <typename CT>
__device__ __forceinline__ void coarse_words(float (&c)[9][2], const CT* __restrict__ uc, uint32_t idx, size_t Nc) {
    if constexpr (CasesPerWord<CT>::value == 2) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(uc);
        const uint4 a = *reinterpret_cast<const uint4*>(w + 4 * (size_t)idx);
        const uint4 b = *reinterpret_cast<const uint4*>(w + 4 * Nc + 4 * (size_t)idx);
        const uint32_t wv[9] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, w[8 * Nc + idx]};
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const half2_t hv = __builtin_bit_cast(half2_t, wv[q]);
            c[q][0] = (float)hv.x;
            c[q][1] = (float)hv.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 9; q++)
#pragma unroll
            for (int h = 0; h < 2; h++) c[q][h] = uc[(size_t)((2 * (q / 3) + h) * 3 + q % 3) * Nc + idx];
    }
}

template <typename CT>
__global__ __launch_bounds__(kFullThreads) void k_prolong_h(Lay Lf, Lay Lc, const CT* __restrict__ uc,
                                                           half2_t* __restrict__ uf, int pin) {
    uint32_t v;
    if (!full_vertex(Lf, xcd_chunk(blockIdx.x, gridDim.x), v)) return;
    if (pin && v == 0) return;
    int x, y, z;
    Lf.pos(v, x, y, z);
    const int nc = Lc.n;
    const int px = x & 1, py = y & 1, pz = z & 1;
    const int X0 = x >> 1, X1 = X0 + 1 == nc ? 0 : X0 + 1;
    const int Ys[2] = {y >> 1, (y >> 1) + 1 == nc ? 0 : (y >> 1) + 1};
    const int Zs[2] = {z >> 1, (z >> 1) + 1 == nc ? 0 : (z >> 1) + 1};
    const float wy[2] = {py ? 0.5f : 1.f, 0.5f}, wz[2] = {pz ? 0.5f : 1.f, 0.5f};
    const size_t Nf = (size_t)Lf.n * Lf.n * Lf.n, Nc = (size_t)nc * nc * nc;
    float a[9][2];
#pragma unroll
    for (int q = 0; q < 9; q++) a[q][0] = a[q][1] = 0.f;
    for (int iz = 0; iz <= pz; iz++)
        for (int iy = 0; iy <= py; iy++) {
            const float w = wy[iy] * wz[iz];
            float c0[9][2], c1[9][2];
            coarse_words<CT>(c0, uc, Lc.idx(X0, Ys[iy], Zs[iz]), Nc);
            if (px) coarse_words<CT>(c1, uc, Lc.idx(X1, Ys[iy], Zs[iz]), Nc);
#pragma unroll
            for (int q = 0; q < 9; q++)
#pragma unroll
                for (int h = 0; h < 2; h++) a[q][h] = fmaf(w, px ? 0.5f * (c0[q][h] + c1[q][h]) : c0[q][h], a[q][h]);
        }
    uint32_t* fw = reinterpret_cast<uint32_t*>(uf);
    uint4* f0 = reinterpret_cast<uint4*>(fw + 4 * (size_t)v);
    uint4* f1 = reinterpret_cast<uint4*>(fw + 4 * Nf + 4 * (size_t)v);
    const uint4 o0 = *f0, o1 = *f1;
    uint32_t wv[9] = {o0.x, o0.y, o0.z, o0.w, o1.x, o1.y, o1.z, o1.w, fw[8 * Nf + v]};
#pragma unroll
    for (int q = 0; q < 9; q++) {
        half2_t hv = __builtin_bit_cast(half2_t, wv[q]);
        hv.x = (_Float16)((float)hv.x + a[q][0]);
        hv.y = (_Float16)((float)hv.y + a[q][1]);
        wv[q] = __builtin_bit_cast(uint32_t, hv);
    }
    *f0 = make_uint4(wv[0], wv[1], wv[2], wv[3]);
    *f1 = make_uint4(wv[4], wv[5], wv[6], wv[7]);
    fw[8 * Nf + v] = wv[8];
}

// ---------------------------------------------------------------------------------------------
// Galerkin coarse operators (restrict_stencil_otf_aos_kernel_1 grid_ex.cu:936-1061;
// restrict_stencil_aos_kernel_1 grid_ex.cu:1065-1160; pinned-origin rows grid.cpp:1184-1325)
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ float hat(int o) { return o == 0 ? 1.f : ((o == 1 || o == -1) ? 0.5f : 0.f); }

// Level-1 stencil away from the pinned origin: st[V][D] = sum_{c} w_{2V+c} G[c][D], c in {-2..1}^3.
// G (64 x 27 x 9) is the density-independent table of P-weighted KE blocks.
template <typename SO>
__global__ __launch_bounds__(256) void k_galerkin_fine(Lay Lf, Lay Lc, const float* __restrict__ G,
                                                       const float* __restrict__ w, SO* __restrict__ stc,
                                                       float clamp) {
    const int nf = Lf.n;
    const size_t Nc = (size_t)Lc.n * Lc.n * Lc.n;
    int X, Y, Z;
    uint32_t V;
    if (!coarse_lex_vertex(Lc, X, Y, Z, V)) return;
    float wc[64];
#pragma unroll
    for (int c = 0; c < 64; c++) {
        int cx = 2 * X + (c & 3) - 2, cy = 2 * Y + ((c >> 2) & 3) - 2, cz = 2 * Z + (c >> 4) - 2;
        cx = cx < 0 ? cx + nf : (cx >= nf ? cx - nf : cx);
        cy = cy < 0 ? cy + nf : (cy >= nf ? cy - nf : cy);
        cz = cz < 0 ? cz + nf : (cz >= nf ? cz - nf : cz);
        wc[c] = w[Lf.idx(cx, cy, cz)];
    }
    for (int D = 0; D < 27; D++) {
        const int Dx = D % 3 - 1, Dy = D / 3 % 3 - 1, Dz = D / 9 - 1;
        float S[9];
#pragma unroll
        for (int k = 0; k < 9; k++) S[k] = 0.f;
#pragma unroll
        for (int c = 0; c < 64; c++) {
            // cell offset o from 2V per axis (-2..1): G[c][D] is zero unless the cell also touches the
            // support of V + D, i.e. o >= 0 where D is +1 and o < 0 where D is -1 (a wave-uniform test:
            // 8 to 64 of the 64 cells per direction, 512 of 1728 in all)
            const int ox = (c & 3) - 2, oy = ((c >> 2) & 3) - 2, oz = (c >> 4) - 2;
            if ((Dx > 0 && ox < 0) || (Dx < 0 && ox >= 0) || (Dy > 0 && oy < 0) || (Dy < 0 && oy >= 0) ||
                (Dz > 0 && oz < 0) || (Dz < 0 && oz >= 0))
                continue;
            const float* g = G + (c * 27 + D) * 9;
#pragma unroll
            for (int k = 0; k < 9; k++) S[k] = fmaf(wc[c], g[k], S[k]);
        }
        if (D == 13) {
            if (fabsf(S[0]) < clamp) S[0] = clamp;
            if (fabsf(S[4]) < clamp) S[4] = clamp;
            if (fabsf(S[8]) < clamp) S[8] = clamp;
        }
#pragma unroll
        for (int k = 0; k < 9; k++) StencilIO<SO>::put(stc, Nc, V, D * 9 + k, S[k]);
    }
}

// Level-1 stencil of the 27 coarse vertices around the pinned origin: node pairs touching the
// origin contribute identity when they are the same node and nothing otherwise.
template <typename SO>
__global__ void k_galerkin_fine_pin(Lay Lf, Lay Lc, const float* __restrict__ ke, const float* __restrict__ w,
                                    SO* __restrict__ stc, float clamp) {
    const int nf = Lf.n, nc = Lc.n;
    const size_t Nc = (size_t)nc * nc * nc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;  // (t%3-1, t/3%3-1, t/9-1) around the origin
    if (t >= 27 * 27) return;
    const int vi = t / 27, D = t % 27;
    const int X = (vi % 3) - 1, Y = (vi / 3 % 3) - 1, Z = (vi / 9) - 1;
    const int Dx = D % 3 - 1, Dy = D / 3 % 3 - 1, Dz = D / 9 - 1;
    float S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 64; c++) {
        const int ox = (c & 3) - 2, oy = ((c >> 2) & 3) - 2, oz = (c >> 4) - 2;  // cell offset from 2V
        // only the cells in the supports of both V and V + D contribute (as in k_galerkin_fine)
        if ((Dx > 0 && ox < 0) || (Dx < 0 && ox >= 0) || (Dy > 0 && oy < 0) || (Dy < 0 && oy >= 0) ||
            (Dz > 0 && oz < 0) || (Dz < 0 && oz >= 0))
            continue;
        int cx = 2 * X + ox, cy = 2 * Y + oy, cz = 2 * Z + oz;
        const int cxw = (cx % nf + nf) % nf, cyw = (cy % nf + nf) % nf, czw = (cz % nf + nf) % nf;
        const float we = w[Lf.idx(cxw, cyw, czw)];
        for (int a = 0; a < 8; a++) {
            const int ax = ox + (a & 1), ay = oy + ((a >> 1) & 1), az = oz + (a >> 2);
            const float wa = hat(ax) * hat(ay) * hat(az);
            if (wa == 0.f) continue;
            const bool pa = ((cx + (a & 1)) % nf == 0) && ((cy + ((a >> 1) & 1)) % nf == 0) &&
                            ((cz + (a >> 2)) % nf == 0);
            for (int b = 0; b < 8; b++) {
                const int bx = ox + (b & 1) - 2 * Dx, by = oy + ((b >> 1) & 1) - 2 * Dy, bz = oz + (b >> 2) - 2 * Dz;
                const float wb = hat(bx) * hat(by) * hat(bz);
                if (wb == 0.f) continue;
                const bool pb = ((cx + (b & 1)) % nf == 0) && ((cy + ((b >> 1) & 1)) % nf == 0) &&
                                ((cz + (b >> 2)) % nf == 0);
                if (pa || pb) {
                    if (a == b) {
                        S[0] += wa * wb;
                        S[4] += wa * wb;
                        S[8] += wa * wb;
                    }
                } else {
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) S[i * 3 + j] += wa * wb * we * ke[(3 * a + i) * 24 + 3 * b + j];
                }
            }
        }
    }
    if (D == 13) {
        if (fabsf(S[0]) < clamp) S[0] = clamp;
        if (fabsf(S[4]) < clamp) S[4] = clamp;
        if (fabsf(S[8]) < clamp) S[8] = clamp;
    }
    const uint32_t V = Lc.idx((X + nc) % nc, (Y + nc) % nc, (Z + nc) % nc);
    for (int k = 0; k < 9; k++) StencilIO<SO>::put(stc, Nc, V, D * 9 + k, S[k]);
}

// Coarser stencil from a stored one: st_c[V][D] = sum_{o,d} hat(o) A_f[2V+o][d] hat3(o+d-2D).
// One thread per coarse vertex and plane Dz (blockIdx.y): every fine block (o, d) is read once and
// added into the nine directions (Dx, Dy) of the plane it reaches, with the weights hat3(o + d - 2D).
template <typename SI, typename SO>
__global__ __launch_bounds__(256) void k_galerkin_coarse(Lay Lf, Lay Lc, const SI* __restrict__ stf,
                                                         SO* __restrict__ stc, float clamp) {
    const int nf = Lf.n;
    const size_t Nc = (size_t)Lc.n * Lc.n * Lc.n, Nf = (size_t)nf * nf * nf;
    int X, Y, Z;
    uint32_t V;
    if (!coarse_lex_vertex(Lc, X, Y, Z, V)) return;
    const uint32_t xs[3] = {Lf.part(wrap_dec(2 * X, nf), 0), Lf.part(2 * X, 0), Lf.part(2 * X + 1, 0)};
    const uint32_t ys[3] = {Lf.part(wrap_dec(2 * Y, nf), 1), Lf.part(2 * Y, 1), Lf.part(2 * Y + 1, 1)};
    const uint32_t zs[3] = {Lf.part(wrap_dec(2 * Z, nf), 2), Lf.part(2 * Z, 2), Lf.part(2 * Z + 1, 2)};
    const Rsrc rf = make_rsrc(stf, (size_t)StencilIO<SI>::words * Nf * 4);
    const int Dz = (int)blockIdx.y - 1;
    float S[9][9];  // [(Dy + 1) * 3 + Dx + 1][entry]
#pragma unroll
    for (int j = 0; j < 9; j++)
#pragma unroll
        for (int k = 0; k < 9; k++) S[j][k] = 0.f;
    for (int o = 0; o < 27; o++) {
        const int ox = o % 3 - 1, oy = o / 3 % 3 - 1, oz = o / 9 - 1;
        const float wo = hat(ox) * hat(oy) * hat(oz);
        const uint32_t vf = xs[ox + 1] + ys[oy + 1] + zs[oz + 1];
        for (int d = 0; d < 27; d++) {
            const int kx = ox + d % 3 - 1, ky = oy + d / 3 % 3 - 1, kz = oz + d / 9 - 1;
            const float wz = wo * hat(kz - 2 * Dz);
            if (wz == 0.f) continue;
            float e[9];
#pragma unroll
            for (int k = 0; k < 9; k++) e[k] = StencilIO<SI>::get(rf, vf * 4u, (uint32_t)(Nf * 4), d * 9 + k);
#pragma unroll
            for (int j = 0; j < 9; j++) {
                // o and d are the same for the whole wave, so is wj: about two of the nine coarse
                // directions (Dx, Dy) get a non-zero weight from a fine pair
                const float wj = wz * hat(ky - 2 * (j / 3 - 1)) * hat(kx - 2 * (j % 3 - 1));
                if (wj == 0.f) continue;
#pragma unroll
                for (int k = 0; k < 9; k++) S[j][k] = fmaf(wj, e[k], S[j][k]);
            }
        }
    }
    if (Dz == 0) {
        if (fabsf(S[4][0]) < clamp) S[4][0] = clamp;
        if (fabsf(S[4][4]) < clamp) S[4][4] = clamp;
        if (fabsf(S[4][8]) < clamp) S[4][8] = clamp;
    }
#pragma unroll
    for (int j = 0; j < 9; j++)
#pragma unroll
        for (int k = 0; k < 9; k++) StencilIO<SO>::put(stc, Nc, V, ((Dz + 1) * 9 + j) * 9 + k, S[j][k]);
}

// ---------------------------------------------------------------------------------------------
// Coarsest level: FP64 dense solve u = A^{-1} f with the inverse formed from a Cholesky
// factorisation of the pinned Galerkin system (replaces the host BiCGSTAB of grid.cpp:349-374).
// One block per load case; dof order 3*v + i.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_coarse_solve(int m, size_t Nc, const double* __restrict__ Ainv,
                                                      const float* __restrict__ f, float* __restrict__ u) {
    extern __shared__ double fb[];
    const int r = blockIdx.x;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        const int v = k / 3, i = k % 3;
        fb[k] = (double)f[(size_t)(r * 3 + i) * Nc + v];
    }
    __syncthreads();
    for (int row = threadIdx.x; row < m; row += blockDim.x) {
        double s = 0.0;
        const double* a = Ainv + (size_t)row * m;
        for (int k = 0; k < m; k++) s = fma(a[k], fb[k], s);
        const int v = row / 3, i = row % 3;
        u[(size_t)(r * 3 + i) * Nc + v] = (float)s;
    }
}

// ---------------------------------------------------------------------------------------------
// Coarsest level: its dense FP64 operator factorised and inverted on the GPU (the reference does it
// on the host, grid.cpp:349-424).  m = 3 n_c^3 <= 1029 (n_c = 4..7); A and T are m x m, row-major.
// ---------------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_fill64(size_t n, double v, double* __restrict__ a) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

// A += the coarsest stencil (lexicographic level; entry k = 9d + 3i + j of vertex v couples (v, i) with
// (v + d, j)), one thread per row, so neighbours that coincide on the periodic grid add up.
__global__ __launch_bounds__(256) void k_coarse_assemble(int nc, const float* __restrict__ st, double* __restrict__ A) {
    const int Nc = nc * nc * nc, m = 3 * Nc;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m) return;
    const int v = row / 3, i = row % 3;
    const int x0 = v % nc, y0 = (v / nc) % nc, z0 = v / (nc * nc);
    double* a = A + (size_t)row * m;
    for (int d = 0; d < 27; d++) {
        const int xq = (x0 + d % 3 - 1 + nc) % nc, yq = (y0 + d / 3 % 3 - 1 + nc) % nc, zq = (z0 + d / 9 - 1 + nc) % nc;
        const int q = xq + nc * (yq + nc * zq);
        for (int j = 0; j < 3; j++) a[3 * q + j] += (double)st[(size_t)(d * 9 + i * 3 + j) * Nc + v];
    }
}

// A = (A + A^T) / 2, one thread per pair i < j
__global__ __launch_bounds__(256) void k_symmetrize64(int m, double* __restrict__ A) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)m * m) return;
    const int i = (int)(t / m), j = (int)(t % m);
    if (j <= i) return;
    const double s = 0.5 * (A[(size_t)i * m + j] + A[(size_t)j * m + i]);
    A[(size_t)i * m + j] = s;
    A[(size_t)j * m + i] = s;
}

// In-place inverse of the symmetric positive definite A by Gauss-Jordan elimination on the diagonal,
// one launch per step k over the whole matrix (one thread per element).  With p = A[k][k] (the k-th
// Schur complement's leading entry: positive exactly when A is positive definite, the square of the
// Cholesky pivot):
//   A[k][j] /= p (j != k), A[k][k] = 1 / p, A[i][k] = -A[i][k] / p (i != k),
//   A[i][j] -= A[i][k] A[k][j] / p (i, j != k).
// Step k reads its pivot row and column from prow / pcol (their values before step k) and writes the
// new row and column k + 1 to nrow / ncol for the next step, so no launch reads what it writes.
// status = 0, or 1 + the first k whose pivot is not positive (every block reads the same pivot and
// stops; later steps see the status and do nothing).
__global__ __launch_bounds__(256) void k_gj_pivot0(int m, const double* __restrict__ A, double* __restrict__ prow,
                                                   double* __restrict__ pcol, int* __restrict__ status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *status = 0;
    if (t < m) {
        prow[t] = A[t];
        pcol[t] = A[(size_t)t * m];
    }
}

__global__ __launch_bounds__(256) void k_gj_step64(int m, int k, double* __restrict__ A,
                                                   const double* __restrict__ prow, const double* __restrict__ pcol,
                                                   double* __restrict__ nrow, double* __restrict__ ncol,
                                                   int* __restrict__ status) {
    if (*status != 0) return;
    const double piv = prow[k];
    if (!(piv > 0.0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *status = 1 + k;
        return;
    }
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)m * m) return;
    const int i = (int)(t / m), j = (int)(t - (size_t)i * m);
    const double inv = 1.0 / piv;
    const double rj = (j == k ? 1.0 : prow[j]) * inv;
    const double v = i == k ? rj : (j == k ? -pcol[i] * inv : fma(-pcol[i], rj, A[t]));
    A[t] = v;
    if (i == k + 1) nrow[j] = v;
    if (j == k + 1) ncol[i] = v;
}

// ---------------------------------------------------------------------------------------------
// FP64 reductions and vector updates (per load case)
// ---------------------------------------------------------------------------------------------
// partial[r * gridDim.x + b] = block b's grid-stride share of a_r . b_r, r = blockIdx.y (3N values
// per load case); products are exact in FP64 and summed in FP64.
__global__ __launch_bounds__(256) void k_dot(size_t M, const float* __restrict__ a, const float* __restrict__ b,
                                             double* __restrict__ partial) {
    __shared__ double sh[4];
    const int r = blockIdx.y;
    const float* ar = a + (size_t)r * M;
    const float* br = b + (size_t)r * M;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    double s = 0.0;
    if ((M & 3) == 0) {
        const float4* a4 = reinterpret_cast<const float4*>(ar);
        const float4* b4 = reinterpret_cast<const float4*>(br);
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < M / 4; i += stride) {
            const float4 x = a4[i], y = b4[i];
            s += ((double)x.x * y.x + (double)x.y * y.y) + ((double)x.z * y.z + (double)x.w * y.w);
        }
    } else {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += stride) s += (double)ar[i] * br[i];
    }
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) partial[r * gridDim.x + blockIdx.x] = s;
}

// out[r] = sum of partial[r * nb .. r * nb + nb), one workgroup per row r = blockIdx.x
// out[r] = sum of partial[r * nb + i] over i < nb; one workgroup per row, four independent loads in
// flight per thread.
__global__ __launch_bounds__(kSumThreads) void k_sum_partials(int nb, const double* __restrict__ partial,
                                                              double* __restrict__ out) {
    __shared__ double sh[kSumThreads / 64];
    const double* p = partial + (size_t)blockIdx.x * nb;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * kSumThreads < nb; i += 4 * kSumThreads) {
        s0 += p[i];
        s1 += p[i + kSumThreads];
        s2 += p[i + 2 * kSumThreads];
        s3 += p[i + 3 * kSumThreads];
    }
    for (; i < nb; i += kSumThreads) s0 += p[i];
    const double s = block_sum<kSumThreads>((s0 + s1) + (s2 + s3), sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// PCG scalars per load case.  sc layout: [0..R) rz, [R..2R) pq, [2R..3R) alpha, [3R..4R) rz_new,
// [4R..5R) beta, [5R..6R) active flag.
__global__ void k_pcg_alpha(int R, double* sc) {
    const int r = threadIdx.x;
    if (r >= R) return;
    const double pq = sc[R + r];
    sc[2 * R + r] = (sc[5 * R + r] != 0.0 && pq != 0.0) ? sc[r] / pq : 0.0;
}

__global__ void k_pcg_beta(int R, double* sc) {
    const int r = threadIdx.x;
    if (r >= R) return;
    const double rz = sc[r], rzn = sc[3 * R + r];
    sc[4 * R + r] = (sc[5 * R + r] != 0.0 && rz != 0.0) ? rzn / rz : 0.0;
    sc[r] = rzn;
}

// CG vector updates, one grid row per load case k = blockIdx.y, grid-stride over the M values of a
// load case in float4 steps (M % 4 == 0; the caller checks it).
//
// r -= alpha q, with per-block FP64 partial sums of |r|^2 in partial[k * gridDim.x + block].
__global__ __launch_bounds__(256) void k_pcg_update_r(int R, size_t M, const double* __restrict__ sc,
                                                      const float* __restrict__ q, float* __restrict__ r,
                                                      double* __restrict__ partial) {
    __shared__ double sh[4];
    const int k = blockIdx.y;
    const float a = (float)sc[2 * R + k];
    const float4* q4 = reinterpret_cast<const float4*>(q + (size_t)k * M);
    float4* r4 = reinterpret_cast<float4*>(r + (size_t)k * M);
    double s = 0.0;
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f* qn = reinterpret_cast<const v4f*>(q4);
    v4f* rn = reinterpret_cast<v4f*>(r4);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < M / 4; i += (size_t)gridDim.x * blockDim.x) {
        const v4f qv = __builtin_nontemporal_load(qn + i);
        v4f rv = __builtin_nontemporal_load(rn + i);
        rv.x = fmaf(-a, qv.x, rv.x);
        rv.y = fmaf(-a, qv.y, rv.y);
        rv.z = fmaf(-a, qv.z, rv.z);
        rv.w = fmaf(-a, qv.w, rv.w);
        __builtin_nontemporal_store(rv, rn + i);
        s += ((double)rv.x * rv.x + (double)rv.y * rv.y) + ((double)rv.z * rv.z + (double)rv.w * rv.w);
    }
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) partial[k * gridDim.x + blockIdx.x] = s;
}

// x += alpha p (the step of this iteration, all load cases: alpha is zero for finished ones), then,
// when do_p, p = z + beta p for the active load cases.
__global__ __launch_bounds__(256) void k_pcg_update_px(int R, size_t M, const double* __restrict__ sc,
                                                       const float* __restrict__ z, float* __restrict__ p,
                                                       float* __restrict__ x, int do_p) {
    const int k = blockIdx.y;
    const float a = (float)sc[2 * R + k];
    const float b = (float)sc[4 * R + k];
    const bool upd = do_p && sc[5 * R + k] != 0.0;
    const float4* z4 = reinterpret_cast<const float4*>(z + (size_t)k * M);
    float4* p4 = reinterpret_cast<float4*>(p + (size_t)k * M);
    float4* x4 = reinterpret_cast<float4*>(x + (size_t)k * M);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < M / 4; i += (size_t)gridDim.x * blockDim.x) {
        float4 pv = p4[i];
        if (a != 0.f) {
            float4 xv = x4[i];
            xv.x = fmaf(a, pv.x, xv.x);
            xv.y = fmaf(a, pv.y, xv.y);
            xv.z = fmaf(a, pv.z, xv.z);
            xv.w = fmaf(a, pv.w, xv.w);
            x4[i] = xv;
        }
        if (upd) {
            const float4 zv = z4[i];
            pv.x = fmaf(b, pv.x, zv.x);
            pv.y = fmaf(b, pv.y, zv.y);
            pv.z = fmaf(b, pv.z, zv.z);
            pv.w = fmaf(b, pv.w, zv.w);
            p4[i] = pv;
        }
    }
}

// The same CG steps with the preconditioned vector z in packed FP16 (hw_index words; load-case pair
// pr = blockIdx.y holds cases 2pr (.x) and 2pr + 1 (.y); M = 3 N values per load case, N % 4 == 0).
//
// words q = pr * 3 + c (c = e / N, the entry e = 4 i of the pair's M = 3 N values) of the four vertices
// e % N .. e % N + 3 of a packed FP16 field (hw_index; N % 4 == 0)
__device__ __forceinline__ void hwords4(half2_t (&o)[4], const half2_t* __restrict__ z, int pr, size_t i, size_t N) {
    const size_t e = 4 * i, c = e / N, v0 = e - c * N;
    const int q = pr * 3 + (int)c;
    if (q == 8) {
        const uint4 w = *reinterpret_cast<const uint4*>(z + 8 * N + v0);
        o[0] = __builtin_bit_cast(half2_t, w.x);
        o[1] = __builtin_bit_cast(half2_t, w.y);
        o[2] = __builtin_bit_cast(half2_t, w.z);
        o[3] = __builtin_bit_cast(half2_t, w.w);
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = z[hw_index(q, v0 + k, N)];
    }
}

// partial[(2pr + h) * gridDim.x + block] = block's share of r_{2pr+h} . z_{2pr+h}
__global__ __launch_bounds__(256) void k_dot_fh(size_t M, const float* __restrict__ r, const half2_t* __restrict__ z,
                                                double* __restrict__ partial) {
    __shared__ double sh[4];
    const int pr = blockIdx.y;
    const size_t N = M / 3;
    const float4* r0 = reinterpret_cast<const float4*>(r + (size_t)(2 * pr) * M);
    const float4* r1 = reinterpret_cast<const float4*>(r + (size_t)(2 * pr + 1) * M);
    double s0 = 0.0, s1 = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < M / 4; i += (size_t)gridDim.x * blockDim.x) {
        const float4 a = r0[i], b = r1[i];
        half2_t zq[4];
        hwords4(zq, z, pr, i, N);
        const half2_t h0 = zq[0], h1 = zq[1], h2 = zq[2], h3 = zq[3];
        s0 += ((double)a.x * (double)h0.x + (double)a.y * (double)h1.x) +
              ((double)a.z * (double)h2.x + (double)a.w * (double)h3.x);
        s1 += ((double)b.x * (double)h0.y + (double)b.y * (double)h1.y) +
              ((double)b.z * (double)h2.y + (double)b.w * (double)h3.y);
    }
    s0 = block_sum<256>(s0, sh);
    if (threadIdx.x == 0) partial[(2 * pr) * gridDim.x + blockIdx.x] = s0;
    s1 = block_sum<256>(s1, sh);
    if (threadIdx.x == 0) partial[(2 * pr + 1) * gridDim.x + blockIdx.x] = s1;
}

// k_pcg_update_px with z in packed FP16; set_p: p = z (start of the iteration, x untouched).
__global__ __launch_bounds__(256) void k_pcg_update_px_h(int R, size_t M, const double* __restrict__ sc,
                                                         const half2_t* __restrict__ z, float* __restrict__ p,
                                                         float* __restrict__ x, int do_p, int set_p) {
    const int pr = blockIdx.y;
    float a[2], b[2];
    bool upd[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int k = 2 * pr + h;
        a[h] = set_p ? 0.f : (float)sc[2 * R + k];
        b[h] = set_p ? 0.f : (float)sc[4 * R + k];
        upd[h] = set_p || (do_p && sc[5 * R + k] != 0.0);
    }
    const size_t N = M / 3;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < M / 4; i += (size_t)gridDim.x * blockDim.x) {
        half2_t zh[4];
        hwords4(zh, z, pr, i, N);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            float4* p4 = reinterpret_cast<float4*>(p + (size_t)(2 * pr + h) * M);
            float4* x4 = reinterpret_cast<float4*>(x + (size_t)(2 * pr + h) * M);
            float4 pv = set_p ? make_float4(0.f, 0.f, 0.f, 0.f) : p4[i];
            if (a[h] != 0.f) {
                float4 xv = x4[i];
                xv.x = fmaf(a[h], pv.x, xv.x);
                xv.y = fmaf(a[h], pv.y, xv.y);
                xv.z = fmaf(a[h], pv.z, xv.z);
                xv.w = fmaf(a[h], pv.w, xv.w);
                x4[i] = xv;
            }
            if (upd[h]) {
                pv.x = fmaf(b[h], pv.x, (float)zh[0][h]);
                pv.y = fmaf(b[h], pv.y, (float)zh[1][h]);
                pv.z = fmaf(b[h], pv.z, (float)zh[2][h]);
                pv.w = fmaf(b[h], pv.w, (float)zh[3][h]);
                p4[i] = pv;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// C^H and sensitivity (elasticMatrix_kernel_opt homogenization.cu:208-346;
// Sensitivity_kernel_opt_2 homogenization.cu:697-858).  One thread per cell; the six solution
// fields of the cell's 8 nodes are read in FP32 and the cell energies accumulated in FP32.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float macro_disp(int node, int s, int c) {
    const float px = (float)(node & 1), py = (float)((node >> 1) & 1), pz = (float)(node >> 2);
    switch (s) {
        case 0: return c == 0 ? px : 0.f;
        case 1: return c == 1 ? py : 0.f;
        case 2: return c == 2 ? pz : 0.f;
        case 3: return c == 1 ? 0.5f * pz : (c == 2 ? 0.5f * py : 0.f);
        case 4: return c == 0 ? 0.5f * pz : (c == 2 ? 0.5f * px : 0.f);
        default: return c == 0 ? 0.5f * py : (c == 1 ? 0.5f * px : 0.f);
    }
}

__device__ __forceinline__ void cell_energies(const Lay& L, size_t N, uint32_t e, const float* __restrict__ ke,
                                              const float* __restrict__ chi, float E[21], int& x, int& y, int& z) {
    L.pos(e, x, y, z);
    const uint32_t xs[2] = {L.part(x, 0), L.part(wrap_inc(x, L.n), 0)};
    const uint32_t ys[2] = {L.part(y, 1), L.part(wrap_inc(y, L.n), 1)};
    const uint32_t zs[2] = {L.part(z, 2), L.part(wrap_inc(z, L.n), 2)};
    float U[6][24];
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const uint32_t q = xs[b & 1] + ys[(b >> 1) & 1] + zs[b >> 2];
#pragma unroll
        for (int s = 0; s < 6; s++)
#pragma unroll
            for (int c = 0; c < 3; c++) U[s][3 * b + c] = macro_disp(b, s, c) - chi[(size_t)(s * 3 + c) * N + q];
    }
#pragma unroll
    for (int k = 0; k < 21; k++) E[k] = 0.f;
#pragma unroll
    for (int i = 0; i < 24; i++) {
        float KU[6];
#pragma unroll
        for (int s = 0; s < 6; s++) KU[s] = 0.f;
#pragma unroll
        for (int j = 0; j < 24; j++) {
            const float kij = ke[i * 24 + j];
#pragma unroll
            for (int s = 0; s < 6; s++) KU[s] = fmaf(kij, U[s][j], KU[s]);
        }
        int k = 0;
#pragma unroll
        for (int a = 0; a < 6; a++)
#pragma unroll
            for (int b = a; b < 6; b++) {
                E[k] = fmaf(U[a][i], KU[b], E[k]);
                k++;
            }
    }
}

// partial[k * nblocks + block] = sum of w_e E_k,e over the block's cells (FP64)
__global__ __launch_bounds__(kFullThreads) void k_elastic_matrix(Lay L, size_t N, uint32_t nchunks,
                                                                 const float* __restrict__ ke,
                                                                 const float* __restrict__ w,
                                                                 const float* __restrict__ chi,
                                                                 double* __restrict__ partial) {
    __shared__ double sh[kFullThreads / 64];
    double acc[21];
#pragma unroll
    for (int k = 0; k < 21; k++) acc[k] = 0.0;
    for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        uint32_t e;
        if (!full_vertex(L, chunk, e)) continue;
        float E[21];
        int x, y, z;
        cell_energies(L, N, e, ke, chi, E, x, y, z);
        const double we = w[e];
#pragma unroll
        for (int k = 0; k < 21; k++) acc[k] += we * (double)E[k];
    }
    for (int k = 0; k < 21; k++) {
        const double s = block_sum<kFullThreads>(acc[k], sh);
        if (threadIdx.x == 0) partial[k * gridDim.x + blockIdx.x] = s;
    }
}

// sens[x + (y + z*n)*pitch] = (1/N) sum_{i<=j} E_ij (dC_ij + [i!=j] dC_ji)
__global__ __launch_bounds__(kFullThreads) void k_sensitivity(Lay L, size_t N, int pitch,
                                                              const float* __restrict__ ke,
                                                              const float* __restrict__ chi,
                                                              const float* __restrict__ wdc,
                                                              float* __restrict__ sens) {
    uint32_t e;
    if (!full_vertex(L, xcd_chunk(blockIdx.x, gridDim.x), e)) return;
    float E[21];
    int x, y, z;
    cell_energies(L, N, e, ke, chi, E, x, y, z);
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 21; k++) s = fmaf(E[k], wdc[k], s);
    // pitch 0: the level's own (parity-blocked) cell order, unpitched
    sens[pitch ? x + ((size_t)y + (size_t)z * L.n) * pitch : (size_t)e] = s / (float)N;
}

// ---------------------------------------------------------------------------------------------
// Density utilities
// ---------------------------------------------------------------------------------------------

// Symmetry projection (symetrize_tensor_kernel, TensorExpression.cuh:72-262): each cell receives
// the FP64 average over the group applied to its orbit's smallest-index member, so every member of
// an orbit receives the same bits.  gperm/gflip list the group elements; n cubic.
__global__ __launch_bounds__(256) void k_symmetrize(int n, int pitch, int ng, const int* __restrict__ gperm,
                                                    const int* __restrict__ gflip, const float* __restrict__ in,
                                                    float* __restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)n * n * n;
    if (e >= N) return;
    const int p[3] = {(int)(e % n), (int)((e / n) % n), (int)(e / ((long)n * n))};
    long best = e;
    int bq[3] = {p[0], p[1], p[2]};
    for (int g = 0; g < ng; g++) {
        int q[3];
        for (int k = 0; k < 3; k++) {
            const int a = p[(gperm[g] >> (2 * k)) & 3];
            q[k] = ((gflip[g] >> k) & 1) ? n - 1 - a : a;
        }
        const long id = q[0] + (long)n * (q[1] + (long)n * q[2]);
        if (id < best) {
            best = id;
            bq[0] = q[0]; bq[1] = q[1]; bq[2] = q[2];
        }
    }
    double s = 0.0;
    for (int g = 0; g < ng; g++) {
        int q[3];
        for (int k = 0; k < 3; k++) {
            const int a = bq[(gperm[g] >> (2 * k)) & 3];
            q[k] = ((gflip[g] >> k) & 1) ? n - 1 - a : a;
        }
        s += (double)in[q[0] + ((size_t)q[1] + (size_t)q[2] * n) * pitch];
    }
    out[p[0] + ((size_t)p[1] + (size_t)p[2] * n) * pitch] = (float)(s / (double)ng);
}

// Sigmund filter (filterSens_kernel, OCOptimizer.cu:76-153): wk holds the normalised weights of the
// (2h+1)^3 box.
__global__ __launch_bounds__(256) void k_oc_filter(int n, int pitch, int h, int periodic, const float* __restrict__ wk,
                                                   const float* __restrict__ sens, const float* __restrict__ rho,
                                                   float* __restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)n * n * n;
    if (e >= N) return;
    const int p[3] = {(int)(e % n), (int)((e / n) % n), (int)(e / ((long)n * n))};
    const int side = 2 * h + 1;
    double sum = 0.0, ws = 0.0;
    for (int i = 0; i < side * side * side; i++) {
        int q[3] = {p[0] + i % side - h, p[1] + (i / side) % side - h, p[2] + i / (side * side) - h};
        bool inb = true;
        for (int k = 0; k < 3; k++) {
            if (periodic) q[k] = ((q[k] % n) + n) % n;
            if (q[k] < 0 || q[k] >= n) inb = false;
        }
        if (!inb) continue;
        const size_t id = q[0] + ((size_t)q[1] + (size_t)q[2] * n) * pitch;
        sum += (double)sens[id] * (double)rho[id] * (double)wk[i];
        ws += (double)wk[i];
    }
    const size_t me = p[0] + ((size_t)p[1] + (size_t)p[2] * n) * pitch;
    out[me] = (float)(sum / (ws * (double)rho[me]));
}

// Density filter conv with radial_convker_t (TensorExpression.cuh:288-368): out(i) = sum_off w(off) in(i + off)
// / W(i), W(i) = sum of w over the neighbours inside the grid (every neighbour when periodic); wk holds the
// normalised weights of the (2h+1)^3 box.  FP64 sums.  W is written when wsum is not null.
__global__ __launch_bounds__(256) void k_conv_filter(int n, int pitch, int h, int periodic, const float* __restrict__ wk,
                                                     const float* __restrict__ in, float* __restrict__ out,
                                                     float* __restrict__ wsum) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)n * n * n;
    if (e >= N) return;
    const int p[3] = {(int)(e % n), (int)((e / n) % n), (int)(e / ((long)n * n))};
    const int side = 2 * h + 1;
    double sum = 0.0, ws = 0.0;
    for (int i = 0; i < side * side * side; i++) {
        int q[3] = {p[0] + i % side - h, p[1] + (i / side) % side - h, p[2] + i / (side * side) - h};
        bool inb = true;
        for (int k = 0; k < 3; k++) {
            if (periodic) q[k] = ((q[k] % n) + n) % n;
            if (q[k] < 0 || q[k] >= n) inb = false;
        }
        if (!inb) continue;
        sum += (double)wk[i] * (double)in[q[0] + ((size_t)q[1] + (size_t)q[2] * n) * pitch];
        ws += (double)wk[i];
    }
    const size_t me = p[0] + ((size_t)p[1] + (size_t)p[2] * n) * pitch;
    out[me] = (float)(sum / ws);
    if (wsum) wsum[me] = (float)ws;
}

// Its adjoint (conv_backward_kernel, TensorExpression.cuh:331-366): din(j) = sum_off w(off) dout(j - off) / W(j - off)
// over the in-grid (or wrapped) j - off; wsum from k_conv_filter.
__global__ __launch_bounds__(256) void k_conv_filter_back(int n, int pitch, int h, int periodic,
                                                          const float* __restrict__ wk, const float* __restrict__ dout,
                                                          const float* __restrict__ wsum, float* __restrict__ din) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)n * n * n;
    if (e >= N) return;
    const int p[3] = {(int)(e % n), (int)((e / n) % n), (int)(e / ((long)n * n))};
    const int side = 2 * h + 1;
    double sum = 0.0;
    for (int i = 0; i < side * side * side; i++) {
        int q[3] = {p[0] - (i % side - h), p[1] - ((i / side) % side - h), p[2] - (i / (side * side) - h)};
        bool inb = true;
        for (int k = 0; k < 3; k++) {
            if (periodic) q[k] = ((q[k] % n) + n) % n;
            if (q[k] < 0 || q[k] >= n) inb = false;
        }
        if (!inb) continue;
        const size_t id = q[0] + ((size_t)q[1] + (size_t)q[2] * n) * pitch;
        sum += (double)wk[i] * (double)dout[id] / (double)wsum[id];
    }
    din[p[0] + ((size_t)p[1] + (size_t)p[2] * n) * pitch] = (float)sum;
}

// sgm projection (sgm_exp_t, AutoDiff.h:808-823): y = 1 / (1 + exp(-k (x - c))); backward gx = gy k y (1 - y).
__global__ __launch_bounds__(256) void k_sgm(int n, int pitch, double k, double c, const float* __restrict__ x,
                                             float* __restrict__ y) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)n * n * n;
    if (e >= N) return;
    const size_t id = (e % n) + (size_t)(e / n) * pitch;
    y[id] = (float)(1.0 / (1.0 + exp(-k * ((double)x[id] - c))));
}

__global__ __launch_bounds__(256) void k_sgm_back(int n, int pitch, double k, const float* __restrict__ y,
                                                  const float* __restrict__ gy, float* __restrict__ gx) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)n * n * n;
    if (e >= N) return;
    const size_t id = (e % n) + (size_t)(e / n) * pitch;
    const double yv = y[id];
    gx[id] = (float)((double)gy[id] * k * yv * (1.0 - yv));
}

// OC update of one cell for multiplier g (update_kernel, OCOptimizer.cu:12-34), FP64 arithmetic rounded
// to FP32.
__device__ __forceinline__ float oc_cell(double r0, double sv, double g, double min_rho, double step, double damp) {
    double B = -sv / g;
    if (B < 0) B = 0.01;
    double xv = (damp == 0.5 ? sqrt(B) : pow(B, damp)) * r0;
    if (xv - r0 < -step) xv = r0 - step;
    if (xv - r0 > step) xv = r0 + step;
    if (xv < min_rho) xv = min_rho;
    if (xv > 1.0) xv = 1.0;
    return (float)xv;
}

// OC trial for multiplier g: per-block FP64 partial sums of the new density; the new density is written
// to out unless out is null (sum only).
__global__ __launch_bounds__(256) void k_oc_trial(int n, int pitch, double g, double min_rho, double step,
                                                  double damp, const float* __restrict__ sens,
                                                  const float* __restrict__ rho, float* __restrict__ out,
                                                  double* __restrict__ partial) {
    __shared__ double sh[4];
    const long N = (long)n * n * n;
    double s = 0.0;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const int x = (int)(e % n), y = (int)((e / n) % n), z = (int)(e / ((long)n * n));
        const size_t id = x + ((size_t)y + (size_t)z * n) * pitch;
        const float o = oc_cell(rho[id], sens[id], g, min_rho, step, damp);
        if (out) out[id] = o;
        s += (double)o;
    }
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// The OC update for the final multiplier, in place (rho is read and written by the same thread).
__global__ __launch_bounds__(256) void k_oc_apply(int n, int pitch, double g, double min_rho, double step, double damp,
                                                  const float* __restrict__ sens, float* rho) {
    const long N = (long)n * n * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const int x = (int)(e % n), y = (int)((e / n) % n), z = (int)(e / ((long)n * n));
        const size_t id = x + ((size_t)y + (size_t)z * n) * pitch;
        rho[id] = oc_cell(rho[id], sens[id], g, min_rho, step, damp);
    }
}

// Elementwise density maps of the optimisation loop (rho.pow(p) of the reference's expression templates
// and its chain rule), over the n^3 live cells of pitched lexicographic arrays.  p = 3 is evaluated as
// r * r * r and 3 r r d, the products the host loops of the drivers form.
__device__ __forceinline__ size_t pitched(long e, int n, int pitch) {
    const int x = (int)(e % n), y = (int)((e / n) % n), z = (int)(e / ((long)n * n));
    return x + ((size_t)y + (size_t)z * n) * pitch;
}
__global__ __launch_bounds__(256) void k_density_pow(int n, int pitch, float p, const float* __restrict__ rho,
                                                     float* __restrict__ x) {
    const long N = (long)n * n * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const size_t id = pitched(e, n, pitch);
        const float r = rho[id];
        x[id] = p == 3.f ? r * r * r : powf(r, p);
    }
}
__global__ __launch_bounds__(256) void k_density_pow_grad(int n, int pitch, float p, const float* __restrict__ rho,
                                                          const float* dx, float* drho) {
    const long N = (long)n * n * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const size_t id = pitched(e, n, pitch);
        const float r = rho[id];
        drho[id] = p == 3.f ? 3.f * r * r * dx[id] : p * powf(r, p - 1.f) * dx[id];
    }
}
__global__ __launch_bounds__(256) void k_field_sum(int n, int pitch, const float* __restrict__ a,
                                                   double* __restrict__ partial) {
    __shared__ double sh[4];
    const long N = (long)n * n * n;
    double s = 0.0;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x)
        s += (double)a[pitched(e, n, pitch)];
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_field_fill(int n, int pitch, float value, float* __restrict__ a) {
    const long N = (long)n * n * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x)
        a[pitched(e, n, pitch)] = value;
}

__global__ __launch_bounds__(256) void k_absmax(long N, int n, int pitch, const float* __restrict__ a,
                                                float* __restrict__ partial) {
    __shared__ float sh[256];
    float m = 0.f;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const int x = (int)(e % n), y = (int)((e / n) % n), z = (int)(e / ((long)n * n));
        m = fmaxf(m, fabsf(a[x + ((size_t)y + (size_t)z * n) * pitch]));
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// pitched lexicographic density -> dense cell weights w = emin + x
// pitched lexicographic density -> cell weights w = emin + x and raw x in level storage order
__global__ __launch_bounds__(256) void k_set_weights(Lay L, int pitch, double emin, const float* __restrict__ x,
                                                     float* __restrict__ w, float* __restrict__ xc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = (uint32_t)L.n;
    if (i >= n * n * n) return;
    const int xx = (int)(i % n), y = (int)((i / n) % n), z = (int)(i / (n * n));
    const uint32_t e = L.idx(xx, y, z);
    const float xv = x[xx + ((size_t)y + (size_t)z * L.n) * pitch];
    xc[e] = xv;
    w[e] = (float)(emin + (double)xv);
}

__global__ __launch_bounds__(256) void k_copy_pitched(int n, int pitch_in, int pitch_out, const float* __restrict__ in,
                                                      float* __restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)n * n * n;
    if (e >= N) return;
    const int x = (int)(e % n), y = (int)((e / n) % n), z = (int)(e / ((long)n * n));
    out[x + ((size_t)y + (size_t)z * n) * pitch_out] = in[x + ((size_t)y + (size_t)z * n) * pitch_in];
}

inline unsigned nblk(size_t work, int bs = 256) { return (unsigned)((work + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------------------------
static Lay make_lay(int n, int lex) {
    Lay L;
    L.n = n;
    L.h = n / 2;
    L.h3 = (uint32_t)L.h * L.h * L.h;
    L.lex = lex;
    return L;
}

// workgroups of a full-level kernel (see full_vertex)
static unsigned full_grid(const Lay& L) {
    const size_t N = (size_t)L.n * L.n * L.n;
    return L.lex ? nblk(N, kFullThreads) : nblk(L.h3, 64);
}

int ke_mfma_table(const double* kl, const double* km, double lam, double mu, void* table, float* c0, float* c1) {
    // One exact table T = r Kl + Km (KE = mu/72 T) when r = lambda / mu is a short binary fraction
    // (nu = 0.3 gives r = 3/2) and every entry of T is an FP16 value; else Kl and Km separately.
    std::vector<double> T(576);
    const double r = std::nearbyint(lam / mu * 256.0) / 256.0;
    bool one = std::fabs(r - lam / mu) <= 1e-12 * std::fabs(lam / mu);
    for (int e = 0; one && e < 576; e++) {
        T[e] = r * kl[e] + km[e];
        one = (double)(_Float16)T[e] == T[e];
    }
    // lane l of tile t holds row 16t + l % 16, K entries 8(l / 16) .. + 7.  Row 16t + 4g + j is slot
    // 4t + j of lane group g: one table: slots 0..5 its rows k = 0..5 of the group; two tables: slots
    // 0..5 the Kl rows, 6..11 the Km rows.  Row k of group g is component k % 3 of node
    // (k / 3, g & 1, g >> 1).  K entry 8g' + e (e < 6) is component e % 3 of node (e / 3, g' & 1, g' >> 1);
    // entries 6 and 7 are zero.
    _Float16* A = static_cast<_Float16*>(table);
    for (int t = 0; t < 3; t++)
        for (int l = 0; l < 64; l++) {
            const int m = l % 16, gp = l / 16, g = m / 4, slot = 4 * t + m % 4;
            const double* src = one ? (slot < 6 ? T.data() : nullptr) : (slot < 6 ? kl : km);
            const int k = slot % 6, o = k / 3 + 2 * (g & 1) + 4 * (g >> 1), i = k % 3;
            for (int e = 0; e < 8; e++) {
                const int op = e / 3 + 2 * (gp & 1) + 4 * (gp >> 1), ip = e % 3;
                A[(t * 64 + l) * 8 + e] = (_Float16)(src && e < 6 ? src[(3 * o + i) * 24 + 3 * op + ip] : 0.0);
            }
        }
    *c0 = (float)((one ? mu : lam) / 72.0);
    *c1 = one ? 0.f : (float)(mu / 72.0);
    return one ? 1 : 2;
}

int fine_apply_mfma_ok(int n) { return n % 32 == 0 && 72.0 * n * n * n < 4294967296.0; }

static unsigned ef_grid(int n) {
    const unsigned h = (unsigned)n / 2;
    return (h / kEfTX) * (h / kEfTY) * (h / kEfTZ);
}

template <int MODE, bool DOT, bool BETA, int PX = 0>
static void fine_apply_ef(hipStream_t s, int n, const FineMfma* mf, const float* w, const void* u, const float* f,
                          float* out, double* partial, const float* dotv, const double* beta, float* pp = nullptr,
                          float* xx = nullptr, const double* alpha = nullptr) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    const half8_t* am = static_cast<const half8_t*>(mf->table);
    const uint32_t* uw = static_cast<const uint32_t*>(u);
    if (mf->nt == 1)
        hipLaunchKernelGGL((k_fine_apply_ef<MODE, DOT, BETA, 1, PX>), dim3(ef_grid(n)), dim3(kEfThreads), 0, s, L, N,
                           am, mf->c0, mf->c1, w, uw, f, out, partial, dotv, beta, pp, xx, alpha);
    else
        hipLaunchKernelGGL((k_fine_apply_ef<MODE, DOT, BETA, 2, PX>), dim3(ef_grid(n)), dim3(kEfThreads), 0, s, L, N,
                           am, mf->c0, mf->c1, w, uw, f, out, partial, dotv, beta, pp, xx, alpha);
}

void launch_fine_apply(hipStream_t s, int n, const float* ke, const FineMfma* mf, const float* w, const void* u,
                       int u_half, const float* f, float* out) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    const dim3 g(full_grid(L)), b(kFullThreads);
    if (u_half && mf && fine_apply_mfma_ok(n)) {
        if (f)
            fine_apply_ef<1, false, false>(s, n, mf, w, u, f, out, nullptr, nullptr, nullptr);
        else
            fine_apply_ef<0, false, false>(s, n, mf, w, u, f, out, nullptr, nullptr, nullptr);
        return;
    }
    if (u_half) {
        const half2_t* uh = static_cast<const half2_t*>(u);
        if (f)
            hipLaunchKernelGGL((k_fine_apply<6, 1, false, half2_t>), g, b, 0, s, L, N, ke, w, uh, f, out, nullptr,
                               nullptr, nullptr, nullptr, nullptr, nullptr);
        else
            hipLaunchKernelGGL((k_fine_apply<6, 0, false, half2_t>), g, b, 0, s, L, N, ke, w, uh, f, out, nullptr,
                               nullptr, nullptr, nullptr, nullptr, nullptr);
        return;
    }
    const float* uf = static_cast<const float*>(u);
    if (f)
        hipLaunchKernelGGL((k_fine_apply<6, 1, false, float>), g, b, 0, s, L, N, ke, w, uf, f, out, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_fine_apply<6, 0, false, float>), g, b, 0, s, L, N, ke, w, uf, f, out, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr);
}

// neighbours whose loads are in flight ahead of the CG pass's block products (FP32 z: one, as two
// spill there)
template <typename A>
constexpr int kCgPrefetch = CasesPerWord<A>::value == 2 ? 2 : 1;
template <typename A, bool FIRST>
static void fine_apply_cg_row(hipStream_t s, const Lay& L, size_t N, const float* ke, const float* w, const A* z,
                              float* p, float* x, const double* sc, float* q, double* partial) {
    constexpr int R = 6;
    hipLaunchKernelGGL((k_fine_apply<R, 0, true, A, !FIRST, FIRST ? 1 : 2, kCgPrefetch<A>>), dim3(full_grid(L)),
                       dim3(kFullThreads), 0, s, L, N, ke, w, z, nullptr, q, partial, nullptr, sc + 4 * R, p, x,
                       sc + 2 * R);
}

int fine_cg_staged_ok(int n) { return n % 128 == 0; }

void launch_fine_apply_cg(hipStream_t s, int n, const float* ke, const FineMfma* mf, const float* w, const void* z,
                          int z_half, int staged, float* p, float* x, const double* sc, int first, float* q,
                          double* partial, int dot_blocks, double* out6) {
    constexpr int R = 6;
    const size_t N = (size_t)n * n * n, M = 3 * N;
    const Lay L = make_lay(n, 0);
    unsigned nb = full_grid(L);
    if (z_half && mf && fine_apply_mfma_ok(n)) {
        if (first)
            fine_apply_ef<0, true, false, 1>(s, n, mf, w, z, nullptr, q, partial, nullptr, nullptr, p, x, nullptr);
        else
            fine_apply_ef<0, true, true, 2>(s, n, mf, w, z, nullptr, q, partial, nullptr, sc + 4 * R, p, x, sc + 2 * R);
        nb = ef_grid(n);
    } else if (L.h3 % 64) {
        // partial waves (FP32 z only: packed FP16 needs n % 8 == 0): the same step from separate passes
        if (first) {
            (void)launch_set_p(s, M, z, 0, p, dot_blocks);
        } else {
            launch_pcg_update_px(s, M, sc, z, 0, p, x, 1, dot_blocks);
        }
        launch_fine_apply(s, n, ke, nullptr, w, p, 0, nullptr, q);
        launch_dot6(s, M, p, q, partial, dot_blocks, out6);
        return;
    } else if (z_half && staged && fine_cg_staged_ok(n)) {
        const uint32_t* zw = static_cast<const uint32_t*>(z);
        if (first)
            hipLaunchKernelGGL((k_fine_cg_staged<false, 1>), dim3(nb), dim3(kFullThreads), 0, s, L, N, ke, w, zw, q,
                               partial, nullptr, p, x, nullptr);
        else
            hipLaunchKernelGGL((k_fine_cg_staged<true, 2>), dim3(nb), dim3(kFullThreads), 0, s, L, N, ke, w, zw, q,
                               partial, sc + 4 * R, p, x, sc + 2 * R);
    } else if (z_half) {
        const half2_t* zh = static_cast<const half2_t*>(z);
        if (first)
            fine_apply_cg_row<half2_t, true>(s, L, N, ke, w, zh, p, x, sc, q, partial);
        else
            fine_apply_cg_row<half2_t, false>(s, L, N, ke, w, zh, p, x, sc, q, partial);
    } else {
        const float* zf = static_cast<const float*>(z);
        if (first)
            fine_apply_cg_row<float, true>(s, L, N, ke, w, zf, p, x, sc, q, partial);
        else
            fine_apply_cg_row<float, false>(s, L, N, ke, w, zf, p, x, sc, q, partial);
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(6), dim3(kSumThreads), 0, s, (int)nb, partial, out6);
}

template <bool REV, int ZC, typename A>
static void fine_gs_colour(hipStream_t s, const Lay& L, size_t N, int colour, float omega, const float* ke,
                           const float* w, A* u, const float* f) {
    hipLaunchKernelGGL((k_fine_gs<6, REV, ZC, A>), dim3(nblk(L.h3)), dim3(256), 0, s, L, N, colour, omega, ke, w, u,
                       f);
}

template <typename A>
static void fine_gs_pick(hipStream_t s, const Lay& L, size_t N, int colour, float omega, int reverse, int zero_start,
                         const float* ke, const float* w, A* u, const float* f) {
    if (reverse) {
        fine_gs_colour<true, -1>(s, L, N, colour, omega, ke, w, u, f);
        return;
    }
    switch (zero_start ? colour : -1) {
        case 0: fine_gs_colour<false, 0>(s, L, N, colour, omega, ke, w, u, f); break;
        case 1: fine_gs_colour<false, 1>(s, L, N, colour, omega, ke, w, u, f); break;
        case 2: fine_gs_colour<false, 2>(s, L, N, colour, omega, ke, w, u, f); break;
        case 3: fine_gs_colour<false, 3>(s, L, N, colour, omega, ke, w, u, f); break;
        case 4: fine_gs_colour<false, 4>(s, L, N, colour, omega, ke, w, u, f); break;
        case 5: fine_gs_colour<false, 5>(s, L, N, colour, omega, ke, w, u, f); break;
        case 6: fine_gs_colour<false, 6>(s, L, N, colour, omega, ke, w, u, f); break;
        case 7: fine_gs_colour<false, 7>(s, L, N, colour, omega, ke, w, u, f); break;
        default: fine_gs_colour<false, -1>(s, L, N, colour, omega, ke, w, u, f); break;
    }
}

void launch_fine_gs(hipStream_t s, int n, int colour, float omega, int reverse, int zero_start, const float* ke,
                    const float* w, void* u, int u_half, const float* f) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    if (u_half)
        fine_gs_pick(s, L, N, colour, omega, reverse, zero_start, ke, w, static_cast<half2_t*>(u), f);
    else
        fine_gs_pick(s, L, N, colour, omega, reverse, zero_start, ke, w, static_cast<float*>(u), f);
}

// fused colour pairs need every workgroup to own whole rows of one parity block
int fine_gs_pairs_ok(int n) {
    const int h = n / 2;
    const long h3 = (long)h * h * h;
    return n % 2 == 0 && h <= 256 && 256 % h == 0 && h3 % 256 == 0;
}

template <bool REV, int ZC, typename A, bool DOT = false, bool ROWN = false>
static void fine_gs2_pair(hipStream_t s, const Lay& L, size_t N, int cfirst, float omega, const float* ke,
                          const float* w, A* u, const float* f, double* partial = nullptr, int pstride = 0,
                          float* rown = nullptr) {
    hipLaunchKernelGGL((k_fine_gs2<6, REV, ZC, A, DOT, ROWN>), dim3(L.h3 / 256), dim3(256), 0, s, L, N, cfirst, omega,
                       ke, w, u, f, partial, pstride, rown);
}

template <int ZC, typename A>
static void fine_gs2_zero(hipStream_t s, const Lay& L, size_t N, float omega, const float* ke, const float* w, A* u,
                          const float* f, float* rown) {
    if (rown)
        fine_gs2_pair<false, ZC, A, false, true>(s, L, N, ZC, omega, ke, w, u, f, nullptr, 0, rown);
    else
        fine_gs2_pair<false, ZC>(s, L, N, ZC, omega, ke, w, u, f);
}

template <typename A>
static void fine_gs2_pick(hipStream_t s, const Lay& L, size_t N, int cfirst, float omega, int reverse,
                          int zero_start, const float* ke, const float* w, A* u, const float* f, float* rown) {
    if (reverse) {
        fine_gs2_pair<true, -1>(s, L, N, cfirst, omega, ke, w, u, f);
        return;
    }
    switch (zero_start ? cfirst : -1) {
        case 0: fine_gs2_zero<0>(s, L, N, omega, ke, w, u, f, rown); break;
        case 2: fine_gs2_zero<2>(s, L, N, omega, ke, w, u, f, rown); break;
        case 4: fine_gs2_zero<4>(s, L, N, omega, ke, w, u, f, rown); break;
        case 6: fine_gs2_zero<6>(s, L, N, omega, ke, w, u, f, rown); break;
        default: fine_gs2_pair<false, -1>(s, L, N, cfirst, omega, ke, w, u, f); break;
    }
}

void launch_fine_gs2(hipStream_t s, int n, int cfirst, float omega, int reverse, int zero_start, const float* ke,
                     const float* w, void* u, int u_half, const float* f, double* partial, int pstride, float* rown) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    if (reverse || !zero_start) rown = nullptr;  // own-colour residuals only for zero-start passes
    if (partial && reverse && u_half) {  // f . u_new per workgroup (the r . z of a CG step)
        fine_gs2_pair<true, -1, half2_t, true>(s, L, N, cfirst, omega, ke, w, static_cast<half2_t*>(u), f, partial,
                                               pstride);
        return;
    }
    if (u_half)
        fine_gs2_pick(s, L, N, cfirst, omega, reverse, zero_start, ke, w, static_cast<half2_t*>(u), f, rown);
    else
        fine_gs2_pick(s, L, N, cfirst, omega, reverse, zero_start, ke, w, static_cast<float*>(u), f, rown);
}

void launch_fine_residual_later(hipStream_t s, int n, const float* ke, const float* w, const void* u, int u_half,
                                float* r) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    const dim3 g(nblk(L.h3, 64)), b(256);
    if (u_half)
        hipLaunchKernelGGL((k_fine_resid_later<6, half2_t>), g, b, 0, s, L, N, ke, w, static_cast<const half2_t*>(u),
                           r, 0.f);
    else
        hipLaunchKernelGGL((k_fine_resid_later<6, float>), g, b, 0, s, L, N, ke, w, static_cast<const float*>(u), r,
                           0.f);
}

void launch_macro_rhs(hipStream_t s, int n, const float* kchi, const float* w, float* f) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    hipLaunchKernelGGL(k_macro_rhs, dim3(full_grid(L)), dim3(kFullThreads), 0, s, L, N, kchi, w, f);
}

// One load case per FP64 residual launch: with two, the FP64 gather pipeline no longer fits the
// register file.
void launch_fine_residual64(hipStream_t s, int n, const double* ke, const double* kchi, const float* xc, double emin,
                            const double* u, float* r32, double* partial, double* out6, int do_k) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    const dim3 g(full_grid(L)), b(kFullThreads);
    if (!do_k)
        hipLaunchKernelGGL((k_fine_residual64<false>), g, b, 0, s, L, N, ke, kchi, xc, emin, u, r32, partial);
    else
        hipLaunchKernelGGL((k_fine_residual64<true>), g, b, 0, s, L, N, ke, kchi, xc, emin, u, r32, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(6), dim3(kSumThreads), 0, s, (int)full_grid(L), partial, out6);
}

void launch_rhs_norm64(hipStream_t s, int n, const double* kchi, const float* xc, double emin, double* partial,
                       double* out6) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    const unsigned nb = full_grid(L);
    hipLaunchKernelGGL(k_rhs_norm64, dim3(nb), dim3(kFullThreads), 0, s, L, N, kchi, xc, emin, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(6), dim3(kSumThreads), 0, s, (int)nb, partial, out6);
}

void launch_refine_accumulate(hipStream_t s, size_t N, const double* sc, int pending, const float* x, const float* p,
                              double* u, int add) {
    hipLaunchKernelGGL(k_refine_accumulate, dim3(2048, 9), dim3(256), 0, s, N, sc, pending, x, p, u, add);
}

void launch_round_f32(hipStream_t s, size_t N, const double* u, float* x) {
    hipLaunchKernelGGL(k_round_f32, dim3(nblk(9 * N)), dim3(256), 0, s, N, u, x);
}

void launch_widen_f64(hipStream_t s, size_t N, const float* x, double* u) {
    hipLaunchKernelGGL(k_widen_f64, dim3(nblk(9 * N)), dim3(256), 0, s, N, x, u);
}

void launch_pack_half(hipStream_t s, size_t N, const float* u, void* z) {
    hipLaunchKernelGGL(k_pack_half, dim3(nblk(9 * N)), dim3(256), 0, s, N, u, static_cast<half2_t*>(z));
}

void launch_unpack_half(hipStream_t s, size_t N, const void* z, float* u) {
    hipLaunchKernelGGL(k_unpack_half, dim3(nblk(9 * N)), dim3(256), 0, s, N, static_cast<const half2_t*>(z), u);
}

void launch_zero_pin(hipStream_t s, size_t N, void* u, int u_half) {
    if (u_half)
        hipLaunchKernelGGL((k_zero_pin<uint32_t, true>), dim3(1), dim3(64), 0, s, N, 9, static_cast<uint32_t*>(u));
    else
        hipLaunchKernelGGL((k_zero_pin<float>), dim3(1), dim3(64), 0, s, N, 18, static_cast<float*>(u));
}

void launch_zero_pin64(hipStream_t s, size_t N, double* u) {
    hipLaunchKernelGGL(k_zero_pin64, dim3(1), dim3(64), 0, s, N, u);
}

void launch_zero_case64(hipStream_t s, size_t N, int r, double* u) {
    hipLaunchKernelGGL(k_zero_case64, dim3(nblk(3 * N)), dim3(256), 0, s, N, r, u);
}

// Levels up to this size run the coarse-stencil kernels with one load case per thread (six grid
// rows): their stencils sit in cache, and the extra threads hide the gather latency that dominates
// there.  Larger levels keep all six load cases per thread so the stored stencil is read once.
constexpr int kSplitLoadCasesMaxN = 32;

template <int R, typename SA, typename A>
static void stencil_apply_case(hipStream_t s, const Lay& L, size_t N, const SA* st, const A* u, const float* f,
                               float* out) {
    const dim3 g(full_grid(L), 6 / R), b(kFullThreads);
    if (f)
        hipLaunchKernelGGL((k_stencil_apply<R, 1, SA, A>), g, b, 0, s, L, N, st, u, f, out);
    else
        hipLaunchKernelGGL((k_stencil_apply<R, 0, SA, A>), g, b, 0, s, L, N, st, u, f, out);
}

template <typename SA>
static void stencil_apply_pick(hipStream_t s, const Lay& L, size_t N, bool split, const SA* st, const void* u,
                               int u_half, const float* f, float* out) {
    if (u_half)
        stencil_apply_case<6>(s, L, N, st, static_cast<const half2_t*>(u), f, out);
    else if (split)
        stencil_apply_case<1>(s, L, N, st, static_cast<const float*>(u), f, out);
    else
        stencil_apply_case<6>(s, L, N, st, static_cast<const float*>(u), f, out);
}

int stencil_u_half_ok(int n, int lex) { return !lex && n > kSplitLoadCasesMaxN && n % 8 == 0; }

void launch_stencil_apply(hipStream_t s, int n, int lex, const void* st, int st_half, const void* u, int u_half,
                          const float* f, float* out) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, lex);
    const bool split = n <= kSplitLoadCasesMaxN;
    if (u_half && !stencil_u_half_ok(n, lex)) return;  // the caller checks stencil_u_half_ok
    if (st_half)
        stencil_apply_pick(s, L, N, split, static_cast<const half2_t*>(st), u, u_half, f, out);
    else
        stencil_apply_pick(s, L, N, split, static_cast<const float*>(st), u, u_half, f, out);
}

template <int R, bool REV, int ZC, typename SA, typename A>
static void stencil_gs_colour(hipStream_t s, const Lay& L, size_t N, int colour, float omega, const SA* st, A* u,
                              const float* f) {
    hipLaunchKernelGGL((k_stencil_gs<R, REV, ZC, SA, A>), dim3(nblk(L.h3), 6 / R), dim3(256), 0, s, L, N, colour, omega,
                       st, u, f);
}

template <int R, typename SA, typename A>
static void stencil_gs_pick(hipStream_t s, const Lay& L, size_t N, int colour, float omega, int reverse,
                            int zero_start, const SA* st, A* u, const float* f) {
    if (reverse) {
        stencil_gs_colour<R, true, -1>(s, L, N, colour, omega, st, u, f);
        return;
    }
    switch (zero_start ? colour : -1) {
        case 0: stencil_gs_colour<R, false, 0>(s, L, N, colour, omega, st, u, f); break;
        case 1: stencil_gs_colour<R, false, 1>(s, L, N, colour, omega, st, u, f); break;
        case 2: stencil_gs_colour<R, false, 2>(s, L, N, colour, omega, st, u, f); break;
        case 3: stencil_gs_colour<R, false, 3>(s, L, N, colour, omega, st, u, f); break;
        case 4: stencil_gs_colour<R, false, 4>(s, L, N, colour, omega, st, u, f); break;
        case 5: stencil_gs_colour<R, false, 5>(s, L, N, colour, omega, st, u, f); break;
        case 6: stencil_gs_colour<R, false, 6>(s, L, N, colour, omega, st, u, f); break;
        case 7: stencil_gs_colour<R, false, 7>(s, L, N, colour, omega, st, u, f); break;
        default: stencil_gs_colour<R, false, -1>(s, L, N, colour, omega, st, u, f); break;
    }
}

template <typename SA>
static void stencil_gs_st(hipStream_t s, const Lay& L, size_t N, bool split, int colour, float omega, int reverse,
                          int zero_start, const SA* st, void* u, int u_half, const float* f) {
    if (u_half)
        stencil_gs_pick<6>(s, L, N, colour, omega, reverse, zero_start, st, static_cast<half2_t*>(u), f);
    else if (split)
        stencil_gs_pick<1>(s, L, N, colour, omega, reverse, zero_start, st, static_cast<float*>(u), f);
    else
        stencil_gs_pick<6>(s, L, N, colour, omega, reverse, zero_start, st, static_cast<float*>(u), f);
}

template <int R, bool REV, int ZC, typename SA, typename A, bool ROWN = false>
static void stencil_gs2_colour(hipStream_t s, const Lay& L, size_t N, int cfirst, float omega, const SA* st, A* u,
                               const float* f, float* rown = nullptr) {
    hipLaunchKernelGGL((k_stencil_gs2<R, REV, ZC, SA, A, ROWN>), dim3(L.h3 / 256, 6 / R), dim3(256), 0, s, L, N, cfirst,
                       omega, st, u, f, rown);
}

template <int R, int ZC, typename SA, typename A>
static void stencil_gs2_zero(hipStream_t s, const Lay& L, size_t N, float omega, const SA* st, A* u, const float* f,
                             float* rown) {
    if (rown)
        stencil_gs2_colour<R, false, ZC, SA, A, true>(s, L, N, ZC, omega, st, u, f, rown);
    else
        stencil_gs2_colour<R, false, ZC>(s, L, N, ZC, omega, st, u, f);
}

template <int R, typename SA, typename A>
static void stencil_gs2_pick(hipStream_t s, const Lay& L, size_t N, int cfirst, float omega, int reverse,
                             int zero_start, const SA* st, A* u, const float* f, float* rown) {
    if (reverse) {
        stencil_gs2_colour<R, true, -1>(s, L, N, cfirst, omega, st, u, f);
        return;
    }
    switch (zero_start ? cfirst : -1) {
        case 0: stencil_gs2_zero<R, 0>(s, L, N, omega, st, u, f, rown); break;
        case 2: stencil_gs2_zero<R, 2>(s, L, N, omega, st, u, f, rown); break;
        case 4: stencil_gs2_zero<R, 4>(s, L, N, omega, st, u, f, rown); break;
        case 6: stencil_gs2_zero<R, 6>(s, L, N, omega, st, u, f, rown); break;
        default: stencil_gs2_colour<R, false, -1>(s, L, N, cfirst, omega, st, u, f); break;
    }
}

template <typename SA>
static void stencil_gs2_st(hipStream_t s, const Lay& L, size_t N, bool split, int cfirst, float omega, int reverse,
                           int zero_start, const SA* st, void* u, int u_half, const float* f, float* rown) {
    if (u_half)
        stencil_gs2_pick<6>(s, L, N, cfirst, omega, reverse, zero_start, st, static_cast<half2_t*>(u), f, rown);
    else if (split)
        stencil_gs2_pick<1>(s, L, N, cfirst, omega, reverse, zero_start, st, static_cast<float*>(u), f, rown);
    else
        stencil_gs2_pick<6>(s, L, N, cfirst, omega, reverse, zero_start, st, static_cast<float*>(u), f, rown);
}

// pairs only where a launch runs six load cases per thread: on the one-case levels (<= 32^3) the
// launch is latency-bound and two colours in one cost more than two launches (14 vs 11 us at 32^3)
int stencil_gs_pairs_ok(int n, int lex) { return !lex && n > kSplitLoadCasesMaxN && fine_gs_pairs_ok(n); }

void launch_stencil_gs2(hipStream_t s, int n, int cfirst, float omega, int reverse, int zero_start, const void* st,
                        int st_half, void* u, int u_half, const float* f, float* rown) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    const bool split = n <= kSplitLoadCasesMaxN;
    if (!stencil_gs_pairs_ok(n, 0)) return;            // the caller checks stencil_gs_pairs_ok
    if (u_half && !stencil_u_half_ok(n, 0)) return;    // ... and stencil_u_half_ok
    if (reverse || !zero_start) rown = nullptr;        // own-colour residuals only for zero-start passes
    if (st_half)
        stencil_gs2_st(s, L, N, split, cfirst, omega, reverse, zero_start, static_cast<const half2_t*>(st), u, u_half,
                       f, rown);
    else
        stencil_gs2_st(s, L, N, split, cfirst, omega, reverse, zero_start, static_cast<const float*>(st), u, u_half, f,
                       rown);
}

template <int R, typename SA, typename A>
static void stencil_resid_later(hipStream_t s, const Lay& L, size_t N, const SA* st, const A* u, float* r) {
    hipLaunchKernelGGL((k_stencil_resid_later<R, SA, A>), dim3(nblk(L.h3, 64), 6 / R), dim3(256), 0, s, L, N, st, u,
                       r, 0.f);
}

template <typename SA>
static void stencil_resid_later_st(hipStream_t s, const Lay& L, size_t N, const SA* st, const void* u, int u_half,
                                   float* r) {
    if (u_half)
        stencil_resid_later<6>(s, L, N, st, static_cast<const half2_t*>(u), r);
    else
        stencil_resid_later<6>(s, L, N, st, static_cast<const float*>(u), r);
}

void launch_stencil_residual_later(hipStream_t s, int n, const void* st, int st_half, const void* u, int u_half,
                                   float* r) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    if (!stencil_gs_pairs_ok(n, 0)) return;          // the caller checks stencil_gs_pairs_ok
    if (u_half && !stencil_u_half_ok(n, 0)) return;  // ... and stencil_u_half_ok
    if (st_half)
        stencil_resid_later_st(s, L, N, static_cast<const half2_t*>(st), u, u_half, r);
    else
        stencil_resid_later_st(s, L, N, static_cast<const float*>(st), u, u_half, r);
}

void launch_stencil_gs(hipStream_t s, int n, int colour, float omega, int reverse, int zero_start, const void* st,
                       int st_half, void* u, int u_half, const float* f) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    const bool split = n <= kSplitLoadCasesMaxN;
    if (u_half && !stencil_u_half_ok(n, 0)) return;  // the caller checks stencil_u_half_ok
    if (st_half)
        stencil_gs_st(s, L, N, split, colour, omega, reverse, zero_start, static_cast<const half2_t*>(st), u, u_half, f);
    else
        stencil_gs_st(s, L, N, split, colour, omega, reverse, zero_start, static_cast<const float*>(st), u, u_half, f);
}

// Transfers take one load case per thread (six grid rows): the load cases' planes are disjoint, so
// this costs no extra traffic and shortens each thread's chain of dependent loads.
void launch_restrict(hipStream_t s, int nf, int lexc, const float* rf, float* fc) {
    const size_t Nc = (size_t)(nf / 2) * (nf / 2) * (nf / 2);
    const Lay Lf = make_lay(nf, 0), Lc = make_lay(nf / 2, lexc);
    if (nf % 8 == 0)
        hipLaunchKernelGGL(k_restrict4, dim3(nblk(Nc / 4), 6), dim3(256), 0, s, Lf, Lc, rf, fc);
    else
        hipLaunchKernelGGL((k_restrict<1>), dim3(nblk(Nc), 6), dim3(256), 0, s, Lf, Lc, rf, fc);
}

void launch_prolong(hipStream_t s, int nf, int lexc, const void* uc, int uc_half, void* uf, int u_half, int pin) {
    const Lay Lf = make_lay(nf, 0), Lc = make_lay(nf / 2, lexc);
    if (uc_half && (!u_half || lexc)) return;  // packed FP16 coarse u: parity-blocked, into a packed FP16 fine u
    if (u_half && uc_half)
        hipLaunchKernelGGL((k_prolong_h<half2_t>), dim3(full_grid(Lf)), dim3(kFullThreads), 0, s, Lf, Lc,
                           static_cast<const half2_t*>(uc), static_cast<half2_t*>(uf), pin);
    else if (u_half)  // packed FP16 fine fields are only used with nf % 8 == 0
        hipLaunchKernelGGL((k_prolong_h<float>), dim3(full_grid(Lf)), dim3(kFullThreads), 0, s, Lf, Lc,
                           static_cast<const float*>(uc), static_cast<half2_t*>(uf), pin);
    else if (nf % 8 == 0)
        hipLaunchKernelGGL(k_prolong4, dim3(nblk(Lf.h3 / 4, 64), 6), dim3(kFullThreads), 0, s, Lf, Lc,
                           static_cast<const float*>(uc), static_cast<float*>(uf), pin);
    else
        hipLaunchKernelGGL((k_prolong<1>), dim3(full_grid(Lf), 6), dim3(kFullThreads), 0, s, Lf, Lc,
                           static_cast<const float*>(uc), static_cast<float*>(uf), pin);
}

void launch_galerkin_fine(hipStream_t s, int nf, int lexc, const float* G, const float* ke, const float* w, void* stc,
                          int out_half, float clamp) {
    const size_t Nc = (size_t)(nf / 2) * (nf / 2) * (nf / 2);
    const Lay Lf = make_lay(nf, 0), Lc = make_lay(nf / 2, lexc);
    if (out_half) {
        half2_t* o = static_cast<half2_t*>(stc);
        hipLaunchKernelGGL(k_galerkin_fine<half2_t>, dim3(nblk(Nc)), dim3(256), 0, s, Lf, Lc, G, w, o, clamp);
        hipLaunchKernelGGL(k_galerkin_fine_pin<half2_t>, dim3(nblk(27 * 27)), dim3(256), 0, s, Lf, Lc, ke, w, o, clamp);
    } else {
        float* o = static_cast<float*>(stc);
        hipLaunchKernelGGL(k_galerkin_fine<float>, dim3(nblk(Nc)), dim3(256), 0, s, Lf, Lc, G, w, o, clamp);
        hipLaunchKernelGGL(k_galerkin_fine_pin<float>, dim3(nblk(27 * 27)), dim3(256), 0, s, Lf, Lc, ke, w, o, clamp);
    }
}

template <typename SI, typename SO>
static void galerkin_coarse_case(hipStream_t s, const Lay& Lf, const Lay& Lc, const void* stf, void* stc, float clamp) {
    const size_t Nc = (size_t)Lc.n * Lc.n * Lc.n;
    hipLaunchKernelGGL((k_galerkin_coarse<SI, SO>), dim3(nblk(Nc), 3), dim3(256), 0, s, Lf, Lc,
                       static_cast<const SI*>(stf), static_cast<SO*>(stc), clamp);
}

void launch_galerkin_coarse(hipStream_t s, int nf, int lexc, const void* stf, int in_half, void* stc, int out_half,
                            float clamp) {
    const Lay Lf = make_lay(nf, 0), Lc = make_lay(nf / 2, lexc);
    if (in_half && out_half)
        galerkin_coarse_case<half2_t, half2_t>(s, Lf, Lc, stf, stc, clamp);
    else if (in_half)
        galerkin_coarse_case<half2_t, float>(s, Lf, Lc, stf, stc, clamp);
    else if (out_half)
        galerkin_coarse_case<float, half2_t>(s, Lf, Lc, stf, stc, clamp);
    else
        galerkin_coarse_case<float, float>(s, Lf, Lc, stf, stc, clamp);
}

void launch_coarse_solve(hipStream_t s, int nc, const double* Ainv, const float* f, float* u) {
    const size_t Nc = (size_t)nc * nc * nc;
    const int m = (int)(3 * Nc);
    hipLaunchKernelGGL(k_coarse_solve, dim3(6), dim3(256), sizeof(double) * m, s, m, Nc, Ainv, f, u);
}

void launch_coarse_factor(hipStream_t s, int nc, const float* st, double* ainv, double* piv, int* status) {
    const int m = 3 * nc * nc * nc;
    const size_t mm = (size_t)m * m;
    hipLaunchKernelGGL(k_fill64, dim3(nblk(mm) < 1024 ? nblk(mm) : 1024), dim3(256), 0, s, mm, 0.0, ainv);
    hipLaunchKernelGGL(k_coarse_assemble, dim3(nblk((size_t)m)), dim3(256), 0, s, nc, st, ainv);
    hipLaunchKernelGGL(k_symmetrize64, dim3(nblk(mm)), dim3(256), 0, s, m, ainv);
    // pivot rows and columns, double-buffered: piv = [row 0 | col 0 | row 1 | col 1], m values each
    double* row[2] = {piv, piv + 2 * (size_t)m};
    double* col[2] = {piv + m, piv + 3 * (size_t)m};
    hipLaunchKernelGGL(k_gj_pivot0, dim3(nblk((size_t)m)), dim3(256), 0, s, m, ainv, row[0], col[0], status);
    for (int k = 0; k < m; k++)
        hipLaunchKernelGGL(k_gj_step64, dim3(nblk(mm)), dim3(256), 0, s, m, k, ainv, row[k & 1], col[k & 1],
                           row[(k + 1) & 1], col[(k + 1) & 1], status);
}

void launch_dot6(hipStream_t s, size_t M, const float* a, const float* b, double* partial, int nblocks,
                 double* out) {
    hipLaunchKernelGGL(k_dot, dim3(nblocks, 6), dim3(256), 0, s, M, a, b, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(6), dim3(kSumThreads), 0, s, nblocks, partial, out);
}

void launch_sum_partials(hipStream_t s, int R, int nblocks, const double* partial, double* out) {
    hipLaunchKernelGGL(k_sum_partials, dim3(R), dim3(kSumThreads), 0, s, nblocks, partial, out);
}

void launch_pcg_alpha(hipStream_t s, double* sc) { hipLaunchKernelGGL(k_pcg_alpha, dim3(1), dim3(64), 0, s, 6, sc); }
void launch_pcg_beta(hipStream_t s, double* sc) { hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(64), 0, s, 6, sc); }

void launch_pcg_update_r(hipStream_t s, size_t M, const double* sc, const float* q, float* r, double* partial,
                         int nblocks, double* out6) {
    hipLaunchKernelGGL(k_pcg_update_r, dim3(nblocks, 6), dim3(256), 0, s, 6, M, sc, q, r, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(6), dim3(kSumThreads), 0, s, nblocks, partial, out6);
}

void launch_pcg_update_px(hipStream_t s, size_t M, const double* sc, const void* z, int z_half, float* p, float* x,
                          int do_p, int nblocks) {
    if (z_half)
        hipLaunchKernelGGL(k_pcg_update_px_h, dim3(nblocks, 3), dim3(256), 0, s, 6, M, sc,
                           static_cast<const half2_t*>(z), p, x, do_p, 0);
    else
        hipLaunchKernelGGL(k_pcg_update_px, dim3(nblocks, 6), dim3(256), 0, s, 6, M, sc, static_cast<const float*>(z),
                           p, x, do_p);
}

hipError_t launch_set_p(hipStream_t s, size_t M, const void* z, int z_half, float* p, int nblocks) {
    if (!z_half) return hipMemcpyAsync(p, z, sizeof(float) * 6 * M, hipMemcpyDeviceToDevice, s);
    hipLaunchKernelGGL(k_pcg_update_px_h, dim3(nblocks, 3), dim3(256), 0, s, 6, M, nullptr,
                       static_cast<const half2_t*>(z), p, nullptr, 0, 1);
    return hipGetLastError();
}

void launch_dot_rz(hipStream_t s, size_t M, const float* r, const void* z, int z_half, double* partial, int nblocks,
                   double* out6) {
    if (!z_half) {
        launch_dot6(s, M, r, static_cast<const float*>(z), partial, nblocks, out6);
        return;
    }
    hipLaunchKernelGGL(k_dot_fh, dim3(nblocks, 3), dim3(256), 0, s, M, r, static_cast<const half2_t*>(z), partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(6), dim3(kSumThreads), 0, s, nblocks, partial, out6);
}

void launch_elastic_matrix(hipStream_t s, int n, const float* ke, const float* w, const float* chi, double* partial,
                           int nblocks, double* out21) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    const unsigned nchunks = full_grid(L);
    const int nb = (int)std::min<unsigned>((unsigned)nblocks, nchunks);
    hipLaunchKernelGGL(k_elastic_matrix, dim3(nb), dim3(kFullThreads), 0, s, L, N, nchunks, ke, w, chi, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(21), dim3(kSumThreads), 0, s, nb, partial, out21);
}

void launch_sensitivity(hipStream_t s, int n, int pitch, const float* ke, const float* chi, const float* wdc,
                        float* sens) {
    const size_t N = (size_t)n * n * n;
    const Lay L = make_lay(n, 0);
    hipLaunchKernelGGL(k_sensitivity, dim3(full_grid(L)), dim3(kFullThreads), 0, s, L, N, pitch, ke, chi, wdc, sens);
}

void launch_symmetrize(hipStream_t s, int n, int pitch, int ng, const int* gperm, const int* gflip, const float* in,
                       float* out) {
    const size_t N = (size_t)n * n * n;
    hipLaunchKernelGGL(k_symmetrize, dim3(nblk(N)), dim3(256), 0, s, n, pitch, ng, gperm, gflip, in, out);
}

void launch_oc_filter(hipStream_t s, int n, int pitch, int h, int periodic, const float* wk, const float* sens,
                      const float* rho, float* out) {
    const size_t N = (size_t)n * n * n;
    hipLaunchKernelGGL(k_oc_filter, dim3(nblk(N)), dim3(256), 0, s, n, pitch, h, periodic, wk, sens, rho, out);
}

void launch_conv_filter(hipStream_t s, int n, int pitch, int h, int periodic, const float* wk, const float* in,
                        float* out, float* wsum) {
    const size_t N = (size_t)n * n * n;
    hipLaunchKernelGGL(k_conv_filter, dim3(nblk(N)), dim3(256), 0, s, n, pitch, h, periodic, wk, in, out, wsum);
}

void launch_conv_filter_back(hipStream_t s, int n, int pitch, int h, int periodic, const float* wk, const float* dout,
                             const float* wsum, float* din) {
    const size_t N = (size_t)n * n * n;
    hipLaunchKernelGGL(k_conv_filter_back, dim3(nblk(N)), dim3(256), 0, s, n, pitch, h, periodic, wk, dout, wsum, din);
}

void launch_sgm(hipStream_t s, int n, int pitch, double k, double c, const float* x, float* y) {
    const size_t N = (size_t)n * n * n;
    hipLaunchKernelGGL(k_sgm, dim3(nblk(N)), dim3(256), 0, s, n, pitch, k, c, x, y);
}

void launch_sgm_back(hipStream_t s, int n, int pitch, double k, const float* y, const float* gy, float* gx) {
    const size_t N = (size_t)n * n * n;
    hipLaunchKernelGGL(k_sgm_back, dim3(nblk(N)), dim3(256), 0, s, n, pitch, k, y, gy, gx);
}

void launch_oc_trial(hipStream_t s, int n, int pitch, double g, double min_rho, double step, double damp,
                     const float* sens, const float* rho, float* out, double* partial, int nblocks, double* sum) {
    hipLaunchKernelGGL(k_oc_trial, dim3(nblocks), dim3(256), 0, s, n, pitch, g, min_rho, step, damp, sens, rho, out,
                       partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kSumThreads), 0, s, nblocks, partial, sum);
}

void launch_density_pow(hipStream_t s, int n, int pitch, float p, const float* rho, float* x) {
    hipLaunchKernelGGL(k_density_pow, dim3(1024), dim3(256), 0, s, n, pitch, p, rho, x);
}
void launch_density_pow_grad(hipStream_t s, int n, int pitch, float p, const float* rho, const float* dx, float* drho) {
    hipLaunchKernelGGL(k_density_pow_grad, dim3(1024), dim3(256), 0, s, n, pitch, p, rho, dx, drho);
}
void launch_field_sum(hipStream_t s, int n, int pitch, const float* a, double* partial, int nblocks) {
    hipLaunchKernelGGL(k_field_sum, dim3(nblocks), dim3(256), 0, s, n, pitch, a, partial);
}
void launch_field_fill(hipStream_t s, int n, int pitch, float value, float* a) {
    hipLaunchKernelGGL(k_field_fill, dim3(1024), dim3(256), 0, s, n, pitch, value, a);
}

void launch_absmax(hipStream_t s, int n, int pitch, const float* a, float* partial, int nblocks) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_absmax, dim3(nblocks), dim3(256), 0, s, N, n, pitch, a, partial);
}

void launch_set_weights(hipStream_t s, int n, int pitch, double emin, const float* x, float* w, float* xc) {
    const size_t N = (size_t)n * n * n;
    hipLaunchKernelGGL(k_set_weights, dim3(nblk(N)), dim3(256), 0, s, make_lay(n, 0), pitch, emin, x, w, xc);
}

// ---------------------------------------------------------------------------------------------
// Bridges between the solver's fields and the learned warm-start network (net.hip): the network runs
// on a g^3 grid (g | n) with bf16 channels-last tensors.
// ---------------------------------------------------------------------------------------------
namespace {

// Network input: channel 0 of cell (i, j, k) = mean of rho = x^(1/3) over the f^3 fine cells it covers
// (f = n / g), channels 1..7 = 0.  x holds the raw x = rho^3 per cell in level storage order.
__global__ __launch_bounds__(256) void k_net_input(Lay L, int g, const float* __restrict__ x, uint16_t* __restrict__ in) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long g3 = (long)g * g * g;
    if (v >= g3) return;
    const int f = L.n / g;
    const int i = (int)(v % g), j = (int)((v / g) % g), k = (int)(v / ((long)g * g));
    float s = 0.f;
    for (int c = 0; c < f * f * f; c++) {
        const int xx = i * f + c % f, yy = j * f + (c / f) % f, zz = k * f + c / (f * f);
        s += cbrtf(fmaxf(x[L.idx(xx, yy, zz)], 0.f));
    }
    *reinterpret_cast<uint4*>(in + v * 8) = make_uint4(f2bf(s / (float)(f * f * f)), 0u, 0u, 0u);
}

// Per-channel mean of the solution sampled at the coarse vertices (i f, j f, k f), weighted by the
// stiffness rho^3 of the network input at the same index (the network's loss weights).
__global__ __launch_bounds__(256) void k_net_target_mean(Lay L, int g, const double* __restrict__ u,
                                                         const uint16_t* __restrict__ in, double* __restrict__ mean) {
    __shared__ double sh[2][256];
    const int c = blockIdx.x;
    const long g3 = (long)g * g * g;
    const size_t N = (size_t)L.n * L.n * L.n;
    const int f = L.n / g;
    double s = 0.0, ws = 0.0;
    for (long v = threadIdx.x; v < g3; v += blockDim.x) {
        const int i = (int)(v % g), j = (int)((v / g) % g), k = (int)(v / ((long)g * g));
        const double r = bf2f(in[v * 8]);
        s += r * r * r * u[u64_index(c, L.idx(i * f, j * f, k * f), N)];
        ws += r * r * r;
    }
    sh[0][threadIdx.x] = s;
    sh[1][threadIdx.x] = ws;
    __syncthreads();
    for (int t = 128; t > 0; t >>= 1) {
        if ((int)threadIdx.x < t) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + t];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + t];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) mean[c] = sh[1][0] > 0 ? sh[0][0] / sh[1][0] : 0.0;
}

// Network target: channel 3 s + k of coarse vertex v = scale (u_sk(i f, j f, k f) - mean) / n.
__global__ __launch_bounds__(256) void k_net_target(Lay L, int g, const double* __restrict__ u,
                                                    const double* __restrict__ mean, double scale,
                                                    uint16_t* __restrict__ tg) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long g3 = (long)g * g * g;
    if (v >= g3) return;
    const size_t N = (size_t)L.n * L.n * L.n;
    const int f = L.n / g;
    const int i = (int)(v % g), j = (int)((v / g) % g), k = (int)(v / ((long)g * g));
    const uint32_t e = L.idx(i * f, j * f, k * f);
    const double sc = scale / (double)L.n;
    for (int c = 0; c < 18; c++) tg[v * 18 + c] = f2bf((float)(sc * (u[u64_index(c, e, N)] - mean[c])));
}

// Warm start from the prediction P ([g^3][ldp], channel 3 s + k = scale u_sk / n): periodic trilinear
// interpolation to the n^3 vertices, minus the value at vertex 0 (the pinned vertex keeps u = 0),
// times n / scale, into the FP64 iterate U64 and its FP32 copy X (level storage order).
__global__ __launch_bounds__(256) void k_net_warm_start(Lay L, int g, const float* __restrict__ P, int ldp, double inv,
                                                        double* __restrict__ U64, float* __restrict__ X) {
    const size_t N = (size_t)L.n * L.n * L.n;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    int x, y, z;
    L.pos((uint32_t)t, x, y, z);
    const int f = L.n / g;
    const int i0 = x / f, j0 = y / f, k0 = z / f;
    const float tx = (float)(x - i0 * f) / f, ty = (float)(y - j0 * f) / f, tz = (float)(z - k0 * f) / f;
    long cv[8];
    float wt[8];
    for (int c = 0; c < 8; c++) {
        const int i = (i0 + (c & 1)) % g, j = (j0 + ((c >> 1) & 1)) % g, k = (k0 + (c >> 2)) % g;
        cv[c] = i + (long)g * (j + (long)g * k);
        wt[c] = ((c & 1) ? tx : 1.f - tx) * (((c >> 1) & 1) ? ty : 1.f - ty) * ((c >> 2) ? tz : 1.f - tz);
    }
    for (int ch = 0; ch < 18; ch++) {
        float s = 0.f;
        for (int c = 0; c < 8; c++) s += wt[c] * P[cv[c] * ldp + ch];
        const double u = ((double)s - (double)P[ch]) * inv;
        U64[u64_index(ch, t, N)] = u;
        X[ch * N + t] = (float)u;
    }
}

}  // namespace

void launch_net_input(hipStream_t s, int n, int g, const float* x, uint16_t* in) {
    const long g3 = (long)g * g * g;
    hipLaunchKernelGGL(k_net_input, dim3(nblk(g3)), dim3(256), 0, s, make_lay(n, 0), g, x, in);
}

void launch_net_target(hipStream_t s, int n, int g, const double* u64, const uint16_t* in, double scale, double* mean18,
                       uint16_t* tg) {
    const long g3 = (long)g * g * g;
    hipLaunchKernelGGL(k_net_target_mean, dim3(18), dim3(256), 0, s, make_lay(n, 0), g, u64, in, mean18);
    hipLaunchKernelGGL(k_net_target, dim3(nblk(g3)), dim3(256), 0, s, make_lay(n, 0), g, u64, mean18, scale, tg);
}

void launch_net_warm_start(hipStream_t s, int n, int g, const float* P, int ldp, double scale, double* u64, float* x) {
    const size_t N = (size_t)n * n * n;
    hipLaunchKernelGGL(k_net_warm_start, dim3(nblk(N)), dim3(256), 0, s, make_lay(n, 0), g, P, ldp, (double)n / scale,
                       u64, x);
}

void launch_copy_pitched(hipStream_t s, int n, int pitch_in, int pitch_out, const float* in, float* out) {
    const size_t N = (size_t)n * n * n;
    hipLaunchKernelGGL(k_copy_pitched, dim3(nblk(N)), dim3(256), 0, s, n, pitch_in, pitch_out, in, out);
}

}  // namespace homo3d

// =============================================================================================
// MMA (Svanberg's method of moving asymptotes; mmasub_ker + subsolv_g, mmaOpt.cu:429-1019)
// FP64 per-variable kernels; m <= kMmaMaxM constraints; dfdx is m x n row-major.
// =============================================================================================
namespace homo3d {
namespace {

struct MmaDual {          // small dual vectors of subsolv, read by every thread
    double lam[kMmaMaxM];
    double epsi;
};

template <int BS>
__device__ __forceinline__ double block_max(double v, double* sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = -1e300;
    if (threadIdx.x == 0)
        for (int i = 0; i < BS / 64; i++) s = fmax(s, sh[i]);
    __syncthreads();
    return s;
}

// asymptotes, move limits and the convex approximation terms (mmasub_ker, mmaOpt.cu:864-989);
// part[i*nb + b] = block sums of P_ij/(upp-x) + Q_ij/(x-low)
__global__ __launch_bounds__(256) void k_mma_setup(int m, long n, int iter, double move, const double* __restrict__ x,
                                                   const double* __restrict__ xmin, const double* __restrict__ xmax,
                                                   const double* __restrict__ xold1, const double* __restrict__ xold2,
                                                   const double* __restrict__ df0, const double* __restrict__ dfdx,
                                                   double* __restrict__ low, double* __restrict__ upp,
                                                   double* __restrict__ alfa, double* __restrict__ beta,
                                                   double* __restrict__ p0, double* __restrict__ q0,
                                                   double* __restrict__ P, double* __restrict__ Q,
                                                   double* __restrict__ part) {
    __shared__ double sh[4];
    const double raa0 = 1e-5, albefa = 0.1, asyinit = 0.5, asyincr = 1.2, asydecr = 0.7;
    double bs[kMmaMaxM] = {0, 0, 0, 0};
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        const double xj = x[j], dx = xmax[j] - xmin[j];
        double lo, up;
        if (iter <= 2) {
            lo = xj - asyinit * dx;
            up = xj + asyinit * dx;
        } else {
            const double zzz = (xj - xold1[j]) * (xold1[j] - xold2[j]);
            const double fac = zzz > 0 ? asyincr : (zzz < 0 ? asydecr : 1.0);
            lo = xj - fac * (xold1[j] - low[j]);
            up = xj + fac * (upp[j] - xold1[j]);
            lo = fmax(lo, xj - 10 * dx);
            lo = fmin(lo, xj - 0.01 * dx);
            up = fmin(up, xj + 10 * dx);
            up = fmax(up, xj + 0.01 * dx);
        }
        low[j] = lo;
        upp[j] = up;
        alfa[j] = fmax(fmax(lo + albefa * (xj - lo), xj - move * dx), xmin[j]);
        beta[j] = fmin(fmin(up - albefa * (up - xj), xj + move * dx), xmax[j]);
        const double xmamiinv = 1.0 / fmax(dx, 1e-5);
        const double ux1 = up - xj, xl1 = xj - lo, ux2 = ux1 * ux1, xl2 = xl1 * xl1;
        const double dp = fmax(df0[j], 0.0), dm = fmax(-df0[j], 0.0);
        const double pq0 = 1e-3 * (dp + dm) + raa0 * xmamiinv;
        p0[j] = (dp + pq0) * ux2;
        q0[j] = (dm + pq0) * xl2;
        for (int i = 0; i < m; i++) {
            const double g = dfdx[(size_t)i * n + j];
            const double pp = fmax(g, 0.0), qq = fmax(-g, 0.0), pq = 1e-3 * (pp + qq) + raa0 * xmamiinv;
            const double Pv = (pp + pq) * ux2, Qv = (qq + pq) * xl2;
            P[(size_t)i * n + j] = Pv;
            Q[(size_t)i * n + j] = Qv;
            bs[i] += Pv / ux1 + Qv / xl1;
        }
    }
    for (int i = 0; i < m; i++) {
        const double s = block_sum<256>(bs[i], sh);
        if (threadIdx.x == 0) part[i * gridDim.x + blockIdx.x] = s;
    }
}

// per-variable primal residual terms of subsolv at (x, xsi, eta; lam, epsi):
// part layout per block b: [0] sum sq, [1] max abs (in part2), [2..2+m) gvec partials
__global__ __launch_bounds__(256) void k_mma_resid(int m, long n, const MmaDual dual, const double* __restrict__ x,
                                                   const double* __restrict__ xsi, const double* __restrict__ eta,
                                                   const double* __restrict__ low, const double* __restrict__ upp,
                                                   const double* __restrict__ alfa, const double* __restrict__ beta,
                                                   const double* __restrict__ p0, const double* __restrict__ q0,
                                                   const double* __restrict__ P, const double* __restrict__ Q,
                                                   double* __restrict__ part, double* __restrict__ partmax) {
    __shared__ double sh[4];
    double ssq = 0, smax = 0;
    double gv[kMmaMaxM] = {0, 0, 0, 0};
    const double epsi = dual.epsi;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        const double xj = x[j], ux1 = upp[j] - xj, xl1 = xj - low[j];
        double pl = p0[j], ql = q0[j];
        for (int i = 0; i < m; i++) {
            const double Pv = P[(size_t)i * n + j], Qv = Q[(size_t)i * n + j];
            pl += Pv * dual.lam[i];
            ql += Qv * dual.lam[i];
            gv[i] += Pv / ux1 + Qv / xl1;
        }
        const double dpsidx = pl / (ux1 * ux1) - ql / (xl1 * xl1);
        const double rex = dpsidx - xsi[j] + eta[j];
        const double rexsi = xsi[j] * (xj - alfa[j]) - epsi;
        const double reeta = eta[j] * (beta[j] - xj) - epsi;
        ssq += rex * rex + rexsi * rexsi + reeta * reeta;
        smax = fmax(smax, fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
    }
    double t = block_sum<256>(ssq, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
    for (int i = 0; i < m; i++) {
        t = block_sum<256>(gv[i], sh);
        if (threadIdx.x == 0) part[(1 + i) * gridDim.x + blockIdx.x] = t;
    }
    t = block_max<256>(smax, sh);
    if (threadIdx.x == 0) partmax[blockIdx.x] = t;
}

// Newton system terms: delx, diagx per variable; block sums of GG_i*delx/diagx and
// GG_i*GG_k/diagx (k >= i) for the (m+1) x (m+1) system (subsolv_g, mmaOpt.cu:560-640)
__global__ __launch_bounds__(256) void k_mma_newton(int m, long n, const MmaDual dual, const double* __restrict__ x,
                                                    const double* __restrict__ xsi, const double* __restrict__ eta,
                                                    const double* __restrict__ low, const double* __restrict__ upp,
                                                    const double* __restrict__ alfa, const double* __restrict__ beta,
                                                    const double* __restrict__ p0, const double* __restrict__ q0,
                                                    const double* __restrict__ P, const double* __restrict__ Q,
                                                    double* __restrict__ delx, double* __restrict__ diagx,
                                                    double* __restrict__ part) {
    __shared__ double sh[4];
    constexpr int NS = kMmaMaxM + kMmaMaxM * (kMmaMaxM + 1) / 2;
    double acc[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) acc[k] = 0;
    const double epsi = dual.epsi;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        const double xj = x[j], ux1 = upp[j] - xj, xl1 = xj - low[j];
        const double ux2 = ux1 * ux1, xl2 = xl1 * xl1, ux3 = ux1 * ux2, xl3 = xl1 * xl2;
        double pl = p0[j], ql = q0[j], gg[kMmaMaxM];
        for (int i = 0; i < m; i++) {
            const double Pv = P[(size_t)i * n + j], Qv = Q[(size_t)i * n + j];
            pl += Pv * dual.lam[i];
            ql += Qv * dual.lam[i];
            gg[i] = Pv / ux2 - Qv / xl2;
        }
        const double dpsidx = pl / ux2 - ql / xl2;
        const double dl = dpsidx - epsi / (xj - alfa[j]) + epsi / (beta[j] - xj);
        const double dg = 2 * (pl / ux3 + ql / xl3) + xsi[j] / (xj - alfa[j]) + eta[j] / (beta[j] - xj);
        delx[j] = dl;
        diagx[j] = dg;
        int k = kMmaMaxM;
        for (int i = 0; i < m; i++) {
            acc[i] += gg[i] * dl / dg;
            for (int q = i; q < m; q++) acc[k + i * kMmaMaxM - i * (i - 1) / 2 + (q - i)] += gg[i] * gg[q] / dg;
        }
    }
    for (int k = 0; k < NS; k++) {
        const double t = block_sum<256>(acc[k], sh);
        if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = t;
    }
}

// Newton direction per variable and block maxima of the step-length ratios
// (stepxx/stepalfa/stepbeta of subsolv_g, mmaOpt.cu:680-700)
__global__ __launch_bounds__(256) void k_mma_dir(int m, long n, const MmaDual dual, const double* __restrict__ dlam,
                                                 const double* __restrict__ x, const double* __restrict__ xsi,
                                                 const double* __restrict__ eta, const double* __restrict__ low,
                                                 const double* __restrict__ upp, const double* __restrict__ alfa,
                                                 const double* __restrict__ beta, const double* __restrict__ P,
                                                 const double* __restrict__ Q, const double* __restrict__ delx,
                                                 const double* __restrict__ diagx, double* __restrict__ dx,
                                                 double* __restrict__ dxsi, double* __restrict__ deta,
                                                 double* __restrict__ partmax) {
    __shared__ double sh[4];
    double stm = -1e300;
    const double epsi = dual.epsi;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        const double xj = x[j], ux2 = (upp[j] - xj) * (upp[j] - xj), xl2 = (xj - low[j]) * (xj - low[j]);
        double gt = 0;
        for (int i = 0; i < m; i++) gt += (P[(size_t)i * n + j] / ux2 - Q[(size_t)i * n + j] / xl2) * dlam[i];
        const double d = -delx[j] / diagx[j] - gt / diagx[j];
        const double xa = xj - alfa[j], bx = beta[j] - xj;
        const double ds = -xsi[j] + epsi / xa - (xsi[j] * d) / xa;
        const double de = -eta[j] + epsi / bx + (eta[j] * d) / bx;
        dx[j] = d;
        dxsi[j] = ds;
        deta[j] = de;
        stm = fmax(stm, fmax(fmax(-1.01 * ds / xsi[j], -1.01 * de / eta[j]), fmax(-1.01 * d / xa, 1.01 * d / bx)));
    }
    const double t = block_max<256>(stm, sh);
    if (threadIdx.x == 0) partmax[blockIdx.x] = t;
}

// x = xold + steg dx (and xsi, eta)
__global__ __launch_bounds__(256) void k_mma_step(long n, double steg, const double* __restrict__ xo,
                                                  const double* __restrict__ xsio, const double* __restrict__ etao,
                                                  const double* __restrict__ dx, const double* __restrict__ dxsi,
                                                  const double* __restrict__ deta, double* __restrict__ x,
                                                  double* __restrict__ xsi, double* __restrict__ eta) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    x[j] = xo[j] + steg * dx[j];
    xsi[j] = xsio[j] + steg * dxsi[j];
    eta[j] = etao[j] + steg * deta[j];
}

__global__ __launch_bounds__(256) void k_mma_init(long n, const double* __restrict__ alfa, const double* __restrict__ beta,
                                                  double* __restrict__ x, double* __restrict__ xsi,
                                                  double* __restrict__ eta) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double xv = 0.5 * (alfa[j] + beta[j]);
    x[j] = xv;
    xsi[j] = fmax(1.0 / (xv - alfa[j]), 1.0);
    eta[j] = fmax(1.0 / (beta[j] - xv), 1.0);
}

// pitched FP32 design array <-> dense FP64 vector
__global__ __launch_bounds__(256) void k_f32_to_f64(long n, const float* __restrict__ a, double* __restrict__ b) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) b[j] = (double)a[j];
}
__global__ __launch_bounds__(256) void k_f64_to_f32(long n, const double* __restrict__ a, float* __restrict__ b) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) b[j] = (float)a[j];
}

__global__ void k_max_partials(int nb, const double* __restrict__ partial, double* __restrict__ out) {
    __shared__ double sh[4];
    double s = -1e300;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s = fmax(s, partial[i]);
    s = block_max<256>(s, sh);
    if (threadIdx.x == 0) out[0] = s;
}

}  // namespace

void launch_mma_setup(hipStream_t s, int m, long n, int iter, double move, const double* x, const double* xmin,
                      const double* xmax, const double* xold1, const double* xold2, const double* df0,
                      const double* dfdx, double* low, double* upp, double* alfa, double* beta, double* p0, double* q0,
                      double* P, double* Q, double* part, int nb, double* out) {
    hipLaunchKernelGGL(k_mma_setup, dim3(nb), dim3(256), 0, s, m, n, iter, move, x, xmin, xmax, xold1, xold2, df0, dfdx,
                       low, upp, alfa, beta, p0, q0, P, Q, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(m), dim3(kSumThreads), 0, s, nb, part, out);
}

void launch_mma_resid(hipStream_t s, int m, long n, const double* lam, double epsi, const double* x,
                      const double* xsi, const double* eta, const double* low, const double* upp, const double* alfa,
                      const double* beta, const double* p0, const double* q0, const double* P, const double* Q,
                      double* part, double* partmax, int nb, double* out) {
    MmaDual d{};
    for (int i = 0; i < m; i++) d.lam[i] = lam[i];
    d.epsi = epsi;
    hipLaunchKernelGGL(k_mma_resid, dim3(nb), dim3(256), 0, s, m, n, d, x, xsi, eta, low, upp, alfa, beta, p0, q0, P,
                       Q, part, partmax);
    hipLaunchKernelGGL(k_sum_partials, dim3(1 + m), dim3(kSumThreads), 0, s, nb, part, out);
    hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, s, nb, partmax, out + 1 + m);
}

void launch_mma_newton(hipStream_t s, int m, long n, const double* lam, double epsi, const double* x,
                       const double* xsi, const double* eta, const double* low, const double* upp, const double* alfa,
                       const double* beta, const double* p0, const double* q0, const double* P, const double* Q,
                       double* delx, double* diagx, double* part, int nb, double* out) {
    MmaDual d{};
    for (int i = 0; i < m; i++) d.lam[i] = lam[i];
    d.epsi = epsi;
    const int NS = kMmaMaxM + kMmaMaxM * (kMmaMaxM + 1) / 2;
    hipLaunchKernelGGL(k_mma_newton, dim3(nb), dim3(256), 0, s, m, n, d, x, xsi, eta, low, upp, alfa, beta, p0, q0, P,
                       Q, delx, diagx, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(NS), dim3(kSumThreads), 0, s, nb, part, out);
}

void launch_mma_dir(hipStream_t s, int m, long n, double epsi, const double* dlam, const double* x, const double* xsi,
                    const double* eta, const double* low, const double* upp, const double* alfa, const double* beta,
                    const double* P, const double* Q, const double* delx, const double* diagx, double* dx,
                    double* dxsi, double* deta, double* partmax, int nb, double* out) {
    MmaDual d{};
    d.epsi = epsi;
    hipLaunchKernelGGL(k_mma_dir, dim3(nb), dim3(256), 0, s, m, n, d, dlam, x, xsi, eta, low, upp, alfa, beta, P, Q,
                       delx, diagx, dx, dxsi, deta, partmax);
    hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, s, nb, partmax, out);
}

void launch_mma_step(hipStream_t s, long n, double steg, const double* xo, const double* xsio, const double* etao,
                     const double* dx, const double* dxsi, const double* deta, double* x, double* xsi, double* eta) {
    hipLaunchKernelGGL(k_mma_step, dim3(nblk(n)), dim3(256), 0, s, n, steg, xo, xsio, etao, dx, dxsi, deta, x, xsi,
                       eta);
}

void launch_mma_init(hipStream_t s, long n, const double* alfa, const double* beta, double* x, double* xsi,
                     double* eta) {
    hipLaunchKernelGGL(k_mma_init, dim3(nblk(n)), dim3(256), 0, s, n, alfa, beta, x, xsi, eta);
}

void launch_f32_to_f64(hipStream_t s, long n, const float* a, double* b) {
    hipLaunchKernelGGL(k_f32_to_f64, dim3(nblk(n)), dim3(256), 0, s, n, a, b);
}

void launch_f64_to_f32(hipStream_t s, long n, const double* a, float* b) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(nblk(n)), dim3(256), 0, s, n, a, b);
}

}  // namespace homo3d

// =============================================================================================
// OC update without allocation: homo3d_oc_update_noalloc
// =============================================================================================
// The bisection's reduction scratch (|s| maxima of kOcBlocks workgroups, kOcBlocks FP64 partial volume
// sums and their total) lives in this file-scope device array: every GPU the library runs on holds one
// copy until the process ends.  Its device address is looked up once per device and used for every
// launch; every launch rewrites what it reads, so nothing needs clearing between calls.  Calls that
// overlap on one GPU share the array and race.
namespace {
constexpr int kOcBlocks = 1024;
constexpr int kOcMaxReso = 1024;  // largest grid accepted (as the density-only handle)
constexpr int kOcScratchDoubles = kOcBlocks + kOcBlocks / 2 + 8;  // partials, float maxima, total
constexpr int kOcMaxDevices = 64;
static __device__ __align__(256) double scratch[kOcScratchDoubles];
double* g_scratch[kOcMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_ptr(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kOcMaxDevices) return HOMO3D_HIP_ERROR;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HOMO3D_HIP_ERROR;
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return HOMO3D_OK;
}
}  // namespace

extern "C" int homo3d_oc_update_noalloc(const float* sens, float* rho, int n, int pitch, double vol, double min_rho,
                                        double step, double damp, void* stream) {
    using namespace homo3d;
    if (!sens || !rho || n < 1 || pitch < n) return HOMO3D_BAD_ARG;
    if (n > kOcMaxReso) return HOMO3D_BAD_ARG;  // beyond the stated maximum grid
    double* sc = nullptr;
    int rc = scratch_ptr(&sc);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* partial = sc;
    float* pmax = reinterpret_cast<float*>(sc + kOcBlocks);
    double* total = sc + kOcBlocks + kOcBlocks / 2;
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_absmax, dim3(kOcBlocks), dim3(256), 0, s, N, n, pitch, sens, pmax);
    std::vector<float> pm(kOcBlocks);
    if (hipMemcpyAsync(pm.data(), pmax, sizeof(float) * kOcBlocks, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return HOMO3D_HIP_ERROR;
    double hi = 0, lo = 0, g = 0;
    for (float v : pm) hi = std::max(hi, (double)v);
    const double Ne = (double)N;
    // the same bisection as homo3d_oc_update (OCOptimizer.cu:218-247): the trials only sum, and the
    // last trial's multiplier is applied in place at the end
    for (int it = 0; it < 20; it++) {
        g = (hi + lo) / 2;
        double sum = 0;
        hipLaunchKernelGGL(k_oc_trial, dim3(kOcBlocks), dim3(256), 0, s, n, pitch, g, min_rho, step, damp, sens,
                           (const float*)rho, (float*)nullptr, partial);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kSumThreads), 0, s, kOcBlocks, (const double*)partial, total);
        if (hipMemcpyAsync(&sum, total, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return HOMO3D_HIP_ERROR;
        const double cur = sum / Ne;
        if (cur < vol - 1e-4)
            hi = g;
        else if (cur > vol + 1e-4)
            lo = g;
        else
            break;
    }
    hipLaunchKernelGGL(k_oc_apply, dim3(kOcBlocks), dim3(256), 0, s, n, pitch, g, min_rho, step, damp, sens, rho);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return HOMO3D_HIP_ERROR;
    return HOMO3D_OK;
}
