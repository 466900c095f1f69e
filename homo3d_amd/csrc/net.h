// Internal interface of the learned warm-start network (net.hip) for the solver handle's entry points
// (api.cpp).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct homo3d_net;

namespace homo3d {

// Network grid g (cells per axis the network runs at) and the output scale: the network predicts
// target_scale * u / n, mean-free per channel.
int net_grid(const homo3d_net* net);
int net_pred_ld(const homo3d_net* net);  // row stride (floats) of the prediction net_forward_one returns
float net_target_scale(const homo3d_net* net);
hipStream_t net_stream(const homo3d_net* net);
// Input slot of batch element 0 (bf16, [g^3][8], channel 0 = density, channels 1..7 = 0).
uint16_t* net_input_slot(homo3d_net* net);
// Forward pass of batch element 0 only, on the net's stream; returns the prediction [g^3][18] (FP32).
int net_forward_one(homo3d_net* net, const float** pred);
// Dataset slot for a new pair (input [g^3][8] and target [g^3][18], bf16), or nullptr when full.
int net_dataset_push(homo3d_net* net, uint16_t** input, uint16_t** target);
int net_set_error(homo3d_net* net, int code, const char* msg);

// bf16 helpers shared by the bridge kernels (kernels.hip)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

}  // namespace homo3d
