// The learned face closure's network, 16 -> 64 -> 64 -> 2 with ReLU, on CDNA matrix cores.
//
// Weights are BF16, stored transposed (out x in: the rows a B operand reads are contiguous); biases and
// accumulation are FP32. Activations between layers are rounded to BF16. A wave evaluates 16 samples. The first
// layer (K = 16) runs on v_mfma_f32_16x16x16_bf16: lane l holds A[row l%16][k 4(l/16)..+3] and B[k 4(l/16)..+3][col
// l%16] (one 8-byte access each). The two K = 64 layers run on gfx950's v_mfma_f32_16x16x32_bf16, two issues per
// output tile instead of four: lane l holds A[row l%16][k 8(l/16)..+7] and the same k of B (one 16-byte access
// each; any assignment of k to lanes gives the same product so long as A and B share it). Either way lane l receives
// D[row 4(l/16)+j][col l%16], j = 0..3.
//
// FP32 parameter vector (the optimiser's master copy), in this order:
//   W1 [64][16], b1 [64], W2 [64][64], b2 [64], W3 [2][64], b3 [2]            (MLP_NPARAM floats)
// BF16 weight block: W1 [64][16], W2 [64][64], W3 padded to [16][64] (rows 2..15 zero)   (MLP_NBF16 values)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace w2d_mlp {

constexpr int IN = 16, HID = 64, OUT = 2;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + HID * IN, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
              OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + OUT * HID, NPARAM = OFF_B3 + OUT;
constexpr int BF_W1 = 0, BF_W2 = BF_W1 + HID * IN, BF_W3 = BF_W2 + HID * HID, NBF16 = BF_W3 + 16 * HID;
constexpr int WAVE_SAMPLES = 16;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> BF16 bits, round to nearest even (finite inputs; the network's values are bounded).
__host__ __device__ inline uint16_t to_bf16(float f) {
    union {
        float f;
        uint32_t u;
    } v{f};
    const uint32_t u = v.u;
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t b) {
    union {
        uint32_t u;
        float f;
    } v{(uint32_t)b << 16};
    return v.f;
}

// LDS image of the network for one workgroup.
struct alignas(16) Net {
    uint16_t w[NBF16];
    float b1[HID], b2[HID], b3[OUT];
};

// Stage the BF16 weights and FP32 biases into LDS (every thread of the workgroup takes part; no barrier here).
__device__ inline void load_net(Net& n, const uint16_t* __restrict__ wbf, const float* __restrict__ params) {
    for (int q = threadIdx.x; q < NBF16 / 4; q += blockDim.x)
        reinterpret_cast<uint2*>(n.w)[q] = reinterpret_cast<const uint2*>(wbf)[q];
    for (int q = threadIdx.x; q < HID; q += blockDim.x) {
        n.b1[q] = params[OFF_B1 + q];
        n.b2[q] = params[OFF_B2 + q];
    }
    if (threadIdx.x < OUT) n.b3[threadIdx.x] = params[OFF_B3 + threadIdx.x];
}

__device__ inline bf16x4 ld4(const uint16_t* p) {
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFF);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFF);
    r[3] = (short)(v.y >> 16);
    return r;
}

// Eight BF16 values from a 16-byte aligned address.
__device__ inline bf16x8 ld8(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p)); }

// One wave, 16 samples: x [16][IN] BF16 (LDS) -> h1, h2 [16][HID] BF16 (LDS, post-ReLU) and y [16][OUT] FP32
// (LDS). The wave's lanes all take part; the caller orders LDS between waves.
__device__ inline void forward_wave(const Net& n, const uint16_t* x, uint16_t* h1, uint16_t* h2, float* y) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, k0 = 4 * (lane >> 4);
    const bf16x4 a1 = ld4(x + r * IN + k0);
#pragma unroll
    for (int nt = 0; nt < HID / 16; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, ld4(n.w + BF_W1 + (nt * 16 + r) * IN + k0), acc, 0, 0, 0);
        const int col = nt * 16 + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float z = acc[j] + n.b1[col];
            h1[(k0 + j) * HID + col] = to_bf16(z > 0.f ? z : 0.f);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    const int k8 = 8 * (lane >> 4);
#pragma unroll
    for (int nt = 0; nt < HID / 16; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < HID / 32; ++kt)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(h1 + r * HID + kt * 32 + k8),
                                                          ld8(n.w + BF_W2 + (nt * 16 + r) * HID + kt * 32 + k8), acc,
                                                          0, 0, 0);
        const int col = nt * 16 + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float z = acc[j] + n.b2[col];
            h2[(k0 + j) * HID + col] = to_bf16(z > 0.f ? z : 0.f);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < HID / 32; ++kt)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(h2 + r * HID + kt * 32 + k8),
                                                      ld8(n.w + BF_W3 + r * HID + kt * 32 + k8), acc, 0, 0, 0);
    if (r < OUT)
#pragma unroll
        for (int j = 0; j < 4; ++j) y[(k0 + j) * OUT + r] = acc[j] + n.b3[r];
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
}

}  // namespace w2d_mlp
