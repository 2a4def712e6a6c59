// Training and batch inference of the learned face closure's network (mlp.hpp): the forward pass on matrix cores,
// the backward pass and Adam in FP32, the weight gradient reduced over workgroups in a fixed order. The reference
// has no learned component; these entry points serve the closure's training driver (closure/train.py).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mlp.hpp"
#include "wadi2d_hip.h"

using namespace w2d_mlp;

int w2d_report(int code, const char* msg);  // csrc/step.hip: the message w2d_last_error returns

namespace {

constexpr int MT = 256;             // threads per workgroup: 4 waves
constexpr int CHUNK = 4 * WAVE_SAMPLES;  // samples per workgroup pass

struct alignas(16) Work {
    uint16_t x[CHUNK * IN];
    uint16_t h1[CHUNK * HID];
    uint16_t h2[CHUNK * HID];
    float y[CHUNK * OUT];
    float dy[CHUNK * OUT];
    float dz2[CHUNK * HID];
    float dz1[CHUNK * HID];
};

// Forward of one chunk of samples [s0, s0 + CHUNK) (rows past n are zero inputs; their outputs are ignored).
__device__ void forward_chunk(const Net& net, Work& w, const uint16_t* __restrict__ x, long long s0, long long n) {
    for (int q = threadIdx.x; q < CHUNK * IN / 4; q += MT) {
        const long long s = s0 + (q * 4) / IN;
        uint2 v = make_uint2(0u, 0u);
        if (s < n) v = reinterpret_cast<const uint2*>(x + s0 * IN)[q];
        reinterpret_cast<uint2*>(w.x)[q] = v;
    }
    __syncthreads();
    const int wv = threadIdx.x >> 6;
    forward_wave(net, w.x + wv * WAVE_SAMPLES * IN, w.h1 + wv * WAVE_SAMPLES * HID, w.h2 + wv * WAVE_SAMPLES * HID,
                 w.y + wv * WAVE_SAMPLES * OUT);
    __syncthreads();
}

__global__ __launch_bounds__(MT) void forward_kernel(const uint16_t* __restrict__ wbf, const float* __restrict__ params,
                                                     const uint16_t* __restrict__ x, long long n,
                                                     float* __restrict__ y) {
    __shared__ Net net;
    __shared__ Work w;
    load_net(net, wbf, params);
    __syncthreads();
    for (long long s0 = (long long)blockIdx.x * CHUNK; s0 < n; s0 += (long long)gridDim.x * CHUNK) {
        forward_chunk(net, w, x, s0, n);
        for (int q = threadIdx.x; q < CHUNK * OUT; q += MT)
            if (s0 + q / OUT < n) y[s0 * OUT + q] = w.y[q];
        __syncthreads();
    }
}

// Masked mean-square loss over n samples, sum over outputs of mean over samples of mask * (y - t)^2, divided by
// `norm` (the caller's sample count). Each workgroup leaves its gradient sum and loss sum in partial[block].
__global__ __launch_bounds__(MT) void train_kernel(const uint16_t* __restrict__ wbf, const float* __restrict__ params,
                                                   const uint16_t* __restrict__ x, const float* __restrict__ t,
                                                   const float* __restrict__ mask, long long n, float norm,
                                                   float* __restrict__ partial) {
    __shared__ Net net;
    __shared__ Work w;
    __shared__ float red[MT / 64];
    load_net(net, wbf, params);
    __syncthreads();
    const int tid = threadIdx.x;
    // gradient entries this thread owns: W2 16, W1 4, W3 and biases one each below 128 / 64 / 2
    float g_w2[16], g_w1[4], g_w3 = 0.f, g_b1 = 0.f, g_b2 = 0.f, g_b3 = 0.f, loss = 0.f;
#pragma unroll
    for (int q = 0; q < 16; ++q) g_w2[q] = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) g_w1[q] = 0.f;
    const float inv = 1.f / norm;
    for (long long s0 = (long long)blockIdx.x * CHUNK; s0 < n; s0 += (long long)gridDim.x * CHUNK) {
        forward_chunk(net, w, x, s0, n);
        if (tid < CHUNK * OUT) {
            const long long s = s0 + tid / OUT;
            float d = 0.f;
            if (s < n) {
                const float e = w.y[tid] - t[s0 * OUT + tid];
                const float m = mask[s0 * OUT + tid];
                loss += m * e * e * inv;
                d = 2.f * m * e * inv;
            }
            w.dy[tid] = d;
        }
        __syncthreads();
        // dz2[s][k] = relu'(h2) * sum_c W3[c][k] dy[s][c]
        for (int q = tid; q < CHUNK * HID; q += MT) {
            const int s = q / HID, k = q % HID;
            float a = 0.f;
#pragma unroll
            for (int c = 0; c < OUT; ++c) a += from_bf16(net.w[BF_W3 + c * HID + k]) * w.dy[s * OUT + c];
            w.dz2[q] = from_bf16(w.h2[q]) > 0.f ? a : 0.f;
        }
        __syncthreads();
        // dz1[s][k] = relu'(h1) * sum_n W2[n][k] dz2[s][n]
        for (int q = tid; q < CHUNK * HID; q += MT) {
            const int s = q / HID, k = q % HID;
            float a = 0.f;
            for (int m = 0; m < HID; ++m) a += from_bf16(net.w[BF_W2 + m * HID + k]) * w.dz2[s * HID + m];
            w.dz1[q] = from_bf16(w.h1[q]) > 0.f ? a : 0.f;
        }
        __syncthreads();
        // weight gradients, summed over the chunk's samples in sample order
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = tid * 16 + q, o = e / HID, k = e % HID;
            float a = 0.f;
            for (int s = 0; s < CHUNK; ++s) a += w.dz2[s * HID + o] * from_bf16(w.h1[s * HID + k]);
            g_w2[q] += a;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid * 4 + q, o = e / IN, k = e % IN;
            float a = 0.f;
            for (int s = 0; s < CHUNK; ++s) a += w.dz1[s * HID + o] * from_bf16(w.x[s * IN + k]);
            g_w1[q] += a;
        }
        if (tid < OUT * HID) {
            const int o = tid / HID, k = tid % HID;
            float a = 0.f;
            for (int s = 0; s < CHUNK; ++s) a += w.dy[s * OUT + o] * from_bf16(w.h2[s * HID + k]);
            g_w3 += a;
        }
        if (tid < HID) {
            float a1 = 0.f, a2 = 0.f;
            for (int s = 0; s < CHUNK; ++s) {
                a1 += w.dz1[s * HID + tid];
                a2 += w.dz2[s * HID + tid];
            }
            g_b1 += a1;
            g_b2 += a2;
        }
        if (tid < OUT) {
            float a = 0.f;
            for (int s = 0; s < CHUNK; ++s) a += w.dy[s * OUT + tid];
            g_b3 += a;
        }
        __syncthreads();
    }
    float* p = partial + (size_t)blockIdx.x * (NPARAM + 1);
#pragma unroll
    for (int q = 0; q < 16; ++q) p[OFF_W2 + tid * 16 + q] = g_w2[q];
#pragma unroll
    for (int q = 0; q < 4; ++q) p[OFF_W1 + tid * 4 + q] = g_w1[q];
    if (tid < OUT * HID) p[OFF_W3 + tid] = g_w3;
    if (tid < HID) {
        p[OFF_B1 + tid] = g_b1;
        p[OFF_B2 + tid] = g_b2;
    }
    if (tid < OUT) p[OFF_B3 + tid] = g_b3;
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if ((tid & 63) == 0) red[tid >> 6] = loss;
    __syncthreads();
    if (tid == 0) p[NPARAM] = red[0] + red[1] + red[2] + red[3];
}

// grad[e] = sum over workgroups, in workgroup order, of partial[b][e] (e = NPARAM is the loss).
__global__ void reduce_kernel(const float* __restrict__ partial, int blocks, float* __restrict__ grad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > NPARAM) return;
    float a = 0.f;
    for (int b = 0; b < blocks; ++b) a += partial[(size_t)b * (NPARAM + 1) + e];
    grad[e] = a;
}

// Adam on the FP32 master copy, then the BF16 weight block from it.
__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2,
                            uint16_t* __restrict__ wbf) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < NPARAM) {
        const float ge = g[e];
        const float me = b1 * m[e] + (1.f - b1) * ge;
        const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
        m[e] = me;
        v[e] = ve;
        p[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
    }
}

__global__ void pack_kernel(const float* __restrict__ p, uint16_t* __restrict__ wbf) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < HID * IN) wbf[BF_W1 + e] = to_bf16(p[OFF_W1 + e]);
    if (e < HID * HID) wbf[BF_W2 + e] = to_bf16(p[OFF_W2 + e]);
    if (e < 16 * HID) wbf[BF_W3 + e] = e < OUT * HID ? to_bf16(p[OFF_W3 + e]) : (uint16_t)0;
}

// Rows idx[0..n) of x, t and mask into dense arrays (a minibatch).
__global__ void take_kernel(const uint16_t* __restrict__ x, const float* __restrict__ t, const float* __restrict__ m,
                            const int64_t* __restrict__ idx, long long n, uint16_t* __restrict__ xo,
                            float* __restrict__ to, float* __restrict__ mo) {
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long long s = idx[q];
    const uint2* xs = reinterpret_cast<const uint2*>(x + s * IN);
    uint2* xd = reinterpret_cast<uint2*>(xo + q * IN);
#pragma unroll
    for (int k = 0; k < IN / 4; ++k) xd[k] = xs[k];
#pragma unroll
    for (int c = 0; c < OUT; ++c) {
        to[q * OUT + c] = t[s * OUT + c];
        mo[q * OUT + c] = m[s * OUT + c];
    }
}

int blocks_for(long long n, int max_blocks) {
    long long b = (n + CHUNK - 1) / CHUNK;
    if (b > max_blocks) b = max_blocks;
    return b < 1 ? 1 : (int)b;
}

int check_launch() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : w2d_report(3, hipGetErrorString(e));
}

}  // namespace

extern "C" {

int w2d_dev_alloc(int64_t bytes, void** out) {
    if (!out || bytes < 0) return w2d_report(1, "dev_alloc: bad arguments");
    *out = nullptr;
    if (bytes == 0) return 0;
    if (hipMalloc(out, (size_t)bytes) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return w2d_report(2, "dev_alloc: out of device memory");
    }
    return 0;
}

int w2d_dev_free(void* p) {
    if (p && hipFree(p) != hipSuccess) return w2d_report(3, "dev_free: hipFree failed");
    return 0;
}

int w2d_dev_copy(void* dst, const void* src, int64_t bytes, int32_t kind) {
    if (((!dst || !src) && bytes) || bytes < 0 || kind < 0 || kind > 2) return w2d_report(1, "dev_copy: bad arguments");
    if (bytes == 0) return 0;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    const hipError_t e = hipMemcpy(dst, src, (size_t)bytes, k);
    return e == hipSuccess ? 0 : w2d_report(3, hipGetErrorString(e));
}

int w2d_dev_synchronize(void) {
    const hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? 0 : w2d_report(3, hipGetErrorString(e));
}

int w2d_mlp_take(const uint16_t* x, const float* target, const float* mask, const int64_t* idx, int64_t n,
                 uint16_t* x_out, float* target_out, float* mask_out, void* stream) {
    if (n < 0 || (n && (!x || !target || !mask || !idx || !x_out || !target_out || !mask_out)))
        return w2d_report(1, "mlp_take: bad arguments");
    if (n == 0) return 0;
    take_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(x, target, mask, idx, n, x_out,
                                                                             target_out, mask_out);
    return check_launch();
}

int w2d_mlp_sizes(int32_t* nparam, int32_t* nbf16) {
    if (!nparam || !nbf16) return w2d_report(1, "mlp_sizes: bad arguments");
    *nparam = NPARAM;
    *nbf16 = NBF16;
    return 0;
}

int w2d_mlp_pack(const float* params, uint16_t* wbf16, void* stream) {
    if (!params || !wbf16) return w2d_report(1, "mlp_pack: bad arguments");
    pack_kernel<<<(HID * HID + 255) / 256, 256, 0, (hipStream_t)stream>>>(params, wbf16);
    return check_launch();
}

int w2d_mlp_forward(const uint16_t* wbf16, const float* params, const uint16_t* x, int64_t n, float* y,
                    void* stream) {
    if (!wbf16 || !params || (n && (!x || !y)) || n < 0) return w2d_report(1, "mlp_forward: bad arguments");
    if (n == 0) return 0;
    forward_kernel<<<blocks_for(n, 4096), MT, 0, (hipStream_t)stream>>>(wbf16, params, x, n, y);
    return check_launch();
}

int w2d_mlp_train_blocks(int64_t n, int32_t* blocks) {
    if (!blocks || n < 0) return w2d_report(1, "mlp_train_blocks: bad arguments");
    *blocks = blocks_for(n, 512);
    return 0;
}

int w2d_mlp_gradient(const uint16_t* wbf16, const float* params, const uint16_t* x, const float* target,
                     const float* mask, int64_t n, float norm, float* partial, int32_t blocks, float* grad,
                     void* stream) {
    if (!wbf16 || !params || !x || !target || !mask || !partial || !grad || n < 1 || blocks < 1 || !(norm > 0.f))
        return w2d_report(1, "mlp_gradient: bad arguments");
    if (blocks != blocks_for(n, 512)) return w2d_report(1, "mlp_gradient: blocks is not w2d_mlp_train_blocks(n)");
    hipStream_t st = (hipStream_t)stream;
    train_kernel<<<blocks, MT, 0, st>>>(wbf16, params, x, target, mask, n, norm, partial);
    if (int s = check_launch()) return s;
    reduce_kernel<<<(NPARAM + 1 + 255) / 256, 256, 0, st>>>(partial, blocks, grad);
    return check_launch();
}

int w2d_mlp_adam(float* params, float* m, float* v, const float* grad, int32_t step, float lr, float beta1,
                 float beta2, float eps, uint16_t* wbf16, void* stream) {
    if (!params || !m || !v || !grad || !wbf16 || step < 1) return w2d_report(1, "mlp_adam: bad arguments");
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipStream_t st = (hipStream_t)stream;
    adam_kernel<<<(NPARAM + 255) / 256, 256, 0, st>>>(params, m, v, grad, lr, beta1, beta2, eps, c1, c2, wbf16);
    if (int s = check_launch()) return s;
    pack_kernel<<<(HID * HID + 255) / 256, 256, 0, st>>>(params, wbf16);
    return check_launch();
}

}  // extern "C"
