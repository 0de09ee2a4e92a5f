// Learned absorbing boundary: the network and its training on the GPU.
//
// Network: three 3x3 convolutions, 5 -> 16 -> 16 -> 2 channels, zero padding,
// no bias, identity activations.  The boundary operator must be linear in the
// wavefield (superposition, amplitude scaling), so the layers are linear; the
// three layers then compose into one 7x7 filter from 5 fields to 2 velocity
// corrections, which is what the propagation kernel applies.
// Weights are BF16 (round-to-nearest-even of an FP32 master copy) and every sum
// accumulates in FP32.  Training: mean squared error over masked pixels, Adam.
//
// Layouts: activations [M][C][H][W] (W fastest), weights [Cout][Cin][3][3].
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

using namespace ouc;

namespace {

constexpr int kC[4] = {5, 16, 16, 2};
constexpr int kWSizes[3] = {16 * 5 * 9, 16 * 16 * 9, 2 * 16 * 9};

__device__ __forceinline__ float bf(const __hip_bfloat16 v) { return __bfloat162float(v); }

// out[m][co][y][x] = sum_ci sum_ky,kx w[co][ci][ky][kx] * in[m][ci][y+ky-1][x+kx-1]
__global__ void k_conv_fwd(int M, int Ci, int Co, int H, int W, const float *__restrict__ in,
                           const __hip_bfloat16 *__restrict__ w, float *__restrict__ out) {
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)M * Co * H * W;
    if (idx >= total) return;
    int x = idx % W, y = (idx / W) % H, co = (idx / ((long)W * H)) % Co;
    long m = idx / ((long)W * H * Co);
    float acc = 0.f;
    for (int ci = 0; ci < Ci; ++ci) {
        const float *ip = in + ((m * Ci + ci) * H) * W;
        const __hip_bfloat16 *wp = w + ((long)co * Ci + ci) * 9;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            int yy = y + ky - 1;
            if (yy < 0 || yy >= H) continue;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                int xx = x + kx - 1;
                if (xx < 0 || xx >= W) continue;
                acc += bf(wp[ky * 3 + kx]) * ip[yy * W + xx];
            }
        }
    }
    out[idx] = acc;
}

// din[m][ci][y][x] = sum_co sum_k w[co][ci][k] * dout[m][co][y-ky+1][x-kx+1]
__global__ void k_conv_bwd_data(int M, int Ci, int Co, int H, int W, const float *__restrict__ dout,
                                const __hip_bfloat16 *__restrict__ w, float *__restrict__ din) {
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)M * Ci * H * W;
    if (idx >= total) return;
    int x = idx % W, y = (idx / W) % H, ci = (idx / ((long)W * H)) % Ci;
    long m = idx / ((long)W * H * Ci);
    float acc = 0.f;
    for (int co = 0; co < Co; ++co) {
        const float *dp = dout + ((m * Co + co) * H) * W;
        const __hip_bfloat16 *wp = w + ((long)co * Ci + ci) * 9;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            int yy = y - ky + 1;
            if (yy < 0 || yy >= H) continue;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                int xx = x - kx + 1;
                if (xx < 0 || xx >= W) continue;
                acc += bf(wp[ky * 3 + kx]) * dp[yy * W + xx];
            }
        }
    }
    din[idx] = acc;
}

// dw[co][ci][ky][kx] = sum_m,y,x dout[m][co][y][x] * in[m][ci][y+ky-1][x+kx-1];
// one block per weight, a block-wide reduction, no atomics (deterministic).
__global__ void k_conv_bwd_weight(int M, int Ci, int Co, int H, int W, const float *__restrict__ in,
                                  const float *__restrict__ dout, float *__restrict__ dw) {
    const int wi = blockIdx.x; // over Co*Ci*9
    const int k = wi % 9, ci = (wi / 9) % Ci, co = wi / (9 * Ci);
    const int ky = k / 3, kx = k % 3;
    const long npos = (long)M * H * W;
    float acc = 0.f;
    for (long p = threadIdx.x; p < npos; p += blockDim.x) {
        int x = p % W, y = (p / W) % H;
        long m = p / ((long)W * H);
        int yy = y + ky - 1, xx = x + kx - 1;
        if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
        acc += dout[((m * Co + co) * H + y) * W + x] * in[((m * Ci + ci) * H + yy) * W + xx];
    }
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) dw[wi] = red[0];
}

// Masked MSE: loss = sum mask*(out-y)^2 / sum mask; dout = 2 mask (out-y) / sum mask.
__global__ void k_mse(long n, int HW, int C, const float *__restrict__ out, const float *__restrict__ y,
                      const float *__restrict__ mask, float inv_count, float *__restrict__ dout,
                      float *__restrict__ partial) {
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (idx < n) {
        long m = idx / ((long)C * HW);
        int pix = idx % HW;
        float mk = mask[m * HW + pix];
        float d = out[idx] - y[idx];
        dout[idx] = 2.f * mk * d * inv_count;
        e = mk * d * d;
    }
    __shared__ float red[256];
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_round_bf16(int n, const float *__restrict__ w, __hip_bfloat16 *__restrict__ wb) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = __float2bfloat16(w[i]);
}

__global__ void k_adam(int n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float a = b1 * m1[i] + (1.f - b1) * gi, b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

constexpr int kNW = kWSizes[0] + kWSizes[1] + kWSizes[2];

// The forward pass: BF16 rounding of the FP32 master weights w32 into wb16,
// then the three convolutions act[0] -> act[1] -> act[2] -> act[3].
void forward_layers(hipStream_t s, int M, int H, int W, const float *w32, __hip_bfloat16 *wb16,
                    const float *const act[4]) {
    const long HW = (long)H * W;
    const int off[3] = {0, kWSizes[0], kWSizes[0] + kWSizes[1]};
    k_round_bf16<<<(kNW + 255) / 256, 256, 0, s>>>(kNW, w32, wb16);
    for (int l = 0; l < 3; ++l) {
        long tot = (long)M * kC[l + 1] * HW;
        k_conv_fwd<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(M, kC[l], kC[l + 1], H, W, act[l], wb16 + off[l],
                                                               (float *)act[l + 1]);
    }
    OUC_HIP(hipGetLastError());
}

struct Net {
    int M = 0, H = 0, W = 0;
    DevBuf w32, wb16, g32, m1, m2;      // all layers concatenated
    DevBuf act[4], dact[4];             // activations and their gradients
    DevBuf x, y, mask, dout, partial;
    int off[3] = {0, kWSizes[0], kWSizes[0] + kWSizes[1]};
    int nw = kNW;

    void alloc(int M_, int H_, int W_) {
        M = M_; H = H_; W = W_;
        if ((long)M * 16 * H * W > (1L << 31)) throw Error(OUC_INVALID_ARG, "training set too large");
        w32.alloc(nw * 4); wb16.alloc(nw * 2); g32.alloc(nw * 4); m1.alloc(nw * 4); m2.alloc(nw * 4);
        for (int l = 0; l < 4; ++l) {
            act[l].alloc((size_t)M * kC[l] * H * W * 4);
            dact[l].alloc((size_t)M * kC[l] * H * W * 4);
        }
        y.alloc((size_t)M * 2 * H * W * 4);
        mask.alloc((size_t)M * H * W * 4);
        partial.alloc(((size_t)M * 2 * H * W / 256 + 2) * 4);
    }
    __hip_bfloat16 *wl(int l) const { return wb16.as<__hip_bfloat16>() + off[l]; }

    void forward(hipStream_t s) {
        const float *acts[4] = {act[0].f(), act[1].f(), act[2].f(), act[3].f()};
        forward_layers(s, M, H, W, w32.f(), wb16.as<__hip_bfloat16>(), acts);
    }

    // forward + loss + backward; returns the loss
    float pass(hipStream_t s, bool backward, float mask_count) {
        const int HW = H * W;
        forward(s);
        long n = (long)M * 2 * HW;
        unsigned nb = (unsigned)((n + 255) / 256);
        k_mse<<<nb, 256, 0, s>>>(n, HW, 2, act[3].f(), y.f(), mask.f(), 1.f / mask_count, dact[3].f(), partial.f());
        std::vector<float> part(nb);
        OUC_HIP(hipMemcpyAsync(part.data(), partial.ptr, nb * 4, hipMemcpyDeviceToHost, s));
        if (backward) {
            for (int l = 2; l >= 0; --l) {
                k_conv_bwd_weight<<<kC[l + 1] * kC[l] * 9, 256, 0, s>>>(M, kC[l], kC[l + 1], H, W, act[l].f(),
                                                                       dact[l + 1].f(), g32.f() + off[l]);
                if (l > 0) {
                    long tot = (long)M * kC[l] * HW;
                    k_conv_bwd_data<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(M, kC[l], kC[l + 1], H, W,
                                                                                dact[l + 1].f(), wl(l), dact[l].f());
                }
            }
        }
        OUC_HIP(hipStreamSynchronize(s));
        double loss = 0;
        for (float v : part) loss += v;
        return (float)(loss / mask_count);
    }
};

ouc_status lb_fail(const std::string &m, ouc_status c) { return ouc_fail_global(c, m); }

float mask_total(const float *mask, long n) {
    double s = 0;
    for (long i = 0; i < n; ++i) s += mask[i];
    return (float)s;
}

// Scratch of the allocation-free forward pass (ouc_lb_forward_device): the
// BF16 copy of the weights in the first kLbWeightBytes, then the two hidden
// activation maps [M][16][H][W] of at most kLbMaxPixels = M*H*W pixels each.
// A file-scope device array: every GPU the library runs on holds its own copy
// (16 MiB + 8 KiB) until the process ends, and every call on that GPU shares it.
constexpr long kLbMaxPixels = 131072;
constexpr size_t kLbWeightBytes = 8192;
static_assert(kLbWeightBytes >= kNW * sizeof(__hip_bfloat16), "weight part of the scratch too small");
constexpr size_t kLbScratchBytes = kLbWeightBytes + 2 * sizeof(float) * 16 * (size_t)kLbMaxPixels;
static __device__ __align__(256) unsigned char scratch[kLbScratchBytes];

// Device address of `scratch` on each GPU, looked up once per device with that
// device current.  The array's name is a host-side stand-in in host code;
// only these addresses are passed to kernels.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

void *scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(OUC_INVALID_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        OUC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(OUC_DEVICE, "scratch has no device address");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

} // namespace

extern "C" {

// Forward pass of the layered network (BF16 weights, FP32 accumulation):
// x [M][5][H][W] -> out [M][2][H][W], zero padding at the window edges.
ouc_status ouc_lb_forward(int32_t device, int32_t M, int32_t H, int32_t W, const float *x, const float *weights,
                          float *out) {
    if (M < 1 || H < 1 || W < 1 || !x || !weights || !out) return lb_fail("bad arguments", OUC_INVALID_ARG);
    try {
        OUC_HIP(hipSetDevice(device));
        Net net;
        net.alloc(M, H, W);
        const size_t HW = (size_t)H * W;
        OUC_HIP(hipMemcpy(net.act[0].ptr, x, (size_t)M * 5 * HW * 4, hipMemcpyHostToDevice));
        OUC_HIP(hipMemcpy(net.w32.ptr, weights, net.nw * 4, hipMemcpyHostToDevice));
        net.forward(nullptr);
        OUC_HIP(hipMemcpy(out, net.act[3].ptr, (size_t)M * 2 * HW * 4, hipMemcpyDeviceToHost));
        return OUC_OK;
    } catch (const Error &e) {
        return lb_fail(e.what(), e.code);
    }
}

// ouc_lb_forward on device buffers, with no allocation: the weight copy and
// the hidden activations live in the per-GPU `scratch` array; the four
// launches are queued on `stream` (NULL: the null stream) and the call returns
// without waiting for them.
ouc_status ouc_lb_forward_device(int32_t device, int32_t M, int32_t H, int32_t W, const float *x,
                                 const float *weights, float *out, void *stream) {
    if (M < 1 || H < 1 || W < 1 || !x || !weights || !out) return lb_fail("bad arguments", OUC_INVALID_ARG);
    if ((long)M * H * W > kLbMaxPixels)
        return lb_fail("M*H*W = " + std::to_string((long)M * H * W) + " exceeds the scratch maximum of " +
                           std::to_string(kLbMaxPixels) + " pixels",
                       OUC_INVALID_ARG);
    try {
        OUC_HIP(hipSetDevice(device));
        unsigned char *base = (unsigned char *)scratch_on(device);
        __hip_bfloat16 *wb16 = (__hip_bfloat16 *)base;
        float *h1 = (float *)(base + kLbWeightBytes);
        float *h2 = h1 + (size_t)16 * M * H * W;
        const float *acts[4] = {x, h1, h2, out};
        forward_layers((hipStream_t)stream, M, H, W, weights, wb16, acts);
        return OUC_OK;
    } catch (const Error &e) {
        return lb_fail(e.what(), e.code);
    }
}

// Loss and gradient of the boundary network (the BF16-rounded weights are used
// in the forward pass; the gradient is with respect to them).  x: [M][5][H][W],
// y: [M][2][H][W], mask: [M][H][W], weights: FP32 [16*5*9 + 16*16*9 + 2*16*9].
ouc_status ouc_lb_loss_grad(int32_t device, int32_t M, int32_t H, int32_t W, const float *x, const float *y,
                            const float *mask, const float *weights, float *loss, float *grad) {
    if (M < 1 || H < 1 || W < 1 || !x || !y || !mask || !weights || !loss) return lb_fail("bad arguments", OUC_INVALID_ARG);
    try {
        OUC_HIP(hipSetDevice(device));
        Net net;
        net.alloc(M, H, W);
        const size_t HW = (size_t)H * W;
        OUC_HIP(hipMemcpy(net.act[0].ptr, x, (size_t)M * 5 * HW * 4, hipMemcpyHostToDevice));
        OUC_HIP(hipMemcpy(net.y.ptr, y, (size_t)M * 2 * HW * 4, hipMemcpyHostToDevice));
        OUC_HIP(hipMemcpy(net.mask.ptr, mask, (size_t)M * HW * 4, hipMemcpyHostToDevice));
        OUC_HIP(hipMemcpy(net.w32.ptr, weights, net.nw * 4, hipMemcpyHostToDevice));
        float cnt = mask_total(mask, (long)M * HW) * 2.f;
        if (!(cnt > 0)) return lb_fail("empty mask", OUC_INVALID_ARG);
        *loss = net.pass(nullptr, grad != nullptr, cnt);
        if (grad) OUC_HIP(hipMemcpy(grad, net.g32.ptr, net.nw * 4, hipMemcpyDeviceToHost));
        return OUC_OK;
    } catch (const Error &e) {
        return lb_fail(e.what(), e.code);
    }
}

// Train the boundary network with Adam on the GPU.  weights: in/out FP32
// master weights; history: `epochs` losses (full-batch gradient steps).
ouc_status ouc_lb_train(int32_t device, int32_t M, int32_t H, int32_t W, const float *x, const float *y,
                        const float *mask, int32_t epochs, float lr, float *weights, float *history) {
    if (M < 1 || H < 1 || W < 1 || epochs < 1 || !x || !y || !mask || !weights) return lb_fail("bad arguments", OUC_INVALID_ARG);
    try {
        OUC_HIP(hipSetDevice(device));
        Net net;
        net.alloc(M, H, W);
        const size_t HW = (size_t)H * W;
        OUC_HIP(hipMemcpy(net.act[0].ptr, x, (size_t)M * 5 * HW * 4, hipMemcpyHostToDevice));
        OUC_HIP(hipMemcpy(net.y.ptr, y, (size_t)M * 2 * HW * 4, hipMemcpyHostToDevice));
        OUC_HIP(hipMemcpy(net.mask.ptr, mask, (size_t)M * HW * 4, hipMemcpyHostToDevice));
        OUC_HIP(hipMemcpy(net.w32.ptr, weights, net.nw * 4, hipMemcpyHostToDevice));
        OUC_HIP(hipMemset(net.m1.ptr, 0, net.nw * 4));
        OUC_HIP(hipMemset(net.m2.ptr, 0, net.nw * 4));
        float cnt = mask_total(mask, (long)M * HW) * 2.f;
        if (!(cnt > 0)) return lb_fail("empty mask", OUC_INVALID_ARG);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        double p1 = 1, p2 = 1;
        for (int e = 0; e < epochs; ++e) {
            float L = net.pass(nullptr, true, cnt);
            if (history) history[e] = L;
            p1 *= b1;
            p2 *= b2;
            k_adam<<<(net.nw + 255) / 256, 256>>>(net.nw, net.w32.f(), net.g32.f(), net.m1.f(), net.m2.f(), lr, b1, b2,
                                                  eps, (float)(1 - p1), (float)(1 - p2));
        }
        OUC_HIP(hipMemcpy(weights, net.w32.ptr, net.nw * 4, hipMemcpyDeviceToHost));
        return OUC_OK;
    } catch (const Error &e) {
        return lb_fail(e.what(), e.code);
    }
}

} // extern "C"
