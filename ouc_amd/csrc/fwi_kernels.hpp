// Elastic FWI gradient kernels of the fused engine: the adjoint residual
// (adsource, P-S-seperating.cu:1484-1498), the gradient and illumination
// (gradfun, P-S-seperating.cu:1518-1575), the per-shot sum and the model
// update (revise_model_end, P-S-seperating.cu:1500-1517).
#pragma once

#include "fused_kernels.hpp"

namespace ouc {

// Records of a batch [shot][2][nt][nx] become residuals (modelled - observed)
// in place; each block leaves its sum of squared residuals in part[shot][block].
__global__ __launch_bounds__(256) void k_fresidual(long per_shot, float *__restrict__ rec,
                                                   const float *__restrict__ obs, double *__restrict__ part) {
    __shared__ double red[256];
    const long base = (long)blockIdx.y * per_shot;
    double acc = 0;
    for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < per_shot; k += (long)gridDim.x * 256) {
        const float r = rec[base + k] - obs[base + k];
        rec[base + k] = r;
        acc += (double)r * (double)r;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// Tile of the gradient kernel: 64 rows along z by 16 columns along x.
constexpr int kGradTZ = 64, kGradTX = 16;

// Blocks along x of the gradient grid for an interior of nx x nz: a multiple
// of 8, so that block b of a shot runs on XCD b mod 8 whatever the shot.
__host__ __device__ inline int grad_tiles(int nx, int nz) {
    return ((nz + kGradTZ - 1) / kGradTZ) * ((nx + kGradTX - 1) / kGradTX);
}
__host__ __device__ inline int grad_blocks(int nx, int nz) { return (grad_tiles(nx, nz) + 7) / 8 * 8; }

// One backward step of gradfun on the interior [L, Xn-L) x [L, Zn-L).
// Source: the reconstructed totals V^{t+1} as the reconstruction's stress pass
// of this step reads them (ring cells from the strips at rc.slot_v); receiver:
// this step's V before its update, with the injected residual row, so that its
// stress fields (no CPML memory in the interior) are theta = vp^2 (gxx + gzz),
// omega = vs^2 (gxz - gzx) and du = vs^2 g.  Derivatives use the engine's
// stencil (the reference's N = 6 duxdz at P-S-seperating.cu:1556 pairs f[j+4]
// with f[j-6]; this is the plain Dp_z).  The live physics has no density, so
// the reference's factor 2 P vp is 2 vp.  Accumulators are [shot][plane].
// Blocks are dispatched round-robin over the 8 XCDs, so block b takes tile
// (b mod 8) * T + b / 8 (z fastest): each XCD streams a contiguous run of
// tiles and the halos its neighbours read are in its own L2.
template <int N>
__global__ __launch_bounds__(256) void k_fgrad(FGrid g, FWave src, int sbuf, FRecon rc, FWave rcv, FStep a,
                                               const float *__restrict__ p2, const float *__restrict__ q2,
                                               float *__restrict__ gvp, float *__restrict__ gvs,
                                               float *__restrict__ ill) {
    constexpr int HX = kGradTX + 2 * N, HZ = kGradTZ + 2 * N;
    __shared__ float sx_[HX][HZ], sz_[HX][HZ], rx_[HX][HZ], rz_[HX][HZ];
    const int ntz = (g.nz + kGradTZ - 1) / kGradTZ, per_xcd = gridDim.x / 8;
    const int tile = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (tile >= grad_tiles(g.nx, g.nz)) return;
    const int i0 = g.L + (tile / ntz) * kGradTX, j0 = g.L + (tile % ntz) * kGradTZ;
    const long so = (long)blockIdx.z * g.plane;
    const KStrips rs = shot_strips(rc, blockIdx.z);
    const int tid = threadIdx.y * 64 + threadIdx.x;
    for (int e = tid; e < HX * HZ; e += 256) {
        const int xi = e / HZ, zj = e - xi * HZ;
        const int i = i0 - N + xi, j = j0 - N + zj;
        float vx = 0.f, vz = 0.f, ux = 0.f, uz = 0.f;
        if (i < g.Xn && j < g.Zn) {
            recon_vin(g, src, rc, rs, sbuf, so, i, j, vx, vz);
            load_vin(g, rcv, a, blockIdx.z, so, i, j, ux, uz);
        }
        sx_[xi][zj] = vx;
        sz_[xi][zj] = vz;
        rx_[xi][zj] = ux;
        rz_[xi][zj] = uz;
    }
    __syncthreads();
    const int zl = threadIdx.x + N, j = j0 + threadIdx.x;
    if (j >= g.Zn - g.L) return;
    const float *c = g.c;
    for (int xr = threadIdx.y; xr < kGradTX; xr += 4) {
        const int i = i0 + xr, xl = xr + N;
        if (i >= g.Xn - g.L) break;
        float dxx = 0.f, dzz = 0.f, dzx = 0.f, dxz = 0.f, gxx = 0.f, gzz = 0.f, gzx = 0.f, gxz = 0.f;
#pragma unroll
        for (int m = 1; m <= N; ++m) {
            const float cm = c[m - 1];
            dxx += cm * (sx_[xl + m - 1][zl] - sx_[xl - m][zl]);
            dzz += cm * (sz_[xl][zl + m - 1] - sz_[xl][zl - m]);
            dzx += cm * (sz_[xl + m][zl] - sz_[xl - m + 1][zl]);
            dxz += cm * (sx_[xl][zl + m] - sx_[xl][zl - m + 1]);
            gxx += cm * (rx_[xl + m - 1][zl] - rx_[xl - m][zl]);
            gzz += cm * (rz_[xl][zl + m - 1] - rz_[xl][zl - m]);
            gzx += cm * (rz_[xl + m][zl] - rz_[xl - m + 1][zl]);
            gxz += cm * (rx_[xl][zl + m] - rx_[xl][zl - m + 1]);
        }
        dxx *= g.idx; dzx *= g.idx; gxx *= g.idx; gzx *= g.idx;
        dzz *= g.idz; dxz *= g.idz; gzz *= g.idz; gxz *= g.idz;
        const long km = (long)i * g.P + j, k = so + km;
        const float p = p2[km], q = q2[km];
        const float th = p * (gxx + gzz), om = q * (gxz - gzx);
        const float uzx = q * gzx, uzz = q * gzz, uxz = q * gxz, uxx = q * gxx;
        const float vp = sqrtf(p), vs = sqrtf(q);
        gvp[k] -= 2.f * vp * ((dxx + dzz) * th);
        gvs[k] -= 2.f * vs * (om * (dxz - dzx) + uzx * dzx + uzz * dzz + uxz * dxz + uxx * dxx);
        const float a0 = sx_[xl][zl], b0 = sz_[xl][zl];
        ill[k] += a0 * a0 + b0 * b0;
    }
}

// Sum over the B shots of a [shot][plane] accumulator onto an x-major
// interior image (added to what out holds).
__global__ void k_shot_sum(FGrid g, int B, const float *__restrict__ acc, float *__restrict__ out) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x, ix = blockIdx.y;
    if (iz >= g.nz || ix >= g.nx) return;
    const long k = (long)(ix + g.L) * g.P + iz + g.L;
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += acc[(long)b * g.plane + k];
    out[(long)ix * g.nz + iz] += s;
}

// revise_model_end: v - step * 0.95^iter * grad, in double as the reference
// evaluates a_2 * pow(0.95, k) * Grad (P-S-seperating.cu:1511-1512).
__global__ void k_model_update(long n, const float *__restrict__ v, const float *__restrict__ grad, double scale,
                               float *__restrict__ out) {
#pragma clang fp contract(off) // a product, then a difference, each rounded
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = (float)((double)v[k] - scale * (double)grad[k]);
}

} // namespace ouc
