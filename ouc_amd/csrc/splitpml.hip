// First-order velocity-stress engine with split PML: the reference's `stress`
// (P-S-seperating.cu:487-555) and `velocity` (P-S-seperating.cu:390-485)
// kernels, which its main() defines but never launches.  Forward modelling of
// one shot with the P/S split of the velocity; one stress and one velocity
// launch per step on 64 x 4 tiles.
#include <vector>

#include "engine.hpp"
#include "splitpml.hpp"

namespace ouc {
namespace {

// The split-PML kernels write their 12th-order coefficients with five
// significant digits (P-S-seperating.cu:401, 498); other orders use the
// live scheme's closed form.
void sp_coeffs(int N, const KGrid &g, float *c) {
    static const float a6[6] = {+1.2213e+0f, -9.6931e-2f, 1.7448e-2f, -2.9673e-3f, +3.5901e-4f, -2.1848e-5f};
    for (int k = 0; k < N; ++k) c[k] = N == 6 ? a6[k] : g.c[k];
}

struct SpWave {
    float *vx, *vz, *vpx, *vpz, *vsx, *vsz;
    float *th, *om, *thx, *thz, *omx, *omz, *uxx, *uxz, *uzx, *uzz;
};
constexpr int kSpArrays = 16;

struct SpGrid {
    int Xn, Zn, P, N;
    float idx, idz, dt;
    float c[8];
};

// Tiles of 64 z x 4 x over the update region [N, Xn-N) x [N, Zn-N).  Blocks
// are dispatched round-robin over the 8 XCDs, so block b takes tile
// (b mod 8) * T + b / 8 (z fastest, T = blocks / 8): each XCD streams a
// contiguous run of tiles and the x halos its tiles read are in its own L2.
struct SpTile {
    int i, j;
    bool ok;
};
__device__ __forceinline__ SpTile sp_tile(const SpGrid &g) {
    const int ntz = (g.Zn - 2 * g.N + 63) / 64, ntx = (g.Xn - 2 * g.N + 3) / 4;
    const int tile = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    const int i = g.N + (tile / ntz) * 4 + threadIdx.y, j = g.N + (tile % ntz) * 64 + threadIdx.x;
    return {i, j, tile < ntz * ntx && i < g.Xn - g.N && j < g.Zn - g.N};
}
inline unsigned sp_blocks(const SpGrid &g) {
    const unsigned n = (unsigned)((g.Zn - 2 * g.N + 63) / 64) * (unsigned)((g.Xn - 2 * g.N + 3) / 4);
    return (n + 7) / 8 * 8;
}

// Damping of one axis: the update is (1 / (1 + h)) ((1 - h) f + dt rhs), h = dt a / 2.
struct Damp {
    float a, b;
};
__device__ __forceinline__ Damp damp(float dt, float a) {
    const float h = 0.5f * dt * a;
    return {1.f / (1.f + h), 1.f - h};
}

template <int N>
__global__ __launch_bounds__(256) void k_sp_stress(SpGrid g, SpWave w, const float *__restrict__ p2,
                                                   const float *__restrict__ q2, const float *__restrict__ ax,
                                                   const float *__restrict__ az, int si, int sj, float s) {
    const SpTile c = sp_tile(g);
    if (!c.ok) return;
    const int i = c.i, j = c.j;
    const long k = (long)i * g.P + j;
    const float dvxdx = dmx<N>(w.vx, k, g.P, g.c) * g.idx, dvxdz = dpz<N>(w.vx, k, g.c) * g.idz;
    const float dvzdz = dmz<N>(w.vz, k, g.c) * g.idz, dvzdx = dpx<N>(w.vz, k, g.P, g.c) * g.idx;
    const Damp dx = damp(g.dt, ax[k]), dz = damp(g.dt, az[k]);
    const float p = p2[k], q = q2[k], dt = g.dt;
    const float thx = dx.a * (dt * p * dvxdx + dx.b * w.thx[k]);
    const float thz = dz.a * (dt * p * dvzdz + dz.b * w.thz[k]);
    const float omx = dx.a * (-dt * q * dvzdx + dx.b * w.omx[k]);
    const float omz = dz.a * (dt * q * dvxdz + dz.b * w.omz[k]);
    w.thx[k] = thx; w.thz[k] = thz; w.omx[k] = omx; w.omz[k] = omz;
    w.uxx[k] = dx.a * (dt * q * dvxdx + dx.b * w.uxx[k]);
    w.uxz[k] = dz.a * (dt * q * dvxdz + dz.b * w.uxz[k]);
    w.uzx[k] = dx.a * (dt * q * dvzdx + dx.b * w.uzx[k]);
    w.uzz[k] = dz.a * (dt * q * dvzdz + dz.b * w.uzz[k]);
    w.th[k] = thx + thz + ((i == si && j == sj) ? s : 0.f);
    w.om[k] = omx + omz;
}

template <int N>
__global__ __launch_bounds__(256) void k_sp_velocity(SpGrid g, SpWave w, const float *__restrict__ ax,
                                                     const float *__restrict__ az) {
    const SpTile c = sp_tile(g);
    if (!c.ok) return;
    const long k = (long)c.i * g.P + c.j;
    const int P = g.P;
    const float dthdx = dpx<N>(w.th, k, P, g.c) * g.idx, dthdz = dpz<N>(w.th, k, g.c) * g.idz;
    const float domdx = dmx<N>(w.om, k, P, g.c) * g.idx, domdz = dmz<N>(w.om, k, g.c) * g.idz;
    const float duxxdz = dpz<N>(w.uxx, k, g.c) * g.idz, duxzdx = dmx<N>(w.uxz, k, P, g.c) * g.idx;
    const float duzxdz = dmz<N>(w.uzx, k, g.c) * g.idz, duzzdx = dpx<N>(w.uzz, k, P, g.c) * g.idx;
    const Damp dx = damp(g.dt, ax[k]), dz = damp(g.dt, az[k]);
    const float dt = g.dt;
    const float cx = duzxdz - duzzdx, cz = duxzdx - duxxdz;
    const float vpx = dx.a * (dx.b * w.vpx[k] + dt * (dthdx + cx));
    const float vpz = dz.a * (dz.b * w.vpz[k] + dt * (dthdz + cz));
    const float vsx = dx.a * (dx.b * w.vsx[k] + dt * (domdz + cx));
    const float vsz = dz.a * (dz.b * w.vsz[k] + dt * (-domdx + cz));
    w.vpx[k] = vpx; w.vpz[k] = vpz; w.vsx[k] = vsx; w.vsz[k] = vsz;
    w.vx[k] = vpx + vsx;
    w.vz[k] = vpz + vsz;
}

// Receivers: every interior column of row zr, after the velocity update.
__global__ void k_sp_record(SpGrid g, SpWave w, int L, int nx, int zr, int t, int nt, float *__restrict__ rec) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= nx) return;
    const long k = (long)(ix + L) * g.P + zr;
    const float *f[6] = {w.vx, w.vz, w.vpx, w.vpz, w.vsx, w.vsz};
#pragma unroll
    for (int q = 0; q < 6; ++q) rec[((long)q * nt + t) * nx + ix] = f[q][k];
}

#define OUC_SP_DISPATCH(NV, ...)                                                 \
    switch (NV) {                                                                \
    case 1: { constexpr int NN = 1; __VA_ARGS__; } break;                        \
    case 2: { constexpr int NN = 2; __VA_ARGS__; } break;                        \
    case 3: { constexpr int NN = 3; __VA_ARGS__; } break;                        \
    case 4: { constexpr int NN = 4; __VA_ARGS__; } break;                        \
    case 5: { constexpr int NN = 5; __VA_ARGS__; } break;                        \
    case 6: { constexpr int NN = 6; __VA_ARGS__; } break;                        \
    case 7: { constexpr int NN = 7; __VA_ARGS__; } break;                        \
    default: { constexpr int NN = 8; __VA_ARGS__; } break;                       \
    }

} // namespace

void splitpml_forward(ouc_ctx *c, const ouc_shot *s, const std::vector<float> &wav, float *seis) {
    const Grid &gr = c->g;
    const KGrid kg = c->kgrid();
    SpGrid g{gr.Xn, gr.Zn, gr.P, gr.N, kg.idx, kg.idz, kg.dt, {}};
    sp_coeffs(gr.N, kg, g.c);
    const size_t plane = (size_t)gr.Xn * gr.P;
    DevBuf mem, rec;
    mem.alloc(kSpArrays * plane * sizeof(float));
    rec.alloc((size_t)6 * gr.nt * gr.nx * sizeof(float));
    hipStream_t st = c->stream;
    OUC_HIP(hipMemsetAsync(mem.ptr, 0, mem.bytes, st));
    SpWave w{};
    float **order[kSpArrays] = {&w.vx, &w.vz, &w.vpx, &w.vpz, &w.vsx, &w.vsz, &w.th, &w.om,
                                &w.thx, &w.thz, &w.omx, &w.omz, &w.uxx, &w.uxz, &w.uzx, &w.uzz};
    for (int a = 0; a < kSpArrays; ++a) *order[a] = mem.f() + a * plane;
    const int N = gr.N, Sx = s->sx + gr.L, Sz = s->sz + gr.L, Zr = s->rec_z + gr.L;
    const dim3 upd(sp_blocks(g)), blk(64, 4);
    const float *p2 = c->p2.f(), *q2 = c->q2.f(), *ax = c->ax.f(), *az = c->az.f();
    OUC_HIP(hipEventRecord(c->ev[0], st));
    for (int t = 0; t < gr.nt; ++t) {
        OUC_SP_DISPATCH(N, k_sp_stress<NN><<<upd, blk, 0, st>>>(g, w, p2, q2, ax, az, Sx, Sz, wav[t]));
        OUC_SP_DISPATCH(N, k_sp_velocity<NN><<<upd, blk, 0, st>>>(g, w, ax, az));
        k_sp_record<<<(gr.nx + 255) / 256, 256, 0, st>>>(g, w, gr.L, gr.nx, Zr, t, gr.nt, rec.f());
    }
    OUC_HIP(hipGetLastError());
    OUC_HIP(hipEventRecord(c->ev[1], st));
    std::vector<float> tmp((size_t)6 * gr.nt * gr.nx);
    OUC_HIP(hipMemcpyAsync(tmp.data(), rec.ptr, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    OUC_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    OUC_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->last = ouc_timings{};
    c->last.forward_ms = c->last.total_ms = ms;
    c->last.step_launches = 2 * (int64_t)gr.nt;
    c->last.cell_updates = (double)(gr.Xn - 2 * N) * (gr.Zn - 2 * N) * gr.nt;
    if (seis)
        for (int q = 0; q < 6; ++q)
            for (int t = 0; t < gr.nt; ++t)
                for (int ix = 0; ix < gr.nx; ++ix)
                    seis[((size_t)q * gr.nx + ix) * gr.nt + t] = tmp[((size_t)q * gr.nt + t) * gr.nx + ix];
}

} // namespace ouc
