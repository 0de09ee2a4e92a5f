// First-order velocity-stress engine with split PML: the reference's `stress`
// (P-S-seperating.cu:487-555) and `velocity` (P-S-seperating.cu:390-485)
// kernels, which its main() defines but never launches.  Forward modelling of
// one shot with the P/S split of the velocity; one stress and one velocity
// launch per step on 64 x 4 tiles.
#include <functional>
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
// The same tiling over [lo, Xn-lo) x [lo, Zn-lo): lo = N for the update
// region, lo = L for the interior the reconstruction and imaging run on.
__device__ __forceinline__ SpTile sp_tile(const SpGrid &g, int lo) {
    const int ntz = (g.Zn - 2 * lo + 63) / 64, ntx = (g.Xn - 2 * lo + 3) / 4;
    const int tile = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
    const int i = lo + (tile / ntz) * 4 + threadIdx.y, j = lo + (tile % ntz) * 64 + threadIdx.x;
    return {i, j, tile < ntz * ntx && i < g.Xn - lo && j < g.Zn - lo};
}
__device__ __forceinline__ SpTile sp_tile(const SpGrid &g) { return sp_tile(g, g.N); }
inline unsigned sp_blocks(const SpGrid &g, int lo) {
    const unsigned n = (unsigned)((g.Zn - 2 * lo + 63) / 64) * (unsigned)((g.Xn - 2 * lo + 3) / 4);
    return (n + 7) / 8 * 8;
}
inline unsigned sp_blocks(const SpGrid &g) { return sp_blocks(g, g.N); }

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

// ---- RTM with this engine (its backward half, P-S-seperating.cu:728-887 and
// 1066-1229, which the reference defines but never drives) ------------------

// Reverse of the interior velocity pass (velocity_wavefield_reconstruction,
// P-S-seperating.cu:1066-1167): V^t = V^{t+1} - dt f(stress^{t+1}) on the
// interior, where the damping is zero.
template <int N>
__global__ __launch_bounds__(256) void k_sp_recon_velocity(SpGrid g, SpWave w, int L) {
    const SpTile c = sp_tile(g, L);
    if (!c.ok) return;
    const long k = (long)c.i * g.P + c.j;
    const int P = g.P;
    const float dthdx = dpx<N>(w.th, k, P, g.c) * g.idx, dthdz = dpz<N>(w.th, k, g.c) * g.idz;
    const float domdx = dmx<N>(w.om, k, P, g.c) * g.idx, domdz = dmz<N>(w.om, k, g.c) * g.idz;
    const float duxxdz = dpz<N>(w.uxx, k, g.c) * g.idz, duxzdx = dmx<N>(w.uxz, k, P, g.c) * g.idx;
    const float duzxdz = dmz<N>(w.uzx, k, g.c) * g.idz, duzzdx = dpx<N>(w.uzz, k, P, g.c) * g.idx;
    const float dt = g.dt;
    const float cx = duzxdz - duzzdx, cz = duxzdx - duxxdz;
    const float vpx = w.vpx[k] - dt * (dthdx + cx);
    const float vpz = w.vpz[k] - dt * (dthdz + cz);
    const float vsx = w.vsx[k] - dt * (domdz + cx);
    const float vsz = w.vsz[k] - dt * (-domdx + cz);
    w.vpx[k] = vpx; w.vpz[k] = vpz; w.vsx[k] = vsx; w.vsz[k] = vsz;
    w.vx[k] = vpx + vsx;
    w.vz[k] = vpz + vsz;
}

// Reverse of the interior stress pass (stress_wavefield_reconstruction,
// P-S-seperating.cu:1168-1229): stress^t = stress^{t+1} - dt g(V^t); theta
// carries the source sample s of forward step t-1, as the forward theta^t does.
template <int N>
__global__ __launch_bounds__(256) void k_sp_recon_stress(SpGrid g, SpWave w, const float *__restrict__ p2,
                                                         const float *__restrict__ q2, int L, int si, int sj, float s) {
    const SpTile c = sp_tile(g, L);
    if (!c.ok) return;
    const int i = c.i, j = c.j;
    const long k = (long)i * g.P + j;
    const float dvxdx = dmx<N>(w.vx, k, g.P, g.c) * g.idx, dvxdz = dpz<N>(w.vx, k, g.c) * g.idz;
    const float dvzdz = dmz<N>(w.vz, k, g.c) * g.idz, dvzdx = dpx<N>(w.vz, k, g.P, g.c) * g.idx;
    const float p = p2[k], q = q2[k], dt = g.dt;
    const float thx = w.thx[k] - dt * p * dvxdx;
    const float thz = w.thz[k] - dt * p * dvzdz;
    const float omx = w.omx[k] + dt * q * dvzdx;
    const float omz = w.omz[k] - dt * q * dvxdz;
    w.thx[k] = thx; w.thz[k] = thz; w.omx[k] = omx; w.omz[k] = omz;
    w.uxx[k] = w.uxx[k] - dt * q * dvxdx;
    w.uxz[k] = w.uxz[k] - dt * q * dvxdz;
    w.uzx[k] = w.uzx[k] - dt * q * dvzdx;
    w.uzz[k] = w.uzz[k] - dt * q * dvzdz;
    w.th[k] = thx + thz + ((i == si && j == sj) ? s : 0.f);
    w.om[k] = omx + omz;
}

// Ring store of the eight strip fields (vx, vz, theta, omega, duzdx, duzdz,
// duxdx, duxdz: save_wavefiled, P-S-seperating.cu:1230-1287), FP32,
// [slot][field][ring cell]: save slot `slot` of fields [f0, f1), or restore
// them (slot < 0 restores zeros).
__global__ __launch_bounds__(256) void k_sp_strips(KRing r, SpGrid g, SpWave w, float *__restrict__ store, int slot,
                                                   int f0, int f1, int save) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= r.ncells) return;
    int i, j;
    ring_cell(r, q, i, j);
    const long k = (long)i * g.P + j;
    float *const f[8] = {w.vx, w.vz, w.th, w.om, w.uzx, w.uzz, w.uxx, w.uxz};
    for (int a = f0; a < f1; ++a) {
        float *const e = store + ((long)slot * 8 + a) * r.ncells + q;
        if (save) *e = f[a][k];
        else f[a][k] = slot >= 0 ? *e : 0.f;
    }
}

// load_record (P-S-seperating.cu:1471-1483): the records of step t replace vx,
// vz on the receiver row; rec is [component][nt][nx].
__global__ void k_sp_inject(SpGrid g, SpWave w, int L, int nx, int zr, int t, int nt, const float *__restrict__ rec) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= nx) return;
    const long k = (long)(ix + L) * g.P + zr;
    w.vx[k] = rec[(long)t * nx + ix];
    w.vz[k] = rec[((long)nt + t) * nx + ix];
}

// corr_v (P-S-seperating.cu:1426-1439) on the interior: the source P velocity
// of level t+1 against the receiver's P and S velocities after its step t.
__global__ __launch_bounds__(256) void k_sp_correlate(SpGrid g, SpWave src, SpWave rcv, int L, float *__restrict__ npp,
                                                      float *__restrict__ nps, float *__restrict__ den) {
    const SpTile c = sp_tile(g, L);
    if (!c.ok) return;
    const long k = (long)c.i * g.P + c.j;
    const float sa = src.vpx[k], sb = src.vpz[k];
    npp[k] += sa * rcv.vpx[k] + sb * rcv.vpz[k];
    den[k] += sa * sa + sb * sb;
    nps[k] += sa * rcv.vsx[k] + sb * rcv.vsz[k];
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

namespace {
struct SpState {
    SpGrid g{};
    DevBuf mem;
    SpWave w{};
};

void sp_state(ouc_ctx *c, SpState &S) {
    const Grid &gr = c->g;
    const KGrid kg = c->kgrid();
    S.g = SpGrid{gr.Xn, gr.Zn, gr.P, gr.N, kg.idx, kg.idz, kg.dt, {}};
    sp_coeffs(gr.N, kg, S.g.c);
    const size_t plane = (size_t)gr.Xn * gr.P;
    S.mem.alloc(kSpArrays * plane * sizeof(float));
    OUC_HIP(hipMemsetAsync(S.mem.ptr, 0, S.mem.bytes, c->stream));
    float **order[kSpArrays] = {&S.w.vx, &S.w.vz, &S.w.vpx, &S.w.vpz, &S.w.vsx, &S.w.vsz, &S.w.th, &S.w.om,
                                &S.w.thx, &S.w.thz, &S.w.omx, &S.w.omz, &S.w.uxx, &S.w.uxz, &S.w.uzx, &S.w.uzz};
    for (int a = 0; a < kSpArrays; ++a) *order[a] = S.mem.f() + a * plane;
}
} // namespace

void splitpml_rtm(ouc_ctx *c, const ouc_shot *s, const std::vector<float> &wav, float *rec,
                  const std::function<void()> &between, float *npp, float *nps, float *den) {
    const Grid &gr = c->g;
    if (c->p.semantics != OUC_SEM_EXACT)
        throw Error(OUC_UNSUPPORTED, "RTM with the split-PML engine is defined with EXACT time alignment only");
    if (c->p.strip_precision != OUC_STRIP_FP32)
        throw Error(OUC_UNSUPPORTED, "RTM with the split-PML engine stores FP32 strips only");
    SpState src, rcv;
    sp_state(c, src);
    sp_state(c, rcv);
    const SpGrid &g = src.g;
    const KRing kr = c->kring();
    DevBuf store;
    store.alloc((size_t)gr.nt * 8 * kr.ncells * sizeof(float));
    hipStream_t st = c->stream;
    const int N = gr.N, L = gr.L, nt = gr.nt, Sx = s->sx + L, Sz = s->sz + L, Zr = s->rec_z + L;
    const dim3 upd(sp_blocks(g)), inter(sp_blocks(g, L)), blk(64, 4);
    const unsigned rb = (unsigned)((kr.ncells + 255) / 256);
    const float *p2 = c->p2.f(), *q2 = c->q2.f(), *ax = c->ax.f(), *az = c->az.f();
    const size_t pn = gr.pitched() * sizeof(float);
    OUC_HIP(hipMemsetAsync(npp, 0, pn, st));
    OUC_HIP(hipMemsetAsync(nps, 0, pn, st));
    OUC_HIP(hipMemsetAsync(den, 0, pn, st));
    OUC_HIP(hipEventRecord(c->ev[0], st));
    for (int t = 0; t < nt; ++t) {
        OUC_SP_DISPATCH(N, k_sp_stress<NN><<<upd, blk, 0, st>>>(g, src.w, p2, q2, ax, az, Sx, Sz, wav[t]));
        OUC_SP_DISPATCH(N, k_sp_velocity<NN><<<upd, blk, 0, st>>>(g, src.w, ax, az));
        k_sp_record<<<(gr.nx + 255) / 256, 256, 0, st>>>(g, src.w, L, gr.nx, Zr, t, nt, rec);
        k_sp_strips<<<rb, 256, 0, st>>>(kr, g, src.w, store.f(), t, 0, 8, 1);
    }
    OUC_HIP(hipGetLastError());
    between(); // records replaced by the caller's, and muted
    OUC_HIP(hipEventRecord(c->ev[1], st));
    for (int t = nt - 1; t >= 0; --t) {
        k_sp_inject<<<(gr.nx + 255) / 256, 256, 0, st>>>(g, rcv.w, L, gr.nx, Zr, t, nt, rec);
        OUC_SP_DISPATCH(N, k_sp_stress<NN><<<upd, blk, 0, st>>>(g, rcv.w, p2, q2, ax, az, -1, -1, 0.f));
        OUC_SP_DISPATCH(N, k_sp_velocity<NN><<<upd, blk, 0, st>>>(g, rcv.w, ax, az));
        k_sp_correlate<<<inter, blk, 0, st>>>(g, src.w, rcv.w, L, npp, nps, den);
        // the source from level t+1 to level t
        k_sp_strips<<<rb, 256, 0, st>>>(kr, g, src.w, store.f(), t, 2, 8, 0);
        OUC_SP_DISPATCH(N, k_sp_recon_velocity<NN><<<inter, blk, 0, st>>>(g, src.w, L));
        k_sp_strips<<<rb, 256, 0, st>>>(kr, g, src.w, store.f(), t - 1, 0, 2, 0);
        OUC_SP_DISPATCH(N, k_sp_recon_stress<NN><<<inter, blk, 0, st>>>(g, src.w, p2, q2, L, Sx, Sz,
                                                                        t >= 1 ? wav[t - 1] : 0.f));
    }
    OUC_HIP(hipGetLastError());
    OUC_HIP(hipEventRecord(c->ev[2], st));
    OUC_HIP(hipStreamSynchronize(st));
    c->last = ouc_timings{};
    c->last.step_launches = 2 * (int64_t)nt;
    c->last.cell_updates = (double)nt * (2.0 * (gr.Xn - 2 * N) * (gr.Zn - 2 * N) + (double)gr.nx * gr.nz);
}

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
