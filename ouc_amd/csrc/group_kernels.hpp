// Shots of several grid sizes in one launch.
//
// A grouped run advances the batches of several contexts (each its own grid,
// model and shots) through their time loops together.  Every kernel of a time
// step is one launch over the blocks of all contexts active at that step:
// block b belongs to the i-th listed context for first[i] <= b < first[i+1],
// and within that context it takes the block index the context's own launch
// would give it, so each context computes exactly what its own batch would.
// The contexts' static arguments sit in a device table (GCtx); the per-step
// ones (step, V buffer parities, strip slots, source P level) follow from the
// group step tau as the single-context loops derive them on the host.  The
// contexts' loops end together: context c runs its steps 0 .. nt_c - 1 at
// tau = start_c .. T - 1, start_c = T - nt_c, forward and (reversed) backward.
#pragma once

#include "edge_kernels.hpp"
#include "lpad_kernels.hpp"

namespace ouc {

constexpr int kGroupMax = 8; // contexts in one grouped launch

struct GCtx {
    FGrid g, rg;       // the grid; the grid with the reconstruction core bounds
    FWave S, R;        // source and receiver wavefields
    KModel km;
    const float2 *pq;
    const float *p2, *q2;
    FStep fa, ra, sa;  // step templates: source forward run, receiver run, reconstruction
    KRing ring;
    KStrips ks0;
    KStripStride sstride;
    float2 *vp[3];     // source P levels, by level mod 3
    EPlan ep[2];       // CPML band plans: source forward run (saving the strip ring's stress), receiver run
    int eb[2];         // their blocks per shot
    const int2 *tiles_s, *tiles_v, *tiles_rs, *tiles_rv, *tiles_p, *tiles_m;
    int nts, ntv, ntrs, ntrv, ntp, ntm;
    const LbCell *lbcells;
    int nlbcells, lbrb; // network ring cells, their blocks per shot
    const float *lbw;  // composed filter, or NULL (ring held at zero)
    float *lbd;
    float *lbc;        // the network's staged input rows, and their length
    int lbE;
    float lbsc;
    unsigned int *mon_f, *mon_r;
    int nt, B, start, exact, edge, rblocks;
    int cbx, cby, rbx, rby; // core and reconstruction-core launch grids (x, y)
};

struct GLaunch {
    int n;                      // contexts in this launch
    int ctx[kGroupMax];         // their indices in tab
    int first[kGroupMax + 1];   // their first blocks
    int tau;                    // group step
};

// Context of block b and the block's index within it.  The table is a
// read-only kernel argument apart from everything the kernels write, so its
// fields load when used, as a plain launch's arguments do.
__device__ __forceinline__ const GCtx &gctx(const GCtx *__restrict__ tab, const GLaunch &L, int b, int &local) {
    int i = 0;
    while (i + 1 < L.n && b >= L.first[i + 1]) ++i;
    local = b - L.first[i];
    return tab[L.ctx[i]];
}
__device__ __forceinline__ uint3 gbid(int local, int nx, int ny) {
    return make_uint3(local % nx, (local / nx) % ny, local / (nx * ny));
}

// Per-step arguments, as the single-context loops set them: the forward run's
// V buffer alternates from 0; the receiver run's from 0 at its first step
// t = nt - 1; the reconstruction starts from V^{nt-1}, handed off into buffer
// 1 - (nt mod 2), and alternates from there.
__device__ __forceinline__ FStep g_fwd(const GCtx &c, int t) {
    FStep a = c.fa;
    a.t = t;
    a.rb = t & 1;
    return a;
}
__device__ __forceinline__ FStep g_rcv(const GCtx &c, int t) {
    FStep a = c.ra;
    a.t = t;
    a.rb = (c.nt - 1 - t) & 1;
    const int lev = c.exact ? t + 1 : min(t + 2, c.nt);
    a.sp = c.vp[lev % 3];
    return a;
}
// The reconstruction of backward iteration t (t < nt - 1): theta^{t+1} from
// V^{t+1}, ring totals from slot t+1 (FAITHFUL) or t (EXACT), the hand-off
// step from the forward run's own state.
__device__ __forceinline__ void g_recon(const GCtx &c, int t, FStep &ra, FRecon &rc) {
    ra = c.sa;
    ra.t = t + 1;
    ra.rb = ((1 - (c.nt & 1)) + (c.nt - 2 - t)) & 1;
    rc = FRecon{};
    rc.ring = c.ring;
    rc.strips = c.ks0;
    rc.stride = c.sstride;
    const bool handoff = t == c.nt - 2;
    rc.slot_v = handoff ? -1 : (c.exact ? t : t + 1);
    rc.slot_th = t + 1;
    rc.exact = (c.exact || handoff) ? 1 : 0;
    rc.vp = c.vp[t % 3];
}
__device__ __forceinline__ FStep g_wave(const GCtx &c, bool rcv, int t) { return rcv ? g_rcv(c, t) : g_fwd(c, t); }

template <int N, int MODE, int PF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_core_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    const int t = L.tau - c.start;
    if constexpr (MODE == CORE_RECON) {
        FStep ra;
        FRecon rc;
        g_recon(c, t, ra, rc);
        core_body<N, MODE, PF>(c.rg, c.S, c.pq, ra, rc, gbid(local, c.rbx, c.rby));
    } else {
        core_body<N, MODE, PF>(c.g, MODE == CORE_IMAGE ? c.R : c.S, c.pq, g_wave(c, MODE == CORE_IMAGE, t), FRecon{},
                               gbid(local, c.cbx, c.cby));
    }
}

template <int N, bool IMG>
__global__ __launch_bounds__(kEdgeW) __attribute__((amdgpu_waves_per_eu(2))) void k_edge_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    const int t = L.tau - c.start, v = IMG ? 1 : 0; // the receiver run images
    edge_body<N, IMG>(c.g, IMG ? c.R : c.S, c.km, g_wave(c, IMG, t), c.ep[v], gbid(local, c.eb[v], 1));
}

// Strips of the source's forward step t (stress-like fields of step t and,
// but the last step, V^{t+1}).
__global__ __launch_bounds__(256) void k_fstrip_save_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    const int t = L.tau - c.start, rb = t & 1;
    FWave ws = c.S; // the stress memory variables of this step (the CPML band kernel double-buffers them)
    if (c.edge && rb == 0)
        for (int q = 0; q < 4; ++q) ws.F[q] = c.S.F2[q];
    fstrip_save_body(c.g, ws, c.ring, c.ks0, c.sstride, t, 1 - rb, t < c.nt - 1 ? 2 : 0, gbid(local, c.rblocks, 1));
}

__global__ __launch_bounds__(256) void k_strip_decode_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    strip_decode_body(c.ks0, L.tau - c.start, c.sstride, gbid(local, c.rblocks, RS_NF));
}

template <int N>
__global__ __launch_bounds__(256) void k_recon_stress_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    FStep ra;
    FRecon rc;
    g_recon(c, L.tau - c.start, ra, rc);
    recon_stress_body<N>(c.g, c.S, c.p2, c.q2, rc, ra, c.tiles_rs, gbid(local, c.ntrs, 1));
}

template <int N>
__global__ __launch_bounds__(256) void k_recon_velocity_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    FStep ra;
    FRecon rc;
    g_recon(c, L.tau - c.start, ra, rc);
    recon_velocity_body<N>(c.g, c.S, c.q2, rc, ra, c.tiles_rv, gbid(local, c.ntrv, 1));
}

// Learned-boundary contexts: the pad, the tile kernels of the band between the
// pad and the core, the network ring and the divergence monitor.
template <int N, bool RCV>
__global__ __launch_bounds__(256) void k_band_stress_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    band_stress_body<N>(c.g, RCV ? c.R : c.S, c.km, c.tiles_s, g_wave(c, RCV, L.tau - c.start),
                        gbid(local, c.nts, 1));
}

template <int N, bool RCV>
__global__ __launch_bounds__(256) void k_band_velocity_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    band_velocity_body<N>(c.g, RCV ? c.R : c.S, c.km, c.tiles_v, g_wave(c, RCV, L.tau - c.start),
                          gbid(local, c.ntv, 1));
}

template <bool RCV>
__global__ __launch_bounds__(256) void k_lpad_stress_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    lpad_stress_body(c.g, RCV ? c.R : c.S, c.km, c.tiles_p, g_wave(c, RCV, L.tau - c.start),
                     gbid(local, c.ntp, 1));
}

template <bool RCV>
__global__ __launch_bounds__(256) void k_lb_canon_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    lb_canon_body(c.g, RCV ? c.R : c.S, c.km, c.lbsc, c.lbc, c.lbE, gbid(local, (c.lbE + 255) / 256, 4 * kLbD));
}

template <bool RCV>
__global__ __launch_bounds__(256) void k_lb_ring_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    const FStep a = g_wave(c, RCV, L.tau - c.start);
    lb_ring_body(c.g, RCV ? c.R : c.S, c.lbcells, c.nlbcells, c.lbw, a.rb, c.lbc, c.lbE, c.lbd, gbid(local, c.lbrb, 1));
}

template <bool RCV>
__global__ __launch_bounds__(256) void k_lb_apply_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    const FStep a = g_wave(c, RCV, L.tau - c.start);
    lb_apply_body(c.g, RCV ? c.R : c.S, c.lbcells, c.nlbcells, c.lbd, a.rb, gbid(local, c.lbrb, 1));
}

template <bool RCV>
__global__ __launch_bounds__(256) void k_lpad_velocity_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    const FStep a = g_wave(c, RCV, L.tau - c.start);
    lpad_velocity_body(c.g, RCV ? c.R : c.S, c.km, c.tiles_p, a.rb, gbid(local, c.ntp, 1));
}

template <bool RCV>
__global__ __launch_bounds__(256) void k_lb_monitor_grp(const GCtx *__restrict__ tab, GLaunch L) {
    int local;
    const GCtx &c = gctx(tab, L, blockIdx.x, local);
    const FStep a = g_wave(c, RCV, L.tau - c.start);
    lb_monitor_body(c.g, RCV ? c.R : c.S, c.tiles_m, 1 - a.rb, RCV ? c.mon_r : c.mon_f, gbid(local, c.ntm, 1));
}

} // namespace ouc
