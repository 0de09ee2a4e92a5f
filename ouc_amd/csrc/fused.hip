// Fused engine: batched shots, one launch per kernel per step for the whole
// batch (grid z = shot), band/core split, stencil state double-buffered.
//
// Step structure (per wavefield): k_band_stress, k_core, k_band_velocity; the
// source reconstruction is k_recon_stress_f + k_recon_velocity_f.  Strips are
// FP32 or scaled FP16, [shot][slot][field][ring]; slot t holds V^{t+1} and the
// stress-like fields of step t (the reference's save after forward_u,
// P-S-seperating.cu:2264-2272), and slot nt-1 holds the last stress fields,
// which the reference hands over with read_last_wavefiled2
// (P-S-seperating.cu:1318-1342).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "engine.hpp"
#include "fused.hpp"
#include "edge_kernels.hpp"
#include "fused_kernels.hpp"
#include "fwi_kernels.hpp"
#include "group_kernels.hpp"
#include "lapfft.hpp"
#include "lpad_kernels.hpp"

using namespace ouc;

namespace {

#define OUC_FDISPATCH(NV, ...)                                                   \
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

// k_core with the prefetch depth chosen at run time (1..3 for N = 5, 6; the
// default elsewhere, to bound the number of instantiations).
template <int N, int MODE>
void launch_core_mode(int pf, dim3 grid, hipStream_t s, const FGrid &f, const FWave &w, const float2 *pq,
                      const FStep &a, const FRecon &rc) {
    const uint3 dims = make_uint3(grid.x, grid.y, grid.z); // x-tiles, z blocks, shots (XCD-ordered in the kernel)
    const dim3 lin(xcd_blocks(dims));
    if constexpr (N == 5 || N == 6) {
        switch (pf) {
        case 1: k_core<N, MODE, 1><<<lin, 256, 0, s>>>(f, w, pq, a, rc, dims); return;
        case 3: k_core<N, MODE, 3><<<lin, 256, 0, s>>>(f, w, pq, a, rc, dims); return;
        default: break;
        }
    }
    k_core<N, MODE, kCorePF><<<lin, 256, 0, s>>>(f, w, pq, a, rc, dims);
}

// RECON = 1: the reconstruction; otherwise a wavefield step, with the imaging
// condition when the step carries imaging operands (a.sp).  The tuned
// prefetch depth applies to the plain step; the imaging mode carries more
// operands per column and runs with one column in flight (rotating queues, no
// register spills), the reconstruction with two (the deepest without spills).
template <int N, int RECON>
void launch_core(int pf, dim3 grid, hipStream_t s, const FGrid &f, const FWave &w, const float2 *pq, const FStep &a,
                 const FRecon &rc) {
    if constexpr (RECON) launch_core_mode<N, CORE_RECON>(2, grid, s, f, w, pq, a, rc);
    else if (a.sp) launch_core_mode<N, CORE_IMAGE>(1, grid, s, f, w, pq, a, rc);
    else launch_core_mode<N, CORE_STEP>(pf, grid, s, f, w, pq, a, rc);
}

constexpr int kWaveArrays = 32; // V[2][2], DV[2], S[2], DS[2], 6 stress, 4 F, 8 O, 4 F2

struct GroupCache; // captured loops of grouped runs, kept by their first context
void drop_group(GroupCache *g);

struct FusedState {
    int cap = 0;           // shots the buffers hold
    long plane = 0;
    FGrid g{};
    DevBuf mem_s, mem_r;   // wavefields
    FWave S{}, R{};
    DevBuf vp[3];          // source P velocity levels, (x, z) pairs [cap][plane]
    DevBuf acc[3];         // npp, nps, den
    DevBuf strips32, strips16, e8;   // FP32 strips, or FP16 residuals and chunk exponents
    DevBuf keyb, ef, decb;           // FP16: keyframes, encoder state, decoded slots ([cap][2][RS_NF][ring])
    DevBuf in32;                     // FP16: inner ring part of the FP32-kept fields
    long shot_strip = 0, shot_chunks = 0, shot_key = 0, shot_aux = 0, shot_in32 = 0, nchunks = 0;
    int nslots = 0, fp16 = 0;
    DevBuf rec;            // [cap][2][nt][nx]
    DevBuf srcpos, wav, counts, t0;
    DevBuf tiles_s, tiles_v, tiles_rs, tiles_rv;
    int ntiles_s = 0, ntiles_v = 0, ntiles_rs = 0, ntiles_rv = 0;
    bool edge = false;     // CPML band by the streaming k_edge (CPML contexts) instead of the tile kernels
    // streaming band plans: [0] the source's forward run, [1] the receiver run
    EPlan eplan[2]{};
    int eblocks[2] = {0, 0};
    DevBuf img;            // [4][nx*nz] scratch for one shot's images
    // learned boundary
    DevBuf tiles_p, tiles_m, lbcells, lbw, mon;
    DevBuf lbd; // [cap][nlbcells]: the network ring's new increments of a step
    DevBuf lbc; // [cap][lb_canon_floats(lbE)]: the network's input rows, staged per side
    int lbE = 0; // positions per staged row: max(Xn, Zn)
    int ntiles_p = 0, ntiles_m = 0, nlbcells = 0;
    bool lb_have = false;  // composed filter present (else the ring is held at zero)
    int csx = kCoreSX, pf = kCorePF; // core tile
    float lb_beta = 0.01f;           // learned pad damping (the default keeps 8000-step runs stable)
    DevBuf rec6;           // [6][nt][nx] forward seismograms of one shot
    // FWI: gradient accumulators [3][cap][plane] (vp, vs, illumination), observed
    // records [cap][2][nt][nx], residual partial sums, shot-summed interior outputs
    DevBuf gacc, obs, part, gout;
    // time loops captured as HIP graphs, replayed while the key matches
    // The key holds every host value baked into the captured kernel arguments
    // that can change between calls on one context: the batch size, the loop
    // kind, the recorded components and the receiver row (per-shot positions,
    // wavelets and start steps travel through device memory).
    struct LoopGraph {
        int B = -1, rtm = -1, ncomp = -1, zr = -1;
        std::vector<hipGraphExec_t> execs; // one per chunk of kGraphSteps steps, replayed in order
        void drop() {
            for (hipGraphExec_t e : execs) (void)hipGraphExecDestroy(e);
            execs.clear();
            B = -1;
        }
        bool matches(int b, int r, int n, int z) const {
            return !execs.empty() && B == b && rtm == r && ncomp == n && zr == z;
        }
    } gfwd, gbwd;
    void drop_graphs() { gfwd.drop(); gbwd.drop(); }
    GroupCache *grp = nullptr;
    // side stream (default priority) for the reconstruction, which runs
    // beside the receiver step of the same backward iteration; fork/join events
    hipStream_t s2 = nullptr;
    hipEvent_t evf = nullptr, evj = nullptr;
    void ensure_side() {
        if (s2) return;
        OUC_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        OUC_HIP(hipEventCreateWithFlags(&evf, hipEventDisableTiming));
        OUC_HIP(hipEventCreateWithFlags(&evj, hipEventDisableTiming));
    }
    ~FusedState() {
        drop_graphs();
        drop_group(grp);
        if (evf) (void)hipEventDestroy(evf);
        if (evj) (void)hipEventDestroy(evj);
        if (s2) (void)hipStreamDestroy(s2);
    }
};

// Time loops are captured in chunks of at most kGraphSteps steps, one graph
// each, replayed in order on the same stream.  A loop captured whole is one
// dependency chain of up to ten launches per step; runs whose captured chain
// exceeded about 74,000 launches (a learned-boundary RTM loop of 7,996 steps:
// 81,000) ended their process with a host segmentation fault while the graph
// was built, and every run below that (73,000 and fewer) completed.  Chunks
// bound the chain at about 21,000 launches; the kernels and their order are
// those of the whole loop.
constexpr int kGraphSteps = 2048;

// Enqueues body(i0, i1) -- iterations [i0, i1) of an nsteps-iteration loop --
// on stream s directly, or replays the graphs it was captured into, chunk by
// chunk, the first time for this key (HIP stream capture).  body carries the
// loop's running state across calls, so the chunks are captured in order.
template <class F>
void run_loop(hipStream_t s, bool graph, FusedState::LoopGraph &lg, int B, int rtm, int ncomp, int zr, int nsteps,
              F &&body) {
    if (!graph) {
        body(0, nsteps);
        return;
    }
    if (!lg.matches(B, rtm, ncomp, zr)) {
        lg.drop();
        try {
            for (int i0 = 0; i0 < nsteps; i0 += kGraphSteps) {
                const int i1 = std::min(nsteps, i0 + kGraphSteps);
                hipGraph_t gr = nullptr;
                OUC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try {
                    body(i0, i1);
                } catch (...) {
                    (void)hipStreamEndCapture(s, &gr);
                    if (gr) (void)hipGraphDestroy(gr);
                    throw;
                }
                OUC_HIP(hipStreamEndCapture(s, &gr));
                hipGraphExec_t ex = nullptr;
                hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                OUC_HIP(e);
                lg.execs.push_back(ex);
            }
        } catch (...) {
            lg.drop();
            throw;
        }
        lg.B = B; lg.rtm = rtm; lg.ncomp = ncomp; lg.zr = zr;
    }
    for (hipGraphExec_t ex : lg.execs) OUC_HIP(hipGraphLaunch(ex, s));
}

constexpr int kMonStride = 8; // steps between boundary-monitor launches

FusedState *fs(ouc_ctx *c) { return (FusedState *)c->fused; }

// The first five pair arrays (V[0], V[1], DV, S, DS: ten planes) hold (x, z)
// interleaved; the 22 single arrays follow.
void wave_layout(DevBuf &mem, FWave &w, long plane, int cap) {
    mem.alloc((size_t)kWaveArrays * plane * cap * sizeof(float));
    float *b = mem.f();
    const size_t span = (size_t)plane * cap;
    Pair *pairs[5] = {&w.V[0], &w.V[1], &w.DV, &w.S, &w.DS};
    for (int a = 0; a < 5; ++a) pairs[a]->p = reinterpret_cast<float2 *>(b + (size_t)2 * a * span);
    float **slots[kWaveArrays - 10] = {&w.th,    &w.om,    &w.gxx,   &w.gxz,   &w.gzx,   &w.gzz,   &w.F[0],  &w.F[1],
                                       &w.F[2],  &w.F[3],  &w.O[0],  &w.O[1],  &w.O[2],  &w.O[3],  &w.O[4],  &w.O[5],
                                       &w.O[6],  &w.O[7],  &w.F2[0], &w.F2[1], &w.F2[2], &w.F2[3]};
    for (int a = 0; a < kWaveArrays - 10; ++a) *slots[a] = b + (size_t)(10 + a) * span;
}

FGrid make_fgrid(const ouc_ctx *c) {
    const Grid &g = c->g;
    FGrid f{};
    f.Xn = g.Xn; f.Zn = g.Zn; f.P = g.P; f.L = g.L; f.N = g.N; f.nx = g.nx; f.nz = g.nz; f.nt = g.nt;
    f.plane = (long)g.pitched();
    f.idx = 1.0f / g.dx; f.idz = 1.0f / g.dz; f.dt = g.dt; f.dt2 = g.dt * g.dt;
    stencil_coeffs(g.N, f.c);
    for (int a = g.N; a < 8; ++a) f.c[a] = 0.f;
    f.cx0 = g.L + g.N; f.cx1 = g.Xn - g.L - g.N;
    f.cz0 = g.L + g.N; f.cz1 = g.Zn - g.L - g.N;
    if (f.cx1 <= f.cx0 || f.cz1 <= f.cz0) { f.cx1 = f.cx0; f.cz1 = f.cz0; } // no core
    f.csx = kCoreSX;
    f.lpad = g.learned ? g.L : 0;
    f.vlo = g.learned ? 0 : g.N;
    return f;
}

// 64 x 4 tiles (origin i0, j0) over the update region holding at least one cell
// of `want`.
template <class Pred>
std::vector<int2> tile_list(const FGrid &g, Pred want) {
    std::vector<int2> out;
    for (int i0 = g.N; i0 < g.Xn - g.N; i0 += 4)
        for (int j0 = g.N; j0 < g.Zn - g.N; j0 += 64) {
            bool any = false;
            for (int i = i0; i < std::min(i0 + 4, g.Xn - g.N) && !any; ++i)
                for (int j = j0; j < std::min(j0 + 64, g.Zn - g.N) && !any; ++j) any = want(i, j);
            if (any) out.push_back(make_int2(i0, j0));
        }
    return out;
}

// Orders a tile list for the round-robin dispatch of workgroups over the 8
// XCDs (block b runs on XCD b mod 8): XCD x gets a contiguous run of the
// spatially ordered list, so tiles that share stencil halos share that XCD's
// L2 instead of fetching the halos once per XCD.  The list is padded to a
// multiple of 8 with tiles far outside the grid, on which the band kernels
// return at once.
std::vector<int2> xcd_order(const std::vector<int2> &in) {
    const size_t n8 = (in.size() + 7) / 8 * 8, per = n8 / 8;
    std::vector<int2> out(n8, make_int2(-(1 << 20), -(1 << 20)));
    for (size_t p = 0; p < in.size(); ++p) out[8 * (p % per) + p / per] = in[p];
    return out;
}

// 64 x 4 tiles over the whole padded grid holding a cell within `depth` of the
// edge.
std::vector<int2> edge_tiles(const Grid &g, int depth) {
    std::vector<int2> out;
    for (int i0 = 0; i0 < g.Xn; i0 += 4)
        for (int j0 = 0; j0 < g.Zn; j0 += 64) {
            bool any = false;
            for (int i = i0; i < std::min(i0 + 4, g.Xn) && !any; ++i)
                for (int j = j0; j < std::min(j0 + 64, g.Zn) && !any; ++j)
                    any = std::min(std::min(i, g.Xn - 1 - i), std::min(j, g.Zn - 1 - j)) < depth;
            if (any) out.push_back(make_int2(i0, j0));
        }
    return out;
}

void build_learned(FusedState *st, const Grid &g, hipStream_t s) {
    std::vector<int2> tp = edge_tiles(g, g.L), tm = edge_tiles(g, 2 * g.L);
    st->ntiles_p = (int)tp.size();
    st->ntiles_m = (int)tm.size();
    st->tiles_p.alloc(tp.size() * sizeof(int2));
    st->tiles_m.alloc(tm.size() * sizeof(int2));
    copy_h2d(s, st->tiles_p.ptr, tp.data(), tp.size() * sizeof(int2));
    copy_h2d(s, st->tiles_m.ptr, tm.data(), tm.size() * sizeof(int2));
    // the network ring: velocity cells of the pad that the low-order scheme
    // cannot update and that are not the always-zero ghost line
    std::vector<LbCell> ring;
    for (int i = 0; i < g.Xn; ++i)
        for (int j = 0; j < g.Zn; ++j) {
            const bool gx = i == g.Xn - 1, gz = j == g.Zn - 1;
            const bool live_x = i >= 1 && i < g.Xn - 2 && j >= 1 && j < g.Zn - 1;
            const bool live_z = i >= 1 && i < g.Xn - 1 && j >= 1 && j < g.Zn - 2;
            if (!gx && !live_x) ring.push_back(LbCell{i, j, 0, 0});
            if (!gz && !live_z) ring.push_back(LbCell{i, j, 1, 0});
        }
    // k_lb_ring reads its window from rows staged at canonical depths 0..kLbD-1,
    // which covers every ring cell's window only at canonical depth 0
    for (const LbCell &r : ring) {
        const int SX = r.f == 0 ? 1 : 0, SZ = r.f == 1 ? 1 : 0;
        const int a = std::min(std::min(r.i, g.Xn - 1 - SX - r.i), std::min(r.j, g.Zn - 1 - SZ - r.j));
        if (a != 0) throw Error(OUC_UNSUPPORTED, "learned ring cell off the outermost ring");
    }
    st->lbE = std::max(g.Xn, g.Zn);
    st->nlbcells = (int)ring.size();
    st->lbcells.alloc(ring.size() * sizeof(LbCell));
    copy_h2d(s, st->lbcells.ptr, ring.data(), ring.size() * sizeof(LbCell));
}

size_t per_shot_bytes(const ouc_ctx *c, int fp16) {
    const Grid &g = c->g;
    const size_t plane = g.pitched() * sizeof(float);
    size_t b = 2 * kWaveArrays * plane + 6 * plane + 3 * plane;
    const size_t cells = (size_t)c->ring.ncells, outer = (size_t)c->ring.nouter;
    if (fp16)
        b += (size_t)g.nt * (strip16_offset(cells, outer, RS_NF) * sizeof(__half) +
                             kNInner32 * (cells - outer) * sizeof(float));
    else
        b += (size_t)g.nt * RS_NF * cells * sizeof(float);
    if (fp16)
        b += (size_t)g.nt * RS_NF * ((cells + 63) / 64) +
             (size_t)((g.nt + kStripKey - 1) / kStripKey) * RS_NF * cells * sizeof(float) +
             (size_t)4 * RS_NF * cells * sizeof(float);
    b += (size_t)2 * g.nt * g.nx * sizeof(float) + (size_t)g.nt * sizeof(float) + g.nx * sizeof(int32_t);
    if (g.learned) b += (size_t)lb_canon_floats(std::max(g.Xn, g.Zn)) * sizeof(float);
    return b;
}

void ensure(ouc_ctx *c, int shots) {
    FusedState *st = fs(c);
    const int want16 = c->p.strip_precision == OUC_STRIP_FP16;
    if (st->cap >= shots && st->fp16 == want16) return;
    st->drop_graphs();
    const Grid &g = c->g;
    st->g = make_fgrid(c);
    st->g.csx = st->csx;
    st->g.lb_beta = st->lb_beta;
    st->plane = st->g.plane;
    st->cap = shots;
    st->fp16 = want16;
    wave_layout(st->mem_s, st->S, st->plane, shots);
    wave_layout(st->mem_r, st->R, st->plane, shots);
    const size_t span = (size_t)st->plane * shots * sizeof(float);
    for (auto &lv : st->vp) lv.alloc(2 * span);
    for (auto &b : st->acc) b.alloc(span);
    st->nslots = g.nt;
    st->nchunks = (c->ring.ncells + 63) / 64;
    st->shot_strip = (long)st->nslots * (want16 ? strip16_offset(c->ring.ncells, c->ring.nouter, RS_NF)
                                                : (long)RS_NF * c->ring.ncells);
    st->shot_in32 = want16 ? (long)st->nslots * kNInner32 * (c->ring.ncells - c->ring.nouter) : 0;
    st->shot_chunks = (long)st->nslots * RS_NF * st->nchunks;
    st->shot_key = (long)((st->nslots + kStripKey - 1) / kStripKey) * RS_NF * c->ring.ncells;
    st->shot_aux = (long)2 * RS_NF * c->ring.ncells;
    st->strips32.release(); st->strips16.release(); st->e8.release();
    st->keyb.release(); st->ef.release(); st->decb.release(); st->in32.release();
    if (want16) {
        st->strips16.alloc((size_t)st->shot_strip * shots * sizeof(__half));
        st->in32.alloc((size_t)st->shot_in32 * shots * sizeof(float));
        st->e8.alloc((size_t)st->shot_chunks * shots);
        st->keyb.alloc((size_t)st->shot_key * shots * sizeof(float));
        st->ef.alloc((size_t)st->shot_aux * shots * sizeof(float));
        st->decb.alloc((size_t)st->shot_aux * shots * sizeof(float));
    } else {
        st->strips32.alloc((size_t)st->shot_strip * shots * sizeof(float));
    }
    st->rec.alloc((size_t)shots * 2 * g.nt * g.nx * sizeof(float));
    st->srcpos.alloc((size_t)shots * 2 * sizeof(int));
    st->t0.alloc((size_t)shots * sizeof(int));
    st->wav.alloc((size_t)shots * g.nt * sizeof(float));
    st->counts.alloc((size_t)shots * g.nx * sizeof(int32_t));
    st->img.alloc((size_t)4 * g.nx * g.nz * sizeof(float));
    const FGrid &f = st->g;
    auto core = [&](int i, int j) { return i >= f.cx0 && i < f.cx1 && j >= f.cz0 && j < f.cz1; };
    auto deep = [&](int i, int j) {
        return i >= f.cx0 + g.N && i < f.cx1 - g.N && j >= f.cz0 + g.N && j < f.cz1 - g.N;
    };
    std::vector<int2> ts = xcd_order(tile_list(f, [&](int i, int j) { return !deep(i, j); }));
    std::vector<int2> tv = xcd_order(tile_list(f, [&](int i, int j) { return !core(i, j); }));
    st->ntiles_s = (int)ts.size();
    st->ntiles_v = (int)tv.size();
    st->edge = f.lpad == 0;
    if (st->edge) {
        // band rectangles: top and bottom full width, left and right between them
        const int N = g.N;
        const int rx[4][4] = {{N, g.Xn - N, N, f.cz0},
                              {N, g.Xn - N, f.cz1, g.Zn - N},
                              {N, f.cx0, f.cz0, f.cz1},
                              {f.cx1, g.Xn - N, f.cz0, f.cz1}};
        // The source's forward steps stream 32-column chunks: at a few shots
        // per batch the band has under one wave per SIMD and its latency is
        // exposed, and twice the blocks of half the length run it about 10 %
        // faster.  The receiver's band runs beside the reconstruction on the
        // other stream and keeps 64 (fewer halo columns re-read).
        const int chunk[2] = {32, 64};
        for (int v = 0; v < 2; ++v) {
            EPlan &pl = st->eplan[v];
            pl.xchunk = chunk[v];
            const int TZ = kEdgeW - 4 * N;
            int nb = 0;
            for (int r = 0; r < 4; ++r) {
                ERect &R = pl.r[r];
                R.x0 = rx[r][0]; R.x1 = rx[r][1]; R.z0 = rx[r][2]; R.z1 = rx[r][3];
                const bool any = R.x1 > R.x0 && R.z1 > R.z0;
                R.nzb = any ? (R.z1 - R.z0 + TZ - 1) / TZ : 1;
                R.nxb = any ? (R.x1 - R.x0 + pl.xchunk - 1) / pl.xchunk : 0;
                R.first = nb;
                nb += R.nzb * R.nxb;
            }
            st->eblocks[v] = nb;
        }
    }
    st->tiles_s.alloc(ts.size() * sizeof(int2));
    st->tiles_v.alloc(tv.size() * sizeof(int2));
    copy_h2d(c->stream, st->tiles_s.ptr, ts.data(), ts.size() * sizeof(int2));
    copy_h2d(c->stream, st->tiles_v.ptr, tv.data(), tv.size() * sizeof(int2));
    // reconstruction band: stress on the ring frame minus [L+3N, Xn-L-3N)^2,
    // velocity on the interior minus the recon core [L+2N, Xn-L-2N)^2
    const int L = g.L, N = g.N;
    std::vector<int2> trs, trv;
    for (int i0 = L - N; i0 < g.Xn - L + N; i0 += 4)
        for (int j0 = L - N; j0 < g.Zn - L + N; j0 += 64) {
            bool s_any = false, v_any = false;
            for (int i = i0; i < std::min(i0 + 4, g.Xn - L + N); ++i)
                for (int j = j0; j < std::min(j0 + 64, g.Zn - L + N); ++j) {
                    bool deep3 = i >= L + 3 * N && i < g.Xn - L - 3 * N && j >= L + 3 * N && j < g.Zn - L - 3 * N;
                    bool interior = i >= L && i < g.Xn - L && j >= L && j < g.Zn - L;
                    bool core2 = i >= L + 2 * N && i < g.Xn - L - 2 * N && j >= L + 2 * N && j < g.Zn - L - 2 * N;
                    s_any |= !deep3;
                    v_any |= interior && !core2;
                }
            if (s_any) trs.push_back(make_int2(i0, j0));
            if (v_any) trv.push_back(make_int2(i0, j0));
        }
    trs = xcd_order(trs);
    trv = xcd_order(trv);
    st->ntiles_rs = (int)trs.size();
    st->ntiles_rv = (int)trv.size();
    st->tiles_rs.alloc(trs.size() * sizeof(int2));
    st->tiles_rv.alloc(trv.size() * sizeof(int2));
    copy_h2d(c->stream, st->tiles_rs.ptr, trs.data(), trs.size() * sizeof(int2));
    copy_h2d(c->stream, st->tiles_rv.ptr, trv.data(), trv.size() * sizeof(int2));
    st->mon.alloc((size_t)2 * shots * 2 * sizeof(unsigned int)); // [pass][shot][pad, band]
    if (g.learned) {
        build_learned(st, g, c->stream);
        st->lbd.alloc((size_t)shots * st->nlbcells * sizeof(float));
        st->lbc.alloc((size_t)shots * lb_canon_floats(st->lbE) * sizeof(float));
    }
}

KStrips base_strips(const FusedState *st, const ouc_ctx *c) {
    KStrips s{};
    s.f32 = st->strips32.f();
    s.f16 = st->strips16.as<__half>();
    s.in32 = st->in32.f();
    s.e8 = st->e8.as<int8_t>();
    s.key = st->keyb.f();
    s.err = st->ef.f();
    s.dec = st->decb.f();
    s.nouter = c->ring.nouter;
    s.t0 = st->t0.as<int>();
    s.last_v = st->nslots - 2;  // velocity totals V^{t+1} are saved for t < nt-1
    s.last_th = st->nslots - 1; // the stress-like fields for every t
    s.ncells = c->ring.ncells;
    s.nchunks = st->nchunks;
    s.fp16 = st->fp16;
    return s;
}

// Direct-wave mute for a batch: counts[shot][nx].
__global__ void k_fmute(int nx, int nt, const int32_t *__restrict__ counts, float *__restrict__ rec) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y, s = blockIdx.z;
    if (ix >= nx || t >= nt) return;
    if (t < counts[(long)s * nx + ix]) {
        long o = ((long)s * 2 * nt + t) * nx + ix;
        rec[o] = 0.f;
        rec[o + (long)nt * nx] = 0.f;
    }
}

struct Launch {
    dim3 core, band_s, band_v, recon_s, recon_v, recon_core, full;
    bool has_core, has_recon_core;
    FGrid rg; // the grid with the reconstruction core bounds
};

Launch launches(const FusedState *st, int B) {
    const FGrid &f = st->g;
    Launch L{};
    L.has_core = f.cx1 > f.cx0 && f.cz1 > f.cz0;
    L.core = dim3((f.cx1 - f.cx0 + f.csx - 1) / f.csx, core_zblocks(f.cz0, f.cz1), B);
    L.band_s = dim3(st->ntiles_s, 1, B);
    L.band_v = dim3(st->ntiles_v, 1, B);
    L.recon_s = dim3(st->ntiles_rs, 1, B);
    L.recon_v = dim3(st->ntiles_rv, 1, B);
    L.rg = f;
    L.rg.cx0 = f.L + 2 * f.N; L.rg.cx1 = f.Xn - f.L - 2 * f.N;
    L.rg.cz0 = f.L + 2 * f.N; L.rg.cz1 = f.Zn - f.L - 2 * f.N;
    L.has_recon_core = L.rg.cx1 > L.rg.cx0 && L.rg.cz1 > L.rg.cz0;
    if (!L.has_recon_core) { L.rg.cx1 = L.rg.cx0; L.rg.cz1 = L.rg.cz0; }
    L.recon_core = dim3((L.rg.cx1 - L.rg.cx0 + f.csx - 1) / f.csx, core_zblocks(L.rg.cz0, L.rg.cz1), B);
    L.full = dim3((f.Zn + 63) / 64, (f.Xn + 3) / 4, B);
    return L;
}

// Sampled timing of k_core in profiling mode: event pairs around every
// `stride`-th launch.
struct CoreProbe {
    std::vector<hipEvent_t> ev;
    int stride = 1, used = 0;
    void init(int launches, bool on) {
        if (!on) return;
        int slots = std::min(launches, 512);
        stride = std::max(1, launches / slots);
        ev.resize(2 * slots);
        for (auto &e : ev) OUC_HIP(hipEventCreate(&e));
    }
    bool want(int n) const { return !ev.empty() && n % stride == 0 && used < (int)ev.size() / 2; }
    double total_ms() {
        double t = 0;
        for (int a = 0; a < used; ++a) {
            float ms = 0;
            OUC_HIP(hipEventElapsedTime(&ms, ev[2 * a], ev[2 * a + 1]));
            t += ms;
        }
        return t;
    }
    ~CoreProbe() {
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
};

// One wavefield step.  save_ring: this step's strips are saved afterwards (the
// streaming band kernel then stores the stress of the ring cells).
void wave_step_probed(ouc_ctx *c, FusedState *st, const Launch &L, FWave &w, const FStep &a, CoreProbe &pr,
                      int &nlaunch, unsigned int *mon, bool save_ring = false) {
    const FGrid &f = st->g;
    const KModel km = c->kmodel();
    hipStream_t s = c->stream;
    const int B = (int)L.core.z;
    if (f.lpad)
        k_lpad_stress<<<dim3(st->ntiles_p, 1, B), dim3(64, 4), 0, s>>>(f, w, km, st->tiles_p.as<int2>(), a);
    if (st->edge) {
        const int v = a.inj ? 1 : 0; // the receiver run injects records
        EPlan pl = st->eplan[v];
        pl.save_ring = save_ring ? 1 : 0;
        if (a.sp) {
            OUC_FDISPATCH(f.N, k_edge<NN, true><<<dim3(st->eblocks[v], 1, B), kEdgeW, 0, s>>>(f, w, km, a, pl));
        } else {
            OUC_FDISPATCH(f.N, k_edge<NN, false><<<dim3(st->eblocks[v], 1, B), kEdgeW, 0, s>>>(f, w, km, a, pl));
        }
    } else {
        OUC_FDISPATCH(f.N, k_band_stress<NN><<<L.band_s, dim3(64, 4), 0, s>>>(f, w, km, st->tiles_s.as<int2>(), a));
    }
    if (L.has_core) {
        bool p = pr.want(nlaunch++);
        if (p) OUC_HIP(hipEventRecord(pr.ev[2 * pr.used], s));
        OUC_FDISPATCH(f.N, launch_core<NN, 0>(st->pf, L.core, s, f, w, c->pq.as<float2>(), a, FRecon{}));
        if (p) { OUC_HIP(hipEventRecord(pr.ev[2 * pr.used + 1], s)); ++pr.used; }
    }
    if (!st->edge) {
        OUC_FDISPATCH(f.N, k_band_velocity<NN><<<L.band_v, dim3(64, 4), 0, s>>>(f, w, km, st->tiles_v.as<int2>(), a));
    }
    if (f.lpad) {
        // the ring reads the pad's increments before the pad update advances them
        const float sc = f.dt2 * f.idx;
        const dim3 rg((st->nlbcells + 255) / 256, 1, B);
        if (st->lb_have)
            k_lb_canon<<<dim3((st->lbE + 255) / 256, 4 * kLbD, B), 256, 0, s>>>(f, w, km, sc, st->lbc.f(), st->lbE);
        k_lb_ring<<<rg, 256, 0, s>>>(f, w, st->lbcells.as<LbCell>(), st->nlbcells, st->lb_have ? st->lbw.f() : nullptr,
                                     a.rb, st->lbc.f(), st->lbE, st->lbd.f());
        if (st->lb_have) k_lb_apply<<<rg, 256, 0, s>>>(f, w, st->lbcells.as<LbCell>(), st->nlbcells, st->lbd.f(), a.rb);
        k_lpad_velocity<<<dim3(st->ntiles_p, 1, B), dim3(64, 4), 0, s>>>(f, w, km, st->tiles_p.as<int2>(), a.rb);
        if (a.t % kMonStride == 0)
            k_lb_monitor<<<dim3(st->ntiles_m, 1, B), dim3(64, 4), 0, s>>>(f, w, st->tiles_m.as<int2>(), 1 - a.rb, mon);
    }
}

// Crop of V (both components) of shot 0 into out[2][w][h] (x-major).
__global__ void k_crop_v(FGrid g, const float2 *__restrict__ v, int x0, int z0, int w, int h, float *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= h || i >= w) return;
    const long k = (long)(x0 + i) * g.P + (z0 + j);
    const float2 vk = v[k];
    out[(long)i * h + j] = vk.x;
    out[(long)w * h + (long)i * h + j] = vk.y;
}

// Wavefield snapshots taken after the listed forward steps (increasing).
struct Snap {
    const int *steps = nullptr;
    int n = 0;
    int x0 = 0, z0 = 0, w = 0, h = 0; // padded-grid rectangle
    float *dev = nullptr;             // [n][2][w][h]
};

// Runs B shots through forward (+ strips) and, with rtm, the backward loop;
// leaves the accumulators of every shot in st->acc.  With the learned
// boundary, st->mon holds each shot's boundary monitor afterwards.
// FWI inputs of a batch: observed records [B][2][nx][nt] on the host; the
// backward loop is driven by the residual and accumulates the gradient.
struct FwiBatch {
    const float *observed;
    int rblocks;          // residual blocks per shot (partial sums in st->part)
};

// A batch's per-shot inputs on the device and its source state zeroed: what
// the time loops of run_batch and of a grouped run start from.  `mute`: apply
// the shots' direct-wave mute to the records.
struct BatchIn {
    int B = 0, zr = -1, rblocks = 0;
    bool exact = false;
    Launch L{};
    KRing kr{};
    KStrips ks0{};
    KStripStride sstride{};
    unsigned int *mon_f = nullptr, *mon_r = nullptr;
    size_t span = 0; // bytes of one [shot][plane] array of the batch
};

BatchIn batch_begin(ouc_ctx *c, const ouc_shot *shots, int B, bool rtm, bool mute) {
    FusedState *st = fs(c);
    const Grid &g = c->g;
    hipStream_t s = c->stream;
    const int nt = g.nt;
    // per-shot inputs
    std::vector<int> pos(2 * B), start(B, 0);
    std::vector<float> wavs((size_t)B * nt, 0.f);
    std::vector<int32_t> counts((size_t)B * g.nx, 0);
    int zr = -1;
    for (int b = 0; b < B; ++b) {
        const ouc_shot &sh = shots[b];
        const int nts = sh.nt > 0 ? sh.nt : nt;
        if (nts < 2 || nts > nt) throw Error(OUC_INVALID_ARG, "shot nt out of range");
        start[b] = nt - nts; // a shorter shot starts late and ends with the others
        pos[2 * b] = sh.sx + g.L;
        pos[2 * b + 1] = sh.sz + g.L;
        std::vector<float> w;
        if (sh.wavelet) w.assign(sh.wavelet, sh.wavelet + nts);
        else ricker(nts, sh.fm, g.dt, w);
        std::copy(w.begin(), w.end(), wavs.begin() + (size_t)b * nt + start[b]);
        if (zr < 0) zr = sh.rec_z + g.L;
        if (sh.rec_z + g.L != zr) throw Error(OUC_INVALID_ARG, "all shots of a batch must share the receiver row");
        if (sh.mute && mute) {
            std::vector<int32_t> cnt;
            mute_counts(g, sh, c->vp_pad[(size_t)(sh.sx + g.L) * g.P + zr], cnt);
            for (int ix = 0; ix < g.nx; ++ix) counts[(size_t)b * g.nx + ix] = std::min(nt, cnt[ix] + start[b]);
        }
    }
    OUC_HIP(hipMemcpyAsync(st->t0.ptr, start.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
    OUC_HIP(hipMemcpyAsync(st->srcpos.ptr, pos.data(), pos.size() * sizeof(int), hipMemcpyHostToDevice, s));
    OUC_HIP(hipMemcpyAsync(st->wav.ptr, wavs.data(), wavs.size() * sizeof(float), hipMemcpyHostToDevice, s));
    OUC_HIP(hipMemcpyAsync(st->counts.ptr, counts.data(), counts.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    const size_t span = (size_t)st->plane * B * sizeof(float);
    OUC_HIP(hipMemsetAsync(st->mem_s.ptr, 0, (size_t)kWaveArrays * st->plane * st->cap * sizeof(float), s));
    OUC_HIP(hipMemsetAsync(st->mon.ptr, 0, st->mon.bytes, s));
    if (rtm && st->ef.ptr) OUC_HIP(hipMemsetAsync(st->ef.ptr, 0, st->ef.bytes, s));
    unsigned int *mon_f = st->mon.as<unsigned int>(), *mon_r = mon_f + 2 * st->cap;
    const KRing kr = c->kring();
    const KStrips ks0 = base_strips(st, c);
    const KStripStride sstride{st->shot_strip, st->shot_chunks, st->shot_key, st->shot_aux, st->shot_in32};
    const int rblocks = (int)((c->ring.ncells + 255) / 256);
    OUC_HIP(hipStreamSynchronize(s)); // the host arrays above go out of scope
    BatchIn bi;
    bi.B = B;
    bi.zr = zr;
    bi.rblocks = rblocks;
    bi.exact = c->p.semantics == OUC_SEM_EXACT;
    bi.L = launches(st, B);
    bi.kr = kr;
    bi.ks0 = ks0;
    bi.sstride = sstride;
    bi.mon_f = mon_f;
    bi.mon_r = mon_r;
    bi.span = span;
    return bi;
}

// After the forward loop: direct-wave mute of the records, hand-off of the
// source state (V^nt in buffer nt mod 2; V^{nt-1} into the other; source P
// levels nt and nt-1), receiver state and imaging accumulators zeroed.
void batch_handoff(ouc_ctx *c, const BatchIn &bi) {
    FusedState *st = fs(c);
    const Grid &g = c->g;
    hipStream_t s = c->stream;
    const int nt = g.nt, rb = nt & 1;
    {
        dim3 gm((g.nx + 255) / 256, nt, bi.B);
        k_fmute<<<gm, 256, 0, s>>>(g.nx, nt, st->counts.as<int32_t>(), st->rec.f());
    }
    k_funstep<<<bi.L.full, dim3(64, 4), 0, s>>>(st->g, st->S, rb, st->vp[nt % 3].as<float2>(),
                                                st->vp[(nt - 1) % 3].as<float2>());
    OUC_HIP(hipMemsetAsync(st->mem_r.ptr, 0, (size_t)kWaveArrays * st->plane * st->cap * sizeof(float), s));
    for (auto &b : st->acc) OUC_HIP(hipMemsetAsync(b.ptr, 0, bi.span, s));
}

void run_batch(ouc_ctx *c, const ouc_shot *shots, int B, bool rtm, int ncomp, CoreProbe &probe,
               const Snap *snap = nullptr, const float *rec_in = nullptr, const Snap *rsnap = nullptr,
               const FwiBatch *fwi = nullptr) {
    FusedState *st = fs(c);
    const Grid &g = c->g;
    const FGrid &f = st->g;
    hipStream_t s = c->stream;
    const BatchIn bi = batch_begin(c, shots, B, rtm, !fwi);
    const Launch &L = bi.L;
    const bool exact = bi.exact;
    const int nt = g.nt, zr = bi.zr, rblocks = bi.rblocks;
    unsigned int *mon_f = bi.mon_f, *mon_r = bi.mon_r;
    const KRing kr = bi.kr;
    const KStrips ks0 = bi.ks0;
    const KStripStride sstride = bi.sstride;
    int nlaunch = 0;

    // forward
    FStep a{};
    a.src = st->srcpos.as<int>();
    a.wav = st->wav.f();
    a.zr = zr;
    a.rec = ncomp == 6 ? st->rec6.f() : st->rec.f();
    a.ncomp = ncomp;
    // The forward loop holds the sampled k_core timing events, so it runs eagerly
    // while profiling; the backward loop is captured either way.
    const bool graph = c->graphs && probe.ev.empty() && !snap && !rsnap;
    const bool graph_b = c->graphs && !snap && !rsnap;
    if (graph_b) st->ensure_side(); // the captured backward loop runs independent work on two streams
    int rb = 0, ksnap = 0;
    run_loop(s, graph, st->gfwd, B, rtm, ncomp, zr, nt, [&](int i0, int i1) {
    for (int t = i0; t < i1; ++t) {
        a.t = t;
        a.rb = rb;
        wave_step_probed(c, st, L, st->S, a, probe, nlaunch, mon_f, rtm);
        if (snap && ksnap < snap->n && snap->steps[ksnap] == t) {
            k_crop_v<<<dim3((snap->h + 255) / 256, snap->w), 256, 0, s>>>(
                f, st->S.V[1 - rb].p, snap->x0, snap->z0, snap->w, snap->h,
                snap->dev + (size_t)ksnap * 2 * snap->w * snap->h);
            ++ksnap;
        }
        if (rtm) {
            FWave ws = st->S; // the stress memory variables of this step (the streaming band kernel double-buffers them)
            if (st->edge && rb == 0)
                for (int q = 0; q < 4; ++q) ws.F[q] = st->S.F2[q];
            // stress-like fields of step t, and (but the last step) V^{t+1}, in one launch
            k_fstrip_save<<<dim3(rblocks, 1, B), 256, 0, s>>>(f, ws, kr, ks0, sstride, t, 1 - rb, t < nt - 1 ? 2 : 0);
        }
        rb = 1 - rb;
    }
    OUC_HIP(hipGetLastError());
    });
    rb = nt & 1; // V^nt is in buffer nt mod 2 (also when the loop was replayed)
    if (!rtm) return;
    if (rec_in) {
        // observed records [2][nx][nt] (receiver-major) replace the modelled ones of shot 0
        std::vector<float> tr((size_t)2 * nt * g.nx);
        for (int q = 0; q < 2; ++q)
            for (int ix = 0; ix < g.nx; ++ix)
                for (int t = 0; t < nt; ++t) tr[((size_t)q * nt + t) * g.nx + ix] = rec_in[((size_t)q * g.nx + ix) * nt + t];
        OUC_HIP(hipMemcpyAsync(st->rec.ptr, tr.data(), tr.size() * sizeof(float), hipMemcpyHostToDevice, s));
        OUC_HIP(hipStreamSynchronize(s));
    }
    if (fwi) { // adsource (P-S-seperating.cu:1484-1498): modelled - observed on every receiver and step
        std::vector<float> tr((size_t)B * 2 * nt * g.nx);
        for (int b = 0; b < B; ++b)
            for (int q = 0; q < 2; ++q)
                for (int ix = 0; ix < g.nx; ++ix) {
                    const float *src = fwi->observed + (((size_t)b * 2 + q) * g.nx + ix) * nt;
                    float *dst = tr.data() + ((size_t)b * 2 + q) * nt * g.nx + ix;
                    for (int t = 0; t < nt; ++t) dst[(size_t)t * g.nx] = src[t];
                }
        OUC_HIP(hipMemcpyAsync(st->obs.ptr, tr.data(), tr.size() * sizeof(float), hipMemcpyHostToDevice, s));
        k_fresidual<<<dim3(fwi->rblocks, B), 256, 0, s>>>((long)2 * nt * g.nx, st->rec.f(), st->obs.f(),
                                                           st->part.as<double>());
        OUC_HIP(hipStreamSynchronize(s)); // the host buffer goes out of scope
    }
    c->last.forward_ms = 0; // filled by the caller from events
    OUC_HIP(hipEventRecord(c->ev[1], s));
    batch_handoff(c, bi);
    int rbs = 1 - rb;
    const size_t gspan = (size_t)st->plane * st->cap; // floats per gradient accumulator
    if (fwi) OUC_HIP(hipMemsetAsync(st->gacc.ptr, 0, 3 * gspan * sizeof(float), s));
    int krs = 0;
    FStep r{};
    r.zr = zr;
    r.inj = st->rec.f();
    r.npp = st->acc[0].f(); r.nps = st->acc[1].f(); r.den = st->acc[2].f();
    r.t0 = st->t0.as<int>();
    int rbr = 0;
    // The reconstruction of iteration t writes source P level t and reads only
    // the source state and the strips; the receiver step of the same iteration
    // reads levels t+1 / t+2 (written in earlier iterations) and its own state.
    // So the two run on two streams, joined at the end of every iteration.
    // Eager launches (profiling) stay on one stream.
    hipStream_t sr = graph_b ? st->s2 : s;
    // FP16 strips decode backwards one slot per step, on the reconstruction's
    // stream: iteration t decodes slot t, and the reconstruction of step t reads
    // slots t+1 and (EXACT) t
    const dim3 dgrid(rblocks, RS_NF, B);
    run_loop(s, graph_b, st->gbwd, B, fwi ? 2 : 1, ncomp, zr, nt, [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i) {
        const int t = nt - 1 - i; // the backward loop runs t = nt-1 .. 0
        if (graph_b) {
            OUC_HIP(hipEventRecord(st->evf, s));
            OUC_HIP(hipStreamWaitEvent(sr, st->evf, 0));
        }
        if (ks0.fp16) k_strip_decode<<<dgrid, 256, 0, sr>>>(ks0, t, sstride); // slot t (t = nt-1: the first)
        if (t < nt - 1) {
            // theta^{t+1} from V^{t+1}: ring totals from slot t+1 (FAITHFUL, one step late as the
            // reference) or slot t (EXACT); the hand-off step uses the forward's own state
            FRecon rc{};
            rc.ring = kr;
            rc.strips = ks0;
            rc.stride = sstride;
            const bool handoff = t == nt - 2;
            rc.slot_v = handoff ? -1 : (exact ? t : t + 1);
            rc.slot_th = t + 1;
            rc.exact = (exact || handoff) ? 1 : 0;
            rc.vp = fwi ? nullptr : st->vp[t % 3].as<float2>();
            FStep ra{};
            ra.t = t + 1;
            ra.rb = rbs;
            ra.src = st->srcpos.as<int>();
            ra.wav = st->wav.f();
            OUC_FDISPATCH(f.N, k_recon_stress_f<NN><<<L.recon_s, dim3(64, 4), 0, sr>>>(
                                   f, st->S, c->p2.f(), c->q2.f(), rc, ra, st->tiles_rs.as<int2>()));
            if (L.has_recon_core)
                OUC_FDISPATCH(f.N, launch_core<NN, 1>(st->pf, L.recon_core, sr, L.rg, st->S, c->pq.as<float2>(), ra, rc));
            OUC_FDISPATCH(f.N, k_recon_velocity_f<NN><<<L.recon_v, dim3(64, 4), 0, sr>>>(f, st->S, c->q2.f(), rc, ra,
                                                                                         st->tiles_rv.as<int2>()));
            rbs = 1 - rbs;
            if (rsnap && krs < rsnap->n && rsnap->steps[krs] == t) { // reconstructed V^t
                k_crop_v<<<dim3((rsnap->h + 255) / 256, rsnap->w), 256, 0, sr>>>(
                    f, st->S.V[rbs].p, rsnap->x0, rsnap->z0, rsnap->w, rsnap->h,
                    rsnap->dev + (size_t)krs * 2 * rsnap->w * rsnap->h);
                ++krs;
            }
        }
        const int lev = exact ? t + 1 : std::min(t + 2, nt);
        r.t = t;
        r.rb = rbr;
        // the FWI gradient needs no image: its receiver step runs without the imaging condition
        r.sp = fwi ? nullptr : st->vp[lev % 3].as<float2>();
        CoreProbe none;
        int nl2 = 0;
        wave_step_probed(c, st, L, st->R, r, none, nl2, mon_r);
        rbr = 1 - rbr;
        if (graph_b) {
            OUC_HIP(hipEventRecord(st->evj, sr));
            OUC_HIP(hipStreamWaitEvent(s, st->evj, 0));
        }
        if (fwi) {
            // gradfun on this step: source V^{t+1} (the buffer this step's reconstruction
            // read, ring totals from the same slot), receiver V before its update with the
            // injected residual; the next iteration's reconstruction overwrites the source
            // buffer only after this launch (it forks from this stream)
            FRecon grc{};
            grc.ring = kr;
            grc.strips = ks0;
            grc.stride = sstride;
            grc.slot_v = t >= nt - 2 ? -1 : (exact ? t : t + 1);
            FStep ga = r;
            ga.rb = 1 - rbr;
            const dim3 gg(grad_blocks(g.nx, g.nz), 1, B);
            float *ga0 = st->gacc.f();
            OUC_FDISPATCH(f.N, k_fgrad<NN><<<gg, dim3(64, 4), 0, s>>>(f, st->S, 1 - rbs, grc, st->R, ga, c->p2.f(),
                                                                     c->q2.f(), ga0, ga0 + gspan, ga0 + 2 * gspan));
        }
    }
    OUC_HIP(hipGetLastError());
    });
}

// Images of shot b of the batch into c->img[0..3].
void shot_images(ouc_ctx *c, FusedState *st, int b) {
    const Grid &g = c->g;
    const KGrid kg = c->kgrid();
    const size_t off = (size_t)b * st->plane;
    dim3 grid((g.nz + 255) / 256, g.nx);
    k_image<<<grid, 256, 0, c->stream>>>(kg, st->acc[0].f() + off, st->acc[1].f() + off, st->acc[2].f() + off,
                                         c->img[0].f(), c->img[1].f());
    if (c->p.laplacian == OUC_LAP_STENCIL) {
        k_laplace5<<<grid, 256, 0, c->stream>>>(g.nx, g.nz, g.dx, c->img[0].f(), c->img[2].f());
        k_laplace5<<<grid, 256, 0, c->stream>>>(g.nx, g.nz, g.dx, c->img[1].f(), c->img[3].f());
    } else {
        lap_fft(c, c->img[0].f(), c->img[2].f());
        lap_fft(c, c->img[1].f(), c->img[3].f());
    }
    OUC_HIP(hipGetLastError());
}

int default_capacity(ouc_ctx *c, int nshots, size_t extra_per_shot = 0) {
    size_t freeb = 0, total = 0;
    OUC_HIP(hipMemGetInfo(&freeb, &total));
    // keep what this context already holds available for reuse
    FusedState *st = fs(c);
    size_t held = st->mem_s.bytes + st->mem_r.bytes + st->strips32.bytes + st->strips16.bytes + st->in32.bytes + st->e8.bytes +
                  st->keyb.bytes + st->ef.bytes + st->decb.bytes + st->gacc.bytes + st->obs.bytes;
    size_t avail = (size_t)((freeb + held) * 0.85);
    size_t per = per_shot_bytes(c, c->p.strip_precision == OUC_STRIP_FP16) + extra_per_shot;
    int cap = (int)std::max<size_t>(1, avail / per);
    if (c->batch_cap > 0) cap = std::min(cap, c->batch_cap);
    return std::min(cap, nshots);
}

} // namespace

void fused_init(ouc_ctx *c) { c->fused = new FusedState(); }

void fused_destroy(ouc_ctx *c) {
    delete fs(c);
    c->fused = nullptr;
}

double fused_own_bytes(const ouc_ctx *c) {
    const Grid &g = c->g;
    const double upd = (double)(g.Xn - 2 * g.N) * (g.Zn - 2 * g.N), interior = (double)g.nx * g.nz;
    const double band = upd - (double)std::max(0, g.Xn - 2 * (g.L + g.N)) * std::max(0, g.Zn - 2 * (g.L + g.N));
    const double core = upd - band;
    // core: state 8 R + 8 W, vp^2, vs^2; band: + 6 stress W/R, F/O, damping;
    // receiver core: + source P level and accumulators; reconstruction: stress
    // store then reverse step
    const double fwd = core * 72 + band * (72 + 4 * (6 + 6 + 2 + 8 + 16));
    const double rcv = fwd + interior * (8 + 24);
    const double rec = interior * 4 * (2 + 2 + 6 + 6 + 1 + 8 + 8 + 2);
    return (double)g.nt * (fwd + rcv + rec);
}

// Boundary verdict per shot of the last run_batch: a shot diverged when its
// pad holds a non-finite value or more than kGrowth times the largest
// amplitude that crossed the band just inside the pad (waves reach the pad only
// through that band, so a stable boundary stays below it).
constexpr float kGrowth = 10.f;

std::vector<char> monitor_verdict(ouc_ctx *c, int B) {
    std::vector<char> bad(B, 0);
    if (!c->g.learned) return bad;
    FusedState *st = fs(c);
    std::vector<unsigned int> m((size_t)4 * st->cap);
    OUC_HIP(hipMemcpyAsync(m.data(), st->mon.ptr, m.size() * sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
    OUC_HIP(hipStreamSynchronize(c->stream));
    double worst = 0;
    for (int b = 0; b < B; ++b)
        for (int pass = 0; pass < 2; ++pass) {
            float pad, band;
            std::memcpy(&pad, &m[(size_t)2 * (pass * st->cap + b)], 4);
            std::memcpy(&band, &m[(size_t)2 * (pass * st->cap + b) + 1], 4);
            const bool ok = std::isfinite(pad) && std::isfinite(band) && pad <= kGrowth * band;
            if (!ok) bad[b] = 1;
            const double ratio = band > 0 ? (double)pad / band : (pad > 0 ? INFINITY : 0.0);
            worst = std::max(worst, std::isfinite(pad) ? ratio : INFINITY);
        }
    c->last.boundary_ratio = std::max(c->last.boundary_ratio, worst);
    if (c->p.boundary == OUC_BOUNDARY_LEARNED)
        for (int b = 0; b < B; ++b)
            if (bad[b]) throw Error(OUC_DEVICE, "learned boundary diverged (no CPML fallback requested)");
    return bad;
}

// Re-runs the listed shots on the CPML fallback context and adds their
// Laplacian images to the host stacks; returns its device time.  With
// shot_lap, each listed shot runs alone and its images go to its slot there
// (a one-shot stack is that shot's image).
double rerun_with_cpml(ouc_ctx *c, const ouc_shot *shots, const std::vector<int> &idx, float *stack_pp,
                       float *stack_ps, float *per_shot_ms, double t0_ms, float *shot_lap = nullptr) {
    ouc_ctx *fb = fallback_ctx(c);
    const size_t ni = (size_t)c->g.nx * c->g.nz;
    if (shot_lap) {
        double t = 0;
        for (int k : idx) {
            float *pp = shot_lap + (size_t)k * 2 * ni;
            float ms = 0;
            ouc_status st = ouc_rtm_batch(fb, &shots[k], 1, pp, pp + ni, &ms);
            if (st != OUC_OK) throw Error(st, std::string("CPML fallback: ") + ouc_last_error(fb));
            for (size_t a = 0; a < ni; ++a) {
                if (stack_pp) stack_pp[a] += pp[a];
                if (stack_ps) stack_ps[a] += pp[ni + a];
            }
            t += fb->last.total_ms;
            if (per_shot_ms) per_shot_ms[k] = (float)(t0_ms + t);
        }
        return t;
    }
    std::vector<ouc_shot> sub;
    for (int k : idx) sub.push_back(shots[k]);
    std::vector<float> pp(ni), ps(ni), ms(sub.size());
    ouc_status st = ouc_rtm_batch(fb, sub.data(), (int32_t)sub.size(), pp.data(), ps.data(), ms.data());
    if (st != OUC_OK) throw Error(st, std::string("CPML fallback: ") + ouc_last_error(fb));
    for (size_t a = 0; a < ni; ++a) {
        if (stack_pp) stack_pp[a] += pp[a];
        if (stack_ps) stack_ps[a] += ps[a];
    }
    if (per_shot_ms)
        for (size_t k = 0; k < idx.size(); ++k) per_shot_ms[idx[k]] = (float)(t0_ms + ms[k]);
    return fb->last.total_ms;
}

void fused_rtm_batch(ouc_ctx *c, const ouc_shot *shots, int nshots, float *stack_pp, float *stack_ps,
                     float *per_shot_ms, float *shot_lap) {
    const Grid &g = c->g;
    for (int k = 0; k < nshots; ++k) {
        const ouc_shot *sh = &shots[k];
        if (sh->sx < 0 || sh->sx >= g.nx || sh->sz < 0 || sh->sz >= g.nz || sh->rec_z < 0 || sh->rec_z >= g.nz)
            throw Error(OUC_INVALID_ARG, "shot " + std::to_string(k) + " outside the interior");
        if (!sh->wavelet && !(sh->fm > 0)) throw Error(OUC_INVALID_ARG, "fm must be positive");
    }
    const int cap = default_capacity(c, nshots);
    ensure(c, cap);
    FusedState *st = fs(c);
    const size_t ni = (size_t)g.nx * g.nz;
    std::vector<int> diverged; // shots whose learned boundary diverged
    DevBuf stack;
    stack.alloc(2 * ni * sizeof(float));
    OUC_HIP(hipMemsetAsync(stack.ptr, 0, 2 * ni * sizeof(float), c->stream));
    c->last = ouc_timings{};
    CoreProbe probe;
    probe.init(g.nt * ((nshots + cap - 1) / cap), c->profiling);
    OUC_HIP(hipEventRecord(c->ev[0], c->stream));
    double fwd_ms = 0, bwd_ms = 0, img_ms = 0;
    for (int b0 = 0; b0 < nshots; b0 += cap) {
        const int B = std::min(cap, nshots - b0);
        OUC_HIP(hipEventRecord(c->ev[4], c->stream));
        run_batch(c, shots + b0, B, true, 2, probe);
        OUC_HIP(hipEventRecord(c->ev[2], c->stream));
        std::vector<char> bad = monitor_verdict(c, B);
        for (int b = 0; b < B; ++b) {
            if (bad[b]) {
                diverged.push_back(b0 + b);
                continue;
            }
            shot_images(c, st, b);
            k_accumulate<<<(unsigned)((ni + 255) / 256), 256, 0, c->stream>>>(ni, c->img[2].f(), stack.f());
            k_accumulate<<<(unsigned)((ni + 255) / 256), 256, 0, c->stream>>>(ni, c->img[3].f(), stack.f() + ni);
            if (shot_lap) { // copied (and synchronized) before the next shot's images overwrite img
                float *dst = shot_lap + (size_t)(b0 + b) * 2 * ni;
                copy_d2h(c->stream, dst, c->img[2].ptr, ni * sizeof(float));
                copy_d2h(c->stream, dst + ni, c->img[3].ptr, ni * sizeof(float));
            }
        }
        OUC_HIP(hipEventRecord(c->ev[3], c->stream));
        OUC_HIP(hipStreamSynchronize(c->stream));
        float a = 0, b = 0, d = 0, e = 0;
        OUC_HIP(hipEventElapsedTime(&a, c->ev[4], c->ev[1]));
        OUC_HIP(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        OUC_HIP(hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
        OUC_HIP(hipEventElapsedTime(&e, c->ev[0], c->ev[3]));
        fwd_ms += a; bwd_ms += b; img_ms += d;
        if (per_shot_ms)
            for (int k = 0; k < B; ++k) per_shot_ms[b0 + k] = e;
    }
    OUC_HIP(hipEventRecord(c->ev[5], c->stream));
    if (stack_pp) OUC_HIP(hipMemcpyAsync(stack_pp, stack.ptr, ni * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (stack_ps)
        OUC_HIP(hipMemcpyAsync(stack_ps, stack.f() + ni, ni * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    OUC_HIP(hipStreamSynchronize(c->stream));
    float tot = 0;
    OUC_HIP(hipEventElapsedTime(&tot, c->ev[0], c->ev[5]));
    double fb_ms = 0;
    if (!diverged.empty())
        fb_ms = rerun_with_cpml(c, shots, diverged, stack_pp, stack_ps, per_shot_ms, tot, shot_lap);
    c->last.total_ms = tot + fb_ms;
    c->last.forward_ms = fwd_ms;
    c->last.backward_ms = bwd_ms;
    c->last.image_ms = img_ms;
    c->last.step_launches = (int64_t)2 * g.nt * ((nshots + cap - 1) / cap);
    c->last.step_kernel_ms = probe.total_ms();
    c->last.step_kernel_launches = probe.used;
    const double padded = (double)g.Xn * g.Zn, interior = (double)g.nx * g.nz, cpml = padded - interior;
    c->last.cell_updates = (double)nshots * g.nt * (2 * padded + interior);
    c->last.canonical_bytes = (double)nshots * g.nt * (2 * (120 * interior + 232 * cpml) + 120 * interior + 48 * interior);
    c->last.own_bytes = nshots * fused_own_bytes(c);
    if (!diverged.empty()) {
        c->last.fallback_shots = (int32_t)diverged.size();
        throw Error(OUC_NUMERIC_FALLBACK, std::to_string(diverged.size()) +
                                              " shot(s) ran again with CPML after the learned boundary diverged");
    }
}

// One job's shots on n contexts of one grid at once.  Batch j (shots
// [j*cap, (j+1)*cap)) runs on context j mod n, each context from its own host
// thread on its own stream, so one batch's latency-bound band kernels run
// beside another batch's streaming core kernels.  Every batch hands its shots'
// Laplacian images to the host, and they are committed in shot order into
// FP32 stacks: the same additions, in the same order, as ouc_rtm_batch's
// device stack, so the stacks are bit for bit those of one context (shots
// that fall back to CPML are added in shot order here, after the others
// there).  A thread starts batch j only when j < next + 2n, where next is
// the first uncommitted batch, which bounds the images held on the host; the
// thread that owns batch next is never held back, so the commits progress.
void fused_rtm_concurrent(ouc_ctx *const *cs, int n, const ouc_shot *shots, int nshots, float *stack_pp,
                          float *stack_ps, float *per_shot_ms) {
    if (n < 1 || n > 8) throw Error(OUC_INVALID_ARG, "1 to 8 contexts");
    const ouc_params &p0 = cs[0]->p;
    for (int i = 0; i < n; ++i) {
        if (!cs[i]) throw Error(OUC_INVALID_ARG, "context " + std::to_string(i) + " is NULL");
        for (int k = 0; k < i; ++k)
            if (cs[k] == cs[i]) throw Error(OUC_INVALID_ARG, "a context appears twice");
        const ouc_params &q = cs[i]->p;
        if (q.engine != OUC_ENGINE_FUSED) throw Error(OUC_INVALID_ARG, "concurrent RTM runs on the fused engine");
        if (cs[i]->device != cs[0]->device) throw Error(OUC_INVALID_ARG, "contexts must share the device");
        if (q.nx != p0.nx || q.nz != p0.nz || q.npml != p0.npml || q.half_order != p0.half_order || q.nt != p0.nt ||
            q.dx != p0.dx || q.dz != p0.dz || q.dt != p0.dt || q.R != p0.R || q.semantics != p0.semantics ||
            q.strip_precision != p0.strip_precision || q.boundary != p0.boundary || q.laplacian != p0.laplacian ||
            q.band != p0.band)
            throw Error(OUC_INVALID_ARG, "contexts must have the same parameters (context " + std::to_string(i) + ")");
    }
    const Grid &g = cs[0]->g;
    const size_t ni = (size_t)g.nx * g.nz;
    // One batch size for all: what n equal shares of the free memory (plus
    // what the contexts already hold) fit, within each context's own cap.
    size_t freeb = 0, total = 0;
    OUC_HIP(hipMemGetInfo(&freeb, &total));
    size_t held = 0;
    for (int i = 0; i < n; ++i) {
        const FusedState *st = fs(cs[i]);
        held += st->mem_s.bytes + st->mem_r.bytes + st->strips32.bytes + st->strips16.bytes + st->in32.bytes +
                st->e8.bytes + st->keyb.bytes + st->ef.bytes + st->decb.bytes + st->gacc.bytes + st->obs.bytes;
    }
    const size_t per = per_shot_bytes(cs[0], p0.strip_precision == OUC_STRIP_FP16);
    int cap = (int)std::max<size_t>(1, (size_t)((freeb + held) * 0.85) / ((size_t)n * per));
    for (int i = 0; i < n; ++i)
        if (cs[i]->batch_cap > 0) cap = std::min(cap, cs[i]->batch_cap);
    cap = std::min(cap, (nshots + n - 1) / n);
    const int nb = (nshots + cap - 1) / cap;
    struct CapPin { // the call's batch size on every context, restored afterwards
        ouc_ctx *const *cs; int n; std::vector<int> saved;
        ~CapPin() { for (int i = 0; i < n; ++i) cs[i]->batch_cap = saved[i]; }
    } pin{cs, n, std::vector<int>(n)};
    for (int i = 0; i < n; ++i) { pin.saved[i] = cs[i]->batch_cap; cs[i]->batch_cap = cap; }

    std::vector<float> spp(ni, 0.f), sps(ni, 0.f);
    std::mutex mu;
    std::condition_variable cv;
    std::map<int, std::vector<float>> pending; // finished, uncommitted batches' images
    int next = 0, fallbacks = 0;
    bool stop = false;
    ouc_status err = OUC_OK;
    std::string msg;
    std::vector<ouc_timings> agg(n);
    const auto t0 = std::chrono::steady_clock::now();
    auto worker = [&](int i) {
        ouc_ctx *c = cs[i];
        for (int j = i; j < nb; j += n) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || j < next + 2 * n; });
                if (stop) return;
            }
            const int b0 = j * cap, B = std::min(cap, nshots - b0);
            std::vector<float> imgs;
            ouc_status st = OUC_OK;
            std::string what;
            try {
                imgs.resize((size_t)B * 2 * ni);
                OUC_HIP(hipSetDevice(c->device));
                fused_rtm_batch(c, shots + b0, B, nullptr, nullptr, nullptr, imgs.data());
            } catch (const Error &e) {
                st = e.code;
                what = e.what();
            } catch (const std::bad_alloc &) {
                st = OUC_OOM;
                what = "host allocation failed";
            } catch (const std::exception &e) {
                st = OUC_DEVICE;
                what = e.what();
            }
            const float done_ms =
                std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
            std::lock_guard<std::mutex> lk(mu);
            if (st == OUC_NUMERIC_FALLBACK) { // valid images; the shots ran again on CPML
                fallbacks += c->last.fallback_shots;
            } else if (st != OUC_OK) {
                if (err == OUC_OK) { err = st; msg = "context " + std::to_string(i) + ": " + what; }
                stop = true;
                cv.notify_all();
                return;
            }
            ouc_timings &a = agg[i];
            a.forward_ms += c->last.forward_ms; a.backward_ms += c->last.backward_ms; a.image_ms += c->last.image_ms;
            a.step_launches += c->last.step_launches; a.step_kernel_ms += c->last.step_kernel_ms;
            a.step_kernel_launches += c->last.step_kernel_launches; a.cell_updates += c->last.cell_updates;
            a.canonical_bytes += c->last.canonical_bytes; a.own_bytes += c->last.own_bytes;
            a.boundary_ratio = std::max(a.boundary_ratio, c->last.boundary_ratio);
            if (per_shot_ms)
                for (int k = 0; k < B; ++k) per_shot_ms[b0 + k] = done_ms;
            try {
                pending.emplace(j, std::move(imgs));
            } catch (const std::bad_alloc &) {
                if (err == OUC_OK) { err = OUC_OOM; msg = "host allocation failed"; }
                stop = true;
                cv.notify_all();
                return;
            }
            for (auto it = pending.find(next); it != pending.end(); it = pending.find(next)) {
                const int c0 = next * cap, cB = std::min(cap, nshots - c0);
                for (int k = 0; k < cB; ++k) {
                    const float *pp = it->second.data() + (size_t)k * 2 * ni, *ps = pp + ni;
                    for (size_t q = 0; q < ni; ++q) { spp[q] += pp[q]; sps[q] += ps[q]; }
                }
                pending.erase(it);
                ++next;
            }
            cv.notify_all();
        }
    };
    std::vector<std::thread> th;
    th.reserve(n);
    try {
        for (int i = 0; i < n; ++i) th.emplace_back(worker, i);
    } catch (...) { // a thread could not start: stop the others before unwinding
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto &t : th) t.join();
        throw;
    }
    for (auto &t : th) t.join();
    if (err != OUC_OK) throw Error(err, msg);
    if (stack_pp) std::memcpy(stack_pp, spp.data(), ni * sizeof(float));
    if (stack_ps) std::memcpy(stack_ps, sps.data(), ni * sizeof(float));
    ouc_timings tot{};
    for (const ouc_timings &a : agg) {
        tot.forward_ms += a.forward_ms; tot.backward_ms += a.backward_ms; tot.image_ms += a.image_ms;
        tot.step_launches += a.step_launches; tot.step_kernel_ms += a.step_kernel_ms;
        tot.step_kernel_launches += a.step_kernel_launches; tot.cell_updates += a.cell_updates;
        tot.canonical_bytes += a.canonical_bytes; tot.own_bytes += a.own_bytes;
        tot.boundary_ratio = std::max(tot.boundary_ratio, a.boundary_ratio);
    }
    tot.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    tot.fallback_shots = fallbacks;
    cs[0]->last = tot;
    if (fallbacks > 0)
        throw Error(OUC_NUMERIC_FALLBACK,
                    std::to_string(fallbacks) + " shot(s) ran again with CPML after the learned boundary diverged");
}

// ---------------------------------------------------------------------------
// Grouped runs: the batches of several contexts (grids of different sizes) in
// shared launches, one per kernel and time step (group_kernels.hpp).
// ---------------------------------------------------------------------------
namespace {

struct GroupGraphs {
    std::vector<unsigned char> key; // the context table and loop shape the graphs were captured for
    DevBuf tab;                     // device copy of the context table
    std::vector<hipGraphExec_t> fwd, bwd; // chunks of kGraphSteps steps, replayed in order
    void drop() {
        for (hipGraphExec_t e : fwd) (void)hipGraphExecDestroy(e);
        for (hipGraphExec_t e : bwd) (void)hipGraphExecDestroy(e);
        fwd.clear();
        bwd.clear();
    }
    ~GroupGraphs() { drop(); }
};
// The loops of the last few group compositions (a run in rounds has one per round).
struct GroupCache {
    static constexpr size_t kKeep = 8;
    std::vector<std::unique_ptr<GroupGraphs>> entries; // most recently used last
    GroupGraphs &get(const std::vector<unsigned char> &key) {
        for (size_t i = 0; i < entries.size(); ++i)
            if (entries[i]->key == key) {
                std::unique_ptr<GroupGraphs> e = std::move(entries[i]);
                entries.erase(entries.begin() + (long)i);
                entries.push_back(std::move(e));
                return *entries.back();
            }
        if (entries.size() >= kKeep) entries.erase(entries.begin());
        entries.push_back(std::make_unique<GroupGraphs>());
        return *entries.back();
    }
};
void drop_group(GroupCache *g) { delete g; }

// Half-orders with grouped kernels.
#define OUC_GDISPATCH(n, ...)                                                               \
    switch (n) {                                                                            \
    case 5: { constexpr int NN = 5; __VA_ARGS__; } break;                                   \
    case 6: { constexpr int NN = 6; __VA_ARGS__; } break;                                   \
    default: throw Error(OUC_UNSUPPORTED, "grouped runs support half-orders 5 and 6");       \
    }

GCtx make_gctx(ouc_ctx *c, const BatchIn &bi, int T) {
    FusedState *st = fs(c);
    const Grid &g = c->g;
    GCtx x;
    std::memset(&x, 0, sizeof x);
    x.g = st->g;
    x.rg = bi.L.rg;
    x.S = st->S;
    x.R = st->R;
    x.km = c->kmodel();
    x.pq = c->pq.as<float2>();
    x.p2 = c->p2.f();
    x.q2 = c->q2.f();
    // step templates, as run_batch sets them for an RTM batch
    x.fa.src = st->srcpos.as<int>();
    x.fa.wav = st->wav.f();
    x.fa.zr = bi.zr;
    x.fa.rec = st->rec.f();
    x.fa.ncomp = 2;
    x.ra.zr = bi.zr;
    x.ra.inj = st->rec.f();
    x.ra.npp = st->acc[0].f();
    x.ra.nps = st->acc[1].f();
    x.ra.den = st->acc[2].f();
    x.ra.t0 = st->t0.as<int>();
    x.sa.src = st->srcpos.as<int>();
    x.sa.wav = st->wav.f();
    x.ring = bi.kr;
    x.ks0 = bi.ks0;
    x.sstride = bi.sstride;
    for (int q = 0; q < 3; ++q) x.vp[q] = st->vp[q].as<float2>();
    for (int v = 0; v < 2; ++v) {
        x.ep[v] = st->eplan[v];
        x.ep[v].save_ring = v == 0 ? 1 : 0; // the source's forward steps save strips
        x.eb[v] = st->eblocks[v];
    }
    x.tiles_s = st->tiles_s.as<int2>(); x.nts = st->ntiles_s;
    x.tiles_v = st->tiles_v.as<int2>(); x.ntv = st->ntiles_v;
    x.tiles_rs = st->tiles_rs.as<int2>(); x.ntrs = st->ntiles_rs;
    x.tiles_rv = st->tiles_rv.as<int2>(); x.ntrv = st->ntiles_rv;
    x.tiles_p = st->tiles_p.as<int2>(); x.ntp = st->ntiles_p;
    x.tiles_m = st->tiles_m.as<int2>(); x.ntm = st->ntiles_m;
    x.lbcells = st->lbcells.as<LbCell>();
    x.nlbcells = st->nlbcells;
    x.lbrb = (st->nlbcells + 255) / 256;
    x.lbw = st->lb_have ? st->lbw.f() : nullptr;
    x.lbd = st->lbd.f();
    x.lbc = st->lbc.f();
    x.lbE = st->lbE;
    x.lbsc = st->g.dt2 * st->g.idx;
    x.mon_f = bi.mon_f;
    x.mon_r = bi.mon_r;
    x.nt = g.nt;
    x.B = bi.B;
    x.start = T - g.nt;
    x.exact = bi.exact ? 1 : 0;
    x.edge = st->edge ? 1 : 0;
    x.rblocks = bi.rblocks;
    x.cbx = bi.L.has_core ? (int)bi.L.core.x : 0;
    x.cby = (int)bi.L.core.y;
    x.rbx = bi.L.has_recon_core ? (int)bi.L.recon_core.x : 0;
    x.rby = (int)bi.L.recon_core.y;
    return x;
}

// One batch per context (B[i] shots of context i) through the RTM path in
// shared launches.  Afterwards each context holds its accumulators, records
// and monitor as after run_batch.
void run_group_batch(const std::vector<ouc_ctx *> &cs, const std::vector<const ouc_shot *> &shots,
                     const std::vector<int> &B) {
    const int n = (int)cs.size();
    ouc_ctx *lead = cs[0];
    FusedState *st0 = fs(lead);
    hipStream_t s = lead->stream;
    int T = 0;
    for (ouc_ctx *c : cs) T = std::max(T, c->g.nt);
    std::vector<BatchIn> bi(n);
    std::vector<GCtx> hc(n);
    for (int i = 0; i < n; ++i) {
        bi[i] = batch_begin(cs[i], shots[i], B[i], true, true); // waits for its context's stream
        hc[i] = make_gctx(cs[i], bi[i], T);
    }
    const int N = lead->g.N;
    const bool learned = lead->g.learned, edge = st0->edge, fp16 = st0->fp16 != 0;
    const int pf = st0->pf >= 2 ? 2 : 1; // forward core: the first context's tuned prefetch (1 or 2)
    const bool graph = lead->graphs && !lead->profiling;
    if (!st0->grp) st0->grp = new GroupCache();
    std::vector<unsigned char> key(n * sizeof(GCtx) + 3 * sizeof(int));
    std::memcpy(key.data(), hc.data(), n * sizeof(GCtx));
    const int tail[3] = {T, pf, graph ? 1 : 0};
    std::memcpy(key.data() + n * sizeof(GCtx), tail, sizeof tail);
    GroupGraphs &G = st0->grp->get(key);
    if (G.key != key) {
        G.drop();
        G.tab.alloc(n * sizeof(GCtx));
        copy_h2d(s, G.tab.ptr, hc.data(), n * sizeof(GCtx));
        G.key = key;
    }
    const GCtx *tab = G.tab.as<GCtx>();
    // launch over the contexts active at group step tau, blocks(i, t) blocks for context i at its step t
    auto mk = [&](int tau, auto blocks, int &nb) {
        GLaunch Lg;
        std::memset(&Lg, 0, sizeof Lg);
        Lg.tau = tau;
        nb = 0;
        for (int i = 0; i < n; ++i) {
            const int t = tau - hc[i].start;
            if (t < 0) continue;
            const int k = blocks(i, t);
            if (k <= 0) continue;
            Lg.ctx[Lg.n] = i;
            Lg.first[Lg.n] = nb;
            nb += k;
            ++Lg.n;
        }
        Lg.first[Lg.n] = nb;
        return Lg;
    };
    // one wavefield step (wave_step_probed): rcv 0 the source's forward run, 1 the receiver run
    auto wave_step = [&](int tau, auto rcvc, hipStream_t q) {
        constexpr bool RCV = decltype(rcvc)::value;
        constexpr int rcv = RCV ? 1 : 0;
        int nb;
        GLaunch Lg;
        if (learned) {
            Lg = mk(tau, [&](int i, int) { return hc[i].ntp * hc[i].B; }, nb);
            if (nb) k_lpad_stress_grp<RCV><<<nb, dim3(64, 4), 0, q>>>(tab, Lg);
        }
        if (edge) {
            Lg = mk(tau, [&](int i, int) { return hc[i].eb[rcv] * hc[i].B; }, nb);
            if (nb) OUC_GDISPATCH(N, k_edge_grp<NN, RCV><<<nb, kEdgeW, 0, q>>>(tab, Lg))
        } else {
            Lg = mk(tau, [&](int i, int) { return hc[i].nts * hc[i].B; }, nb);
            if (nb) OUC_GDISPATCH(N, k_band_stress_grp<NN, RCV><<<nb, dim3(64, 4), 0, q>>>(tab, Lg))
        }
        Lg = mk(tau, [&](int i, int) { return hc[i].cbx * hc[i].cby * hc[i].B; }, nb);
        if (nb) {
            if (RCV) OUC_GDISPATCH(N, k_core_grp<NN, CORE_IMAGE, 1><<<nb, 256, 0, q>>>(tab, Lg))
            else if (pf == 2) OUC_GDISPATCH(N, k_core_grp<NN, CORE_STEP, 2><<<nb, 256, 0, q>>>(tab, Lg))
            else OUC_GDISPATCH(N, k_core_grp<NN, CORE_STEP, 1><<<nb, 256, 0, q>>>(tab, Lg))
        }
        if (!edge) {
            Lg = mk(tau, [&](int i, int) { return hc[i].ntv * hc[i].B; }, nb);
            if (nb) OUC_GDISPATCH(N, k_band_velocity_grp<NN, RCV><<<nb, dim3(64, 4), 0, q>>>(tab, Lg))
        }
        if (learned) {
            Lg = mk(tau, [&](int i, int) { return hc[i].lbw ? (hc[i].lbE + 255) / 256 * 4 * kLbD * hc[i].B : 0; }, nb);
            if (nb) k_lb_canon_grp<RCV><<<nb, 256, 0, q>>>(tab, Lg);
            Lg = mk(tau, [&](int i, int) { return hc[i].lbrb * hc[i].B; }, nb);
            if (nb) k_lb_ring_grp<RCV><<<nb, 256, 0, q>>>(tab, Lg);
            Lg = mk(tau, [&](int i, int) { return hc[i].lbw ? hc[i].lbrb * hc[i].B : 0; }, nb);
            if (nb) k_lb_apply_grp<RCV><<<nb, 256, 0, q>>>(tab, Lg);
            Lg = mk(tau, [&](int i, int) { return hc[i].ntp * hc[i].B; }, nb);
            if (nb) k_lpad_velocity_grp<RCV><<<nb, dim3(64, 4), 0, q>>>(tab, Lg);
            Lg = mk(tau, [&](int i, int t) { return t % kMonStride == 0 ? hc[i].ntm * hc[i].B : 0; }, nb);
            if (nb) k_lb_monitor_grp<RCV><<<nb, dim3(64, 4), 0, q>>>(tab, Lg);
        }
    };
    // body(i0, i1) runs iterations [i0, i1) of the T-iteration loop; captured
    // in chunks of kGraphSteps iterations like run_loop's
    auto run = [&](std::vector<hipGraphExec_t> &execs, auto &&body) {
        if (!graph) {
            body(0, T);
            return;
        }
        if (execs.empty()) {
            try {
                for (int i0 = 0; i0 < T; i0 += kGraphSteps) {
                    const int i1 = std::min(T, i0 + kGraphSteps);
                    hipGraph_t gr = nullptr;
                    OUC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    try {
                        body(i0, i1);
                    } catch (...) {
                        (void)hipStreamEndCapture(s, &gr);
                        if (gr) (void)hipGraphDestroy(gr);
                        throw;
                    }
                    OUC_HIP(hipStreamEndCapture(s, &gr));
                    hipGraphExec_t ex = nullptr;
                    hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(gr);
                    OUC_HIP(e);
                    execs.push_back(ex);
                }
            } catch (...) {
                for (hipGraphExec_t e : execs) (void)hipGraphExecDestroy(e);
                execs.clear();
                throw;
            }
        }
        for (hipGraphExec_t ex : execs) OUC_HIP(hipGraphLaunch(ex, s));
    };
    OUC_HIP(hipEventRecord(lead->ev[4], s));
    run(G.fwd, [&](int i0, int i1) {
        for (int tau = i0; tau < i1; ++tau) {
            wave_step(tau, std::false_type{}, s);
            int nb;
            const GLaunch Lg = mk(tau, [&](int i, int) { return hc[i].rblocks * hc[i].B; }, nb);
            if (nb) k_fstrip_save_grp<<<nb, 256, 0, s>>>(tab, Lg);
        }
        OUC_HIP(hipGetLastError());
    });
    OUC_HIP(hipEventRecord(lead->ev[1], s));
    OUC_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) {
        batch_handoff(cs[i], bi[i]);
        OUC_HIP(hipStreamSynchronize(cs[i]->stream));
    }
    // backward: the reconstruction on a second stream beside the receiver step
    // (captured loops only), joined every step, as in run_batch
    if (graph) st0->ensure_side();
    hipStream_t sr = graph ? st0->s2 : s;
    OUC_HIP(hipEventRecord(lead->ev[6], s));
    run(G.bwd, [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            const int tau = T - 1 - i; // backward: tau = T-1 .. 0
            if (graph) {
                OUC_HIP(hipEventRecord(st0->evf, s));
                OUC_HIP(hipStreamWaitEvent(sr, st0->evf, 0));
            }
            int nb;
            GLaunch Lg;
            if (fp16) {
                Lg = mk(tau, [&](int i, int) { return hc[i].rblocks * RS_NF * hc[i].B; }, nb);
                if (nb) k_strip_decode_grp<<<nb, 256, 0, sr>>>(tab, Lg);
            }
            Lg = mk(tau, [&](int i, int t) { return t < hc[i].nt - 1 ? hc[i].ntrs * hc[i].B : 0; }, nb);
            if (nb) OUC_GDISPATCH(N, k_recon_stress_grp<NN><<<nb, dim3(64, 4), 0, sr>>>(tab, Lg))
            Lg = mk(tau, [&](int i, int t) { return t < hc[i].nt - 1 ? hc[i].rbx * hc[i].rby * hc[i].B : 0; }, nb);
            if (nb) OUC_GDISPATCH(N, k_core_grp<NN, CORE_RECON, 2><<<nb, 256, 0, sr>>>(tab, Lg))
            Lg = mk(tau, [&](int i, int t) { return t < hc[i].nt - 1 ? hc[i].ntrv * hc[i].B : 0; }, nb);
            if (nb) OUC_GDISPATCH(N, k_recon_velocity_grp<NN><<<nb, dim3(64, 4), 0, sr>>>(tab, Lg))
            wave_step(tau, std::true_type{}, s);
            if (graph) {
                OUC_HIP(hipEventRecord(st0->evj, sr));
                OUC_HIP(hipStreamWaitEvent(s, st0->evj, 0));
            }
        }
        OUC_HIP(hipGetLastError());
    });
    OUC_HIP(hipEventRecord(lead->ev[2], s));
    OUC_HIP(hipStreamSynchronize(s));
}

// Capacities of the contexts of a grouped run: each context's shots, or its
// set batch capacity, scaled down together when all do not fit in 85 % of the
// device memory free or held by them.
std::vector<int> group_caps(const std::vector<ouc_ctx *> &cs, const std::vector<int> &nshots) {
    size_t freeb = 0, total = 0;
    OUC_HIP(hipMemGetInfo(&freeb, &total));
    size_t held = 0;
    for (ouc_ctx *c : cs) {
        FusedState *st = fs(c);
        held += st->mem_s.bytes + st->mem_r.bytes + st->strips32.bytes + st->strips16.bytes + st->in32.bytes +
                st->e8.bytes + st->keyb.bytes + st->ef.bytes + st->decb.bytes + st->gacc.bytes + st->obs.bytes;
    }
    const double avail = (double)(freeb + held) * 0.85;
    std::vector<int> want(cs.size());
    double need = 0;
    for (size_t i = 0; i < cs.size(); ++i) {
        want[i] = nshots[i];
        if (cs[i]->batch_cap > 0) want[i] = std::min(want[i], cs[i]->batch_cap);
        need += (double)want[i] * per_shot_bytes(cs[i], cs[i]->p.strip_precision == OUC_STRIP_FP16);
    }
    if (need > avail)
        for (auto &w : want) w = std::max(1, (int)(w * (avail / need)));
    return want;
}

} // namespace

void fused_rtm_group(ouc_ctx *const *ctxs, int n, const ouc_shot *const *shots, const int32_t *nshots,
                     float *const *stack_pp, float *const *stack_ps) {
    if (n < 1 || n > kGroupMax) throw Error(OUC_INVALID_ARG, "a group holds 1 to " + std::to_string(kGroupMax) + " contexts");
    std::vector<ouc_ctx *> cs(ctxs, ctxs + n);
    const ouc_ctx *a = cs[0];
    for (int i = 0; i < n; ++i) {
        const ouc_ctx *c = cs[i];
        if (!c) throw Error(OUC_INVALID_ARG, "context " + std::to_string(i) + " is NULL");
        for (int j = 0; j < i; ++j)
            if (cs[j] == c) throw Error(OUC_INVALID_ARG, "a context appears twice in the group");
        if (c->p.engine != OUC_ENGINE_FUSED) throw Error(OUC_UNSUPPORTED, "grouped runs need the fused engine");
        if (c->device != a->device || c->g.N != a->g.N || c->g.learned != a->g.learned ||
            c->p.strip_precision != a->p.strip_precision || c->p.semantics != a->p.semantics)
            throw Error(OUC_INVALID_ARG,
                        "contexts of a group must share the device, half-order, boundary kind, strip precision and semantics");
        if (nshots[i] < 0 || (nshots[i] > 0 && !shots[i])) throw Error(OUC_INVALID_ARG, "bad shot list");
        const Grid &g = c->g;
        for (int k = 0; k < nshots[i]; ++k) {
            const ouc_shot *sh = &shots[i][k];
            if (sh->sx < 0 || sh->sx >= g.nx || sh->sz < 0 || sh->sz >= g.nz || sh->rec_z < 0 || sh->rec_z >= g.nz)
                throw Error(OUC_INVALID_ARG, "shot " + std::to_string(k) + " of context " + std::to_string(i) +
                                                 " outside the interior");
            if (!sh->wavelet && !(sh->fm > 0)) throw Error(OUC_INVALID_ARG, "fm must be positive");
            if (sh->nt > g.nt) throw Error(OUC_INVALID_ARG, "shot nt beyond its context's");
        }
    }
    if (a->g.N != 5 && a->g.N != 6) throw Error(OUC_UNSUPPORTED, "grouped runs support half-orders 5 and 6");
    const std::vector<int> cap = group_caps(cs, std::vector<int>(nshots, nshots + n));
    for (int i = 0; i < n; ++i) ensure(cs[i], std::max(1, cap[i]));
    std::vector<DevBuf> stack(n);
    std::vector<std::vector<int>> diverged(n);
    for (int i = 0; i < n; ++i) {
        const size_t ni = (size_t)cs[i]->g.nx * cs[i]->g.nz;
        stack[i].alloc(2 * ni * sizeof(float));
        OUC_HIP(hipMemsetAsync(stack[i].ptr, 0, 2 * ni * sizeof(float), cs[i]->stream));
        cs[i]->last = ouc_timings{};
    }
    std::vector<int> off(n, 0);
    double fwd_ms = 0, bwd_ms = 0, tot_ms = 0;
    for (;;) {
        std::vector<ouc_ctx *> rc;
        std::vector<const ouc_shot *> rs;
        std::vector<int> rb, ri;
        for (int i = 0; i < n; ++i)
            if (off[i] < nshots[i]) {
                rc.push_back(cs[i]);
                rs.push_back(shots[i] + off[i]);
                rb.push_back(std::min(cap[i], nshots[i] - off[i]));
                ri.push_back(i);
            }
        if (rc.empty()) break;
        run_group_batch(rc, rs, rb);
        float f = 0, b = 0, w = 0;
        OUC_HIP(hipEventElapsedTime(&f, rc[0]->ev[4], rc[0]->ev[1]));
        OUC_HIP(hipEventElapsedTime(&b, rc[0]->ev[6], rc[0]->ev[2]));
        OUC_HIP(hipEventElapsedTime(&w, rc[0]->ev[4], rc[0]->ev[2]));
        fwd_ms += f;
        bwd_ms += b;
        tot_ms += w;
        for (size_t k = 0; k < rc.size(); ++k) {
            ouc_ctx *c = rc[k];
            const int i = ri[k];
            FusedState *st = fs(c);
            const size_t ni = (size_t)c->g.nx * c->g.nz;
            std::vector<char> bad = monitor_verdict(c, rb[k]);
            for (int b2 = 0; b2 < rb[k]; ++b2) {
                if (bad[b2]) {
                    diverged[i].push_back(off[i] + b2);
                    continue;
                }
                shot_images(c, st, b2);
                k_accumulate<<<(unsigned)((ni + 255) / 256), 256, 0, c->stream>>>(ni, c->img[2].f(), stack[i].f());
                k_accumulate<<<(unsigned)((ni + 255) / 256), 256, 0, c->stream>>>(ni, c->img[3].f(),
                                                                                  stack[i].f() + ni);
            }
            OUC_HIP(hipStreamSynchronize(c->stream));
            off[i] += rb[k];
        }
    }
    int nfb = 0;
    for (int i = 0; i < n; ++i) {
        ouc_ctx *c = cs[i];
        const Grid &g = c->g;
        const size_t ni = (size_t)g.nx * g.nz;
        if (stack_pp && stack_pp[i]) copy_d2h(c->stream, stack_pp[i], stack[i].ptr, ni * sizeof(float));
        if (stack_ps && stack_ps[i]) copy_d2h(c->stream, stack_ps[i], stack[i].f() + ni, ni * sizeof(float));
        double fb_ms = 0;
        if (!diverged[i].empty()) {
            fb_ms = rerun_with_cpml(c, shots[i], diverged[i], stack_pp ? stack_pp[i] : nullptr,
                                    stack_ps ? stack_ps[i] : nullptr, nullptr, tot_ms);
            nfb += (int)diverged[i].size();
            c->last.fallback_shots = (int32_t)diverged[i].size();
        }
        c->last.total_ms = tot_ms + fb_ms;
        c->last.forward_ms = fwd_ms;
        c->last.backward_ms = bwd_ms;
        const double padded = (double)g.Xn * g.Zn, interior = (double)g.nx * g.nz;
        c->last.cell_updates = (double)nshots[i] * g.nt * (2 * padded + interior);
    }
    if (nfb)
        throw Error(OUC_NUMERIC_FALLBACK, std::to_string(nfb) + " shot(s) ran again with CPML after the learned boundary diverged");
}

// Single shot through the fused path (a batch of one), returning its images.
void fused_rtm_one(ouc_ctx *c, const ouc_shot *s, const float *rec_in, float *img_pp, float *img_ps, float *lap_pp,
                   float *lap_ps) {
    ensure(c, std::max(1, fs(c)->cap));
    FusedState *st = fs(c);
    CoreProbe probe;
    probe.init(c->g.nt, c->profiling);
    c->last = ouc_timings{};
    OUC_HIP(hipEventRecord(c->ev[0], c->stream));
    run_batch(c, s, 1, true, 2, probe, nullptr, rec_in);
    OUC_HIP(hipEventRecord(c->ev[2], c->stream));
    if (monitor_verdict(c, 1)[0]) {
        ouc_ctx *fb = fallback_ctx(c);
        ouc_status r = ouc_rtm_shot(fb, s, rec_in, img_pp, img_ps, lap_pp, lap_ps);
        if (r != OUC_OK) throw Error(r, std::string("CPML fallback: ") + ouc_last_error(fb));
        c->last = fb->last;
        c->last.fallback_shots = 1;
        throw Error(OUC_NUMERIC_FALLBACK, "the learned boundary diverged; the shot ran again with CPML");
    }
    shot_images(c, st, 0);
    OUC_HIP(hipEventRecord(c->ev[3], c->stream));
    const size_t ni = (size_t)c->g.nx * c->g.nz;
    float *outs[4] = {img_pp, img_ps, lap_pp, lap_ps};
    for (int a = 0; a < 4; ++a)
        if (outs[a]) OUC_HIP(hipMemcpyAsync(outs[a], c->img[a].ptr, ni * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    OUC_HIP(hipStreamSynchronize(c->stream));
    float fw = 0, bw = 0, im = 0, tot = 0;
    OUC_HIP(hipEventElapsedTime(&fw, c->ev[0], c->ev[1]));
    OUC_HIP(hipEventElapsedTime(&bw, c->ev[1], c->ev[2]));
    OUC_HIP(hipEventElapsedTime(&im, c->ev[2], c->ev[3]));
    OUC_HIP(hipEventElapsedTime(&tot, c->ev[0], c->ev[3]));
    c->last.forward_ms = fw; c->last.backward_ms = bw; c->last.image_ms = im; c->last.total_ms = tot;
    c->last.step_launches = 2 * (int64_t)c->g.nt;
    c->last.step_kernel_ms = probe.total_ms();
    c->last.step_kernel_launches = probe.used;
    const Grid &g = c->g;
    const double padded = (double)g.Xn * g.Zn, interior = (double)g.nx * g.nz, cpml = padded - interior;
    c->last.cell_updates = (double)g.nt * (2 * padded + interior);
    c->last.canonical_bytes = (double)g.nt * (2 * (120 * interior + 232 * cpml) + 120 * interior + 48 * interior);
    c->last.own_bytes = fused_own_bytes(c);
}

// FWI gradient summed over shots (gradfun with the adsource residual); outputs
// are x-major interior arrays, misfit = 0.5 * sum of squared residuals.
void fused_fwi_gradient(ouc_ctx *c, const ouc_shot *shots, int nshots, const float *observed, float *grad_vp,
                        float *grad_vs, float *illum, double *misfit) {
    const Grid &g = c->g;
    if (g.learned) throw Error(OUC_UNSUPPORTED, "the FWI gradient runs with the CPML boundary");
    for (int k = 0; k < nshots; ++k) {
        const ouc_shot *sh = &shots[k];
        if (sh->sx < 0 || sh->sx >= g.nx || sh->sz < 0 || sh->sz >= g.nz || sh->rec_z < 0 || sh->rec_z >= g.nz)
            throw Error(OUC_INVALID_ARG, "shot " + std::to_string(k) + " outside the interior");
        if (!sh->wavelet && !(sh->fm > 0)) throw Error(OUC_INVALID_ARG, "fm must be positive");
        if (sh->nt > 0 && sh->nt != g.nt) throw Error(OUC_UNSUPPORTED, "the FWI gradient needs every shot at the context's nt");
    }
    // capacity: the per-shot state plus the gradient accumulators and observed records
    const size_t extra = (size_t)3 * g.pitched() * sizeof(float) + (size_t)2 * g.nt * g.nx * sizeof(float);
    int cap = default_capacity(c, nshots, extra);
    ensure(c, cap);
    FusedState *st = fs(c);
    cap = std::min(cap, st->cap);
    const size_t gbytes = (size_t)3 * st->plane * st->cap * sizeof(float);
    const size_t obytes = (size_t)st->cap * 2 * g.nt * g.nx * sizeof(float);
    const int rblocks = 64;
    if (st->gacc.bytes < gbytes || st->obs.bytes < obytes) {
        st->drop_graphs(); // captured loops hold the old addresses
        if (st->gacc.bytes < gbytes) st->gacc.alloc(gbytes);
        if (st->obs.bytes < obytes) st->obs.alloc(obytes);
    }
    if (st->part.bytes < (size_t)st->cap * rblocks * sizeof(double)) st->part.alloc((size_t)st->cap * rblocks * sizeof(double));
    const size_t ni = (size_t)g.nx * g.nz;
    if (st->gout.bytes < 3 * ni * sizeof(float)) st->gout.alloc(3 * ni * sizeof(float));
    OUC_HIP(hipMemsetAsync(st->gout.ptr, 0, 3 * ni * sizeof(float), c->stream));
    c->last = ouc_timings{};
    CoreProbe probe;
    OUC_HIP(hipEventRecord(c->ev[0], c->stream));
    double m2 = 0;
    std::vector<double> part;
    const size_t gspan = (size_t)st->plane * st->cap;
    for (int b0 = 0; b0 < nshots; b0 += cap) {
        const int B = std::min(cap, nshots - b0);
        FwiBatch fb{observed + (size_t)b0 * 2 * g.nx * g.nt, rblocks};
        run_batch(c, shots + b0, B, true, 2, probe, nullptr, nullptr, nullptr, &fb);
        const dim3 grid((g.nz + 255) / 256, g.nx);
        for (int q = 0; q < 3; ++q)
            k_shot_sum<<<grid, 256, 0, c->stream>>>(st->g, B, st->gacc.f() + q * gspan, st->gout.f() + q * ni);
        OUC_HIP(hipGetLastError());
        part.resize((size_t)B * rblocks);
        OUC_HIP(hipMemcpyAsync(part.data(), st->part.ptr, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        OUC_HIP(hipStreamSynchronize(c->stream));
        for (double v : part) m2 += v;
    }
    float *outs[3] = {grad_vp, grad_vs, illum};
    for (int q = 0; q < 3; ++q)
        if (outs[q])
            OUC_HIP(hipMemcpyAsync(outs[q], st->gout.f() + q * ni, ni * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    OUC_HIP(hipEventRecord(c->ev[5], c->stream));
    OUC_HIP(hipStreamSynchronize(c->stream));
    float tot = 0;
    OUC_HIP(hipEventElapsedTime(&tot, c->ev[0], c->ev[5]));
    c->last.total_ms = tot;
    if (misfit) *misfit = 0.5 * m2;
}

void fused_model_update(int device, int64_t n, const float *v, const float *grad, double step, int iter, float *out) {
    OUC_HIP(hipSetDevice(device));
    DevBuf dv, dg, dout;
    dv.alloc((size_t)n * sizeof(float));
    dg.alloc((size_t)n * sizeof(float));
    dout.alloc((size_t)n * sizeof(float));
    OUC_HIP(hipMemcpy(dv.ptr, v, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    OUC_HIP(hipMemcpy(dg.ptr, grad, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    const double scale = step * std::pow(0.95, iter); // a_2 * pow(0.95, k), P-S-seperating.cu:1511
    k_model_update<<<(unsigned)((n + 255) / 256), 256>>>((long)n, dv.f(), dg.f(), scale, dout.f());
    OUC_HIP(hipGetLastError());
    OUC_HIP(hipMemcpy(out, dout.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
}

// Forward seismograms [6][nx][nt] of one shot through the fused path.
void fused_forward_one(ouc_ctx *c, const ouc_shot *s, float *seis) {
    ensure(c, std::max(1, fs(c)->cap));
    FusedState *st = fs(c);
    const Grid &g = c->g;
    const size_t n6 = (size_t)6 * g.nt * g.nx;
    if (st->rec6.bytes < n6 * sizeof(float)) st->rec6.alloc(n6 * sizeof(float));
    OUC_HIP(hipMemsetAsync(st->rec6.ptr, 0, n6 * sizeof(float), c->stream));
    CoreProbe probe;
    probe.init(g.nt, c->profiling);
    c->last = ouc_timings{};
    OUC_HIP(hipEventRecord(c->ev[0], c->stream));
    run_batch(c, s, 1, false, 6, probe);
    OUC_HIP(hipEventRecord(c->ev[1], c->stream));
    std::vector<float> tmp(n6);
    OUC_HIP(hipMemcpyAsync(tmp.data(), st->rec6.ptr, n6 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    OUC_HIP(hipStreamSynchronize(c->stream));
    if (monitor_verdict(c, 1)[0]) {
        ouc_ctx *fb = fallback_ctx(c);
        ouc_status r = ouc_forward_shot(fb, s, seis);
        if (r != OUC_OK) throw Error(r, std::string("CPML fallback: ") + ouc_last_error(fb));
        c->last = fb->last;
        c->last.fallback_shots = 1;
        throw Error(OUC_NUMERIC_FALLBACK, "the learned boundary diverged; the shot ran again with CPML");
    }
    float fw = 0;
    OUC_HIP(hipEventElapsedTime(&fw, c->ev[0], c->ev[1]));
    c->last.forward_ms = fw;
    c->last.total_ms = fw;
    c->last.step_launches = g.nt;
    c->last.step_kernel_ms = probe.total_ms();
    c->last.step_kernel_launches = probe.used;
    c->last.cell_updates = (double)g.nt * g.Xn * g.Zn;
    if (s->mute) {
        std::vector<int32_t> cnt;
        mute_counts(g, *s, c->vp_pad[(size_t)(s->sx + g.L) * g.P + s->rec_z + g.L], cnt);
        for (int q = 0; q < 6; ++q)
            for (int ix = 0; ix < g.nx; ++ix)
                for (int t = 0; t < std::min(cnt[ix], g.nt); ++t) tmp[((size_t)q * g.nt + t) * g.nx + ix] = 0.f;
    }
    if (seis)
        for (int q = 0; q < 6; ++q)
            for (int t = 0; t < g.nt; ++t)
                for (int ix = 0; ix < g.nx; ++ix)
                    seis[((size_t)q * g.nx + ix) * g.nt + t] = tmp[((size_t)q * g.nt + t) * g.nx + ix];
}

// Forward run of one shot with snapshots of V over a padded-grid rectangle
// after each listed step: out[n][2][w][h].
void fused_snapshots(ouc_ctx *c, const ouc_shot *s, int n, const int32_t *steps, int x0, int z0, int w, int h,
                     float *out) {
    const Grid &g = c->g;
    if (n < 1 || !steps || !out || w < 1 || h < 1 || x0 < 0 || z0 < 0 || x0 + w > g.Xn || z0 + h > g.Zn)
        throw Error(OUC_INVALID_ARG, "snapshot rectangle or step list invalid");
    for (int k = 0; k < n; ++k)
        if (steps[k] < 0 || steps[k] >= g.nt || (k > 0 && steps[k] <= steps[k - 1]))
            throw Error(OUC_INVALID_ARG, "snapshot steps must increase within [0, nt)");
    ensure(c, std::max(1, fs(c)->cap));
    DevBuf buf;
    buf.alloc((size_t)n * 2 * w * h * sizeof(float));
    std::vector<int> st(steps, steps + n);
    Snap sn;
    sn.steps = st.data();
    sn.n = n;
    sn.x0 = x0; sn.z0 = z0; sn.w = w; sn.h = h;
    sn.dev = buf.f();
    CoreProbe probe;
    c->last = ouc_timings{};
    run_batch(c, s, 1, false, 2, probe, &sn);
    OUC_HIP(hipMemcpyAsync(out, buf.ptr, buf.bytes, hipMemcpyDeviceToHost, c->stream));
    OUC_HIP(hipStreamSynchronize(c->stream));
}

namespace {
float bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) { // inf or NaN: truncate, keep NaN quiet
        u &= 0xffff0000u;
        if (x != x) u |= 0x00400000u;
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}
} // namespace

// Network weights: W1 [16][5][3][3], W2 [16][16][3][3], W3 [2][16][3][3]
// (FP32 masters, rounded to BF16).  The layers are linear, so they compose
// exactly into one 7x7 filter [2][5][7][7]; the composition runs in FP64 over
// the BF16 values and is rounded once to FP32.  weights == NULL: no network.
void fused_set_learned(ouc_ctx *c, const float *weights, int n) {
    FusedState *st = fs(c);
    st->drop_graphs(); // the ring kernel's filter argument changes
    if (!weights) {
        st->lb_have = false;
        return;
    }
    const int n1 = 16 * 5 * 9, n2 = 16 * 16 * 9, n3 = 2 * 16 * 9;
    if (n != n1 + n2 + n3) throw Error(OUC_INVALID_ARG, "expected " + std::to_string(n1 + n2 + n3) + " weights");
    std::vector<double> w(n);
    for (int a = 0; a < n; ++a) {
        if (!std::isfinite(weights[a])) throw Error(OUC_INVALID_ARG, "weights must be finite");
        w[a] = bf16_round(weights[a]);
    }
    const double *W1 = w.data(), *W2 = W1 + n1, *W3 = W2 + n2;
    // A[h2][ci][5][5] = sum_h1 W2[h2][h1] (*) W1[h1][ci]
    std::vector<double> A((size_t)16 * 5 * 25, 0.0), K((size_t)2 * 5 * 49, 0.0);
    for (int h2 = 0; h2 < 16; ++h2)
        for (int h1 = 0; h1 < 16; ++h1)
            for (int ci = 0; ci < 5; ++ci)
                for (int y2 = 0; y2 < 3; ++y2)
                    for (int x2 = 0; x2 < 3; ++x2)
                        for (int y1 = 0; y1 < 3; ++y1)
                            for (int x1 = 0; x1 < 3; ++x1)
                                A[((size_t)h2 * 5 + ci) * 25 + (y2 + y1) * 5 + (x2 + x1)] +=
                                    W2[(h2 * 16 + h1) * 9 + y2 * 3 + x2] * W1[(h1 * 5 + ci) * 9 + y1 * 3 + x1];
    for (int co = 0; co < 2; ++co)
        for (int h2 = 0; h2 < 16; ++h2)
            for (int ci = 0; ci < 5; ++ci)
                for (int y3 = 0; y3 < 3; ++y3)
                    for (int x3 = 0; x3 < 3; ++x3)
                        for (int y = 0; y < 5; ++y)
                            for (int x = 0; x < 5; ++x)
                                K[((size_t)co * 5 + ci) * 49 + (y3 + y) * 7 + (x3 + x)] +=
                                    W3[(co * 16 + h2) * 9 + y3 * 3 + x3] * A[((size_t)h2 * 5 + ci) * 25 + y * 5 + x];
    std::vector<float> kf(K.begin(), K.end());
    if (!st->lbw.ptr) st->lbw.alloc(kLbTaps * sizeof(float));
    copy_h2d(c->stream, st->lbw.ptr, kf.data(), kLbTaps * sizeof(float));
    st->lb_have = true;
}

void fused_get_learned_filter(ouc_ctx *c, float *taps) {
    FusedState *st = fs(c);
    if (!st->lb_have) throw Error(OUC_INVALID_ARG, "no network weights set");
    copy_d2h(c->stream, taps, st->lbw.ptr, kLbTaps * sizeof(float));
}

// RTM of one shot with snapshots of the reconstructed source V^t over a
// padded-grid rectangle at the listed times (decreasing, each in [0, nt-2]).
void fused_recon_snapshots(ouc_ctx *c, const ouc_shot *s, int n, const int32_t *steps, int x0, int z0, int w, int h,
                           float *out) {
    const Grid &g = c->g;
    if (n < 1 || !steps || !out || w < 1 || h < 1 || x0 < 0 || z0 < 0 || x0 + w > g.Xn || z0 + h > g.Zn)
        throw Error(OUC_INVALID_ARG, "snapshot rectangle or step list invalid");
    for (int k = 0; k < n; ++k)
        if (steps[k] < 0 || steps[k] > g.nt - 2 || (k > 0 && steps[k] >= steps[k - 1]))
            throw Error(OUC_INVALID_ARG, "reconstruction steps must decrease within [0, nt-2]");
    ensure(c, std::max(1, fs(c)->cap));
    DevBuf buf;
    buf.alloc((size_t)n * 2 * w * h * sizeof(float));
    std::vector<int> st(steps, steps + n);
    Snap sn;
    sn.steps = st.data();
    sn.n = n;
    sn.x0 = x0; sn.z0 = z0; sn.w = w; sn.h = h;
    sn.dev = buf.f();
    CoreProbe probe;
    c->last = ouc_timings{};
    run_batch(c, s, 1, true, 2, probe, nullptr, nullptr, &sn);
    OUC_HIP(hipMemcpyAsync(out, buf.ptr, buf.bytes, hipMemcpyDeviceToHost, c->stream));
    OUC_HIP(hipStreamSynchronize(c->stream));
}

void fused_drop_graphs(ouc_ctx *c) { fs(c)->drop_graphs(); }

// Core tile (columns per block, prefetch depth).
void fused_set_tile(ouc_ctx *c, int cols, int pf) {
    if (cols < 8 || cols > 1024 || pf < 1 || pf > 3) throw Error(OUC_INVALID_ARG, "tile out of range");
    FusedState *st = fs(c);
    st->csx = cols;
    st->pf = pf;
    st->g.csx = cols;
    st->drop_graphs();
}

void fused_get_tile(ouc_ctx *c, int *cols, int *pf) {
    *cols = fs(c)->csx;
    *pf = fs(c)->pf;
}

// Blocks of the forward core with prefetch depth pf one compute unit holds at once.
template <int N>
int core_blocks_per_cu(int pf) {
    int nb = 0;
    if constexpr (N == 5 || N == 6) {
        if (pf == 1) OUC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_core<N, CORE_STEP, 1>, 256, 0));
        else if (pf == 3) OUC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_core<N, CORE_STEP, 3>, 256, 0));
        else OUC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_core<N, CORE_STEP, kCorePF>, 256, 0));
    } else {
        OUC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_core<N, CORE_STEP, kCorePF>, 256, 0));
    }
    return nb;
}

// Core width at which one round of resident blocks covers the launch: the
// (z block, shot) strips times the tiles per strip fit in the blocks the
// device holds at once (several rounds only when the strips alone exceed it).
// A launch of slightly more blocks than one round waits a whole extra block
// time for the excess, so this width avoids the partly filled last round.
int core_fit_cols(const ouc_ctx *c, const FGrid &f, int nshots) {
    int cus = 0, per_cu = 0;
    OUC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    OUC_FDISPATCH(f.N, per_cu = core_blocks_per_cu<NN>(kCorePF));
    const long strips = (long)core_zblocks(f.cz0, f.cz1) * nshots;
    const long tiles = std::max(1L, (long)cus * std::max(1, per_cu) / std::max(1L, strips));
    const long w = f.cx1 - f.cx0;
    return (int)std::min(1024L, std::max(8L, (w + tiles - 1) / tiles));
}

// Times k_core (forward mode) over `reps` launches for each candidate tile on
// this grid with `nshots` shots; keeps the fastest.  ms[k] gets candidate k's
// mean launch time (candidates: cols {8, 16, 32, 64, 128, one-round width
// (core_fit_cols)} x pf {1, 2, 3}).
void fused_autotune(ouc_ctx *c, int nshots, int reps, int *best_cols, int *best_pf, float *ms) {
    if (nshots < 1 || reps < 1) throw Error(OUC_INVALID_ARG, "nshots and reps must be positive");
    ensure(c, nshots);
    FusedState *st = fs(c);
    hipStream_t s = c->stream;
    OUC_HIP(hipMemsetAsync(st->mem_s.ptr, 0, st->mem_s.bytes, s));
    const int cols[6] = {8, 16, 32, 64, 128, core_fit_cols(c, st->g, nshots)}, pfs[3] = {1, 2, 3};
    hipEvent_t e0, e1;
    OUC_HIP(hipEventCreate(&e0));
    OUC_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int bc = st->csx, bp = st->pf, k = 0;
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 3; ++b, ++k) {
            FGrid f = st->g;
            f.csx = cols[a];
            dim3 grid((f.cx1 - f.cx0 + f.csx - 1) / f.csx, core_zblocks(f.cz0, f.cz1), nshots);
            FStep st0{};
            if (f.cx1 <= f.cx0) { if (ms) ms[k] = 0; continue; }
            OUC_FDISPATCH(f.N, launch_core<NN, 0>(pfs[b], grid, s, f, st->S, c->pq.as<float2>(), st0, FRecon{})); // warm
            OUC_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) {
                st0.rb = r & 1;
                OUC_FDISPATCH(f.N, launch_core<NN, 0>(pfs[b], grid, s, f, st->S, c->pq.as<float2>(), st0, FRecon{}));
            }
            OUC_HIP(hipEventRecord(e1, s));
            OUC_HIP(hipEventSynchronize(e1));
            float t = 0;
            OUC_HIP(hipEventElapsedTime(&t, e0, e1));
            t /= reps;
            if (ms) ms[k] = t;
            if (t < best) { best = t; bc = cols[a]; bp = pfs[b]; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    fused_set_tile(c, bc, bp);
    *best_cols = bc;
    *best_pf = bp;
}

double fused_shot_bytes(const ouc_ctx *c) { return (double)per_shot_bytes(c, c->p.strip_precision == OUC_STRIP_FP16); }

void fused_set_learned_damping(ouc_ctx *c, float beta) {
    if (!(beta >= 0.f && beta < 1.f)) throw Error(OUC_INVALID_ARG, "damping must be in [0, 1)");
    FusedState *st = fs(c);
    st->lb_beta = beta;
    st->g.lb_beta = beta;
    st->drop_graphs();
}
