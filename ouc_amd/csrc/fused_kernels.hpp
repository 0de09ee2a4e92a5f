// Fused-engine HIP kernels.
//
// Arithmetic form: the state of a wavefield is the total velocity V = Vp + Vs
// and the S part Vs, each stepped as (value, increment) with
// V'' = (Tx - OTx, Tz - OTz) + (Wz - OWz, -Wx + OWx) + 2K and
// Vs'' = (Wz - OWz, -Wx + OWx) + K, where K = (C - E - OC + OE, B - A - OB + OA)
// of the reference (P-S-seperating.cu:291-294); Vp = V - Vs.  K's non-psi part
// is evaluated in the product-rule form D((vs^2 - vs^2_0) g), equal in exact
// arithmetic because the x and z staggered operators commute.
//
// Per step and wavefield there are three launches:
//   k_band_stress   stress on the CPML band (every cell that is not "core") plus
//                   an N-cell margin into the core, stored to global, with the
//                   CPML F update (forward_s, P-S-seperating.cu:316-384);
//   k_core          the core [L+N, Xn-L-N) x [L+N, Zn-L-N): stress and velocity
//                   in one pass, nothing but the state read and written;
//   k_band_velocity velocity with CPML O terms on the band (forward_u,
//                   P-S-seperating.cu:205-314).
// The core's stress halo lies in [L, Xn-L)^2, where the damping is zero, so the
// core kernel carries no CPML terms.  V is read through stencils by neighbouring
// blocks, so it is double-buffered (read rb, write 1-rb); every other array is
// read only at its own cell and is updated in place.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "kernels_plain.hpp"

namespace ouc {

// Calls f(std::integral_constant<int, r>) for r = 0 .. Q-1, unrolled.
template <class F, int... Rs>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, Rs...>) {
    (f(std::integral_constant<int, Rs>{}), ...);
}
template <int Q, class F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_seq(f, std::make_integer_sequence<int, Q>{});
}

// Kernel bodies take the block index as an argument: a plain launch passes its
// own, a grouped launch (several grids in one launch) the block's index within
// its context.
__device__ __forceinline__ uint3 blockIdx3() { return make_uint3(blockIdx.x, blockIdx.y, blockIdx.z); }

// Geometry shared by the fused kernels.
struct FGrid {
    int Xn, Zn, P, L, N, nx, nz, nt;
    long plane;     // floats per pitched array of one shot
    float idx, idz, dt, dt2;
    float c[8];
    int cx0, cx1, cz0, cz1; // core
    int csx;        // core tile: columns each k_core block streams
    int vlo;        // V is nonzero on [vlo, Xn-vlo) x [vlo, Zn-vlo): N with CPML, 0 with the learned pad
    int lpad;       // learned boundary: width of the low-order pad (0: CPML)
    float lb_beta;  // learned boundary: increment damping at the outer edge of the pad
};

// Distance of cell (i, j) from the nearest edge of the padded grid.
__device__ __forceinline__ int edge_dist(const FGrid &g, int i, int j) {
    return min(min(i, g.Xn - 1 - i), min(j, g.Zn - 1 - j));
}
__device__ __forceinline__ bool in_lpad(const FGrid &g, int i, int j) {
    return g.lpad > 0 && edge_dist(g, i, j) < g.lpad;
}

// (x, z) component pairs stored interleaved: element k of a plane is
// (x[k], z[k]) at floats 2k and 2k+1, so a lane moves both components of its
// cell with one 8-byte access.  Pair[f][k] is component f of element k.
struct Comp {
    float *b;
    __device__ __forceinline__ float &operator[](long k) const { return b[2 * k]; }
};
struct Pair {
    float2 *p = nullptr;
    __host__ __device__ __forceinline__ Comp operator[](int f) const { return Comp{reinterpret_cast<float *>(p) + f}; }
};

// One wavefield for all shots of a batch; arrays are [shot][plane].
struct FWave {
    Pair V[2];               // [buffer] (vx, vz)
    Pair DV, S, DS;          // increment of V, S part, its increment
    float *th, *om, *gxx, *gxz, *gzx, *gzz; // band stress (stored)
    float *F[4], *O[8];
    float *F2[4];            // second F buffer (the streaming CPML band kernel double-buffers F)
};

// Per-step, per-batch arguments.
struct FStep {
    int t;                   // time index of this step
    int rb;                  // V buffer read this step (written: 1 - rb)
    const int *src;          // [shot][2] padded source cell, or NULL
    const float *wav;        // [shot][nt] wavelet, or NULL
    int zr;                  // padded receiver row
    float *rec;              // FWD: [shot][ncomp][nt][nx] records, or NULL
    int ncomp;
    const float *inj;        // RCV: [shot][2][nt][nx] records to inject, or NULL
    const float2 *sp;        // RCV: source P velocity level (x, z) [shot][plane], or NULL
    float *npp, *nps, *den;  // RCV: imaging accumulators [shot][plane]
    const int *t0;           // RCV: [shot] first step of each shot (later starts image only their own steps)
};

__device__ __forceinline__ bool in_update(const FGrid &g, int i, int j) {
    return i >= g.N && i < g.Xn - g.N && j >= g.N && j < g.Zn - g.N;
}

// V as seen by the stress pass at (i, j): the state, zero outside the update
// region, the injected record on the receiver row (load_record,
// P-S-seperating.cu:1471-1483).
__device__ __forceinline__ void load_vin(const FGrid &g, const FWave &w, const FStep &a, int shot, long so, int i, int j,
                                         float &vx, float &vz) {
    vx = 0.f;
    vz = 0.f;
    if (i < g.vlo || i >= g.Xn - g.vlo || j < g.vlo || j >= g.Zn - g.vlo) return;
    long k = so + (long)i * g.P + j;
    const float2 v = w.V[a.rb].p[k];
    vx = v.x;
    vz = v.y;
    if (a.inj && j == a.zr && i >= g.L && i < g.Xn - g.L) {
        long o = ((long)shot * 2 * g.nt + a.t) * g.nx + (i - g.L);
        vx = a.inj[o];
        vz = a.inj[o + (long)g.nt * g.nx];
    }
}

__device__ __forceinline__ float src_term(const FGrid &g, const FStep &a, int shot, int i, int j) {
    if (!a.src) return 0.f;
    const int s = shot;
    return (i == a.src[2 * s] && j == a.src[2 * s + 1]) ? a.wav[(long)s * g.nt + a.t] : 0.f;
}

// Record totals and P/S parts after the update (P-S-seperating.cu:307-311).
__device__ __forceinline__ void record(const FGrid &g, const FStep &a, int shot, int i, int j, float vx, float vz, float sx,
                                       float sz) {
    if (!a.rec || j != a.zr || i < g.L || i >= g.Xn - g.L) return;
    const long plane = (long)g.nt * g.nx;
    long o = ((long)shot * a.ncomp) * plane + (long)a.t * g.nx + (i - g.L);
    a.rec[o] = vx;
    a.rec[o + plane] = vz;
    if (a.ncomp == 6) {
        a.rec[o + 2 * plane] = vx - sx;
        a.rec[o + 3 * plane] = vz - sz;
        a.rec[o + 4 * plane] = sx;
        a.rec[o + 5 * plane] = sz;
    }
}

// corr_v (P-S-seperating.cu:1426-1439) for an interior cell.
__device__ __forceinline__ void correlate(const FGrid &g, const FStep &a, int shot, long k, int i, int j, float vx, float vz,
                                          float sx, float sz) {
    if (!a.sp || i < g.L || i >= g.Xn - g.L || j < g.L || j >= g.Zn - g.L) return;
    if (a.t0 && a.t < a.t0[shot]) return;
    const float2 spk = a.sp[k];
    float sa = spk.x, sb = spk.y;
    float rpx = vx - sx, rpz = vz - sz;
    a.npp[k] += sa * rpx + sb * rpz;
    a.den[k] += sa * sa + sb * sb;
    a.nps[k] += sa * sx + sb * sz;
}

// ---------------------------------------------------------------------------
// Core kernel.  Block = 256 lanes along z covering TZ = 224 output rows
// plus a 2N halo each side; the block streams along x over SX columns.  At
// stream position c it loads V at column c, forms the stress at column c - N
// (x derivatives from a register queue of V, z derivatives through LDS), and
// updates column c - 2N (x derivatives of the stress from register queues, z
// derivatives computed at stress time and delayed N columns).
// ---------------------------------------------------------------------------
struct FRecon {
    KRing ring;
    KStrips strips;      // base for shot 0; per-shot offset below
    KStripStride stride; // between shots' strip stores
    int slot_v;          // slot of the ring totals (< 0: the state's own, no override)
    int slot_th;         // slot of the ring stress-like fields
    int exact;           // add the source term (EXACT semantics, or the hand-off step)
    float2 *vp;          // [shot][plane] output P level (x, z), or NULL (no imaging condition)
};

__device__ __forceinline__ KStrips shot_strips(const FRecon &rc, int shot) { return strips_of_shot(rc.strips, rc.stride, shot); }

// Totals seen by the reconstruction's stress pass at (i, j).
__device__ __forceinline__ void recon_vin(const FGrid &g, const FWave &w, const FRecon &rc, const KStrips &s, int rb,
                                          long so, int i, int j, float &vx, float &vz) {
    if (i < g.vlo || i >= g.Xn - g.vlo || j < g.vlo || j >= g.Zn - g.vlo) { vx = vz = 0.f; return; }
    long r = rc.slot_v >= 0 ? ring_index(rc.ring, i, j) : -1;
    if (r >= 0) {
        vx = strip_get(s, rc.slot_v, RS_VX, r);
        vz = strip_get(s, rc.slot_v, RS_VZ, r);
        return;
    }
    long k = so + (long)i * g.P + j;
    const float2 v = w.V[rb].p[k];
    vx = v.x;
    vz = v.y;
}

// Output rows of a core block: 224 = 7 lines of 32 floats (128 B), with block
// origins on line boundaries, so every column's state reads and writes cover
// whole lines; the 2N halo rows each side fit in the 256 lanes for N <= 8.
constexpr int kCoreTZ = 224;
__host__ __device__ constexpr int core_zalign(int cz0) { return cz0 & ~31; }
__host__ __device__ constexpr int core_zblocks(int cz0, int cz1) {
    return cz1 > cz0 ? (cz1 - core_zalign(cz0) + kCoreTZ - 1) / kCoreTZ : 0;
}

constexpr int kCoreSX = 64; // default columns per block
constexpr int kCorePF = 2;  // default columns of loads in flight ahead of use

// Loads and stores at a 32-bit byte offset from a block-uniform base.  The core
// addresses every array it streams as base + (column * P + z) * 4 with the
// shot's plane folded into the base, so all arrays of one column share one
// lane offset (scalar base + vector offset addressing, no 64-bit address
// arithmetic per array).  Offsets stay below 4 GiB: one shot's plane.
__device__ __forceinline__ float ldb(const float *base, unsigned boff) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + boff);
}
__device__ __forceinline__ void stb(float *base, unsigned boff, float v) {
    *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + boff) = v;
}
__device__ __forceinline__ float2 ldb2(const float2 *base, unsigned boff) {
    return *reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(base) + boff);
}
__device__ __forceinline__ void stb2(float2 *base, unsigned boff, float2 v) {
    *reinterpret_cast<float2 *>(reinterpret_cast<char *>(base) + boff) = v;
}
// Streaming store (non-temporal hint) for state the launch does not read
// back: the core's new V, increments and S part.  Measured about 3% faster
// than a plain store on the forward core at 2048 x 1024, 8 shots.
typedef float ntf2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void stb2s(float2 *base, unsigned boff, float2 v) {
    ntf2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<ntf2 *>(reinterpret_cast<char *>(base) + boff));
}

// MODE 0 steps a wavefield (the forward source run, or the receiver run
// without imaging); MODE 2 is the receiver run with the imaging condition
// (its five imaging operands are prefetched with the state).  MODE 1 runs the
// source reconstruction (reshot_u + reshot_s, P-S-seperating.cu:890-1042)
// through the same pipeline: ring totals come from the strips, the step is the
// reverse one, and the new source P velocity level is written for the imaging
// condition.  Its core is [L+2N, Xn-L-2N)^2, so every stress cell it forms lies
// off the ring.
enum { CORE_STEP = 0, CORE_RECON = 1, CORE_IMAGE = 2 };
template <int N, int MODE, int PF>
__device__ __forceinline__ void core_body(const FGrid &g, const FWave &w, const float2 *__restrict__ pq,
                                              const FStep &a, const FRecon &rc, const uint3 bid) {
    constexpr bool RECON = MODE == CORE_RECON, IMG = MODE == CORE_IMAGE;
    constexpr int NS = IMG ? 11 : 6; // prefetched per-cell operands: state (+ imaging)
    constexpr int TZ = kCoreTZ;
    static_assert(TZ + 4 * N <= 256, "core block too small for the halo");
    constexpr int Q = 2 * N + 1;
    __shared__ float lv[2][2][256]; // [column parity][vx, vz][lane]
    __shared__ float ls[2][5][256]; // [column parity][theta, omega, gzx, gxx, q][lane]
    const int l = threadIdx.x;
    const int z = core_zalign(g.cz0) + bid.y * TZ - 2 * N + l;
    const int xs = g.cx0 + bid.x * g.csx;
    const int xe = min(xs + g.csx, g.cx1);
    const long so = (long)bid.z * g.plane;
    const float *c = g.c;
    const int wb = 1 - a.rb;
    const bool z_out = z >= g.cz0 && z < g.cz1 && l >= 2 * N && l < 2 * N + TZ;
    const bool z_st = l >= N && l < 256 - N && z >= N && z < g.Zn - N;
    const bool z_vin = z >= g.vlo && z < g.Zn - g.vlo;
    const bool z_inj = a.inj && z == a.zr;
    // per-block bases (shot plane folded in) and lane byte offset
    const unsigned P4 = (unsigned)g.P * 4u, zb = (unsigned)z * 4u;
    auto off = [&](int col) { return (unsigned)col * P4 + zb; };        // float arrays
    auto off2 = [&](int col) { return 2u * ((unsigned)col * P4 + zb); }; // pair arrays
    const float2 *const Vr = w.V[a.rb].p + so;
    float2 *const Vw = w.V[wb].p + so;
    float2 *const DVp = w.DV.p + so, *const Sp = w.S.p + so, *const DSp = w.DS.p + so;
    // source cell and wavelet sample of this shot and step, read once
    int src_i = -1, src_j = -1;
    float src_v = 0.f;
    if (a.src && (!RECON || rc.exact)) {
        src_i = a.src[2 * bid.z];
        src_j = a.src[2 * bid.z + 1];
        src_v = a.wav[(long)bid.z * g.nt + a.t];
    }
    // imaging condition on this step (shots that start late image only their own steps)
    const bool img_on = IMG && !(a.t0 && a.t < a.t0[bid.z]);
    const float2 *const sp = img_on ? a.sp + so : nullptr;
    float *const npp = img_on ? a.npp + so : nullptr, *const nps = img_on ? a.nps + so : nullptr;
    float *const den = img_on ? a.den + so : nullptr;
    KStrips rs{};
    if (RECON) rs = shot_strips(rc, bid.z);
    auto vin = [&](int i, float &vx, float &vz) {
        if (RECON) {
            recon_vin(g, w, rc, rs, a.rb, so, i, z, vx, vz);
            return;
        }
        vx = 0.f;
        vz = 0.f;
        if (!z_vin || i < g.vlo || i >= g.Xn - g.vlo) return;
        const float2 v = ldb2(Vr, off2(i));
        vx = v.x;
        vz = v.y;
        if (z_inj && i >= g.L && i < g.Xn - g.L) { // load_record, P-S-seperating.cu:1471-1483
            long io = ((long)bid.z * 2 * g.nt + a.t) * g.nx + (i - g.L);
            vx = a.inj[io];
            vz = a.inj[io + (long)g.nt * g.nx];
        }
    };

    float qvx[Q], qvz[Q];                         // V at columns cc-2N..cc
    float xth[Q], xom[Q], xgzz[Q], xgxz[Q], xq[Q]; // stress at columns cs-2N..cs
    float zq[N][4];                               // z-terms Tz, Wz, KXz, KZz, delayed N - 1
    float pv[PF][2], ppq[PF][2], pst[PF][NS];     // prefetched V, (p, q), state (+ imaging)
#pragma unroll
    for (int m = 0; m < Q; ++m) { qvx[m] = qvz[m] = 0.f; xth[m] = xom[m] = xgzz[m] = xgxz[m] = xq[m] = 0.f; }
#pragma unroll
    for (int m = 0; m < N; ++m) zq[m][0] = zq[m][1] = zq[m][2] = zq[m][3] = 0.f;
    auto load_pq = [&](int cs, float *o) {
        o[0] = o[1] = 0.f;
        if (z_st && cs >= xs - N && cs < xe + N) {
            const float2 v = ldb2(pq, off2(cs));
            o[0] = v.x;
            o[1] = v.y;
        }
    };
    auto load_state = [&](int cu, float *o) {
        if (z_out && cu >= xs && cu < xe) {
            const unsigned k = off2(cu);
            float2 v = ldb2(DVp, k); o[0] = v.x; o[1] = v.y;
            v = ldb2(Sp, k); o[2] = v.x; o[3] = v.y;
            v = ldb2(DSp, k); o[4] = v.x; o[5] = v.y;
            if (IMG && img_on) {
                v = ldb2(sp, k); o[6] = v.x; o[7] = v.y;
                const unsigned k4 = off(cu);
                o[8] = ldb(npp, k4); o[9] = ldb(den, k4); o[10] = ldb(nps, k4);
            }
        } else {
#pragma unroll
            for (int e = 0; e < NS; ++e) o[e] = 0.f;
        }
    };
#pragma unroll
    for (int f = 0; f < PF; ++f) {
        const int cc = xs - 2 * N + f;
        vin(cc, pv[f][0], pv[f][1]);
        load_pq(cc - N, ppq[f]);
        load_state(cc - 2 * N, pst[f]);
    }

    // Where the registers allow (N <= 5, one column of prefetch) the V and
    // stress queues rotate instead of shifting: the column loop is unrolled by
    // Q and logical entry m (0 oldest, Q-1 newest) of a queue at unrolled step
    // r lives in register slot (r + 1 + m) mod Q.  Elsewhere they shift.
    constexpr bool ROT = N <= 5 && PF == 1;
    const int cbeg = xs - 2 * N, cend = xe + 2 * N;
    auto body = [&](auto rr, const int cc) {
        constexpr int r = decltype(rr)::value;
#define RS(m) (ROT ? ((r + 1 + (m)) % Q) : (m))
        if constexpr (!ROT) {
#pragma unroll
            for (int m = 0; m < Q - 1; ++m) { qvx[m] = qvx[m + 1]; qvz[m] = qvz[m + 1]; }
        }
        // 1. V at column cc (loaded PF columns ago); issue the loads of column cc + PF
        qvx[RS(Q - 1)] = pv[0][0];
        qvz[RS(Q - 1)] = pv[0][1];
        const float pp = ppq[0][0], qq0 = ppq[0][1];
        float st6[NS];
#pragma unroll
        for (int e = 0; e < NS; ++e) st6[e] = pst[0][e];
#pragma unroll
        for (int f = 0; f < PF - 1; ++f) {
            pv[f][0] = pv[f + 1][0]; pv[f][1] = pv[f + 1][1];
            ppq[f][0] = ppq[f + 1][0]; ppq[f][1] = ppq[f + 1][1];
#pragma unroll
            for (int e = 0; e < NS; ++e) pst[f][e] = pst[f + 1][e];
        }
        vin(cc + PF, pv[PF - 1][0], pv[PF - 1][1]);
        load_pq(cc + PF - N, ppq[PF - 1]);
        load_state(cc + PF - 2 * N, pst[PF - 1]);
        // 2. stress at column cs = cc - N.  LDS is double-buffered by column
        // parity: one barrier per column makes this column's V and the previous
        // column's stress visible; this column's stress is read after the next
        // barrier, so the z-derivative terms lag one column.
        const int cs = cc - N;
        const int b = cc & 1;
        lv[b][0][l] = qvx[RS(N)];
        lv[b][1][l] = qvz[RS(N)];
        __syncthreads();
        float th = 0.f, om = 0.f, gzx = 0.f, gxx = 0.f, qq = 0.f, gzz = 0.f, gxz = 0.f;
        if (cc >= xs && z_st) {
            float s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
            for (int m = 1; m <= N; ++m) {
                s1 += c[m - 1] * (qvx[RS(N + m - 1)] - qvx[RS(N - m)]);          // D-x Vx
                s2 += c[m - 1] * (qvz[RS(N + m)] - qvz[RS(N - m + 1)]);          // D+x Vz
                s3 += c[m - 1] * (lv[b][0][l + m] - lv[b][0][l - m + 1]); // D+z Vx
                s4 += c[m - 1] * (lv[b][1][l + m - 1] - lv[b][1][l - m]); // D-z Vz
            }
            gxx = s1 * g.idx; gzx = s2 * g.idx; gxz = s3 * g.idz; gzz = s4 * g.idz;
            qq = qq0;
            th = pp * (gxx + gzz);
            if (!RECON || rc.exact) th += (cs == src_i && z == src_j) ? src_v : 0.f;
            om = qq * (gxz - gzx);
        }
        ls[b][0][l] = th; ls[b][1][l] = om; ls[b][2][l] = gzx; ls[b][3][l] = gxx; ls[b][4][l] = qq;
        // z-derivative terms at column cs - 1 from the previous column's stress
        // (q0 = q at the cell)
        float tz = 0.f, wz = 0.f, kxz = 0.f, kzz = 0.f;
        if (cc > xs && l >= 2 * N && l < 256 - 2 * N) {
            const int p = b ^ 1;
            const float q0 = ls[p][4][l];
#pragma unroll
            for (int m = 1; m <= N; ++m) {
                tz += c[m - 1] * (ls[p][0][l + m] - ls[p][0][l - m + 1]);
                wz += c[m - 1] * (ls[p][1][l + m - 1] - ls[p][1][l - m]);
                kxz += c[m - 1] * (ls[p][2][l + m - 1] * (ls[p][4][l + m - 1] - q0) - ls[p][2][l - m] * (ls[p][4][l - m] - q0));
                kzz += c[m - 1] * (ls[p][3][l + m] * (ls[p][4][l + m] - q0) - ls[p][3][l - m + 1] * (ls[p][4][l - m + 1] - q0));
            }
            tz *= g.idz; wz *= g.idz; kxz *= g.idz; kzz *= g.idz;
        }
        if constexpr (!ROT) {
#pragma unroll
            for (int m = 0; m < Q - 1; ++m) {
                xth[m] = xth[m + 1]; xom[m] = xom[m + 1]; xgzz[m] = xgzz[m + 1]; xgxz[m] = xgxz[m + 1]; xq[m] = xq[m + 1];
            }
        }
        xth[RS(Q - 1)] = th; xom[RS(Q - 1)] = om; xgzz[RS(Q - 1)] = gzz; xgxz[RS(Q - 1)] = gxz; xq[RS(Q - 1)] = qq;
        // z-terms of column cs - 1 enter the queue; zq[0] is column cs - N = cu
#pragma unroll
        for (int m = 0; m < N - 1; ++m) { zq[m][0] = zq[m + 1][0]; zq[m][1] = zq[m + 1][1]; zq[m][2] = zq[m + 1][2]; zq[m][3] = zq[m + 1][3]; }
        zq[N - 1][0] = tz; zq[N - 1][1] = wz; zq[N - 1][2] = kxz; zq[N - 1][3] = kzz;
        // 3. update at column cu = cc - 2N
        const int cu = cc - 2 * N;
        if (cu >= xs && cu < xe && z_out) {
            const float q0 = xq[RS(N)];
            float tx = 0.f, wx = 0.f, kxx = 0.f, kzx = 0.f;
#pragma unroll
            for (int m = 1; m <= N; ++m) {
                tx += c[m - 1] * (xth[RS(N + m)] - xth[RS(N - m + 1)]);
                wx += c[m - 1] * (xom[RS(N + m - 1)] - xom[RS(N - m)]);
                kxx += c[m - 1] * (xgzz[RS(N + m)] * (xq[RS(N + m)] - q0) - xgzz[RS(N - m + 1)] * (xq[RS(N - m + 1)] - q0));
                kzx += c[m - 1] * (xgxz[RS(N + m - 1)] * (xq[RS(N + m - 1)] - q0) - xgxz[RS(N - m)] * (xq[RS(N - m)] - q0));
            }
            tx *= g.idx; wx *= g.idx; kxx *= g.idx; kzx *= g.idx;
            const float Kx = zq[0][2] - kxx, Kz = kzx - zq[0][3];
            const float hx = zq[0][1], hz = -wx;
            const unsigned k = off2(cu);
            float vx = qvx[RS(0)], vz = qvz[RS(0)];
            if (z_inj) { const float2 v = ldb2(Vr, k); vx = v.x; vz = v.y; } // state, not the injected record
            const float dt2 = g.dt2;
            const float sg = RECON ? -1.f : 1.f; // the reverse step runs the increments backwards
            const float dx = st6[0] + sg * dt2 * (tx + hx + 2.f * Kx); vx = vx + sg * dx;
            const float dz = st6[1] + sg * dt2 * (zq[0][0] + hz + 2.f * Kz); vz = vz + sg * dz;
            stb2s(DVp, k, make_float2(dx, dz));
            stb2s(Vw, k, make_float2(vx, vz));
            const float ex = st6[4] + sg * dt2 * (hx + Kx); const float sx = st6[2] + sg * ex;
            const float ez = st6[5] + sg * dt2 * (hz + Kz); const float sz = st6[3] + sg * ez;
            stb2s(DSp, k, make_float2(ex, ez));
            stb2s(Sp, k, make_float2(sx, sz));
            if (RECON) {
                if (rc.vp) // the source P level for the imaging condition (none for the FWI gradient)
                    stb2(rc.vp + so, k, make_float2(vx - sx, vz - sz));
            } else {
                record(g, a, bid.z, cu, z, vx, vz, sx, sz);
                if (IMG && img_on) { // corr_v (P-S-seperating.cu:1426-1439); core cells are interior
                    const float sa = st6[6 % NS], sb = st6[7 % NS];
                    const float rpx = vx - sx, rpz = vz - sz;
                    const unsigned k4 = off(cu);
                    stb(npp, k4, st6[8 % NS] + (sa * rpx + sb * rpz));
                    stb(den, k4, st6[9 % NS] + (sa * sa + sb * sb));
                    stb(nps, k4, st6[10 % NS] + (sa * sx + sb * sz));
                }
            }
        }
#undef RS
    };
    if constexpr (ROT) {
        for (int c0 = cbeg; c0 < cend; c0 += Q)
            static_for<Q>([&](auto rr) {
                const int cc = c0 + decltype(rr)::value;
                if (cc < cend) body(rr, cc);
            });
    } else {
        for (int cc = cbeg; cc < cend; ++cc) body(std::integral_constant<int, 0>{}, cc);
    }
}
// XCD-aware block order for a launch of dims.x x-tiles, dims.y z blocks and
// dims.z shots.  Workgroups go round-robin to the 8 XCDs (workgroup b to XCD
// b mod 8), each with its own L2.  The launch is one-dimensional with
// 8 * ceil(T / 8) blocks (T = tiles x z blocks x shots); XCD k takes the k-th
// eighth of the tiles in (x-tile, z block, shot) order, so the shots and z
// blocks of an x-tile, which share the model planes and their halo rows, and
// its neighbouring x-tiles meet in one L2.  Blocks past T return at once.
constexpr int kXcds = 8;
__host__ __device__ inline unsigned xcd_blocks(const uint3 dims) {
    const unsigned t = dims.x * dims.y * dims.z;
    return kXcds * ((t + kXcds - 1) / kXcds);
}
__device__ __forceinline__ bool xcd_bid(const uint3 dims, uint3 &bid) {
    const unsigned t = dims.x * dims.y * dims.z, per = (t + kXcds - 1) / kXcds;
    const unsigned b = blockIdx.x, q = (b % kXcds) * per + b / kXcds;
    if (q >= t) return false;
    const unsigned yz = dims.y * dims.z, r = q % yz;
    bid = make_uint3(q / yz, r / dims.z, r % dims.z);
    return true;
}

template <int N, int MODE, int PF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_core(FGrid g, FWave w, const float2 *__restrict__ pq,
                                              FStep a, FRecon rc, uint3 dims) {
    uint3 bid;
    if (!xcd_bid(dims, bid)) return;
    core_body<N, MODE, PF>(g, w, pq, a, rc, bid);
}

// ---------------------------------------------------------------------------
// Band kernels: 64 x 4 tiles from a host-built list (tile origin i0, j0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool in_core(const FGrid &g, int i, int j) {
    return i >= g.cx0 && i < g.cx1 && j >= g.cz0 && j < g.cz1;
}

// Stress with CPML on band cells and the core's outer N-cell margin.
template <int N>
__device__ __forceinline__ void band_stress_body(const FGrid &g, const FWave &w, const KModel &m, const int2 *__restrict__ tiles,
                                                     const FStep &a, const uint3 bid) {
    const int2 t0 = tiles[bid.x];
    const int j = t0.y + threadIdx.x, i = t0.x + threadIdx.y;
    if (!in_update(g, i, j) || in_lpad(g, i, j)) return;
    // skip core cells farther than N from the band
    if (i >= g.cx0 + N && i < g.cx1 - N && j >= g.cz0 + N && j < g.cz1 - N) return;
    const long so = (long)bid.z * g.plane;
    const long k0 = (long)i * g.P + j, k = so + k0;
    const float *c = g.c;
    float s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
    for (int mm = 1; mm <= N; ++mm) {
        float a0, b0, a1, b1;
        load_vin(g, w, a, bid.z, so, i + mm - 1, j, a0, b0);
        load_vin(g, w, a, bid.z, so, i - mm, j, a1, b1);
        s1 += c[mm - 1] * (a0 - a1); // D-x Vx
        load_vin(g, w, a, bid.z, so, i + mm, j, a0, b0);
        load_vin(g, w, a, bid.z, so, i - mm + 1, j, a1, b1);
        s2 += c[mm - 1] * (b0 - b1); // D+x Vz
        load_vin(g, w, a, bid.z, so, i, j + mm, a0, b0);
        load_vin(g, w, a, bid.z, so, i, j - mm + 1, a1, b1);
        s3 += c[mm - 1] * (a0 - a1); // D+z Vx
        load_vin(g, w, a, bid.z, so, i, j + mm - 1, a0, b0);
        load_vin(g, w, a, bid.z, so, i, j - mm, a1, b1);
        s4 += c[mm - 1] * (b0 - b1); // D-z Vz
    }
    const float gxx = s1 * g.idx, gzx = s2 * g.idx, gxz = s3 * g.idz, gzz = s4 * g.idz;
    const float a1 = m.ax[k0], a2 = m.az[k0];
    float Fxx = 0.f, Fzz = 0.f, Fxz = 0.f, Fzx = 0.f;
    if (a1 != 0.f || a2 != 0.f) {
        const float dt = g.dt;
        Fxx = w.F[0][k]; Fzz = w.F[1][k]; Fxz = w.F[2][k]; Fzx = w.F[3][k];
        Fxx = Fxx + (a1 * gxx - a1 * Fxx) * dt;
        Fzz = Fzz + (a2 * gzz - a2 * Fzz) * dt;
        Fxz = Fxz + (a2 * gxz - a2 * Fxz) * dt;
        Fzx = Fzx + (a1 * gzx - a1 * Fzx) * dt;
        w.F[0][k] = Fxx; w.F[1][k] = Fzz; w.F[2][k] = Fxz; w.F[3][k] = Fzx;
    }
    w.th[k] = m.p[k0] * (gxx + gzz - Fxx - Fzz) + src_term(g, a, bid.z, i, j);
    w.om[k] = m.q[k0] * (gxz - gzx - Fxz + Fzx);
    w.gxx[k] = gxx; w.gxz[k] = gxz; w.gzx[k] = gzx; w.gzz[k] = gzz;
}
template <int N>
__global__ __launch_bounds__(256) void k_band_stress(FGrid g, FWave w, KModel m, const int2 *__restrict__ tiles,
                                                     FStep a) {
    band_stress_body<N>(g, w, m, tiles, a, blockIdx3());
}

// Velocity with CPML on band cells.  Every array is addressed from a
// block-uniform base (shot plane folded in) with a 32-bit lane byte offset, so
// the z taps are immediate offsets of one address and the x taps share one
// address per column offset across all fields.
template <int N>
__device__ __forceinline__ void band_velocity_body(const FGrid &g, const FWave &w, const KModel &m, const int2 *__restrict__ tiles,
                                                       const FStep &a, const uint3 bid) {
    const int2 t0 = tiles[bid.x];
    const int j = t0.y + threadIdx.x, i = t0.x + threadIdx.y;
    if (!in_update(g, i, j) || in_core(g, i, j) || in_lpad(g, i, j)) return;
    const long so = (long)bid.z * g.plane;
    const long k0 = (long)i * g.P + j, k = so + k0;
    const unsigned o = (unsigned)k0 * 4u; // byte offset of the cell in a plane
    const int P4 = g.P * 4;
    const float *c = g.c;
    const float *const om = w.om + so, *const th = w.th + so, *const gzx = w.gzx + so, *const gzz = w.gzz + so;
    const float *const gxz = w.gxz + so, *const gxx = w.gxx + so;
    const float *const F0 = w.F[0] + so, *const F1 = w.F[1] + so, *const F2 = w.F[2] + so, *const F3 = w.F[3] + so;
    const float *q = m.q; // per grid, no shot offset
    // Cg = D-z(q gzx), Eg = D+x(q gzz), Bg = D-x(q gxz), Ag = D+z(q gxx) and the
    // same of q F (the psi parts of the reference's du fields); KX, KZ are
    // Cg - Eg and Bg - Ag in the product-rule form (q - q0).
    float Wx = 0.f, Wz = 0.f, Tx = 0.f, Tz = 0.f, KX = 0.f, KZ = 0.f;
    float Cg = 0.f, Eg = 0.f, Bg = 0.f, Ag = 0.f, Cf = 0.f, Ef = 0.f, Bf = 0.f, Af = 0.f;
    // The product-rule form subtracts q0 (D-z gzx - D+x gzz), which vanishes
    // because the staggered x and z operators commute -- only if every g came
    // from the same operator.  Within N of the learned boundary's low-order pad
    // the g's do not, so the plain form (q0 = 0) is used there, as the
    // reconstruction uses it next to its ring.
    const float q0 = (g.lpad && edge_dist(g, i, j) < g.lpad + N) ? 0.f : ldb(q, o);
#pragma unroll
    for (int mm = 1; mm <= N; ++mm) {
        const float cm = c[mm - 1];
        const unsigned xp = o + (unsigned)(mm * P4), xm = o - (unsigned)(mm * P4);          // i + mm, i - mm
        const unsigned xp1 = o + (unsigned)((mm - 1) * P4), xm1 = o - (unsigned)((mm - 1) * P4); // i + mm-1, i - (mm-1)
        const unsigned zp = o + 4u * mm, zm = o - 4u * mm, zp1 = o + 4u * (mm - 1), zm1 = o - 4u * (mm - 1);
        Wx += cm * (ldb(om, xp1) - ldb(om, xm));
        Wz += cm * (ldb(om, zp1) - ldb(om, zm));
        Tx += cm * (ldb(th, xp) - ldb(th, xm1));
        Tz += cm * (ldb(th, zp) - ldb(th, zm1));
        {   // D-z along z: cells k+mm-1, k-mm
            const float qa = ldb(q, zp1), qb = ldb(q, zm);
            const float ga = ldb(gzx, zp1), gb = ldb(gzx, zm);
            KX += cm * (ga * (qa - q0) - gb * (qb - q0)) * g.idz;
            Cg += cm * (ga * qa - gb * qb);
            Cf += cm * (ldb(F3, zp1) * qa - ldb(F3, zm) * qb);
        }
        {   // D+x along x: cells k+mm*P, k-(mm-1)*P
            const float qa = ldb(q, xp), qb = ldb(q, xm1);
            const float ga = ldb(gzz, xp), gb = ldb(gzz, xm1);
            KX -= cm * (ga * (qa - q0) - gb * (qb - q0)) * g.idx;
            Eg += cm * (ga * qa - gb * qb);
            Ef += cm * (ldb(F1, xp) * qa - ldb(F1, xm1) * qb);
        }
        {   // D-x along x: cells k+(mm-1)*P, k-mm*P
            const float qa = ldb(q, xp1), qb = ldb(q, xm);
            const float ga = ldb(gxz, xp1), gb = ldb(gxz, xm);
            KZ += cm * (ga * (qa - q0) - gb * (qb - q0)) * g.idx;
            Bg += cm * (ga * qa - gb * qb);
            Bf += cm * (ldb(F2, xp1) * qa - ldb(F2, xm) * qb);
        }
        {   // D+z along z: cells k+mm, k-mm+1
            const float qa = ldb(q, zp), qb = ldb(q, zm1);
            const float ga = ldb(gxx, zp), gb = ldb(gxx, zm1);
            KZ -= cm * (ga * (qa - q0) - gb * (qb - q0)) * g.idz;
            Ag += cm * (ga * qa - gb * qb);
            Af += cm * (ldb(F0, zp) * qa - ldb(F0, zm1) * qb);
        }
    }
    Wx *= g.idx; Wz *= g.idz; Tx *= g.idx; Tz *= g.idz;
    const float Cp = (Cg - Cf) * g.idz, Ep = (Eg - Ef) * g.idx, Bp = (Bg - Bf) * g.idx, Ap = (Ag - Af) * g.idz;
    const float FX = Cf * g.idz - Ef * g.idx, FZ = Bf * g.idx - Af * g.idz;
    const float a1 = ldb(m.ax, o), a2 = ldb(m.az, o);
    float OTx = 0.f, OTz = 0.f, OWz = 0.f, OWx = 0.f, OC = 0.f, OE = 0.f, OB = 0.f, OA = 0.f;
    if (a1 != 0.f || a2 != 0.f) {
        const float dt = g.dt;
        float *const O0 = w.O[0] + so, *const O1 = w.O[1] + so, *const O2 = w.O[2] + so, *const O3 = w.O[3] + so;
        float *const O4 = w.O[4] + so, *const O5 = w.O[5] + so, *const O6 = w.O[6] + so, *const O7 = w.O[7] + so;
        OC = ldb(O0, o); OE = ldb(O1, o); OB = ldb(O2, o); OA = ldb(O3, o);
        OTx = ldb(O4, o); OTz = ldb(O5, o); OWz = ldb(O6, o); OWx = ldb(O7, o);
        OC = OC + (a2 * Cp - a2 * OC) * dt;
        OE = OE + (a1 * Ep - a1 * OE) * dt;
        OB = OB + (a1 * Bp - a1 * OB) * dt;
        OA = OA + (a2 * Ap - a2 * OA) * dt;
        OTx = OTx + (a1 * Tx - a1 * OTx) * dt;
        OTz = OTz + (a2 * Tz - a2 * OTz) * dt;
        OWz = OWz + (a2 * Wz - a2 * OWz) * dt;
        OWx = OWx + (a1 * Wx - a1 * OWx) * dt;
        stb(O0, o, OC); stb(O1, o, OE); stb(O2, o, OB); stb(O3, o, OA);
        stb(O4, o, OTx); stb(O5, o, OTz); stb(O6, o, OWz); stb(O7, o, OWx);
    }
    const float Kx = (KX - FX) - OC + OE, Kz = (KZ - FZ) - OB + OA;
    const float hx = Wz - OWz, hz = -Wx + OWx;
    const float dt2 = g.dt2;
    const int wb = 1 - a.rb;
    const unsigned o2 = 2u * o; // byte offset of the cell in a pair plane
    float2 *const DVp = w.DV.p + so, *const Sp = w.S.p + so, *const DSp = w.DS.p + so;
    const float2 dv = ldb2(DVp, o2), v = ldb2(w.V[a.rb].p + so, o2);
    const float dx = dv.x + dt2 * ((Tx - OTx) + hx + 2.f * Kx); const float vx = v.x + dx;
    const float dz = dv.y + dt2 * ((Tz - OTz) + hz + 2.f * Kz); const float vz = v.y + dz;
    stb2(DVp, o2, make_float2(dx, dz));
    stb2(w.V[wb].p + so, o2, make_float2(vx, vz));
    const float2 ds = ldb2(DSp, o2), sv = ldb2(Sp, o2);
    const float ex = ds.x + dt2 * (hx + Kx); const float sx = sv.x + ex;
    const float ez = ds.y + dt2 * (hz + Kz); const float sz = sv.y + ez;
    stb2(DSp, o2, make_float2(ex, ez));
    stb2(Sp, o2, make_float2(sx, sz));
    record(g, a, bid.z, i, j, vx, vz, sx, sz);
    correlate(g, a, bid.z, k, i, j, vx, vz, sx, sz);
}
template <int N>
__global__ __launch_bounds__(256) void k_band_velocity(FGrid g, FWave w, KModel m, const int2 *__restrict__ tiles,
                                                       FStep a) {
    band_velocity_body<N>(g, w, m, tiles, a, blockIdx3());
}

// ---------------------------------------------------------------------------
// Reconstruction (reshot_u + reshot_s, P-S-seperating.cu:890-1042) fused:
// on the interior, V and the stress-like fields of ring cells come from the
// strips; the reverse step uses theta^{t+1} recomputed from V^{t+1}.  Writes the
// source P velocity of the new level for the imaging condition.
// ---------------------------------------------------------------------------
// Stores stress-like fields of the ring frame and interior for the reverse
// step of the next launch (the stress is a function of V^{t+1} only).  The
// block's 64 x 4 tile and its N-cell halo of totals (ring cells from the
// strips) are staged in LDS once, so each total is looked up once per tile
// rather than once per stencil tap.
template <int N>
__device__ __forceinline__ void recon_stress_body(const FGrid &g, const FWave &w, const float *__restrict__ p2,
                                                        const float *__restrict__ q2, const FRecon &rc, const FStep &a,
                                                        const int2 *__restrict__ tiles, const uint3 bid) {
    constexpr int HX = 4 + 2 * N, HZ = 64 + 2 * N;
    __shared__ float lx[HX][HZ], lz[HX][HZ];
    const int2 t0 = tiles[bid.x];
    const long so = (long)bid.z * g.plane;
    const KStrips s = shot_strips(rc, bid.z);
    for (int e = threadIdx.y * 64 + threadIdx.x; e < HX * HZ; e += 256) {
        const int xi = e / HZ, zj = e - xi * HZ;
        float vx, vz;
        recon_vin(g, w, rc, s, a.rb, so, t0.x - N + xi, t0.y - N + zj, vx, vz);
        lx[xi][zj] = vx;
        lz[xi][zj] = vz;
    }
    __syncthreads();
    const int j = t0.y + threadIdx.x, i = t0.x + threadIdx.y;
    if (i >= g.Xn - g.L + N || j >= g.Zn - g.L + N) return;
    // only the band next to the ring and its N-cell margin: the recon core forms its own
    if (i >= g.L + 3 * N && i < g.Xn - g.L - 3 * N && j >= g.L + 3 * N && j < g.Zn - g.L - 3 * N) return;
    const long r = ring_index(rc.ring, i, j);
    // outer ring cells have no interior neighbours to differentiate; the
    // reconstruction reads them only as strip values
    if (r < 0 && (i < g.L || i >= g.Xn - g.L || j < g.L || j >= g.Zn - g.L)) return;
    float th, om, gxx, gxz, gzx, gzz;
    if (r >= 0) {
        th = strip_get(s, rc.slot_th, RS_TH, r);
        om = strip_get(s, rc.slot_th, RS_OM, r);
        gzx = strip_get(s, rc.slot_th, RS_GZX, r);
        gzz = strip_get(s, rc.slot_th, RS_GZZ, r);
        gxx = strip_get(s, rc.slot_th, RS_GXX, r);
        gxz = strip_get(s, rc.slot_th, RS_GXZ, r);
    } else {
        const float *c = g.c;
        const int xl = threadIdx.y + N, zl = threadIdx.x + N;
        float s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
        for (int mm = 1; mm <= N; ++mm) {
            s1 += c[mm - 1] * (lx[xl + mm - 1][zl] - lx[xl - mm][zl]);
            s2 += c[mm - 1] * (lz[xl + mm][zl] - lz[xl - mm + 1][zl]);
            s3 += c[mm - 1] * (lx[xl][zl + mm] - lx[xl][zl - mm + 1]);
            s4 += c[mm - 1] * (lz[xl][zl + mm - 1] - lz[xl][zl - mm]);
        }
        gxx = s1 * g.idx; gzx = s2 * g.idx; gxz = s3 * g.idz; gzz = s4 * g.idz;
        const long k0 = (long)i * g.P + j;
        th = p2[k0] * (gxx + gzz);
        if (rc.exact) th += src_term(g, a, bid.z, i, j);
        om = q2[k0] * (gxz - gzx);
    }
    const long k = so + (long)i * g.P + j;
    w.th[k] = th; w.om[k] = om; w.gxx[k] = gxx; w.gxz[k] = gxz; w.gzx[k] = gzx; w.gzz[k] = gzz;
}
template <int N>
__global__ __launch_bounds__(256) void k_recon_stress_f(FGrid g, FWave w, const float *__restrict__ p2,
                                                        const float *__restrict__ q2, FRecon rc, FStep a,
                                                        const int2 *__restrict__ tiles) {
    recon_stress_body<N>(g, w, p2, q2, rc, a, tiles, blockIdx3());
}

// Reverse velocity step on the interior from the stored stress-like fields.
template <int N>
__device__ __forceinline__ void recon_velocity_body(const FGrid &g, const FWave &w, const float *__restrict__ q2, const FRecon &rc,
                                                          const FStep &a, const int2 *__restrict__ tiles, const uint3 bid) {
    const int2 t0 = tiles[bid.x];
    const int j = t0.y + threadIdx.x, i = t0.x + threadIdx.y;
    const int L = g.L;
    if (i < L || j < L || i >= g.Xn - L || j >= g.Zn - L) return;
    if (i >= L + 2 * N && i < g.Xn - L - 2 * N && j >= L + 2 * N && j < g.Zn - L - 2 * N) return; // recon core
    const long so = (long)bid.z * g.plane;
    const long k0 = (long)i * g.P + j, k = so + k0;
    const float *c = g.c;
    const int P = g.P;
    float Wx = 0.f, Wz = 0.f, Tx = 0.f, Tz = 0.f, KX = 0.f, KZ = 0.f;
    const bool safe = i >= L + 2 * N && i < g.Xn - L - 2 * N && j >= L + 2 * N && j < g.Zn - L - 2 * N;
    const float q0 = safe ? q2[k0] : 0.f;
#pragma unroll
    for (int mm = 1; mm <= N; ++mm) {
        const float cm = c[mm - 1];
        const long xp = (long)mm * P, xm1 = (long)(mm - 1) * P;
        Wx += cm * (w.om[k + xm1] - w.om[k - xp]);
        Wz += cm * (w.om[k + mm - 1] - w.om[k - mm]);
        Tx += cm * (w.th[k + xp] - w.th[k - xm1]);
        Tz += cm * (w.th[k + mm] - w.th[k - mm + 1]);
        KX += cm * ((w.gzx[k + mm - 1] * (q2[k0 + mm - 1] - q0) - w.gzx[k - mm] * (q2[k0 - mm] - q0)) * g.idz -
                    (w.gzz[k + xp] * (q2[k0 + xp] - q0) - w.gzz[k - xm1] * (q2[k0 - xm1] - q0)) * g.idx);
        KZ += cm * ((w.gxz[k + xm1] * (q2[k0 + xm1] - q0) - w.gxz[k - xp] * (q2[k0 - xp] - q0)) * g.idx -
                    (w.gxx[k + mm] * (q2[k0 + mm] - q0) - w.gxx[k - mm + 1] * (q2[k0 - mm + 1] - q0)) * g.idz);
    }
    Wx *= g.idx; Wz *= g.idz; Tx *= g.idx; Tz *= g.idz;
    const float dt2 = g.dt2;
    const int wb = 1 - a.rb;
    const float2 dv = w.DV.p[k], v = w.V[a.rb].p[k];
    const float dx = dv.x - dt2 * (Tx + Wz + 2.f * KX); const float vx = v.x - dx;
    const float dz = dv.y - dt2 * (Tz - Wx + 2.f * KZ); const float vz = v.y - dz;
    w.DV.p[k] = make_float2(dx, dz);
    w.V[wb].p[k] = make_float2(vx, vz);
    const float2 ds = w.DS.p[k], sv = w.S.p[k];
    const float ex = ds.x - dt2 * (Wz + KX); const float sx = sv.x - ex;
    const float ez = ds.y - dt2 * (-Wx + KZ); const float sz = sv.y - ez;
    w.DS.p[k] = make_float2(ex, ez);
    w.S.p[k] = make_float2(sx, sz);
    if (rc.vp) rc.vp[k] = make_float2(vx - sx, vz - sz);
}
template <int N>
__global__ __launch_bounds__(256) void k_recon_velocity_f(FGrid g, FWave w, const float *__restrict__ q2, FRecon rc,
                                                          FStep a, const int2 *__restrict__ tiles) {
    recon_velocity_body<N>(g, w, q2, rc, a, tiles, blockIdx3());
}

// Strip store from the stored band stress and the written V (forward):
// with_v 0 stores the stress-like fields, 1 the velocities, 2 both.
__device__ __forceinline__ void fstrip_save_body(const FGrid &g, const FWave &w, const KRing &r, const KStrips &s0, const KStripStride &d, int slot,
                                                     int vbuf, int with_v, const uint3 bid) {
    long q = (long)bid.x * blockDim.x + threadIdx.x;
    bool valid = q < r.ncells;
    int i = r.L, j = r.L;
    if (valid) ring_cell(r, q, i, j);
    const KStrips s = strips_of_shot(s0, d, bid.z);
    const long so = (long)bid.z * g.plane;
    const long k = so + (long)i * g.P + j;
    if (with_v) {
        const float2 v = valid ? w.V[vbuf].p[k] : make_float2(0.f, 0.f);
        strip_put(s, slot, RS_VX, q, valid, v.x);
        strip_put(s, slot, RS_VZ, q, valid, v.y);
    }
    if (with_v != 1) {
        strip_put(s, slot, RS_TH, q, valid, valid ? w.th[k] : 0.f);
        strip_put(s, slot, RS_OM, q, valid, valid ? w.om[k] : 0.f);
        strip_put(s, slot, RS_GZX, q, valid, valid ? w.gzx[k] - w.F[3][k] : 0.f);
        strip_put(s, slot, RS_GZZ, q, valid, valid ? w.gzz[k] - w.F[1][k] : 0.f);
        strip_put(s, slot, RS_GXX, q, valid, valid ? w.gxx[k] - w.F[0][k] : 0.f);
        strip_put(s, slot, RS_GXZ, q, valid, valid ? w.gxz[k] - w.F[2][k] : 0.f);
    }
}
__global__ __launch_bounds__(256) void k_fstrip_save(FGrid g, FWave w, KRing r, KStrips s0, KStripStride d, int slot,
                                                     int vbuf, int with_v) {
    fstrip_save_body(g, w, r, s0, d, slot, vbuf, with_v, blockIdx3());
}

// Hand-off: (V^nt, D) -> (V^{nt-1}, D) into the other V buffer; P velocity
// of levels nt and nt-1.
__device__ __forceinline__ void funstep_body(const FGrid &g, const FWave &w, int rb, float2 *__restrict__ vp_nt,
                                                 float2 *__restrict__ vp_m1, const uint3 bid) {
    const int j = bid.x * 64 + threadIdx.x, i = bid.y * 4 + threadIdx.y;
    if (i >= g.Xn || j >= g.Zn) return;
    const long k = (long)bid.z * g.plane + (long)i * g.P + j;
    const float2 v = w.V[rb].p[k], sv = w.S.p[k], dv = w.DV.p[k], ds = w.DS.p[k];
    const float vx = v.x, vz = v.y, sx = sv.x, sz = sv.y;
    vp_nt[k] = make_float2(vx - sx, vz - sz);
    const float vx1 = vx - dv.x, vz1 = vz - dv.y, sx1 = sx - ds.x, sz1 = sz - ds.y;
    w.V[1 - rb].p[k] = make_float2(vx1, vz1);
    w.S.p[k] = make_float2(sx1, sz1);
    vp_m1[k] = make_float2(vx1 - sx1, vz1 - sz1);
}
__global__ __launch_bounds__(256) void k_funstep(FGrid g, FWave w, int rb, float2 *__restrict__ vp_nt,
                                                 float2 *__restrict__ vp_m1) {
    funstep_body(g, w, rb, vp_nt, vp_m1, blockIdx3());
}

} // namespace ouc
