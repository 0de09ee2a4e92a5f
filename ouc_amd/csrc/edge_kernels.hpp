// CPML band of the fused engine in one streaming pass.
//
// The band is every update cell outside the core: four rectangles (top and
// bottom full width, left and right between them).  k_edge walks each with
// one-wave blocks (64 lanes along z, 64 - 4N output rows) that stream along x
// like the core kernel: at stream position cc the block loads V at column cc,
// forms the stress at column cc - N with its CPML memory update (forward_s,
// P-S-seperating.cu:316-384) and updates the velocity at column cc - 2N with
// its CPML memory update (forward_u, P-S-seperating.cu:205-314).  x
// derivatives come from register queues, z derivatives from LDS.
//
// Halo columns and rows of a block's stress overlap its neighbours', so the
// stress memory variables F are double-buffered like V: every block reads F of
// the previous step (buffer rb) and the block that owns a cell writes the new
// value (buffer 1 - rb).  The velocity memory variables O are read and written
// by the owner only.  The stress itself is not stored, except on the strip
// ring when the step saves strips.
#pragma once

#include "fused_kernels.hpp"

namespace ouc {

// Velocity-update cells [x0, x1) x [z0, z1); nzb blocks along z, nxb along x,
// first = index of the rectangle's first block.
struct ERect {
    int x0, x1, z0, z1, nzb, nxb, first;
};
struct EPlan {
    ERect r[4];
    int xchunk;    // columns each block streams
    int save_ring; // store the stress of strip-ring cells (forward steps of an RTM run)
};

constexpr int kEdgeW = 64;

__device__ __forceinline__ bool in_ring_frame(const FGrid &g, int i, int j) {
    const int L = g.L, N = g.N;
    const bool frame = i >= L - N && i < g.Xn - L + N && j >= L - N && j < g.Zn - L + N;
    const bool inner = i >= L + N && i < g.Xn - L - N && j >= L + N && j < g.Zn - L - N;
    return frame && !inner;
}

template <int N, bool IMG>
__device__ __forceinline__ void edge_body(const FGrid &g, const FWave &w, const KModel &m,
                                                                                        const FStep &a, const EPlan &pl, const uint3 bid) {
    constexpr int W = kEdgeW, TZ = W - 4 * N, Q = 2 * N + 1;
    constexpr int NS = IMG ? 21 : 16; // prefetched per-cell operands at the update column
    static_assert(TZ > 0, "edge block too small for the halo");
    __shared__ float lv[2][2][W];        // [column parity][vx, vz][lane]
    __shared__ float ls[N + 1][7][W];    // [column mod N+1][theta, omega, gzx, gxx, F3, F0, q][lane]
    const int l = threadIdx.x;
    const int b0 = bid.x;
    const ERect R = b0 >= pl.r[3].first ? pl.r[3] : b0 >= pl.r[2].first ? pl.r[2] : b0 >= pl.r[1].first ? pl.r[1]
                                                                                                       : pl.r[0];
    const int lb = b0 - R.first, zb = lb % R.nzb, xb = lb / R.nzb;
    const int xs = R.x0 + xb * pl.xchunk, xe = min(xs + pl.xchunk, R.x1);
    const int zbase = R.z0 + zb * TZ, z = zbase - 2 * N + l;
    const bool z_out = z >= zbase && z < min(zbase + TZ, R.z1);
    const bool z_st = l >= N && l < W - N && z >= N && z < g.Zn - N;
    const bool z_vin = z >= g.vlo && z < g.Zn - g.vlo;
    const bool z_inj = a.inj && z == a.zr;
    const long so = (long)bid.z * g.plane;
    const float *c = g.c;
    const int wb = 1 - a.rb;
    const unsigned P4 = (unsigned)g.P * 4u, zb4 = (unsigned)z * 4u;
    auto off = [&](int col) { return (unsigned)col * P4 + zb4; };        // float arrays
    auto off2 = [&](int col) { return 2u * ((unsigned)col * P4 + zb4); }; // pair arrays
    const float2 *const Vr = w.V[a.rb].p + so;
    float2 *const Vw = w.V[wb].p + so;
    float2 *const DVp = w.DV.p + so, *const Sp = w.S.p + so, *const DSp = w.DS.p + so;
    float *const *Fr = a.rb == 0 ? w.F : w.F2; // F of the previous step
    float *const *Fw = a.rb == 0 ? w.F2 : w.F; // F of this step (owner writes)
    const float *const Fr0 = Fr[0] + so, *const Fr1 = Fr[1] + so, *const Fr2 = Fr[2] + so, *const Fr3 = Fr[3] + so;
    float *const Fw0 = Fw[0] + so, *const Fw1 = Fw[1] + so, *const Fw2 = Fw[2] + so, *const Fw3 = Fw[3] + so;
    float *const O0 = w.O[0] + so, *const O1 = w.O[1] + so, *const O2 = w.O[2] + so, *const O3 = w.O[3] + so;
    float *const O4 = w.O[4] + so, *const O5 = w.O[5] + so, *const O6 = w.O[6] + so, *const O7 = w.O[7] + so;
    int src_i = -1, src_j = -1;
    float src_v = 0.f;
    if (a.src) {
        src_i = a.src[2 * bid.z];
        src_j = a.src[2 * bid.z + 1];
        src_v = a.wav[(long)bid.z * g.nt + a.t];
    }
    const bool img_on = IMG && !(a.t0 && a.t < a.t0[bid.z]);
    const bool z_int = z >= g.L && z < g.Zn - g.L;
    const float2 *const sp = img_on ? a.sp + so : nullptr;
    float *const npp = img_on ? a.npp + so : nullptr, *const nps = img_on ? a.nps + so : nullptr;
    float *const den = img_on ? a.den + so : nullptr;

    auto vin = [&](int i, float &vx, float &vz) {
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
    // stress operands at column cs: p, q, ax, az, F (previous step)
    auto load_sop = [&](int cs, float *o) {
        if (z_st && cs >= N && cs < g.Xn - N && cs >= xs - N && cs < xe + N) {
            const unsigned k = off(cs);
            o[0] = ldb(m.p, k); o[1] = ldb(m.q, k); o[2] = ldb(m.ax, k); o[3] = ldb(m.az, k);
            o[4] = ldb(Fr0, k); o[5] = ldb(Fr1, k); o[6] = ldb(Fr2, k); o[7] = ldb(Fr3, k);
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) o[e] = 0.f;
        }
    };
    // update operands at column cu: state, O, ax, az (+ imaging)
    auto load_uop = [&](int cu, float *o) {
        if (z_out && cu >= xs && cu < xe) {
            const unsigned k2 = off2(cu), k = off(cu);
            float2 v = ldb2(DVp, k2); o[0] = v.x; o[1] = v.y;
            v = ldb2(Sp, k2); o[2] = v.x; o[3] = v.y;
            v = ldb2(DSp, k2); o[4] = v.x; o[5] = v.y;
            o[6] = ldb(O0, k); o[7] = ldb(O1, k); o[8] = ldb(O2, k); o[9] = ldb(O3, k);
            o[10] = ldb(O4, k); o[11] = ldb(O5, k); o[12] = ldb(O6, k); o[13] = ldb(O7, k);
            o[14] = ldb(m.ax, k); o[15] = ldb(m.az, k);
            if (IMG) {
                if (img_on && z_int && cu >= g.L && cu < g.Xn - g.L) {
                    v = ldb2(sp, k2); o[16 % NS] = v.x; o[17 % NS] = v.y; o[18 % NS] = ldb(npp, k);
                    o[19 % NS] = ldb(den, k); o[20 % NS] = ldb(nps, k);
                } else {
                    o[16 % NS] = o[17 % NS] = o[18 % NS] = o[19 % NS] = o[20 % NS] = 0.f;
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < NS; ++e) o[e] = 0.f;
        }
    };

    // x queues (entry m: column cs - 2N + m): V and the stress fields whose x
    // derivatives the velocity takes
    float qvx[Q], qvz[Q];
    float xth[Q], xom[Q], xgzz[Q], xgxz[Q], xf1[Q], xf2[Q], xq[Q];
#pragma unroll
    for (int e = 0; e < Q; ++e) {
        qvx[e] = qvz[e] = 0.f;
        xth[e] = xom[e] = xgzz[e] = xgxz[e] = xf1[e] = xf2[e] = xq[e] = 0.f;
    }
    float pv[2], psop[8], puop[NS];
    const int cbeg = xs - 2 * N, cend = xe + 2 * N;
    vin(cbeg, pv[0], pv[1]);
    load_sop(cbeg - N, psop);
    load_uop(cbeg - 2 * N, puop);
    // Rotating queues unroll the column loop 2N + 1 times (about 10,400
    // instructions at N = 5, beyond what the instruction cache holds).  The
    // forward band runs alone and gains from the fewer moves; the receiver's
    // band (IMG) runs beside the reconstruction's kernels on the other stream,
    // and there the shifting queues' 1,700-instruction loop is 9 % faster.
    // The arithmetic is the same either way.
    constexpr bool ROT = N <= 5 && !IMG;
    auto body = [&](auto rr, const int cc) {
        constexpr int r = decltype(rr)::value;
#define RS(e) (ROT ? ((r + 1 + (e)) % Q) : (e))
        if constexpr (!ROT) {
#pragma unroll
            for (int e = 0; e < Q - 1; ++e) {
                qvx[e] = qvx[e + 1]; qvz[e] = qvz[e + 1];
                xth[e] = xth[e + 1]; xom[e] = xom[e + 1]; xgzz[e] = xgzz[e + 1]; xgxz[e] = xgxz[e + 1];
                xf1[e] = xf1[e + 1]; xf2[e] = xf2[e + 1]; xq[e] = xq[e + 1];
            }
        }
        qvx[RS(Q - 1)] = pv[0];
        qvz[RS(Q - 1)] = pv[1];
        float so8[8], uo[NS];
#pragma unroll
        for (int e = 0; e < 8; ++e) so8[e] = psop[e];
#pragma unroll
        for (int e = 0; e < NS; ++e) uo[e] = puop[e];
        vin(cc + 1, pv[0], pv[1]);
        load_sop(cc + 1 - N, psop);
        load_uop(cc + 1 - 2 * N, puop);
        const int cs = cc - N;
        const int bp = cc & 1;
        const int sw = (cc - cbeg) % (N + 1);     // LDS slot of column cs
        const int sr = (cc - cbeg + 1) % (N + 1); // LDS slot of column cu = cs - N
        lv[bp][0][l] = qvx[RS(N)];
        lv[bp][1][l] = qvz[RS(N)];
        __syncthreads();
        // 1. stress at column cs with the F update
        float th = 0.f, om = 0.f, gxx = 0.f, gzx = 0.f, gzz = 0.f, gxz = 0.f;
        float Fxx = 0.f, Fzz = 0.f, Fxz = 0.f, Fzx = 0.f;
        const float qq = so8[1];
        if (cs >= xs - N && cs >= N && cs < g.Xn - N && z_st) {
            float s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
            for (int e = 1; e <= N; ++e) {
                s1 += c[e - 1] * (qvx[RS(N + e - 1)] - qvx[RS(N - e)]);   // D-x Vx
                s2 += c[e - 1] * (qvz[RS(N + e)] - qvz[RS(N - e + 1)]);   // D+x Vz
                s3 += c[e - 1] * (lv[bp][0][l + e] - lv[bp][0][l - e + 1]); // D+z Vx
                s4 += c[e - 1] * (lv[bp][1][l + e - 1] - lv[bp][1][l - e]); // D-z Vz
            }
            gxx = s1 * g.idx; gzx = s2 * g.idx; gxz = s3 * g.idz; gzz = s4 * g.idz;
            const float a1 = so8[2], a2 = so8[3];
            if (a1 != 0.f || a2 != 0.f) {
                const float dt = g.dt;
                Fxx = so8[4] + (a1 * gxx - a1 * so8[4]) * dt;
                Fzz = so8[5] + (a2 * gzz - a2 * so8[5]) * dt;
                Fxz = so8[6] + (a2 * gxz - a2 * so8[6]) * dt;
                Fzx = so8[7] + (a1 * gzx - a1 * so8[7]) * dt;
            }
            th = so8[0] * (gxx + gzz - Fxx - Fzz) + ((cs == src_i && z == src_j) ? src_v : 0.f);
            om = qq * (gxz - gzx - Fxz + Fzx);
            if (cs >= xs && cs < xe && z_out) { // this block owns the cell
                const unsigned k = off(cs);
                if (a1 != 0.f || a2 != 0.f) {
                    stb(Fw0, k, Fxx); stb(Fw1, k, Fzz); stb(Fw2, k, Fxz); stb(Fw3, k, Fzx);
                }
                if (pl.save_ring && in_ring_frame(g, cs, z)) {
                    stb(w.th + so, k, th); stb(w.om + so, k, om);
                    stb(w.gxx + so, k, gxx); stb(w.gxz + so, k, gxz); stb(w.gzx + so, k, gzx); stb(w.gzz + so, k, gzz);
                }
            }
        }
        ls[sw][0][l] = th; ls[sw][1][l] = om; ls[sw][2][l] = gzx; ls[sw][3][l] = gxx;
        ls[sw][4][l] = Fzx; ls[sw][5][l] = Fxx; ls[sw][6][l] = qq;
        xth[RS(Q - 1)] = th; xom[RS(Q - 1)] = om; xgzz[RS(Q - 1)] = gzz; xgxz[RS(Q - 1)] = gxz;
        xf1[RS(Q - 1)] = Fzz; xf2[RS(Q - 1)] = Fxz; xq[RS(Q - 1)] = qq;
        // 2. velocity at column cu = cc - 2N (its stress was formed N columns ago)
        const int cu = cc - 2 * N;
        if (cu >= xs && cu < xe && z_out) {
            const float q0 = xq[RS(N)];
            float Wx = 0.f, Wz = 0.f, Tx = 0.f, Tz = 0.f, KX = 0.f, KZ = 0.f;
            float Cg = 0.f, Eg = 0.f, Bg = 0.f, Ag = 0.f, Cf = 0.f, Ef = 0.f, Bf = 0.f, Af = 0.f;
#pragma unroll
            for (int e = 1; e <= N; ++e) {
                const float cm = c[e - 1];
                Wx += cm * (xom[RS(N + e - 1)] - xom[RS(N - e)]);
                Wz += cm * (ls[sr][1][l + e - 1] - ls[sr][1][l - e]);
                Tx += cm * (xth[RS(N + e)] - xth[RS(N - e + 1)]);
                Tz += cm * (ls[sr][0][l + e] - ls[sr][0][l - e + 1]);
                {   // D-z: rows z+e-1, z-e
                    const float qa = ls[sr][6][l + e - 1], qb = ls[sr][6][l - e];
                    const float ga = ls[sr][2][l + e - 1], gb = ls[sr][2][l - e];
                    KX += cm * (ga * (qa - q0) - gb * (qb - q0)) * g.idz;
                    Cg += cm * (ga * qa - gb * qb);
                    Cf += cm * (ls[sr][4][l + e - 1] * qa - ls[sr][4][l - e] * qb);
                }
                {   // D+x: columns cu+e, cu-e+1
                    const float qa = xq[RS(N + e)], qb = xq[RS(N - e + 1)];
                    const float ga = xgzz[RS(N + e)], gb = xgzz[RS(N - e + 1)];
                    KX -= cm * (ga * (qa - q0) - gb * (qb - q0)) * g.idx;
                    Eg += cm * (ga * qa - gb * qb);
                    Ef += cm * (xf1[RS(N + e)] * qa - xf1[RS(N - e + 1)] * qb);
                }
                {   // D-x: columns cu+e-1, cu-e
                    const float qa = xq[RS(N + e - 1)], qb = xq[RS(N - e)];
                    const float ga = xgxz[RS(N + e - 1)], gb = xgxz[RS(N - e)];
                    KZ += cm * (ga * (qa - q0) - gb * (qb - q0)) * g.idx;
                    Bg += cm * (ga * qa - gb * qb);
                    Bf += cm * (xf2[RS(N + e - 1)] * qa - xf2[RS(N - e)] * qb);
                }
                {   // D+z: rows z+e, z-e+1
                    const float qa = ls[sr][6][l + e], qb = ls[sr][6][l - e + 1];
                    const float ga = ls[sr][3][l + e], gb = ls[sr][3][l - e + 1];
                    KZ -= cm * (ga * (qa - q0) - gb * (qb - q0)) * g.idz;
                    Ag += cm * (ga * qa - gb * qb);
                    Af += cm * (ls[sr][5][l + e] * qa - ls[sr][5][l - e + 1] * qb);
                }
            }
            Wx *= g.idx; Wz *= g.idz; Tx *= g.idx; Tz *= g.idz;
            const float Cp = (Cg - Cf) * g.idz, Ep = (Eg - Ef) * g.idx, Bp = (Bg - Bf) * g.idx, Ap = (Ag - Af) * g.idz;
            const float FX = Cf * g.idz - Ef * g.idx, FZ = Bf * g.idx - Af * g.idz;
            const float a1 = uo[14], a2 = uo[15];
            float OTx = 0.f, OTz = 0.f, OWz = 0.f, OWx = 0.f, OC = 0.f, OE = 0.f, OB = 0.f, OA = 0.f;
            const unsigned k = off(cu);
            if (a1 != 0.f || a2 != 0.f) {
                const float dt = g.dt;
                OC = uo[6]; OE = uo[7]; OB = uo[8]; OA = uo[9];
                OTx = uo[10]; OTz = uo[11]; OWz = uo[12]; OWx = uo[13];
                OC = OC + (a2 * Cp - a2 * OC) * dt;
                OE = OE + (a1 * Ep - a1 * OE) * dt;
                OB = OB + (a1 * Bp - a1 * OB) * dt;
                OA = OA + (a2 * Ap - a2 * OA) * dt;
                OTx = OTx + (a1 * Tx - a1 * OTx) * dt;
                OTz = OTz + (a2 * Tz - a2 * OTz) * dt;
                OWz = OWz + (a2 * Wz - a2 * OWz) * dt;
                OWx = OWx + (a1 * Wx - a1 * OWx) * dt;
                stb(O0, k, OC); stb(O1, k, OE); stb(O2, k, OB); stb(O3, k, OA);
                stb(O4, k, OTx); stb(O5, k, OTz); stb(O6, k, OWz); stb(O7, k, OWx);
            }
            const float Kx = (KX - FX) - OC + OE, Kz = (KZ - FZ) - OB + OA;
            const float hx = Wz - OWz, hz = -Wx + OWx;
            const float dt2 = g.dt2;
            const unsigned k2 = off2(cu);
            float vx = qvx[RS(0)], vz = qvz[RS(0)];
            if (z_inj) { const float2 v = ldb2(Vr, k2); vx = v.x; vz = v.y; } // state, not the injected record
            const float dx = uo[0] + dt2 * ((Tx - OTx) + hx + 2.f * Kx); vx = vx + dx;
            const float dz = uo[1] + dt2 * ((Tz - OTz) + hz + 2.f * Kz); vz = vz + dz;
            stb2(DVp, k2, make_float2(dx, dz));
            stb2(Vw, k2, make_float2(vx, vz));
            const float ex = uo[4] + dt2 * (hx + Kx); const float sx = uo[2] + ex;
            const float ez = uo[5] + dt2 * (hz + Kz); const float sz = uo[3] + ez;
            stb2(DSp, k2, make_float2(ex, ez));
            stb2(Sp, k2, make_float2(sx, sz));
            record(g, a, bid.z, cu, z, vx, vz, sx, sz);
            if (IMG && img_on && z_int && cu >= g.L && cu < g.Xn - g.L) { // corr_v, P-S-seperating.cu:1426-1439
                const float sa = uo[16 % NS], sb = uo[17 % NS];
                const float rpx = vx - sx, rpz = vz - sz;
                stb(npp, k, uo[18 % NS] + (sa * rpx + sb * rpz));
                stb(den, k, uo[19 % NS] + (sa * sa + sb * sb));
                stb(nps, k, uo[20 % NS] + (sa * sx + sb * sz));
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
template <int N, bool IMG>
__global__ __launch_bounds__(kEdgeW) __attribute__((amdgpu_waves_per_eu(2))) void k_edge(FGrid g, FWave w, KModel m,
                                                                                        FStep a, EPlan pl) {
    edge_body<N, IMG>(g, w, m, a, pl, blockIdx3());
}

} // namespace ouc
