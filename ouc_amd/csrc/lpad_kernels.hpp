// Learned absorbing boundary: kernels of the low-order pad.
//
// The pad is the outer `lpad` cells of the padded grid on every side.  There
// the scheme runs with second-order staggered differences (reach one cell), so
// only the outermost ring of velocity cells lacks the data of a full update;
// a small network supplies that ring from a 7x7 window of the local velocity
// increment and stress.  Inside the pad the stress and velocity follow the
// same equations as the interior (theta, omega and the product terms of the
// reference's forward_s / forward_u, P-S-seperating.cu:205-384) without CPML.
//
// Liveness is mirror-symmetric on the staggered grid (x shown; z alike):
//   theta-type stress (theta, gxx, gzz) at i in [1, Xn-1);
//   omega-type stress (omega, gxz, gzx) at i in [0, Xn-1);
//   vx (at i+1/2): scheme on [1, Xn-2), network at 0 and Xn-2, zero at Xn-1;
//   vz (at i):     scheme on [1, Xn-1), network at 0 and Xn-1.
// In the frame of each side (a: cells from the outer edge, b: along it) the
// network sees the same geometry on all four sides.
#pragma once

#include "fused_kernels.hpp"

namespace ouc {

constexpr int kLbIn = 5, kLbOut = 2, kLbK = 7;     // composed filter: 2 x 5 x 7 x 7
constexpr int kLbTaps = kLbOut * kLbIn * kLbK * kLbK;

// Network ring entry: the velocity cell (i, j), field (0: vx, 1: vz).
struct LbCell {
    int i, j, f, pad;
};

// 2nd-order stress on pad cells (those with edge_dist < lpad).  V is read as
// the interior's stress pass sees it (load_vin): on the receiver row of the
// receiver wavefield that is the injected record.
__device__ __forceinline__ void lpad_stress_body(const FGrid &g, const FWave &w, const KModel &m, const int2 *__restrict__ tiles,
                                                     const FStep &a, const uint3 bid) {
    const int2 t0 = tiles[bid.x];
    const int j = t0.y + threadIdx.x, i = t0.x + threadIdx.y;
    if (i >= g.Xn || j >= g.Zn || !in_lpad(g, i, j)) return;
    const long so = (long)bid.z * g.plane;
    const long k0 = (long)i * g.P + j, k = so + k0;
    const bool t_live = i >= 1 && i < g.Xn - 1 && j >= 1 && j < g.Zn - 1;
    const bool w_live = i < g.Xn - 1 && j < g.Zn - 1;
    float vx0, vz0, vxa, vza, vxb, vzb;
    load_vin(g, w, a, bid.z, so, i, j, vx0, vz0);
    float gxx = 0.f, gzz = 0.f, gxz = 0.f, gzx = 0.f;
    if (t_live) {
        load_vin(g, w, a, bid.z, so, i - 1, j, vxa, vza);
        load_vin(g, w, a, bid.z, so, i, j - 1, vxb, vzb);
        gxx = (vx0 - vxa) * g.idx;
        gzz = (vz0 - vzb) * g.idz;
    }
    if (w_live) {
        load_vin(g, w, a, bid.z, so, i, j + 1, vxa, vza);
        load_vin(g, w, a, bid.z, so, i + 1, j, vxb, vzb);
        gxz = (vxa - vx0) * g.idz;
        gzx = (vzb - vz0) * g.idx;
    }
    w.th[k] = m.p[k0] * (gxx + gzz);
    w.om[k] = m.q[k0] * (gxz - gzx);
    w.gxx[k] = gxx; w.gzz[k] = gzz; w.gxz[k] = gxz; w.gzx[k] = gzx;
}
__global__ __launch_bounds__(256) void k_lpad_stress(FGrid g, FWave w, KModel m, const int2 *__restrict__ tiles,
                                                     FStep a) {
    lpad_stress_body(g, w, m, tiles, a, blockIdx3());
}

// Increment damping of the pad, 1 - beta ((lpad - d) / lpad)^2 at depth d: a
// weak sponge that removes the slow growth a one-step-trained ring can show
// over thousands of steps.  1 outside the pad.
__device__ __forceinline__ float lb_gamma(const FGrid &g, int d) {
    if (d >= g.lpad) return 1.f;
    const float r = (float)(g.lpad - d) / (float)g.lpad;
    return 1.f - g.lb_beta * r * r;
}

// 2nd-order velocity update on the live pad cells (the network ring and the
// ghost line are left to k_lb_ring / stay zero).  The P/S split is not carried
// in the pad: nothing reads S there.
__device__ __forceinline__ void lpad_velocity_body(const FGrid &g, const FWave &w, const KModel &m, const int2 *__restrict__ tiles,
                                                       int rb, const uint3 bid) {
    const int2 t0 = tiles[bid.x];
    const int j = t0.y + threadIdx.x, i = t0.x + threadIdx.y;
    if (i >= g.Xn || j >= g.Zn || !in_lpad(g, i, j)) return;
    const long so = (long)bid.z * g.plane;
    const long k0 = (long)i * g.P + j, k = so + k0;
    const long P = g.P;
    const float *q = m.q;
    const int wb = 1 - rb;
    const float dt2 = g.dt2;
    const float gam = lb_gamma(g, edge_dist(g, i, j));
    if (i >= 1 && i < g.Xn - 2 && j >= 1 && j < g.Zn - 1) { // vx
        const float Tx = (w.th[k + P] - w.th[k]) * g.idx;
        const float Wz = (w.om[k] - w.om[k - 1]) * g.idz;
        const float C = (q[k0] * w.gzx[k] - q[k0 - 1] * w.gzx[k - 1]) * g.idz;
        const float E = (q[k0 + P] * w.gzz[k + P] - q[k0] * w.gzz[k]) * g.idx;
        const float d = gam * (w.DV[0][k] + dt2 * (Tx + Wz + 2.f * (C - E)));
        w.DV[0][k] = d;
        w.V[wb][0][k] = w.V[rb][0][k] + d;
    }
    if (i >= 1 && i < g.Xn - 1 && j >= 1 && j < g.Zn - 2) { // vz
        const float Tz = (w.th[k + 1] - w.th[k]) * g.idz;
        const float Wx = (w.om[k] - w.om[k - P]) * g.idx;
        const float B = (q[k0] * w.gxz[k] - q[k0 - P] * w.gxz[k - P]) * g.idx;
        const float A = (q[k0 + 1] * w.gxx[k + 1] - q[k0] * w.gxx[k]) * g.idz;
        const float d = gam * (w.DV[1][k] + dt2 * (Tz - Wx + 2.f * (B - A)));
        w.DV[1][k] = d;
        w.V[wb][1][k] = w.V[rb][1][k] + d;
    }
}
__global__ __launch_bounds__(256) void k_lpad_velocity(FGrid g, FWave w, KModel m, const int2 *__restrict__ tiles,
                                                       int rb) {
    lpad_velocity_body(g, w, m, tiles, rb, blockIdx3());
}

// Input channel `ch` of the network (0: D_x, 1: D_z, 2: txx, 3: tzz, 4: txz,
// stresses scaled by dt^2/dx) at original cell (i, j); zero outside the grid.
// txx = theta - 2 q gzz, tzz = theta - 2 q gxx, txz = q (gxz + gzx).
__device__ __forceinline__ float lb_input(const FGrid &g, const FWave &w, const float *__restrict__ q, long so, int ch,
                                          int i, int j, float sc) {
    if (i < 0 || i >= g.Xn || j < 0 || j >= g.Zn) return 0.f;
    const long k0 = (long)i * g.P + j, k = so + k0;
    switch (ch) {
    case 0: return w.DV[0][k];
    case 1: return w.DV[1][k];
    case 2: return sc * (w.th[k] - 2.f * q[k0] * w.gzz[k]);
    case 3: return sc * (w.th[k] - 2.f * q[k0] * w.gxx[k]);
    default: return sc * q[k0] * (w.gxz[k] + w.gzx[k]);
    }
}

// Field staggering (x, z) and mirror signs of the 5 channels, and the channel
// each one becomes under the x <-> z transpose.
__device__ __constant__ const int kLbSX[5] = {1, 0, 0, 0, 1};
__device__ __constant__ const int kLbSZ[5] = {0, 1, 0, 0, 1};
__device__ __constant__ const float kLbXS[5] = {-1.f, 1.f, 1.f, 1.f, -1.f};
__device__ __constant__ const float kLbZS[5] = {1.f, -1.f, 1.f, 1.f, -1.f};
__device__ __constant__ const int kLbT[5] = {1, 0, 3, 2, 4};

// Canonical -> original cell of canonical channel cc at (a, b) on `side`
// (0: x-low, 1: x-high, 2: z-low, 3: z-high); returns the original channel and
// the sign.
__device__ __forceinline__ int lb_map(const FGrid &g, int side, int cc, int a, int b, int &i, int &j, float &sg) {
    if (side < 2) {
        i = side == 0 ? a : g.Xn - 1 - kLbSX[cc] - a;
        j = b;
        sg = side == 0 ? 1.f : kLbXS[cc];
        return cc;
    }
    const int oc = kLbT[cc];
    j = side == 2 ? a : g.Zn - 1 - kLbSZ[oc] - a;
    i = b;
    sg = side == 2 ? 1.f : kLbZS[oc];
    return oc;
}

// The network's input, staged per side in canonical coordinates: the five
// canonical channels (channel map and mirror sign of lb_map applied) at depths
// a = 0..kLbD-1 along the whole side, buf[shot][side][channel][a][position],
// E = max(Xn, Zn) positions per row.  Every ring cell has canonical depth 0
// (build_learned checks it), so its 7x7 window reads depths 0..3 only; each
// staged value is read by up to 28 windows instead of being formed for each.
constexpr int kLbD = 4;
__host__ __device__ inline long lb_canon_floats(int E) { return 4L * kLbIn * kLbD * E; } // per shot
__device__ __forceinline__ int lb_side_len(const FGrid &g, int side) { return side < 2 ? g.Zn : g.Xn; }

// One staged position per thread: bid.x * blockDim.x + lane along the side,
// bid.y = side * kLbD + depth, bid.z = shot.
__device__ __forceinline__ void lb_canon_body(const FGrid &g, const FWave &w, const KModel &m, float sc,
                                              float *__restrict__ buf, int E, const uint3 bid) {
    const int bb = bid.x * blockDim.x + threadIdx.x;
    if (bb >= E) return;
    const int side = bid.y / kLbD, aa = bid.y % kLbD;
    const long so = (long)bid.z * g.plane;
    float *o = buf + (long)bid.z * lb_canon_floats(E) + ((long)side * kLbIn * kLbD + aa) * E + bb;
    const bool in = bb < lb_side_len(g, side);
#pragma unroll
    for (int cc = 0; cc < kLbIn; ++cc) {
        float v = 0.f;
        if (in) {
            int i, j;
            float sg;
            const int ch = lb_map(g, side, cc, aa, bb, i, j, sg);
            v = sg * lb_input(g, w, m.q, so, ch, i, j, sc);
        }
        o[(long)cc * kLbD * E] = v;
    }
}
__global__ __launch_bounds__(256) void k_lb_canon(FGrid g, FWave w, KModel m, float sc, float *__restrict__ buf, int E) {
    lb_canon_body(g, w, m, sc, buf, E, blockIdx3());
}

// Network on the outer ring: for each ring cell, the side with the smallest
// canonical depth (ties: x sides first, low before high), the 7x7x5 canonical
// window (from the staged rows), the composed filter (staged in LDS), and the
// correction added to the increment: D += delta, V_new = V + D.  Every window
// reads the increments before any ring cell advances them (a window holds
// other ring cells), so the new increments go to dnew[shot][ncells] and
// k_lb_apply stores them.  With wts == NULL the ring is held at zero here (the
// pad without the network).
__device__ __forceinline__ void lb_ring_body(const FGrid &g, const FWave &w, const LbCell *__restrict__ cells,
                                             int ncells, const float *__restrict__ wts, int rb,
                                             const float *__restrict__ canon, int E, float *__restrict__ dnew,
                                             const uint3 bid) {
    __shared__ float wl[kLbTaps];
    if (wts) { // uniform: every lane of the block takes part
        for (int t = threadIdx.x; t < kLbTaps; t += blockDim.x) wl[t] = wts[t];
        __syncthreads();
    }
    const int e = bid.x * blockDim.x + threadIdx.x;
    if (e >= ncells) return;
    const LbCell cl = cells[e];
    const long so = (long)bid.z * g.plane;
    const long k = so + (long)cl.i * g.P + cl.j;
    const int wb = 1 - rb;
    if (!wts) {
        w.DV[cl.f][k] = 0.f;
        w.V[wb][cl.f][k] = 0.f;
        return;
    }
    // canonical depth of this cell on each side
    const int f = cl.f; // original channel 0 (vx) or 1 (vz)
    int da[4];
    da[0] = cl.i;
    da[1] = g.Xn - 1 - kLbSX[f] - cl.i;
    da[2] = cl.j;
    da[3] = g.Zn - 1 - kLbSZ[f] - cl.j;
    int side = 0;
    for (int s2 = 1; s2 < 4; ++s2)
        if (da[s2] < da[side]) side = s2;
    const int a = da[side];
    const int b = side < 2 ? cl.j : cl.i;
    const int oc = side < 2 ? f : kLbT[f]; // canonical output channel of this field
    const int len = lb_side_len(g, side);
    const float *wo = wl + oc * kLbIn * kLbK * kLbK;
    const float *cb = canon + (long)bid.z * lb_canon_floats(E) + (long)side * kLbIn * kLbD * E;
    // taps in the order of the filter's layout; a staged value already carries
    // the mirror sign, and w * (s x) equals (w s) x exactly for s = +-1
    float acc = 0.f;
#pragma unroll
    for (int cc = 0; cc < kLbIn; ++cc) {
#pragma unroll
        for (int dy = 0; dy < kLbK; ++dy) {
            const int aa = a + dy - 3;
            if (aa < 0) continue;
            const float *row = cb + ((long)cc * kLbD + aa) * E;
#pragma unroll
            for (int dx = 0; dx < kLbK; ++dx) {
                const int bb = b + dx - 3;
                const float x = (bb >= 0 && bb < len) ? row[bb] : 0.f;
                acc += wo[(cc * kLbK + dy) * kLbK + dx] * x;
            }
        }
    }
    // back to the original field: canonical output oc maps to field f with sign
    float sg_out = 1.f;
    if (side == 1) sg_out = kLbXS[f];
    if (side == 3) sg_out = kLbZS[f];
    dnew[(long)bid.z * ncells + e] = lb_gamma(g, 0) * (w.DV[f][k] + sg_out * acc);
}
__global__ __launch_bounds__(256) void k_lb_ring(FGrid g, FWave w, const LbCell *__restrict__ cells, int ncells,
                                                 const float *__restrict__ wts, int rb, const float *__restrict__ canon,
                                                 int E, float *__restrict__ dnew) {
    lb_ring_body(g, w, cells, ncells, wts, rb, canon, E, dnew, blockIdx3());
}

// The ring's new increments from k_lb_ring: D = dnew, V_new = V + D.
__device__ __forceinline__ void lb_apply_body(const FGrid &g, const FWave &w, const LbCell *__restrict__ cells, int ncells,
                                                  const float *__restrict__ dnew, int rb, const uint3 bid) {
    const int e = bid.x * blockDim.x + threadIdx.x;
    if (e >= ncells) return;
    const LbCell cl = cells[e];
    const long k = (long)bid.z * g.plane + (long)cl.i * g.P + cl.j;
    const float d = dnew[(long)bid.z * ncells + e];
    w.DV[cl.f][k] = d;
    w.V[1 - rb][cl.f][k] = w.V[rb][cl.f][k] + d;
}
__global__ __launch_bounds__(256) void k_lb_apply(FGrid g, FWave w, const LbCell *__restrict__ cells, int ncells,
                                                  const float *__restrict__ dnew, int rb) {
    lb_apply_body(g, w, cells, ncells, dnew, rb, blockIdx3());
}

// Largest |V| of each shot over the pad (mon[2 shot]) and over the band of the
// same width just inside it (mon[2 shot + 1]), as running maxima.  Float bits
// compare as integers: |x| has a clear sign bit and NaN sorts above +inf.
__device__ __forceinline__ void lb_monitor_body(const FGrid &g, const FWave &w, const int2 *__restrict__ tiles, int rb,
                                                    unsigned int *__restrict__ mon, const uint3 bid) {
    const int2 t0 = tiles[bid.x];
    const int j = t0.y + threadIdx.x, i = t0.x + threadIdx.y;
    unsigned int up = 0u, uf = 0u;
    if (i < g.Xn && j < g.Zn) {
        const int d = edge_dist(g, i, j);
        if (d < 2 * g.lpad) {
            const long k = (long)bid.z * g.plane + (long)i * g.P + j;
            const unsigned int u = max(__float_as_uint(fabsf(w.V[rb][0][k])), __float_as_uint(fabsf(w.V[rb][1][k])));
            if (d < g.lpad) up = u;
            else uf = u;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        up = max(up, (unsigned int)__shfl_xor((int)up, o));
        uf = max(uf, (unsigned int)__shfl_xor((int)uf, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(mon + 2 * bid.z, up);
        atomicMax(mon + 2 * bid.z + 1, uf);
    }
}
__global__ __launch_bounds__(256) void k_lb_monitor(FGrid g, FWave w, const int2 *__restrict__ tiles, int rb,
                                                    unsigned int *__restrict__ mon) {
    lb_monitor_body(g, w, tiles, rb, mon, blockIdx3());
}

} // namespace ouc
