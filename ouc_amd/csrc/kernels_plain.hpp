// Plain-engine HIP kernels: the reference's kernel split (stress pass, velocity
// pass, strip copy, reconstruction pair, imaging), one kernel per pass, lanes
// along z (the contiguous axis) so every wave reads whole 256-byte rows.
//
// Arithmetic form.  The reference steps V^{n+1} = 2V^n - V^{n-1} + dt^2 RHS
// separately for the P and S parts (P-S-seperating.cu:291-294).  In FP32 that
// form loses the S part to rounding: the S drive is the curl of a P-dominated
// field and the difference (C - E) of two P-sized terms that cancel wherever
// vs is constant.  The kernels therefore step an algebraically identical state:
//   Vp = D+ Phi + Qp  (Phi'' = theta, Qp'' = -psi_T + K)
//   Vs                (Vs''  = curl drive + K)
// each as (value, increment) pairs, V^{n+1} = V^n + D^{n+1/2}.  The curl of
// D+ Phi vanishes identically, so omega = vs^2 curl(Qp + Vs); and
// K = D(vs^2 g) - D'(vs^2 g') is evaluated as D((vs^2 - vs^2_0) g) - ...,
// which equals the reference's C - E / B - A in exact arithmetic because the
// staggered operators along x and z commute.  The CPML psi terms are the
// reference's (P-S-seperating.cu:236-289, 369-372).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ouc {

struct KGrid {
    int Xn, Zn, P, L, N, nx, nz;
    float idx, idz, dt, dt2;
    float c[8];
};

// One wavefield's device arrays (pitched Xn x P).
struct KWave {
    float *Phi, *DPhi, *Qx, *Qz, *DQx, *DQz, *Sx, *Sz, *DSx, *DSz; // state
    float *Vx, *Vz, *Ux, *Uz;                                      // assembled
    float *th, *om, *gxx, *gxz, *gzx, *gzz;                        // stress-like
    float *F[4];                                                   // F_xx, F_zz, F_xz, F_zx
    float *O[8];                                                   // O_C, O_E, O_B, O_A, O_Tx, O_Tz, O_Wz, O_Wx
};

struct KModel {
    const float *p, *q, *ax, *az; // vp^2, vs^2, damping
};

// ---- staggered operators (P-S-seperating.cu:224-279, 340-365), unscaled ----
template <int N>
__device__ __forceinline__ float dmx(const float *__restrict__ f, long k, int P, const float *c) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= N; ++m) s += c[m - 1] * (f[k + (long)(m - 1) * P] - f[k - (long)m * P]);
    return s;
}
template <int N>
__device__ __forceinline__ float dpx(const float *__restrict__ f, long k, int P, const float *c) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= N; ++m) s += c[m - 1] * (f[k + (long)m * P] - f[k - (long)(m - 1) * P]);
    return s;
}
template <int N>
__device__ __forceinline__ float dmz(const float *__restrict__ f, long k, const float *c) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= N; ++m) s += c[m - 1] * (f[k + (m - 1)] - f[k - m]);
    return s;
}
template <int N>
__device__ __forceinline__ float dpz(const float *__restrict__ f, long k, const float *c) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= N; ++m) s += c[m - 1] * (f[k + m] - f[k - (m - 1)]);
    return s;
}
// D(f * (q - q0)): the product-rule form of the K terms.
template <int N>
__device__ __forceinline__ float dmxq(const float *__restrict__ f, const float *__restrict__ q, long k, int P,
                                      const float *c, float q0) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= N; ++m) {
        long a = k + (long)(m - 1) * P, b = k - (long)m * P;
        s += c[m - 1] * (f[a] * (q[a] - q0) - f[b] * (q[b] - q0));
    }
    return s;
}
template <int N>
__device__ __forceinline__ float dpxq(const float *__restrict__ f, const float *__restrict__ q, long k, int P,
                                      const float *c, float q0) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= N; ++m) {
        long a = k + (long)m * P, b = k - (long)(m - 1) * P;
        s += c[m - 1] * (f[a] * (q[a] - q0) - f[b] * (q[b] - q0));
    }
    return s;
}
template <int N>
__device__ __forceinline__ float dmzq(const float *__restrict__ f, const float *__restrict__ q, long k,
                                      const float *c, float q0) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= N; ++m) {
        long a = k + (m - 1), b = k - m;
        s += c[m - 1] * (f[a] * (q[a] - q0) - f[b] * (q[b] - q0));
    }
    return s;
}
template <int N>
__device__ __forceinline__ float dpzq(const float *__restrict__ f, const float *__restrict__ q, long k,
                                      const float *c, float q0) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= N; ++m) {
        long a = k + m, b = k - (m - 1);
        s += c[m - 1] * (f[a] * (q[a] - q0) - f[b] * (q[b] - q0));
    }
    return s;
}

// Cell of the block's 64 x 4 tile.
#define OUC_CELL(x0, z0)                                   \
    const int j = (z0) + blockIdx.x * 64 + threadIdx.x;    \
    const int i = (x0) + blockIdx.y * 4 + threadIdx.y;

// Totals and the non-gradient part on the update region [N, Xn-N)^2:
// Vx = D+x Phi + Qx + Sx, Ux = Qx + Sx (and z).  Outside it they stay zero, the
// reference's never-updated frame (P-S-seperating.cu:222).
template <int N>
static __global__ __launch_bounds__(256) void k_assemble(KGrid g, KWave w) {
    OUC_CELL(N, N)
    if (i >= g.Xn - N || j >= g.Zn - N) return;
    const long k = (long)i * g.P + j;
    float ux = w.Qx[k] + w.Sx[k], uz = w.Qz[k] + w.Sz[k];
    w.Ux[k] = ux;
    w.Uz[k] = uz;
    w.Vx[k] = dpx<N>(w.Phi, k, g.P, g.c) * g.idx + ux;
    w.Vz[k] = dpz<N>(w.Phi, k, g.c) * g.idz + uz;
}

// load_record (P-S-seperating.cu:1471-1483): the receiver row's totals are
// replaced by the records of step t; dX/dZ keep (record - own value) so the
// stress pass can add the curl of the replacement.
static __global__ void k_inject(KGrid g, KWave w, int zr, const float *__restrict__ recx, const float *__restrict__ recz,
                         float *__restrict__ dX, float *__restrict__ dZ) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= g.nx) return;
    int i = ix + g.L;
    long k = (long)i * g.P + zr;
    float rx = recx[ix], rz = recz[ix];
    dX[i] = rx - w.Vx[k];
    dZ[i] = rz - w.Vz[k];
    w.Vx[k] = rx;
    w.Vz[k] = rz;
}

// Stress pass with CPML, forward_s / rt_s_res (P-S-seperating.cu:316-384,
// 667-726).  g_* are stored raw; the reference's du fields are vs^2 (g - F).
template <int N>
static __global__ __launch_bounds__(256) void k_stress(KGrid g, KWave w, KModel m, int si, int sj, float s, int zr,
                                               const float *__restrict__ dX, const float *__restrict__ dZ) {
    OUC_CELL(N, N)
    if (i >= g.Xn - N || j >= g.Zn - N) return;
    const long k = (long)i * g.P + j;
    const float gxx = dmx<N>(w.Vx, k, g.P, g.c) * g.idx;
    const float gxz = dpz<N>(w.Vx, k, g.c) * g.idz;
    const float gzz = dmz<N>(w.Vz, k, g.c) * g.idz;
    const float gzx = dpx<N>(w.Vz, k, g.P, g.c) * g.idx;
    const float a1 = m.ax[k], a2 = m.az[k];
    float Fxx = 0.f, Fzz = 0.f, Fxz = 0.f, Fzx = 0.f;
    if (a1 != 0.f || a2 != 0.f) {
        Fxx = w.F[0][k]; Fzz = w.F[1][k]; Fxz = w.F[2][k]; Fzx = w.F[3][k];
        Fxx = Fxx + (a1 * gxx - a1 * Fxx) * g.dt;
        Fzz = Fzz + (a2 * gzz - a2 * Fzz) * g.dt;
        Fxz = Fxz + (a2 * gxz - a2 * Fxz) * g.dt;
        Fzx = Fzx + (a1 * gzx - a1 * Fzx) * g.dt;
        w.F[0][k] = Fxx; w.F[1][k] = Fzz; w.F[2][k] = Fxz; w.F[3][k] = Fzx;
    }
    float cu = dpz<N>(w.Ux, k, g.c) * g.idz - dpx<N>(w.Uz, k, g.P, g.c) * g.idx;
    if (zr >= 0 && j >= zr - N && j <= zr + N) {
        // curl of the replacement d, which lives on row zr only
        float sz_ = 0.f;
#pragma unroll
        for (int mm = 1; mm <= N; ++mm) {
            if (j + mm == zr) sz_ += g.c[mm - 1] * dX[i];
            if (j - (mm - 1) == zr) sz_ -= g.c[mm - 1] * dX[i];
        }
        float sx_ = 0.f;
        if (j == zr) {
#pragma unroll
            for (int mm = 1; mm <= N; ++mm) sx_ += g.c[mm - 1] * (dZ[i + mm] - dZ[i - (mm - 1)]);
        }
        cu += sz_ * g.idz - sx_ * g.idx;
    }
    float th = m.p[k] * (gxx + gzz - Fxx - Fzz);
    if (i == si && j == sj) th = th + s;
    w.th[k] = th;
    w.om[k] = m.q[k] * (cu - Fxz + Fzx);
    w.gxx[k] = gxx; w.gxz[k] = gxz; w.gzx[k] = gzx; w.gzz[k] = gzz;
}

// Velocity pass with CPML, forward_u / rt_u_res (P-S-seperating.cu:205-314,
// 562-665), in the (value, increment) form.
template <int N>
static __global__ __launch_bounds__(256) void k_velocity(KGrid g, KWave w, KModel m) {
    OUC_CELL(N, N)
    if (i >= g.Xn - N || j >= g.Zn - N) return;
    const long k = (long)i * g.P + j;
    const float *c = g.c;
    const int P = g.P;
    const float Wx = dmx<N>(w.om, k, P, c) * g.idx, Wz = dmz<N>(w.om, k, c) * g.idz;
    const float q0 = m.q[k];
    float KX = dmzq<N>(w.gzx, m.q, k, c, q0) * g.idz - dpxq<N>(w.gzz, m.q, k, P, c, q0) * g.idx;
    float KZ = dmxq<N>(w.gxz, m.q, k, P, c, q0) * g.idx - dpzq<N>(w.gxx, m.q, k, c, q0) * g.idz;
    const float a1 = m.ax[k], a2 = m.az[k];
    // psi parts of C - E and B - A: -(D-z(q Fzx) - D+x(q Fzz)), -(D-x(q Fxz) - D+z(q Fxx))
    KX -= dmzq<N>(w.F[3], m.q, k, c, 0.f) * g.idz - dpxq<N>(w.F[1], m.q, k, P, c, 0.f) * g.idx;
    KZ -= dmxq<N>(w.F[2], m.q, k, P, c, 0.f) * g.idx - dpzq<N>(w.F[0], m.q, k, c, 0.f) * g.idz;
    float OTx = 0.f, OTz = 0.f, OWz = 0.f, OWx = 0.f, OC = 0.f, OE = 0.f, OB = 0.f, OA = 0.f;
    if (a1 != 0.f || a2 != 0.f) {
        const float C = (dmzq<N>(w.gzx, m.q, k, c, 0.f) - dmzq<N>(w.F[3], m.q, k, c, 0.f)) * g.idz;
        const float E = (dpxq<N>(w.gzz, m.q, k, P, c, 0.f) - dpxq<N>(w.F[1], m.q, k, P, c, 0.f)) * g.idx;
        const float B = (dmxq<N>(w.gxz, m.q, k, P, c, 0.f) - dmxq<N>(w.F[2], m.q, k, P, c, 0.f)) * g.idx;
        const float A = (dpzq<N>(w.gxx, m.q, k, c, 0.f) - dpzq<N>(w.F[0], m.q, k, c, 0.f)) * g.idz;
        const float Tx = dpx<N>(w.th, k, P, c) * g.idx, Tz = dpz<N>(w.th, k, c) * g.idz;
        const float dt = g.dt;
        OC = w.O[0][k]; OE = w.O[1][k]; OB = w.O[2][k]; OA = w.O[3][k];
        OTx = w.O[4][k]; OTz = w.O[5][k]; OWz = w.O[6][k]; OWx = w.O[7][k];
        OC = OC + (a2 * C - a2 * OC) * dt;
        OE = OE + (a1 * E - a1 * OE) * dt;
        OB = OB + (a1 * B - a1 * OB) * dt;
        OA = OA + (a2 * A - a2 * OA) * dt;
        OTx = OTx + (a1 * Tx - a1 * OTx) * dt;
        OTz = OTz + (a2 * Tz - a2 * OTz) * dt;
        OWz = OWz + (a2 * Wz - a2 * OWz) * dt;
        OWx = OWx + (a1 * Wx - a1 * OWx) * dt;
        w.O[0][k] = OC; w.O[1][k] = OE; w.O[2][k] = OB; w.O[3][k] = OA;
        w.O[4][k] = OTx; w.O[5][k] = OTz; w.O[6][k] = OWz; w.O[7][k] = OWx;
    }
    const float Kx = KX - OC + OE, Kz = KZ - OB + OA;
    const float dt2 = g.dt2;
    float d;
    d = w.DPhi[k] + dt2 * w.th[k]; w.DPhi[k] = d; w.Phi[k] = w.Phi[k] + d;
    d = w.DQx[k] + dt2 * (Kx - OTx); w.DQx[k] = d; w.Qx[k] = w.Qx[k] + d;
    d = w.DQz[k] + dt2 * (Kz - OTz); w.DQz[k] = d; w.Qz[k] = w.Qz[k] + d;
    d = w.DSx[k] + dt2 * (Wz - OWz + Kx); w.DSx[k] = d; w.Sx[k] = w.Sx[k] + d;
    d = w.DSz[k] + dt2 * (-Wx + OWx + Kz); w.DSz[k] = d; w.Sz[k] = w.Sz[k] + d;
}

// Receiver-row recording after the velocity update (P-S-seperating.cu:307-311,
// with the post-update value).  out[c][t][ix] for c = vx, vz, vpx, vpz, vsx, vsz
// (ncomp = 2 keeps only the totals).
template <int N>
static __global__ void k_record(KGrid g, KWave w, int zr, int t, int nt, int ncomp, float *__restrict__ out) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= g.nx) return;
    long k = (long)(ix + g.L) * g.P + zr;
    float vpx = dpx<N>(w.Phi, k, g.P, g.c) * g.idx + w.Qx[k];
    float vpz = dpz<N>(w.Phi, k, g.c) * g.idz + w.Qz[k];
    float vsx = w.Sx[k], vsz = w.Sz[k];
    size_t plane = (size_t)nt * g.nx, o = (size_t)t * g.nx + ix;
    out[o] = vpx + vsx;
    out[plane + o] = vpz + vsz;
    if (ncomp == 6) {
        out[2 * plane + o] = vpx;
        out[3 * plane + o] = vpz;
        out[4 * plane + o] = vsx;
        out[5 * plane + o] = vsz;
    }
}

// remove() (P-S-seperating.cu:1634-1657) with per-trace counts from the host.
static __global__ void k_mute(int nx, int nt, int ncomp, const int32_t *__restrict__ counts, float *__restrict__ rec) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x;
    int t = blockIdx.y;
    if (ix >= nx || t >= nt) return;
    if (t < counts[ix]) {
        size_t plane = (size_t)nt * nx;
        for (int c = 0; c < ncomp; ++c) rec[c * plane + (size_t)t * nx + ix] = 0.f;
    }
}

// ---- strips ----------------------------------------------------------------
// The ring: four strips 2N cells wide straddling the interior's edge (top and
// bottom (nx+2N) x 2N, left and right 2N x (nz-2N)).  Each strip splits into
// an outer part, its N-1 outermost pad rows (columns), and an inner part, the
// N interior rows next to the edge plus the first pad row.  Storage order: the
// outer parts of top, bottom, left, right, then their inner parts; within a
// part z runs fastest (one column of a top/bottom part, one z line of a
// left/right part at a time).
struct KRing {
    int L, N, Xn, Zn, nx, nz;
    long ntop, nside, ncells, nouter;
};

__host__ __device__ __forceinline__ void ring_cell(const KRing &r, long q, int &i, int &j) {
    const int N = r.N, L = r.L, w = r.nz - 2 * N;
    const long cw = r.nx + 2 * N;
    if (q < r.nouter) {
        const int wo = N - 1;
        if (q < cw * wo) { i = L - N + (int)(q / wo); j = L - N + (int)(q % wo); return; }
        q -= cw * wo;
        if (q < cw * wo) { i = L - N + (int)(q / wo); j = r.Zn - L + 1 + (int)(q % wo); return; }
        q -= cw * wo;
        if (q < (long)wo * w) { i = L - N + (int)(q / w); j = L + N + (int)(q % w); return; }
        q -= (long)wo * w;
        i = r.Xn - L + 1 + (int)(q / w);
        j = L + N + (int)(q % w);
        return;
    }
    q -= r.nouter;
    const int wi = N + 1;
    if (q < cw * wi) { i = L - N + (int)(q / wi); j = L - 1 + (int)(q % wi); return; }
    q -= cw * wi;
    if (q < cw * wi) { i = L - N + (int)(q / wi); j = r.Zn - L - N + (int)(q % wi); return; }
    q -= cw * wi;
    if (q < (long)wi * w) { i = L - 1 + (int)(q / w); j = L + N + (int)(q % w); return; }
    q -= (long)wi * w;
    i = r.Xn - L - N + (int)(q / w);
    j = L + N + (int)(q % w);
}

// Storage index of padded cell (i, j) in the ring, or -1 off the ring.
__host__ __device__ __forceinline__ long ring_index(const KRing &r, int i, int j) {
    const int L = r.L, N = r.N;
    if (i < L - N || i >= r.Xn - L + N || j < L - N || j >= r.Zn - L + N) return -1;
    if (i >= L + N && i < r.Xn - L - N && j >= L + N && j < r.Zn - L - N) return -1;
    const int w = r.nz - 2 * N, wo = N - 1, wi = N + 1;
    const long cw = r.nx + 2 * N;
    if (j < L + N) {
        const long c = i - (L - N);
        const int m = j - (L - N);
        return m < wo ? c * wo + m : r.nouter + c * wi + (m - wo);
    }
    if (j >= r.Zn - L - N) {
        const long c = i - (L - N);
        const int m = j - (r.Zn - L - N);
        return m <= N ? r.nouter + cw * wi + c * wi + m : cw * wo + c * wo + (m - wi);
    }
    const long row = j - (L + N);
    if (i < L + N) {
        const int m = i - (L - N);
        return m < wo ? 2 * cw * wo + (long)m * w + row : r.nouter + 2 * cw * wi + (long)(m - wo) * w + row;
    }
    const int m = i - (r.Xn - L - N);
    return m <= N ? r.nouter + 2 * cw * wi + (long)wi * w + (long)m * w + row
                  : 2 * cw * wo + (long)wo * w + (long)(m - wi) * w + row;
}

// Strip slot t holds, in this field order: vx, vz (V^{t+1}, the reference's
// save after forward_u, P-S-seperating.cu:2264-2272), then theta, omega and the
// four gradients g' = g - F (the reference's du / vs^2) of step t.
enum { RS_VX = 0, RS_VZ, RS_TH, RS_OM, RS_GZX, RS_GZZ, RS_GXX, RS_GXZ, RS_NF };

// FP32 and scaled-FP16 strip stores.
//
// FP16 strips are predictive.  Slot s of a field stores, per ring cell, the
// residual x = (v^s - p) + r of the value against p, the previous slot's
// decoded value, with r the previous slot's quantization error fed back
// (error feedback).  Each 64-value chunk of a slot shares a power-of-two scale
// 2^e with max|x| * 2^-e in [2^14, 2^15); the residual is stored as binary16
// of x * 2^-e, so d = decode(x) is x rounded to 11 bits and x - d is exact.
// The encoder then holds p' = p + d and r' = x - d: its decoded sequence is
// v^s + r^{s-1} - r^s, an error that is a first time difference (small at the
// wavefield's frequencies) of a quantity bounded by half a unit of the
// residual, itself about the one-step change of v.  Every kStripKey-th slot,
// and the field's last slot, also keeps p' in FP32 as a keyframe.  The
// reconstruction decodes backwards one slot per step, v^s = keyframe or
// v^{s+1} - d^{s+1}, into a two-slot buffer; the FP32 rounding of that
// subtraction is bounded by the keyframe spacing.  Slots are stored in time
// order, with the encoder state zero before the first.
//
// The FP16 store keeps theta and the four gradients exactly (FP32) on the
// inner part of the ring.  The reconstruction integrates the split source
// velocities on the ring's interior rows open loop from these fields (their
// stress pass never sees the split parts), so the rounding of those values
// accumulates in the imaged source field; residual coding there raises the
// image error on the rows next to the edge about fifteenfold over FP32 strips,
// while FP32 storage of the inner part brings it back to the FP32 strips'
// level.  Omega drives only the S part, which the imaging condition does not
// read from the source side, and the velocities feed closed-loop stress, so
// those three fields stay residual-coded on the whole ring.
constexpr int kStripKey = 32;
constexpr unsigned kInner32 = (1u << RS_TH) | (1u << RS_GZX) | (1u << RS_GZZ) | (1u << RS_GXX) | (1u << RS_GXZ);
constexpr int kNInner32 = 5;
__host__ __device__ constexpr bool inner32_field(int f) { return (kInner32 >> f) & 1u; }
__host__ __device__ constexpr int inner32_rank(int f) { return __builtin_popcount(kInner32 & ((1u << f) - 1u)); }
// FP16 elements before field f within one slot of the FP16 store: fields with
// an FP32 inner part code only the ring's outer part.
__host__ __device__ constexpr long strip16_offset(long ncells, long nouter, int f) {
    return ncells * __builtin_popcount(~kInner32 & ((1u << f) - 1u)) + nouter * __builtin_popcount(kInner32 & ((1u << f) - 1u));
}

struct KStrips {
    float *f32;          // [slot][RS_NF][ncells] (FP32 strips)
    __half *f16;         // [slot][strip16_offset(.., RS_NF)] residuals (FP16 strips), field f at strip16_offset(f)
    float *in32;         // [slot][kNInner32][ncells - nouter] inner ring part of the FP32-kept fields (FP16 strips)
    int8_t *e8;          // [slot][RS_NF][nchunks] chunk exponents
    float *key;          // [(slot - key0) / kStripKey][RS_NF][ncells] keyframes
    float *err;          // [2][RS_NF][ncells] encoder state: previous decoded value, fed-back error
    float *dec;          // [2][RS_NF][ncells] decoded slots, by slot parity
    long ncells, nchunks, nouter;
    int fp16;
    int last_v, last_th; // last slot stored for the two velocity fields / the six stress-like fields
    const int *t0;       // batches: [shot] first step of each shot, or NULL (every shot starts at 0)
    int key0;            // this shot's first slot: keyframes are every kStripKey-th slot from it
};

// Element strides between the strip stores of consecutive shots of a batch.
struct KStripStride {
    long strip;  // f32 / f16 elements
    long chunks; // exponent entries
    long key;    // keyframe elements
    long aux;    // encoder-state and decoded-slot elements
    long in32;   // inner FP32 elements
};

__device__ __forceinline__ KStrips strips_of_shot(KStrips s, const KStripStride &d, long z) {
    s.key0 = s.t0 ? s.t0[z] : 0;
    if (s.f32) s.f32 += z * d.strip;
    if (s.f16) s.f16 += z * d.strip;
    if (s.in32) s.in32 += z * d.in32;
    if (s.e8) s.e8 += z * d.chunks;
    if (s.key) s.key += z * d.key;
    if (s.err) s.err += z * d.aux;
    if (s.dec) s.dec += z * d.aux;
    return s;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int scale_exponent(float mx) {
    if (!(mx > 0.f)) return 0;
    int e;
    frexpf(mx, &e);
    e -= 15;
    return e < -127 ? -127 : (e > 127 ? 127 : e);
}

// Keyframes fall on every kStripKey-th slot counted from the shot's own first
// slot, so a shot that starts late in a batch codes its strips exactly as a
// run of its own step count does, and on each field's last slot.
__device__ __forceinline__ bool strip_keyframe(const KStrips &s, int slot, int f) {
    const int rel = slot - s.key0;
    return (rel >= 0 && rel % kStripKey == kStripKey - 1) || slot == (f < RS_TH ? s.last_v : s.last_th);
}
__device__ __forceinline__ long strip_key_index(const KStrips &s, int slot, int f) {
    return ((long)((slot - s.key0) / kStripKey) * RS_NF + f) * s.ncells;
}

// Store `v` of ring cell q (one lane per cell, 64 consecutive cells per wave;
// every lane of the wave takes part, valid or not).
__device__ __forceinline__ void strip_put(const KStrips &s, int slot, int f, long q, bool valid, float v) {
    if (!s.fp16) {
        if (valid) s.f32[((long)slot * RS_NF + f) * s.ncells + q] = v;
        return;
    }
    const bool kept = inner32_field(f) && q >= s.nouter;
    if (valid && kept) s.in32[((long)slot * kNInner32 + inner32_rank(f)) * (s.ncells - s.nouter) + (q - s.nouter)] = v;
    const bool coded = valid && !kept;
    const long base = (long)slot * strip16_offset(s.ncells, s.nouter, RS_NF) + strip16_offset(s.ncells, s.nouter, f);
    const long fc = (long)f * s.ncells + q;
    float *const pp = s.err + fc, *const rp = s.err + (long)RS_NF * s.ncells + fc;
    const float pred = coded ? *pp : 0.f, fb = coded ? *rp : 0.f;
    const float x = (v - pred) + fb;
    float mx = wave_max(coded ? fabsf(x) : 0.f);
    int e = scale_exponent(mx);
    const __half h = __float2half_rn(ldexpf(x, -e));
    if (coded) {
        s.f16[base + q] = h;
        const float d = ldexpf(__half2float(h), e);
        const float r = x - d; // exact: d is x rounded to 11 bits
        float p1 = pred + d;
        p1 = isfinite(p1) ? p1 : 0.f;
        *pp = p1;
        *rp = isfinite(r) ? r : 0.f;
        if (strip_keyframe(s, slot, f)) s.key[strip_key_index(s, slot, f) + q] = p1;
        if ((q & 63) == 0) s.e8[((long)slot * RS_NF + f) * s.nchunks + (q >> 6)] = (int8_t)e;
    }
}

// FP16: `slot` must be one of the last two slots decoded by k_strip_decode.
__device__ __forceinline__ float strip_get(const KStrips &s, int slot, int f, long q) {
    if (!s.fp16) return s.f32[((long)slot * RS_NF + f) * s.ncells + q];
    return s.dec[((long)(slot & 1) * RS_NF + f) * s.ncells + q];
}

// Backward decoding of slot `slot` of every field (blockIdx.y) into
// dec[slot & 1]: the keyframe, or slot+1's decoded value minus slot+1's
// residual (slot+1 must be the slot decoded before).  blockIdx.z is the shot,
// with the given element strides between shots.
__device__ __forceinline__ void strip_decode_body(const KStrips &s0, int slot, const KStripStride &d, const uint3 bid) {
    const long q = (long)bid.x * blockDim.x + threadIdx.x;
    const int f = bid.y;
    if (q >= s0.ncells) return;
    const KStrips s = strips_of_shot(s0, d, bid.z);
    const __half *f16 = s.f16;
    const int8_t *e8 = s.e8;
    const float *key = s.key;
    float *dec = s.dec;
    const int last = f < RS_TH ? s.last_v : s.last_th;
    float v = 0.f;
    if (slot >= 0 && slot <= last) {
        if (inner32_field(f) && q >= s.nouter) {
            v = s.in32[((long)slot * kNInner32 + inner32_rank(f)) * (s.ncells - s.nouter) + (q - s.nouter)];
        } else if (strip_keyframe(s, slot, f)) {
            v = key[strip_key_index(s, slot, f) + q];
        } else {
            const long b1 = (long)(slot + 1) * strip16_offset(s.ncells, s.nouter, RS_NF) +
                            strip16_offset(s.ncells, s.nouter, f);
            const float d = ldexpf(__half2float(f16[b1 + q]), e8[((long)(slot + 1) * RS_NF + f) * s.nchunks + (q >> 6)]);
            v = dec[((long)((slot + 1) & 1) * RS_NF + f) * s.ncells + q] - d;
        }
    }
    dec[((long)(slot & 1) * RS_NF + f) * s.ncells + q] = v;
}
static __global__ __launch_bounds__(256) void k_strip_decode(KStrips s0, int slot, KStripStride d) {
    strip_decode_body(s0, slot, d, make_uint3(blockIdx.x, blockIdx.y, blockIdx.z));
}

// The strip encoder alone: `steps` consecutive slots of field 0, x as
// [steps][n]; the outputs keep the store's layout (s.last_v = steps - 1).
static __global__ __launch_bounds__(256) void k_codec_fp16(KStrips s, const float *__restrict__ x, long n,
                                                           int steps) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = q < n;
    for (int t = 0; t < steps; ++t) strip_put(s, t, 0, q, valid, valid ? x[(long)t * n + q] : 0.f);
}

// Velocity part of slot `slot` from the assembled totals.  Launched with
// 64-lane blocks so each wave covers one 64-cell chunk.
static __global__ __launch_bounds__(256) void k_strip_save_v(KRing r, KStrips s, int slot, const float *__restrict__ Vx,
                                                      const float *__restrict__ Vz, int P) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = q < r.ncells;
    int i = 0, j = 0;
    if (valid) ring_cell(r, q, i, j);
    long k = (long)i * P + j;
    strip_put(s, slot, RS_VX, q, valid, valid ? Vx[k] : 0.f);
    strip_put(s, slot, RS_VZ, q, valid, valid ? Vz[k] : 0.f);
}

static __global__ __launch_bounds__(256) void k_strip_save_theta(KRing r, KStrips s, int slot, KWave w, int P) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = q < r.ncells;
    int i = 0, j = 0;
    if (valid) ring_cell(r, q, i, j);
    long k = (long)i * P + j;
    strip_put(s, slot, RS_TH, q, valid, valid ? w.th[k] : 0.f);
    strip_put(s, slot, RS_OM, q, valid, valid ? w.om[k] : 0.f);
    strip_put(s, slot, RS_GZX, q, valid, valid ? w.gzx[k] - w.F[3][k] : 0.f);
    strip_put(s, slot, RS_GZZ, q, valid, valid ? w.gzz[k] - w.F[1][k] : 0.f);
    strip_put(s, slot, RS_GXX, q, valid, valid ? w.gxx[k] - w.F[0][k] : 0.f);
    strip_put(s, slot, RS_GXZ, q, valid, valid ? w.gxz[k] - w.F[2][k] : 0.f);
}

// read_wavefiled1 (P-S-seperating.cu:1344-1376); slot < 0 writes zeros (V^0).
static __global__ __launch_bounds__(256) void k_strip_restore_v(KRing r, KStrips s, int slot, float *__restrict__ Vx,
                                                         float *__restrict__ Vz, int P) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= r.ncells) return;
    int i, j;
    ring_cell(r, q, i, j);
    long k = (long)i * P + j;
    Vx[k] = slot >= 0 ? strip_get(s, slot, RS_VX, q) : 0.f;
    Vz[k] = slot >= 0 ? strip_get(s, slot, RS_VZ, q) : 0.f;
}

// read_wavefiled2 (P-S-seperating.cu:1377-1424).
static __global__ __launch_bounds__(256) void k_strip_restore_theta(KRing r, KStrips s, int slot, KWave w, int P) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= r.ncells) return;
    int i, j;
    ring_cell(r, q, i, j);
    long k = (long)i * P + j;
    w.th[k] = strip_get(s, slot, RS_TH, q);
    w.om[k] = strip_get(s, slot, RS_OM, q);
    w.gzx[k] = strip_get(s, slot, RS_GZX, q);
    w.gzz[k] = strip_get(s, slot, RS_GZZ, q);
    w.gxx[k] = strip_get(s, slot, RS_GXX, q);
    w.gxz[k] = strip_get(s, slot, RS_GXZ, q);
}

// ---- source reconstruction (interior only, no CPML) -------------------------
__device__ __forceinline__ bool in_ring(const KGrid &g, int i, int j) {
    const int L = g.L, N = g.N;
    bool frame = i >= L - N && i < g.Xn - L + N && j >= L - N && j < g.Zn - L + N;
    bool inner = i >= L + N && i < g.Xn - L - N && j >= L + N && j < g.Zn - L - N;
    return frame && !inner;
}

// reshot_u (P-S-seperating.cu:890-983) as the reverse of the (value,
// increment) step: D^{t+1/2} = D^{t+3/2} - dt^2 R(theta^{t+1}),
// V^t = V^{t+1} - D^{t+1/2}.  Phi is reversed over the whole ring frame (its
// theta there is restored from the strips) because D+ Phi on the interior reads
// up to N-1 cells outside it.  Launched over [L-N, Xn-L+N)^2.
template <int N>
static __global__ __launch_bounds__(256) void k_recon_velocity(KGrid g, KWave w, KModel m) {
    OUC_CELL(g.L - N, g.L - N)
    if (i >= g.Xn - g.L + N || j >= g.Zn - g.L + N) return;
    const long k = (long)i * g.P + j;
    const float dt2 = g.dt2;
    float d = w.DPhi[k] - dt2 * w.th[k];
    w.DPhi[k] = d;
    w.Phi[k] = w.Phi[k] - d;
    const int L = g.L;
    if (i < L || i >= g.Xn - L || j < L || j >= g.Zn - L) return;
    const float *c = g.c;
    const int P = g.P;
    const float Wx = dmx<N>(w.om, k, P, c) * g.idx, Wz = dmz<N>(w.om, k, c) * g.idz;
    // away from the ring every g is a derivative of one assembled field, so the
    // product-rule form applies; near it the restored strip values do not
    // commute and the direct form is used
    const bool safe = i >= L + 2 * N && i < g.Xn - L - 2 * N && j >= L + 2 * N && j < g.Zn - L - 2 * N;
    const float q0 = safe ? m.q[k] : 0.f;
    const float KX = dmzq<N>(w.gzx, m.q, k, c, q0) * g.idz - dpxq<N>(w.gzz, m.q, k, P, c, q0) * g.idx;
    const float KZ = dmxq<N>(w.gxz, m.q, k, P, c, q0) * g.idx - dpzq<N>(w.gxx, m.q, k, c, q0) * g.idz;
    d = w.DQx[k] - dt2 * KX; w.DQx[k] = d; w.Qx[k] = w.Qx[k] - d;
    d = w.DQz[k] - dt2 * KZ; w.DQz[k] = d; w.Qz[k] = w.Qz[k] - d;
    d = w.DSx[k] - dt2 * (Wz + KX); w.DSx[k] = d; w.Sx[k] = w.Sx[k] - d;
    d = w.DSz[k] - dt2 * (-Wx + KZ); w.DSz[k] = d; w.Sz[k] = w.Sz[k] - d;
}

// Interior assembly for the reconstruction: the source P velocity of this level
// for the imaging condition, and the totals / non-gradient part off the ring
// (the ring's totals come from the strips).
template <int N>
static __global__ __launch_bounds__(256) void k_recon_assemble(KGrid g, KWave w, float *__restrict__ vpx_out,
                                                        float *__restrict__ vpz_out) {
    OUC_CELL(g.L, g.L)
    if (i >= g.Xn - g.L || j >= g.Zn - g.L) return;
    const long k = (long)i * g.P + j;
    float px = dpx<N>(w.Phi, k, g.P, g.c) * g.idx, pz = dpz<N>(w.Phi, k, g.c) * g.idz;
    vpx_out[k] = px + w.Qx[k];
    vpz_out[k] = pz + w.Qz[k];
    if (!in_ring(g, i, j)) {
        float ux = w.Qx[k] + w.Sx[k], uz = w.Qz[k] + w.Sz[k];
        w.Ux[k] = ux; w.Uz[k] = uz;
        w.Vx[k] = px + ux;
        w.Vz[k] = pz + uz;
    }
}

// reshot_s (P-S-seperating.cu:985-1042): theta, omega, g from the totals on the
// interior, no CPML; EXACT semantics adds the source as the forward did.
template <int N>
static __global__ __launch_bounds__(256) void k_recon_stress(KGrid g, KWave w, KModel m, int si, int sj, float s) {
    OUC_CELL(g.L, g.L)
    if (i >= g.Xn - g.L || j >= g.Zn - g.L) return;
    const long k = (long)i * g.P + j;
    const int L = g.L;
    const float gxx = dmx<N>(w.Vx, k, g.P, g.c) * g.idx;
    const float gxz = dpz<N>(w.Vx, k, g.c) * g.idz;
    const float gzz = dmz<N>(w.Vz, k, g.c) * g.idz;
    const float gzx = dpx<N>(w.Vz, k, g.P, g.c) * g.idx;
    // cells whose curl stencil reads ring totals use them directly
    const bool direct = (j - (N - 1) < L + N) || (j + N >= g.Zn - L - N) || (i - (N - 1) < L + N) ||
                        (i + N >= g.Xn - L - N);
    float cu = direct ? (gxz - gzx) : (dpz<N>(w.Ux, k, g.c) * g.idz - dpx<N>(w.Uz, k, g.P, g.c) * g.idx);
    float th = m.p[k] * (gxx + gzz);
    if (i == si && j == sj) th = th + s;
    w.th[k] = th;
    w.om[k] = m.q[k] * cu;
    w.gxx[k] = gxx; w.gxz[k] = gxz; w.gzx[k] = gzx; w.gzz[k] = gzz;
}

// Hand-off: (V^nt, D^{nt-1/2}) -> (V^{nt-1}, D^{nt-1/2}) so the reverse step
// starts from the level the reference's reshot_u first reads.
static __global__ __launch_bounds__(256) void k_unstep(KGrid g, KWave w) {
    OUC_CELL(0, 0)
    if (i >= g.Xn || j >= g.Zn) return;
    const long k = (long)i * g.P + j;
    w.Phi[k] -= w.DPhi[k];
    w.Qx[k] -= w.DQx[k];
    w.Qz[k] -= w.DQz[k];
    w.Sx[k] -= w.DSx[k];
    w.Sz[k] -= w.DSz[k];
}

// Source P velocity of one level on the interior (the hand-off levels).
template <int N>
static __global__ __launch_bounds__(256) void k_asm_vp(KGrid g, KWave w, float *__restrict__ vpx_out,
                                                float *__restrict__ vpz_out) {
    OUC_CELL(g.L, g.L)
    if (i >= g.Xn - g.L || j >= g.Zn - g.L) return;
    const long k = (long)i * g.P + j;
    vpx_out[k] = dpx<N>(w.Phi, k, g.P, g.c) * g.idx + w.Qx[k];
    vpz_out[k] = dpz<N>(w.Phi, k, g.c) * g.idz + w.Qz[k];
}

// corr_v (P-S-seperating.cu:1426-1439) on the interior.
template <int N>
static __global__ __launch_bounds__(256) void k_corr(KGrid g, KWave r, const float *__restrict__ spx,
                                              const float *__restrict__ spz, float *__restrict__ npp,
                                              float *__restrict__ nps, float *__restrict__ den) {
    OUC_CELL(g.L, g.L)
    if (i >= g.Xn - g.L || j >= g.Zn - g.L) return;
    const long k = (long)i * g.P + j;
    float rpx = dpx<N>(r.Phi, k, g.P, g.c) * g.idx + r.Qx[k];
    float rpz = dpz<N>(r.Phi, k, g.c) * g.idz + r.Qz[k];
    float a = spx[k], b = spz[k];
    npp[k] += a * rpx + b * rpz;
    den[k] += a * a + b * b;
    nps[k] += a * r.Sx[k] + b * r.Sz[k];
}

// image_fun (P-S-seperating.cu:1441-1455) into x-major interior images.
static __global__ void k_image(KGrid g, const float *__restrict__ npp, const float *__restrict__ nps,
                        const float *__restrict__ den, float *__restrict__ ipp, float *__restrict__ ips) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x, ix = blockIdx.y;
    if (iz >= g.nz || ix >= g.nx) return;
    long k = (long)(ix + g.L) * g.P + iz + g.L;
    size_t o = (size_t)ix * g.nz + iz;
    ipp[o] = npp[k] / den[k];
    ips[o] = nps[k] / den[k];
}

// Laplace (P-S-seperating.cu:1457-1469) on an x-major interior image; the two
// outermost rows/columns stay zero.
static __global__ void k_laplace5(int nx, int nz, float dh, const float *__restrict__ img, float *__restrict__ out) {
    int iz = blockIdx.x * blockDim.x + threadIdx.x, ix = blockIdx.y;
    if (iz >= nz || ix >= nx) return;
    size_t o = (size_t)ix * nz + iz;
    float v = 0.f;
    if (ix >= 2 && ix < nx - 2 && iz >= 2 && iz < nz - 2)
        v = (1 / dh / dh) * (img[o + nz] + img[o - nz] - 2 * img[o] + img[o + 1] + img[o - 1] - 2 * img[o]);
    out[o] = v;
}

static __global__ void k_accumulate(size_t n, const float *__restrict__ a, float *__restrict__ acc) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) acc[k] += a[k];
}

} // namespace ouc
