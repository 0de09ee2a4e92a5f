/*
 * ouc_oracle.cpp -- CPU restatement of gm340/ouc's elastic P/S RTM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine under
 * ouc_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it, and the product never links or calls it.
 *
 * One template, two instantiations: double is the correctness oracle, float is
 * the reported CPU baseline (OpenMP over x columns).  Inputs that the reference
 * defines as float values (model, dt, dx, dz, stencil coefficients) enter both
 * instantiations as those float values; quantities the reference computes in
 * double (CPML damping P-S-seperating.cu:2067-2105, Ricker wavelet
 * P-S-seperating.cu:1954-1965) are computed in double and kept in double by the
 * FP64 oracle, rounded to float by the FP32 form.
 *
 * Semantics follow the reference exactly where it is well defined, and make the
 * following choices where it is not (each one is pinned by a test):
 *  - recording (P-S-seperating.cu:307-311 race): the record of step t is the
 *    post-update V^{t+1} at the receiver row;
 *  - strips (P-S-seperating.cu:1230-1287 out-of-bounds slot offsets): a
 *    compact ring layout holding the same values;
 *  - FAITHFUL restores the velocity strips one step late and images with the
 *    source velocity two steps ahead, as the reference does; EXACT aligns both.
 *
 * Layout: padded Xn x Zn arrays, f[i*Zn + j], x slow and z fast, as the
 * reference (P-S-seperating.cu:18-19).
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/ouc.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Stencil coefficients c_k, k=1..N, of the staggered first derivative.
 * N = 6 uses the reference's literals (P-S-seperating.cu:221); other orders use
 * the closed form of the Taylor staggered coefficients
 *   c_k = (-1)^{k+1} / (2k-1) * prod_{m!=k} (2m-1)^2 / |(2k-1)^2 - (2m-1)^2|,
 * evaluated in double.  Either way the value used is the float one, as the
 * reference's `float a[N]`. */
void coeffs_float(int N, double *c) {
    if (N == 6) {
        const float a[6] = {1.2213364f, -9.6931458e-2f, 1.7447662e-2f,
                            -2.9672895e-3f, 3.5900540e-4f, -2.1847812e-5f};
        for (int k = 0; k < 6; ++k) c[k] = (double)a[k];
        return;
    }
    for (int k = 1; k <= N; ++k) {
        double num = 1.0, den = (double)(2 * k - 1);
        for (int m = 1; m <= N; ++m) {
            if (m == k) continue;
            double a = (double)((2 * m - 1) * (2 * m - 1));
            double b = (double)((2 * k - 1) * (2 * k - 1));
            num *= a;
            den *= std::fabs(b - a);
        }
        double v = num / den;
        if ((k % 2) == 0) v = -v;
        c[k - 1] = (double)(float)v;
    }
}

struct Geo {
    int nx, nz, L, N, Xn, Zn, nt;
};

/* Edge padding, xiangbian() (P-S-seperating.cu:37-119): every pad cell takes the
 * nearest interior cell; corners take the corner interior cell.  Equivalent to
 * clamping both indices into the interior. */
void pad_clamp(const Geo &g, const float *interior, std::vector<double> &out) {
    out.assign((size_t)g.Xn * g.Zn, 0.0);
    for (int i = 0; i < g.Xn; ++i) {
        int ii = std::min(std::max(i - g.L, 0), g.nx - 1);
        for (int j = 0; j < g.Zn; ++j) {
            int jj = std::min(std::max(j - g.L, 0), g.nz - 1);
            out[(size_t)i * g.Zn + j] = (double)interior[(size_t)ii * g.nz + jj];
        }
    }
}

/* CPML damping profiles, P-S-seperating.cu:2067-2105, in double:
 *   ax = log10(1/R) * 1.5 * vp / L * ((L - i)/L)^4        for 0 <= i < L
 *   ax = log10(1/R) * 1.5 * vp / L * ((i - Xn + L)/L)^4   for Xn-L < i < Xn
 * and zero elsewhere; az likewise along j.  No dx factor, log10 not ln. */
void cpml(const Geo &g, double R, const std::vector<double> &vp, std::vector<double> &ax,
          std::vector<double> &az) {
    ax.assign((size_t)g.Xn * g.Zn, 0.0);
    az.assign((size_t)g.Xn * g.Zn, 0.0);
    const int L = g.L;
    for (int i = 0; i < g.Xn; ++i)
        for (int j = 0; j < g.Zn; ++j) {
            size_t k = (size_t)i * g.Zn + j;
            if (i >= 0 && i < L)
                ax[k] = std::log10(1.0 / R) * 1.5 * vp[k] / (L) * std::pow(1.0 * (L - i) / (L), 4.0);
            else if (i > g.Xn - L && i < g.Xn)
                ax[k] = std::log10(1.0 / R) * 1.5 * vp[k] / (L) * std::pow(1.0 * (i - g.Xn + L) / (L), 4.0);
            if (j >= 0 && j < L)
                az[k] = std::log10(1.0 / R) * 1.5 * vp[k] / (L) * std::pow(1.0 * (L - j) / (L), 4.0);
            else if (j > g.Zn - L && j < g.Zn)
                az[k] = std::log10(1.0 / R) * 1.5 * vp[k] / (L) * std::pow(1.0 * (j - g.Zn + L) / (L), 4.0);
        }
}

/* Ricker wavelet, P-S-seperating.cu:1954-1958, with the truncated PI of
 * P-S-seperating.cu:14 and t0 = ceil(1/(fm*dt)). */
const double kPI = 3.1415926;
int ricker_t0(float fm, float dt) { return (int)std::ceil(1.0 / ((double)fm * (double)dt)); }
void ricker(int nt, float fm, float dt, std::vector<double> &s) {
    double Nk = kPI * kPI * (double)fm * (double)fm * (double)dt * (double)dt;
    int t0 = ricker_t0(fm, dt);
    s.resize(nt);
    for (int t = 0; t < nt; ++t) {
        double u = (double)(t - t0) * (double)(t - t0);
        s[t] = (1.0 - 2.0 * Nk * u) * std::exp(-Nk * u);
    }
}

/* Direct-wave mute threshold, remove() P-S-seperating.cu:1634-1657: samples
 * t < 2*t0 + dist*dh/(dt*vp(Sx,Zr)) are zeroed, dist = sqrtf(dx^2+dz^2) in
 * cells.  The expression keeps the reference's float/double mix. */
double mute_threshold(int Sx, int Sz, int Zr, int i, int t0, float dt, float dh, float vpsrc) {
    int ax = std::abs(Sx - i), az = std::abs(Zr - Sz);
    float dist = std::sqrt((float)(ax * ax + az * az));
    float dd = dist * dh;
    return 2 * t0 + (double)dd * 1.0 / (double)(dt * vpsrc);
}

/* Staggered derivative operators (P-S-seperating.cu:224-279, 340-365):
 *   Dm_x f = (1/dx) sum_k c_k (f[i+k-1] - f[i-k])
 *   Dp_x f = (1/dx) sum_k c_k (f[i+k]   - f[i-k+1])
 * and the same along j.  Summed left to right, then divided, as written. */
template <class R>
inline R dmx(const R *f, size_t k, int Zn, const R *c, int N, R h) {
    R s = 0;
    for (int m = 1; m <= N; ++m) s += c[m - 1] * (f[k + (size_t)(m - 1) * Zn] - f[k - (size_t)m * Zn]);
    return s / h;
}
template <class R>
inline R dpx(const R *f, size_t k, int Zn, const R *c, int N, R h) {
    R s = 0;
    for (int m = 1; m <= N; ++m) s += c[m - 1] * (f[k + (size_t)m * Zn] - f[k - (size_t)(m - 1) * Zn]);
    return s / h;
}
template <class R>
inline R dmz(const R *f, size_t k, const R *c, int N, R h) {
    R s = 0;
    for (int m = 1; m <= N; ++m) s += c[m - 1] * (f[k + (m - 1)] - f[k - m]);
    return s / h;
}
template <class R>
inline R dpz(const R *f, size_t k, const R *c, int N, R h) {
    R s = 0;
    for (int m = 1; m <= N; ++m) s += c[m - 1] * (f[k + m] - f[k - (m - 1)]);
    return s / h;
}

template <class R>
struct Setup {
    Geo g;
    R dx, dz, dt, dh;
    R c[8];
    std::vector<R> vp, vs, ax, az;
    std::vector<R> src;
    std::vector<double> vp_d; /* padded vp in double, for the mute */
};

template <class R>
void cast_vec(const std::vector<double> &a, std::vector<R> &b) {
    b.resize(a.size());
    for (size_t k = 0; k < a.size(); ++k) b[k] = (R)a[k];
}

template <class R>
bool make_setup(const ouc_params *p, const ouc_model *m, const ouc_shot *s, Setup<R> &S) {
    if (!p || !m || !m->vp || !m->vs) return false;
    if (p->half_order < 1 || p->half_order > 8) return false;
    if (p->nx < 2 * p->half_order || p->nz < 2 * p->half_order || p->npml < p->half_order || p->nt < 1) return false;
    Geo &g = S.g;
    g.nx = p->nx; g.nz = p->nz; g.L = p->npml; g.N = p->half_order;
    g.Xn = p->nx + 2 * p->npml; g.Zn = p->nz + 2 * p->npml; g.nt = p->nt;
    S.dx = (R)p->dx; S.dz = (R)p->dz; S.dt = (R)p->dt; S.dh = (R)p->dx;
    double c[8];
    coeffs_float(g.N, c);
    for (int k = 0; k < g.N; ++k) S.c[k] = (R)c[k];
    std::vector<double> vpd, vsd, axd, azd;
    pad_clamp(g, m->vp, vpd);
    pad_clamp(g, m->vs, vsd);
    cpml(g, p->R, vpd, axd, azd);
    cast_vec(vpd, S.vp); cast_vec(vsd, S.vs); cast_vec(axd, S.ax); cast_vec(azd, S.az);
    S.vp_d = vpd;
    std::vector<double> sd;
    if (s && s->wavelet) {
        sd.resize(g.nt);
        for (int t = 0; t < g.nt; ++t) sd[t] = (double)s->wavelet[t];
    } else if (s) {
        ricker(g.nt, s->fm, p->dt, sd);
    }
    cast_vec(sd, S.src);
    return true;
}

/* One wavefield: summed velocities, stress-like fields, P/S velocity at two
 * time levels and the CPML memory variables (F_* of forward_s, O_* of
 * forward_u). */
template <class R>
struct Wave {
    std::vector<R> vx, vz;
    std::vector<R> th, om, uxx, uxz, uzx, uzz; /* theta, omega, duxdx, duxdz, duzdx, duzdz */
    std::vector<R> cur[4], prev[4];            /* px, pz, sx, sz */
    std::vector<R> F[4];                       /* F_xx, F_zz, F_xz, F_zx */
    std::vector<R> O[8];                       /* O_C, O_E, O_B, O_A, O_Tx, O_Tz, O_Wz, O_Wx */
    void init(size_t n) {
        for (auto *v : {&vx, &vz, &th, &om, &uxx, &uxz, &uzx, &uzz}) v->assign(n, (R)0);
        for (int k = 0; k < 4; ++k) { cur[k].assign(n, (R)0); prev[k].assign(n, (R)0); F[k].assign(n, (R)0); }
        for (int k = 0; k < 8; ++k) O[k].assign(n, (R)0);
    }
};

/* Planted-mistake switch for the pins (tests only, default 0 = none).
 * Value u in 1..12 swaps the damping profile (ax <-> az) of one memory-variable
 * update: 1..4 the stress pass's F_xx, F_zz, F_xz, F_zx; 5..12 the velocity
 * pass's O_C, O_E, O_B, O_A, O_Tx, O_Tz, O_Wz, O_Wx.  A test shows that each
 * such mistake breaks the absorbing boundary.  In the split-PML
 * reconstruction, 13 restores the ring velocities one slot late and 14 drops
 * the source sample from the reversed stress pass; each breaks its exact
 * reversibility. */
int g_mutation = 0;

/* Damping profile of memory-variable update u (1-based, as g_mutation). */
template <class R>
inline R damp(int u, R ax, R az, bool x_axis) {
    const bool use_x = (g_mutation == u) ? !x_axis : x_axis;
    return use_x ? ax : az;
}

/* Stress pass with CPML, forward_s / rt_s_res (P-S-seperating.cu:316-384,
 * 667-726) on [N, Xn-N) x [N, Zn-N); adds s at (si, sj) after the update. */
template <class R>
void stress_cpml(const Setup<R> &S, Wave<R> &W, int si, int sj, R s) {
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn;
    const R dt = S.dt;
#pragma omp parallel for schedule(static)
    for (int i = N; i < g.Xn - N; ++i)
        for (int j = N; j < Zn - N; ++j) {
            size_t k = (size_t)i * Zn + j;
            R gxx = dmx(W.vx.data(), k, Zn, S.c, N, S.dx);
            R gxz = dpz(W.vx.data(), k, S.c, N, S.dz);
            R gzz = dmz(W.vz.data(), k, S.c, N, S.dz);
            R gzx = dpx(W.vz.data(), k, Zn, S.c, N, S.dx);
            const R ax = S.ax[k], az = S.az[k];
            const R a1 = damp(1, ax, az, true), a2 = damp(2, ax, az, false), a3 = damp(3, ax, az, false),
                    a4 = damp(4, ax, az, true);
            R &Fxx = W.F[0][k], &Fzz = W.F[1][k], &Fxz = W.F[2][k], &Fzx = W.F[3][k];
            Fxx = Fxx + (a1 * gxx - a1 * Fxx) * dt;
            Fzz = Fzz + (a2 * gzz - a2 * Fzz) * dt;
            Fxz = Fxz + (a3 * gxz - a3 * Fxz) * dt;
            Fzx = Fzx + (a4 * gzx - a4 * Fzx) * dt;
            R vp = S.vp[k], vs = S.vs[k];
            W.th[k] = vp * vp * (gxx + gzz - Fxx - Fzz);
            W.om[k] = vs * vs * (gxz - gzx - Fxz + Fzx);
            W.uzx[k] = vs * vs * (gzx - Fzx);
            W.uzz[k] = vs * vs * (gzz - Fzz);
            W.uxz[k] = vs * vs * (gxz - Fxz);
            W.uxx[k] = vs * vs * (gxx - Fxx);
            if (i == si && j == sj) W.th[k] = W.th[k] + s;
        }
}

/* Velocity pass with CPML, forward_u / rt_u_res (P-S-seperating.cu:205-314,
 * 562-665): psi updated before use, leapfrog of the four P/S components,
 * past <- now <- new, Vx = Vpx + Vsx, Vz = Vpz + Vsz. */
template <class R>
void velocity_cpml(const Setup<R> &S, Wave<R> &W) {
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn;
    const R dt = S.dt;
#pragma omp parallel for schedule(static)
    for (int i = N; i < g.Xn - N; ++i)
        for (int j = N; j < Zn - N; ++j) {
            size_t k = (size_t)i * Zn + j;
            R Tx = dpx(W.th.data(), k, Zn, S.c, N, S.dx);
            R Tz = dpz(W.th.data(), k, S.c, N, S.dz);
            R Wx = dmx(W.om.data(), k, Zn, S.c, N, S.dx);
            R Wz = dmz(W.om.data(), k, S.c, N, S.dz);
            R A = dpz(W.uxx.data(), k, S.c, N, S.dz);
            R B = dmx(W.uxz.data(), k, Zn, S.c, N, S.dx);
            R C = dmz(W.uzx.data(), k, S.c, N, S.dz);
            R E = dpx(W.uzz.data(), k, Zn, S.c, N, S.dx);
            const R ax = S.ax[k], az = S.az[k];
            const R b1 = damp(5, ax, az, false), b2 = damp(6, ax, az, true), b3 = damp(7, ax, az, true),
                    b4 = damp(8, ax, az, false), b5 = damp(9, ax, az, true), b6 = damp(10, ax, az, false),
                    b7 = damp(11, ax, az, false), b8 = damp(12, ax, az, true);
            R &OC = W.O[0][k], &OE = W.O[1][k], &OB = W.O[2][k], &OA = W.O[3][k];
            R &OTx = W.O[4][k], &OTz = W.O[5][k], &OWz = W.O[6][k], &OWx = W.O[7][k];
            OC = OC + (b1 * C - b1 * OC) * dt;
            OE = OE + (b2 * E - b2 * OE) * dt;
            OB = OB + (b3 * B - b3 * OB) * dt;
            OA = OA + (b4 * A - b4 * OA) * dt;
            OTx = OTx + (b5 * Tx - b5 * OTx) * dt;
            OTz = OTz + (b6 * Tz - b6 * OTz) * dt;
            OWz = OWz + (b7 * Wz - b7 * OWz) * dt;
            OWx = OWx + (b8 * Wx - b8 * OWx) * dt;
            R dt2 = dt * dt;
            R npx = 2 * W.cur[0][k] - W.prev[0][k] + dt2 * (Tx - OTx + C - E - OC + OE);
            R npz = 2 * W.cur[1][k] - W.prev[1][k] + dt2 * (Tz - OTz + B - A - OB + OA);
            R nsx = 2 * W.cur[2][k] - W.prev[2][k] + dt2 * (Wz - OWz + C - E - OC + OE);
            R nsz = 2 * W.cur[3][k] - W.prev[3][k] + dt2 * (-Wx + OWx + B - A - OB + OA);
            W.prev[0][k] = W.cur[0][k]; W.cur[0][k] = npx;
            W.prev[1][k] = W.cur[1][k]; W.cur[1][k] = npz;
            W.prev[2][k] = W.cur[2][k]; W.cur[2][k] = nsx;
            W.prev[3][k] = W.cur[3][k]; W.cur[3][k] = nsz;
            W.vx[k] = npx + nsx;
            W.vz[k] = npz + nsz;
        }
}

/* Ring of boundary strips.  The reference saves 2N-wide bands of 8 fields
 * around the interior every step (save_wavefiled P-S-seperating.cu:1230-1287)
 * and restores them in the backward loop (P-S-seperating.cu:1344-1424).  Only
 * cells of the frame [L-N, Xn-L+N)^2 that lie within N of the interior edge are
 * ever read by the interior-only reconstruction, so the ring stores exactly
 * those: four rectangles, top/bottom (nx+2N) x 2N, left/right 2N x (nz-2N).
 * Storage order: each rectangle's outer part (its N-1 outermost pad rows or
 * columns) for top, bottom, left, right, then their inner parts (the N
 * interior rows or columns next to the edge and the first pad row or column);
 * z runs fastest within a part. */
struct Ring {
    int L, N, Xn, Zn, nx, nz;
    size_t ntop, nside, ncells, nouter;
    void init(const Geo &g) {
        L = g.L; N = g.N; Xn = g.Xn; Zn = g.Zn; nx = g.nx; nz = g.nz;
        ntop = (size_t)(nx + 2 * N) * (2 * N);
        nside = (size_t)(2 * N) * (nz - 2 * N);
        ncells = 2 * ntop + 2 * nside;
        nouter = (size_t)(N - 1) * (2 * (nx + 2 * N) + 2 * (nz - 2 * N));
    }
    /* Part (outer, inner) of one side: the first local row (top, bottom) or
     * column (left, right) across the strip's 2N, and the part's width. */
    void cell(size_t r, int &i, int &j) const {
        const size_t cw = (size_t)(nx + 2 * N), w = (size_t)(nz - 2 * N);
        const bool inner = r >= nouter;
        if (inner) r -= nouter;
        const size_t wd = inner ? (size_t)(N + 1) : (size_t)(N - 1);
        /* local offsets of the part within the 2N-wide strip: top/left keep
         * the pad side first (outer rows 0..N-2, inner N-1..2N-1), bottom/right
         * the interior side first (inner 0..N, outer N+1..2N-1) */
        const int lo_tl = inner ? N - 1 : 0, lo_br = inner ? 0 : N + 1;
        if (r < cw * wd) { i = L - N + (int)(r / wd); j = L - N + lo_tl + (int)(r % wd); return; }
        r -= cw * wd;
        if (r < cw * wd) { i = L - N + (int)(r / wd); j = Zn - L - N + lo_br + (int)(r % wd); return; }
        r -= cw * wd;
        if (r < wd * w) { i = L - N + lo_tl + (int)(r / w); j = L + N + (int)(r % w); return; }
        r -= wd * w;
        i = Xn - L - N + lo_br + (int)(r / w);
        j = L + N + (int)(r % w);
    }
};

/* Strip fields in slot order: vx, vz, theta, omega, duzdx, duzdz, duxdx, duxdz
 * (the save order of P-S-seperating.cu:1242-1249). */
enum { RS_VX, RS_VZ, RS_TH, RS_OM, RS_UZX, RS_UZZ, RS_UXX, RS_UXZ, RS_NF };

/* Binary16 round-to-nearest-even of a float, with subnormals and overflow to
 * infinity; written from the IEEE 754 binary16 format definition. */
uint16_t f32_to_f16_rne(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) { /* inf or nan */
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    }
    int e = (int)(ax >> 23) - 127; /* unbiased exponent */
    uint32_t mant = ax & 0x7fffffu;
    if (e > 15) return (uint16_t)(sign | 0x7c00u);
    if (e >= -14) { /* normal half */
        uint32_t m = mant | 0x800000u;                   /* 24-bit significand */
        uint32_t keep = m >> 13, rem = m & 0x1fffu;     /* 11 bits kept */
        if (rem > 0x1000u || (rem == 0x1000u && (keep & 1u))) keep += 1;
        uint32_t he = (uint32_t)(e + 15);
        if (keep == 0x800u) { keep = 0x400u; he += 1; }  /* significand overflow */
        if (he >= 31) return (uint16_t)(sign | 0x7c00u);
        return (uint16_t)(sign | (he << 10) | (keep & 0x3ffu));
    }
    /* subnormal half: value = m * 2^(e-23); unit 2^-24 */
    if (e < -25) return (uint16_t)sign;
    uint32_t m = (ax >> 23) ? (mant | 0x800000u) : mant;
    int shift = -e - 1; /* m*2^(e-23) / 2^-24 = m >> (-(e+1)) */
    if (shift > 31) return (uint16_t)sign;
    uint32_t keep = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (keep & 1u))) keep += 1;
    return (uint16_t)(sign | keep); /* keep == 0x400 becomes the smallest normal */
}

float f16_to_f32(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v;
    if (e == 0) v = std::ldexp((float)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp((float)(m | 0x400u), (int)e - 25);
    uint32_t bits;
    std::memcpy(&bits, &v, 4);
    bits |= sign;
    std::memcpy(&v, &bits, 4);
    return v;
}

/* Scaled-FP16 chunk codec.  A chunk of up to 64 floats shares a power-of-two
 * scale 2^e chosen so that the largest magnitude maps into [2^14, 2^15):
 * e = frexp-exponent(max) - 15.  Scaling by 2^-e is exact, so the stored half
 * is the round-to-nearest-even of x * 2^-e and decoding is exact. */
int chunk_exponent(const float *x, int n) {
    float mx = 0.f;
    for (int k = 0; k < n; ++k) mx = std::max(mx, std::fabs(x[k]));
    if (!(mx > 0.f)) return 0;
    int ex;
    std::frexp(mx, &ex);
    return ex - 15;
}
void encode_chunk(const float *x, int n, uint16_t *h, int8_t *e_out) {
    int e = chunk_exponent(x, n);
    e = std::max(-127, std::min(127, e));
    *e_out = (int8_t)e;
    for (int k = 0; k < n; ++k) h[k] = f32_to_f16_rne(std::ldexp(x[k], -e));
}
void decode_chunk(const uint16_t *h, int n, int8_t e, float *x) {
    for (int k = 0; k < n; ++k) x[k] = std::ldexp(f16_to_f32(h[k]), (int)e);
}

/* Predictive scaled-FP16 strip codec, the definition the GPU strip store
 * implements.  Per field and ring cell, slots in time order: the encoder
 * codes the residual x = (v - p) + r of the value against p, the previous
 * slot's decoded value, with r the previous slot's quantization error fed
 * back; the chunk of 64 cells shares the exponent of max|x| (encode_chunk),
 * d = decode(x) is x rounded to 11 bits, r' = x - d is exact, and the encoder
 * moves to p' = p + d (a non-finite p' or r' restarts at zero).  The decoded
 * sequence is therefore v^s + r^{s-1} - r^s.  Every kStripKey-th slot
 * (s % kStripKey == kStripKey - 1) and the field's last slot keep p' in FP32
 * as a keyframe.  Decoding runs backwards, as the reconstruction uses the
 * strips: v^s = keyframe, or v^{s+1} - d^{s+1} in FP32. */
constexpr int kStripKey = 32;

struct Codec16 {
    size_t ncells = 0, nch = 0;
    int nfields = 0;
    std::vector<int> last;      /* last slot stored, per field */
    std::vector<uint16_t> h;    /* [slot][field][cell] residuals */
    std::vector<int8_t> e;      /* [slot][field][chunk] */
    std::vector<float> key;     /* [slot / kStripKey][field][cell] */
    std::vector<float> p, r;    /* encoder state [field][cell] */
    std::vector<float> dec;     /* [2][field][cell] decoded slots by parity */
    std::vector<int> cur;       /* per field: lowest decoded slot (-1: none yet) */
    void init(size_t cells, int slots, int fields, const std::vector<int> &lastslot) {
        ncells = cells; nch = (cells + 63) / 64; nfields = fields; last = lastslot;
        h.assign((size_t)slots * fields * cells, 0);
        e.assign((size_t)slots * fields * nch, 0);
        key.assign((size_t)((slots + kStripKey - 1) / kStripKey) * fields * cells, 0.f);
        p.assign((size_t)fields * cells, 0.f);
        r.assign((size_t)fields * cells, 0.f);
        dec.assign((size_t)2 * fields * cells, 0.f);
        cur.assign(fields, -1);
    }
    bool keyframe(int s, int f) const { return s % kStripKey == kStripKey - 1 || s == last[f]; }
    void encode(int s, int f, const float *v) {
        float *pf = p.data() + (size_t)f * ncells, *rf = r.data() + (size_t)f * ncells;
        uint16_t *hs = h.data() + ((size_t)s * nfields + f) * ncells;
        int8_t *es = e.data() + ((size_t)s * nfields + f) * nch;
        float *ks = keyframe(s, f) ? key.data() + ((size_t)(s / kStripKey) * nfields + f) * ncells : nullptr;
        float x[64], d[64];
        for (size_t c0 = 0, ci = 0; c0 < ncells; c0 += 64, ++ci) {
            int n = (int)std::min<size_t>(64, ncells - c0);
            for (int k = 0; k < n; ++k) x[k] = (v[c0 + k] - pf[c0 + k]) + rf[c0 + k];
            encode_chunk(x, n, hs + c0, es + ci);
            decode_chunk(hs + c0, n, es[ci], d);
            for (int k = 0; k < n; ++k) {
                float rr = x[k] - d[k];
                float p1 = pf[c0 + k] + d[k];
                rf[c0 + k] = std::isfinite(rr) ? rr : 0.f;
                pf[c0 + k] = std::isfinite(p1) ? p1 : 0.f;
                if (ks) ks[c0 + k] = pf[c0 + k];
            }
        }
    }
    /* Decoded values of slot s of field f; uses must come in non-increasing
     * slot order (s may be the last or the previous decoded slot). */
    const float *decoded(int s, int f) {
        float *d0 = dec.data() + (size_t)f * ncells;
        auto buf = [&](int slot) { return d0 + (size_t)(slot & 1) * nfields * ncells; };
        if (cur[f] < 0) {
            const float *k = key.data() + ((size_t)(last[f] / kStripKey) * nfields + f) * ncells;
            std::copy(k, k + ncells, buf(last[f]));
            cur[f] = last[f];
        }
        while (cur[f] > s) {
            const int t = cur[f] - 1;
            float *o = buf(t);
            if (keyframe(t, f)) {
                const float *k = key.data() + ((size_t)(t / kStripKey) * nfields + f) * ncells;
                std::copy(k, k + ncells, o);
            } else {
                const uint16_t *hs = h.data() + ((size_t)(t + 1) * nfields + f) * ncells;
                const int8_t *es = e.data() + ((size_t)(t + 1) * nfields + f) * nch;
                const float *up = buf(t + 1);
                for (size_t q = 0; q < ncells; ++q)
                    o[q] = up[q] - std::ldexp(f16_to_f32(hs[q]), (int)es[q >> 6]);
            }
            cur[f] = t;
        }
        return buf(s);
    }
};

/* Fields whose inner ring part the FP16 store keeps exactly (in FP32):
 * theta and the four gradients.  The reconstruction integrates the split
 * source velocities of the ring's interior rows open loop from them, so their
 * rounding accumulates in the imaged source field; omega (S part only) and the
 * velocities (closed-loop stress) stay residual-coded on the whole ring. */
inline bool inner32_field(int f) {
    return f == RS_TH || f == RS_UZX || f == RS_UZZ || f == RS_UXX || f == RS_UXZ;
}
inline int inner32_rank(int f) {
    int k = 0;
    for (int a = 0; a < f; ++a) k += inner32_field(a) ? 1 : 0;
    return k;
}
constexpr int kNInner32 = 5;

template <class R>
struct Strips {
    Ring ring;
    int nslots = 0;
    bool fp16 = false;
    std::vector<R> data;          /* FP32 mode: [slot][field][ring] */
    std::vector<Codec16> codec;   /* FP16 mode, one per field (slots saved in time order, restored in reverse):
                                     the whole ring, or its outer part for the FP32-kept fields */
    std::vector<float> in32;      /* FP16 mode: [slot][kNInner32][inner cells] */
    size_t ninner = 0;
    void init(const Geo &g, int slots, bool use_fp16) {
        ring.init(g);
        nslots = slots;
        fp16 = use_fp16;
        ninner = ring.ncells - ring.nouter;
        if (fp16) {
            codec.resize(RS_NF);
            for (int f = 0; f < RS_NF; ++f)
                codec[f].init(inner32_field(f) ? ring.nouter : ring.ncells, slots, 1, std::vector<int>(1, slots - 1));
            in32.assign((size_t)slots * kNInner32 * ninner, 0.f);
        } else {
            data.assign((size_t)slots * RS_NF * ring.ncells, (R)0);
        }
    }
    R *slot(int t, int f) { return data.data() + ((size_t)t * RS_NF + f) * ring.ncells; }
    float *inner(int t, int f) { return in32.data() + ((size_t)t * kNInner32 + inner32_rank(f)) * ninner; }
    void save(int t, const Wave<R> &W) {
        const std::vector<R> *src[RS_NF] = {&W.vx, &W.vz, &W.th, &W.om, &W.uzx, &W.uzz, &W.uxx, &W.uxz};
        std::vector<float> v(fp16 ? ring.ncells : 0);
        for (int f = 0; f < RS_NF; ++f) {
            for (size_t r = 0; r < ring.ncells; ++r) {
                int i, j;
                ring.cell(r, i, j);
                const R x = (*src[f])[(size_t)i * ring.Zn + j];
                if (fp16) v[r] = (float)x;
                else slot(t, f)[r] = x;
            }
            if (fp16) {
                codec[f].encode(t, 0, v.data());
                if (inner32_field(f)) std::copy(v.begin() + ring.nouter, v.end(), inner(t, f));
            }
        }
    }
    void restore(int t, int f0, int f1, std::vector<R> *const *dst) {
        for (int f = f0; f < f1; ++f) {
            std::vector<R> &out = *dst[f - f0];
            const float *d16 = fp16 ? codec[f].decoded(t, 0) : nullptr;
            const float *in = fp16 && inner32_field(f) ? inner(t, f) : nullptr;
            const R *d = fp16 ? nullptr : slot(t, f);
            for (size_t r = 0; r < ring.ncells; ++r) {
                int i, j;
                ring.cell(r, i, j);
                R v;
                if (!fp16) v = d[r];
                else if (in && r >= ring.nouter) v = (R)in[r - ring.nouter];
                else v = (R)d16[r];
                out[(size_t)i * ring.Zn + j] = v;
            }
        }
    }
    void zero_v(std::vector<R> &vx, std::vector<R> &vz) {
        for (size_t r = 0; r < ring.ncells; ++r) {
            int i, j;
            ring.cell(r, i, j);
            vx[(size_t)i * ring.Zn + j] = (R)0;
            vz[(size_t)i * ring.Zn + j] = (R)0;
        }
    }
};

/* Source-wavefield reconstruction on the interior [L, Xn-L)^2, no CPML.
 * reshot_u (P-S-seperating.cu:890-983): Vp*_s <- now; now <- past;
 * past <- 2 now - Vp*_s + dt^2 L(theta_s); Vx_s = past_px + past_sx. */
template <class R>
void recon_velocity(const Setup<R> &S, Wave<R> &Ssrc, std::vector<R> *snew, std::vector<R> *now,
                    std::vector<R> *past) {
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn, L = g.L;
    const R dt = S.dt;
#pragma omp parallel for schedule(static)
    for (int i = L; i < g.Xn - L; ++i)
        for (int j = L; j < Zn - L; ++j) {
            size_t k = (size_t)i * Zn + j;
            R Tx = dpx(Ssrc.th.data(), k, Zn, S.c, N, S.dx);
            R Tz = dpz(Ssrc.th.data(), k, S.c, N, S.dz);
            R Wx = dmx(Ssrc.om.data(), k, Zn, S.c, N, S.dx);
            R Wz = dmz(Ssrc.om.data(), k, S.c, N, S.dz);
            R A = dpz(Ssrc.uxx.data(), k, S.c, N, S.dz);
            R B = dmx(Ssrc.uxz.data(), k, Zn, S.c, N, S.dx);
            R C = dmz(Ssrc.uzx.data(), k, S.c, N, S.dz);
            R E = dpx(Ssrc.uzz.data(), k, Zn, S.c, N, S.dx);
            R rhs[4] = {Tx + C - E, Tz + B - A, Wz + C - E, -Wx + B - A};
            for (int q = 0; q < 4; ++q) {
                snew[q][k] = now[q][k];
                now[q][k] = past[q][k];
                past[q][k] = 2 * now[q][k] - snew[q][k] + (dt * dt) * rhs[q];
            }
            Ssrc.vx[k] = past[0][k] + past[2][k];
            Ssrc.vz[k] = past[1][k] + past[3][k];
        }
}

/* reshot_s (P-S-seperating.cu:985-1042): theta, omega, du from Vx_s, Vz_s on
 * the interior, no CPML and no source. */
template <class R>
void recon_stress(const Setup<R> &S, Wave<R> &W, int si, int sj, R s) {
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn, L = g.L;
#pragma omp parallel for schedule(static)
    for (int i = L; i < g.Xn - L; ++i)
        for (int j = L; j < Zn - L; ++j) {
            size_t k = (size_t)i * Zn + j;
            R gxx = dmx(W.vx.data(), k, Zn, S.c, N, S.dx);
            R gxz = dpz(W.vx.data(), k, S.c, N, S.dz);
            R gzz = dmz(W.vz.data(), k, S.c, N, S.dz);
            R gzx = dpx(W.vz.data(), k, Zn, S.c, N, S.dx);
            R vp = S.vp[k], vs = S.vs[k];
            W.th[k] = vp * vp * (gxx + gzz);
            W.om[k] = vs * vs * (gxz - gzx);
            W.uzx[k] = vs * vs * (gzx);
            W.uzz[k] = vs * vs * (gzz);
            W.uxz[k] = vs * vs * (gxz);
            W.uxx[k] = vs * vs * (gxx);
            if (i == si && j == sj) W.th[k] = W.th[k] + s;
        }
}

template <class R>
void copy_wave_stress(const Wave<R> &a, Wave<R> &b) {
    b.vx = a.vx; b.vz = a.vz; b.th = a.th; b.om = a.om;
    b.uxx = a.uxx; b.uxz = a.uxz; b.uzx = a.uzx; b.uzz = a.uzz;
}

/* gradfun (P-S-seperating.cu:1518-1575), one backward step on the interior:
 * derivatives of the source velocities ux, uz with the engine's staggered
 * operators, against the receiver's stress fields of the same step.  The
 * reference hard-codes N = 6 and, in its duxdz sum, pairs f[j+4] with f[j-6]
 * (P-S-seperating.cu:1556); here duxdz is the plain Dp_z of the stress pass.
 * The live physics has no density (P = 1). */
template <class R>
void gradfun_step(const Setup<R> &S, const R *ux, const R *uz, const R *th, const R *om, const R *uxx, const R *uxz,
                  const R *uzx, const R *uzz, R *gvp, R *gvs, R *ill) {
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn, L = g.L;
#pragma omp parallel for schedule(static)
    for (int i = L; i < g.Xn - L; ++i)
        for (int j = L; j < Zn - L; ++j) {
            size_t k = (size_t)i * Zn + j;
            R dxx = dmx(ux, k, Zn, S.c, N, S.dx);
            R dzz = dmz(uz, k, S.c, N, S.dz);
            R dzx = dpx(uz, k, Zn, S.c, N, S.dx);
            R dxz = dpz(ux, k, S.c, N, S.dz);
            R vp = S.vp[k], vs = S.vs[k];
            gvp[k] -= 2 * vp * ((dxx + dzz) * th[k]);
            gvs[k] -= 2 * vs * (om[k] * (dxz - dzx) + uzx[k] * dzx + uzz[k] * dzz + uxz[k] * dxz + uxx[k] * dxx);
            ill[k] += ux[k] * ux[k] + uz[k] * uz[k];
        }
}

/* First-order velocity-stress engine with split PML: `stress`
 * (P-S-seperating.cu:487-555) then `velocity` (P-S-seperating.cu:390-485) per
 * step, never launched by the reference's main().  The damping factors
 * (1 / (1 + dt a / 2)) and (1 - dt a / 2) are evaluated in R (the reference
 * mixes double literals into float expressions).  At N = 6 the coefficients are
 * these kernels' five-digit literals.  Records: the six components at row Zr
 * of every interior column after each velocity pass (the reference's
 * `if (j = Z_receive)` assigns rather than compares, so every thread writes
 * the receiver row). */
template <class R>
void laplace_interior(int nx, int nz, R dh, const R *img, R *lap);

/* One wavefield of the first-order engine: totals, P/S velocities, the
 * stress-like fields and their x / z split parts. */
template <class R>
struct SplitWave {
    std::vector<R> vx, vz, vpx, vpz, vsx, vsz, th, om, thx, thz, omx, omz, uxx, uxz, uzx, uzz;
    void init(size_t n) {
        for (auto *v : {&vx, &vz, &vpx, &vpz, &vsx, &vsz, &th, &om, &thx, &thz, &omx, &omz, &uxx, &uxz, &uzx, &uzz})
            v->assign(n, (R)0);
    }
};

/* `stress` (P-S-seperating.cu:487-555; as stress_backward_propagation,
 * P-S-seperating.cu:815-886, for the receiver run) on [N, Xn-N)^2, adding s
 * at (si, sj) to theta. */
template <class R>
void sp_stress(const Setup<R> &S, const R *c, SplitWave<R> &W, int si, int sj, R s) {
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn;
    const R dt = S.dt, half = (R)0.5, one = (R)1;
#pragma omp parallel for schedule(static)
    for (int i = N; i < g.Xn - N; ++i)
        for (int j = N; j < Zn - N; ++j) {
            size_t k = (size_t)i * Zn + j;
            R dvxdx = dmx(W.vx.data(), k, Zn, c, N, S.dx), dvxdz = dpz(W.vx.data(), k, c, N, S.dz);
            R dvzdz = dmz(W.vz.data(), k, c, N, S.dz), dvzdx = dpx(W.vz.data(), k, Zn, c, N, S.dx);
            R hx = half * dt * S.ax[k], hz = half * dt * S.az[k];
            R ax1 = one / (one + hx), az1 = one / (one + hz), ax2 = one - hx, az2 = one - hz;
            R p2 = S.vp[k] * S.vp[k], q2 = S.vs[k] * S.vs[k];
            W.thx[k] = ax1 * (dt * p2 * dvxdx + ax2 * W.thx[k]);
            W.thz[k] = az1 * (dt * p2 * dvzdz + az2 * W.thz[k]);
            W.omx[k] = ax1 * (-dt * q2 * dvzdx + ax2 * W.omx[k]);
            W.omz[k] = az1 * (dt * q2 * dvxdz + az2 * W.omz[k]);
            W.uxx[k] = ax1 * (dt * q2 * dvxdx + ax2 * W.uxx[k]);
            W.uxz[k] = az1 * (dt * q2 * dvxdz + az2 * W.uxz[k]);
            W.uzx[k] = ax1 * (dt * q2 * dvzdx + ax2 * W.uzx[k]);
            W.uzz[k] = az1 * (dt * q2 * dvzdz + az2 * W.uzz[k]);
            W.th[k] = W.thx[k] + W.thz[k] + ((i == si && j == sj) ? s : (R)0);
            W.om[k] = W.omx[k] + W.omz[k];
        }
}

/* `velocity` (P-S-seperating.cu:390-485; as velocity_backward_propagation,
 * P-S-seperating.cu:728-813, for the receiver run) on [N, Xn-N)^2. */
template <class R>
void sp_velocity(const Setup<R> &S, const R *c, SplitWave<R> &W) {
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn;
    const R dt = S.dt, half = (R)0.5, one = (R)1;
#pragma omp parallel for schedule(static)
    for (int i = N; i < g.Xn - N; ++i)
        for (int j = N; j < Zn - N; ++j) {
            size_t k = (size_t)i * Zn + j;
            R dthdx = dpx(W.th.data(), k, Zn, c, N, S.dx), dthdz = dpz(W.th.data(), k, c, N, S.dz);
            R domdx = dmx(W.om.data(), k, Zn, c, N, S.dx), domdz = dmz(W.om.data(), k, c, N, S.dz);
            R duxxdz = dpz(W.uxx.data(), k, c, N, S.dz), duxzdx = dmx(W.uxz.data(), k, Zn, c, N, S.dx);
            R duzxdz = dmz(W.uzx.data(), k, c, N, S.dz), duzzdx = dpx(W.uzz.data(), k, Zn, c, N, S.dx);
            R hx = half * dt * S.ax[k], hz = half * dt * S.az[k];
            R ax1 = one / (one + hx), az1 = one / (one + hz), ax2 = one - hx, az2 = one - hz;
            W.vpx[k] = ax1 * (ax2 * W.vpx[k] + dt * (dthdx + duzxdz - duzzdx));
            W.vpz[k] = az1 * (az2 * W.vpz[k] + dt * (dthdz + duxzdx - duxxdz));
            W.vsx[k] = ax1 * (ax2 * W.vsx[k] + dt * (domdz + duzxdz - duzzdx));
            W.vsz[k] = az1 * (az2 * W.vsz[k] + dt * (-domdx + duxzdx - duxxdz));
            W.vx[k] = W.vpx[k] + W.vsx[k];
            W.vz[k] = W.vpz[k] + W.vsz[k];
        }
}

/* Reverse of the interior velocity pass, velocity_wavefield_reconstruction
 * (P-S-seperating.cu:1066-1167): on [L, Xn-L)^2, where the damping is zero,
 * V^t = V^{t+1} - dt f(stress^{t+1}). */
template <class R>
void sp_recon_velocity(const Setup<R> &S, const R *c, SplitWave<R> &W) {
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn, L = g.L;
    const R dt = S.dt;
#pragma omp parallel for schedule(static)
    for (int i = L; i < g.Xn - L; ++i)
        for (int j = L; j < Zn - L; ++j) {
            size_t k = (size_t)i * Zn + j;
            R dthdx = dpx(W.th.data(), k, Zn, c, N, S.dx), dthdz = dpz(W.th.data(), k, c, N, S.dz);
            R domdx = dmx(W.om.data(), k, Zn, c, N, S.dx), domdz = dmz(W.om.data(), k, c, N, S.dz);
            R duxxdz = dpz(W.uxx.data(), k, c, N, S.dz), duxzdx = dmx(W.uxz.data(), k, Zn, c, N, S.dx);
            R duzxdz = dmz(W.uzx.data(), k, c, N, S.dz), duzzdx = dpx(W.uzz.data(), k, Zn, c, N, S.dx);
            W.vpx[k] = W.vpx[k] - dt * (dthdx + duzxdz - duzzdx);
            W.vpz[k] = W.vpz[k] - dt * (dthdz + duxzdx - duxxdz);
            W.vsx[k] = W.vsx[k] - dt * (domdz + duzxdz - duzzdx);
            W.vsz[k] = W.vsz[k] - dt * (-domdx + duxzdx - duxxdz);
            W.vx[k] = W.vpx[k] + W.vsx[k];
            W.vz[k] = W.vpz[k] + W.vsz[k];
        }
}

/* Reverse of the interior stress pass, stress_wavefield_reconstruction
 * (P-S-seperating.cu:1168-1229): stress^t = stress^{t+1} - dt g(V^t) on
 * [L, Xn-L)^2; theta^t carries the source sample s of forward step t-1 at
 * (si, sj), as the forward theta^t does. */
template <class R>
void sp_recon_stress(const Setup<R> &S, const R *c, SplitWave<R> &W, int si, int sj, R s) {
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn, L = g.L;
    const R dt = S.dt;
#pragma omp parallel for schedule(static)
    for (int i = L; i < g.Xn - L; ++i)
        for (int j = L; j < Zn - L; ++j) {
            size_t k = (size_t)i * Zn + j;
            R dvxdx = dmx(W.vx.data(), k, Zn, c, N, S.dx), dvxdz = dpz(W.vx.data(), k, c, N, S.dz);
            R dvzdz = dmz(W.vz.data(), k, c, N, S.dz), dvzdx = dpx(W.vz.data(), k, Zn, c, N, S.dx);
            R p2 = S.vp[k] * S.vp[k], q2 = S.vs[k] * S.vs[k];
            W.thx[k] = W.thx[k] - dt * p2 * dvxdx;
            W.thz[k] = W.thz[k] - dt * p2 * dvzdz;
            W.omx[k] = W.omx[k] + dt * q2 * dvzdx;
            W.omz[k] = W.omz[k] - dt * q2 * dvxdz;
            W.uxx[k] = W.uxx[k] - dt * q2 * dvxdx;
            W.uxz[k] = W.uxz[k] - dt * q2 * dvxdz;
            W.uzx[k] = W.uzx[k] - dt * q2 * dvzdx;
            W.uzz[k] = W.uzz[k] - dt * q2 * dvzdz;
            W.th[k] = W.thx[k] + W.thz[k] + ((i == si && j == sj) ? s : (R)0);
            W.om[k] = W.omx[k] + W.omz[k];
        }
}

/* First-order velocity-stress engine with split PML: `stress`
 * (P-S-seperating.cu:487-555) then `velocity` (P-S-seperating.cu:390-485) per
 * step, never launched by the reference's main().  The damping factors
 * (1 / (1 + dt a / 2)) and (1 - dt a / 2) are evaluated in R (the reference
 * mixes double literals into float expressions).  At N = 6 the coefficients are
 * these kernels' five-digit literals.  Records: the six components at row Zr
 * of every interior column after each velocity pass (the reference's
 * `if (j = Z_receive)` assigns rather than compares, so every thread writes
 * the receiver row).
 *
 * RTM (rtm = true) with this engine, which the reference declares but never
 * drives, defined as the live engine's shot loop with aligned time levels:
 *  - the forward run saves, after each step t, the ring (Ring) of vx, vz,
 *    theta, omega and the four du fields (save_wavefiled,
 *    P-S-seperating.cu:1230-1287), exactly;
 *  - the records (vx, vz on row Zr, muted as the live engine mutes them) are
 *    injected by replacement into the receiver run (load_record,
 *    P-S-seperating.cu:1471-1483), which steps with the damped kernels;
 *  - backward step t correlates the source P velocity of level t+1 with the
 *    receiver's P and S velocities after its step t (corr_v,
 *    P-S-seperating.cu:1426-1439), then reverses the source from level t+1 to
 *    t: ring stress of level t+1 restored, the velocity pass reversed, ring
 *    velocities of level t restored (zero at t = 0), the stress pass
 *    reversed.  In exact arithmetic the reconstruction equals the forward run.
 * recon_err (optional, [nt]) receives max |Vpx_recon^t - Vpx_fwd^t| over the
 * interior divided by the largest |Vpx_fwd| (tiny grids: the forward levels
 * are kept). */
template <class R>
int run_splitpml(const ouc_params *p, const ouc_model *m, const ouc_shot *s, double *seis, bool rtm = false,
                 double *img_pp = nullptr, double *img_ps = nullptr, double *lap_pp = nullptr,
                 double *lap_ps = nullptr, double *recon_err = nullptr) {
    Setup<R> S;
    if (!make_setup(p, m, s, S)) return OUC_INVALID_ARG;
    const Geo &g = S.g;
    const int N = g.N, Zn = g.Zn, nt = g.nt, nx = g.nx, nz = g.nz, L = g.L;
    if (s->sx < 0 || s->sx >= nx || s->sz < 0 || s->sz >= g.nz || s->rec_z < 0 || s->rec_z >= g.nz)
        return OUC_INVALID_ARG;
    R c[8];
    const float a6[6] = {+1.2213e+0f, -9.6931e-2f, 1.7448e-2f, -2.9673e-3f, +3.5901e-4f, -2.1848e-5f};
    for (int k = 0; k < N; ++k) c[k] = N == 6 ? (R)a6[k] : S.c[k];
    const int Sx = s->sx + L, Sz = s->sz + L, Zr = s->rec_z + L;
    const size_t n = (size_t)g.Xn * Zn;
    SplitWave<R> F;
    F.init(n);
    Ring ring;
    ring.init(g);
    std::vector<R> store;
    if (rtm) store.assign((size_t)nt * RS_NF * ring.ncells, (R)0);
    auto fields = [](SplitWave<R> &W) {
        return std::vector<std::vector<R> *>{&W.vx, &W.vz, &W.th, &W.om, &W.uzx, &W.uzz, &W.uxx, &W.uxz};
    };
    std::vector<std::vector<R>> rec(2, std::vector<R>((size_t)nx * nt, (R)0));
    std::vector<std::vector<R>> fwd_px;
    if (rtm && recon_err) fwd_px.assign(nt + 1, std::vector<R>());
    if (rtm && recon_err) fwd_px[0].assign(n, (R)0);
    for (int t = 0; t < nt; ++t) {
        sp_stress(S, c, F, Sx, Sz, S.src[t]);
        sp_velocity(S, c, F);
        for (int ix = 0; ix < nx; ++ix) {
            size_t k = (size_t)(ix + L) * Zn + Zr;
            rec[0][(size_t)ix * nt + t] = F.vx[k];
            rec[1][(size_t)ix * nt + t] = F.vz[k];
            if (seis) {
                const R *f[6] = {F.vx.data(), F.vz.data(), F.vpx.data(), F.vpz.data(), F.vsx.data(), F.vsz.data()};
                for (int q = 0; q < 6; ++q) seis[((size_t)q * nx + ix) * nt + t] = (double)f[q][k];
            }
        }
        if (rtm) {
            auto fs = fields(F);
            for (int f = 0; f < RS_NF; ++f)
                for (size_t r = 0; r < ring.ncells; ++r) {
                    int i, j;
                    ring.cell(r, i, j);
                    store[((size_t)t * RS_NF + f) * ring.ncells + r] = (*fs[f])[(size_t)i * Zn + j];
                }
            if (recon_err) fwd_px[t + 1] = F.vpx;
        }
    }
    if (!rtm) return OUC_OK;
    if (s->mute) {
        int t0 = ricker_t0(s->fm, p->dt);
        float vsrc = (float)S.vp_d[(size_t)Sx * Zn + Zr];
        for (int ix = 0; ix < nx; ++ix) {
            double thr = mute_threshold(Sx, Sz, Zr, ix + L, t0, p->dt, p->dx, vsrc);
            for (int t = 0; t < nt; ++t)
                if (t < thr) { rec[0][(size_t)ix * nt + t] = 0; rec[1][(size_t)ix * nt + t] = 0; }
        }
    }
    auto restore = [&](SplitWave<R> &W, int slot, int f0, int f1) {
        auto fs = fields(W);
        for (int f = f0; f < f1; ++f)
            for (size_t r = 0; r < ring.ncells; ++r) {
                int i, j;
                ring.cell(r, i, j);
                (*fs[f])[(size_t)i * Zn + j] = slot >= 0 ? store[((size_t)slot * RS_NF + f) * ring.ncells + r] : (R)0;
            }
    };
    SplitWave<R> Rw;
    Rw.init(n);
    std::vector<R> num_pp(n, (R)0), num_ps(n, (R)0), den(n, (R)0);
    double pmax = 0;
    if (recon_err) {
        for (int t = 1; t <= nt; ++t)
            for (int i = L; i < g.Xn - L; ++i)
                for (int j = L; j < Zn - L; ++j)
                    pmax = std::max(pmax, std::fabs((double)fwd_px[t][(size_t)i * Zn + j]));
        if (pmax == 0) pmax = 1;
        recon_err[nt - 1 < 0 ? 0 : nt - 1] = 0;
    }
    for (int t = nt - 1; t >= 0; --t) {
        for (int ix = 0; ix < nx; ++ix) { /* load_record */
            size_t k = (size_t)(ix + L) * Zn + Zr;
            Rw.vx[k] = rec[0][(size_t)ix * nt + t];
            Rw.vz[k] = rec[1][(size_t)ix * nt + t];
        }
        sp_stress(S, c, Rw, -1, -1, (R)0);
        sp_velocity(S, c, Rw);
#pragma omp parallel for schedule(static)
        for (int i = L; i < g.Xn - L; ++i)
            for (int j = L; j < Zn - L; ++j) {
                size_t k = (size_t)i * Zn + j;
                R spx = F.vpx[k], spz = F.vpz[k];
                num_pp[k] += spx * Rw.vpx[k] + spz * Rw.vpz[k];
                den[k] += spx * spx + spz * spz;
                num_ps[k] += spx * Rw.vsx[k] + spz * Rw.vsz[k];
            }
        /* the source from level t+1 to level t */
        restore(F, t, RS_TH, RS_NF);
        sp_recon_velocity(S, c, F);
        restore(F, g_mutation == 13 ? t : t - 1, RS_VX, RS_VZ + 1);
        sp_recon_stress(S, c, F, Sx, Sz, (t >= 1 && g_mutation != 14) ? S.src[t - 1] : (R)0);
        if (recon_err) {
            double e = 0;
            for (int i = L; i < g.Xn - L; ++i)
                for (int j = L; j < Zn - L; ++j) {
                    size_t k = (size_t)i * Zn + j;
                    e = std::max(e, std::fabs((double)F.vpx[k] - (double)fwd_px[t][k]));
                }
            recon_err[t] = e / pmax;
        }
    }
    std::vector<R> cpp((size_t)nx * nz), cps((size_t)nx * nz), lpp((size_t)nx * nz), lps((size_t)nx * nz);
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz) {
            const size_t k = (size_t)(ix + L) * Zn + iz + L, oi = (size_t)ix * nz + iz;
            cpp[oi] = num_pp[k] / den[k];
            cps[oi] = num_ps[k] / den[k];
        }
    laplace_interior(nx, nz, S.dh, cpp.data(), lpp.data());
    laplace_interior(nx, nz, S.dh, cps.data(), lps.data());
    for (size_t oi = 0; oi < (size_t)nx * nz; ++oi) {
        if (img_pp) img_pp[oi] = (double)cpp[oi];
        if (img_ps) img_ps[oi] = (double)cps[oi];
        if (lap_pp) lap_pp[oi] = (double)lpp[oi];
        if (lap_ps) lap_ps[oi] = (double)lps[oi];
    }
    return OUC_OK;
}

struct Outputs {
    double *seis = nullptr;      /* [6][nx][nt] */
    double *rec_out = nullptr;   /* [2][nx][nt] records used for RTM (after the mute) */
    double *img_pp = nullptr, *img_ps = nullptr, *lap_pp = nullptr, *lap_ps = nullptr; /* nx*nz */
    double *recon_err = nullptr; /* [nt]: max |V_recon - V_fwd| / max |V_fwd| per step */
    /* FWI: observed records [2][nx][nt]; the receiver wavefield is driven by
     * the residual (modelled - observed, no mute) and the gradient of
     * gradfun accumulates over every backward step into nx*nz arrays */
    const double *observed = nullptr;
    double *grad_vp = nullptr, *grad_vs = nullptr, *illum = nullptr, *misfit = nullptr;
    /* traces of the backward loop (tests on tiny grids), interior arrays x-major:
     * fwd_v [nt+1][4][nx][nz] forward source (Vpx, Vpz, Vsx, Vsz) at levels 0..nt;
     * rcv [nt][4][nx][nz] receiver (Vpx, Vpz, Vsx, Vsz) after the velocity pass
     * of backward step t; acc [3][nx][nz] num_pp, num_ps, den;
     * inj [nt][2][Xn] receiver (Vx, Vz) on row Zr right after load_record */
    double *fwd_v = nullptr, *rcv = nullptr, *acc = nullptr, *inj = nullptr;
    /* src_v [nt+1][2][nx][nz]: the source P velocity (Vpx, Vpz) level by level as
     * the backward loop holds it: levels nt and nt-1 handed over from the
     * forward run, level t <= nt-2 produced by the reconstruction of step t */
    double *src_v = nullptr;
};

/* Laplace (P-S-seperating.cu:1457-1469) on interior-only nx x nz arrays
 * (x-major): the 5-point Laplacian / dh^2 on [2, nx-2) x [2, nz-2), zero on the
 * 2-cell border (the reference never writes it). */
template <class R>
void laplace_interior(int nx, int nz, R dh, const R *img, R *lap) {
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) {
            const size_t k = (size_t)i * nz + j;
            R v = 0;
            if (i >= 2 && i < nx - 2 && j >= 2 && j < nz - 2)
                v = (1 / dh / dh) * (img[k + nz] + img[k - nz] - 2 * img[k] + img[k + 1] + img[k - 1] - 2 * img[k]);
            lap[k] = v;
        }
}

/* The whole shot: forward loop (P-S-seperating.cu:2258-2292), mute
 * (P-S-seperating.cu:2293), backward loop (P-S-seperating.cu:2313-2364) and
 * imaging (P-S-seperating.cu:2367-2373). */
template <class R>
int run_shot(const ouc_params *p, const ouc_model *m, const ouc_shot *s, const double *rec_in, bool rtm,
             Outputs &o) {
    Setup<R> S;
    if (!make_setup(p, m, s, S)) return OUC_INVALID_ARG;
    const Geo &g = S.g;
    const int L = g.L, Zn = g.Zn, nt = g.nt, nx = g.nx, nz = g.nz;
    if (s->sx < 0 || s->sx >= nx || s->sz < 0 || s->sz >= nz || s->rec_z < 0 || s->rec_z >= nz)
        return OUC_INVALID_ARG;
    const int Sx = s->sx + L, Sz = s->sz + L, Zr = s->rec_z + L;
    const size_t n = (size_t)g.Xn * Zn;
    const bool exact = p->semantics == OUC_SEM_EXACT;

    Wave<R> F;
    F.init(n);
    Strips<R> strips;
    if (rtm) strips.init(g, std::max(nt - 1, 1), p->strip_precision == OUC_STRIP_FP16);

    std::vector<R> rec[2];
    rec[0].assign((size_t)nx * nt, (R)0);
    rec[1].assign((size_t)nx * nt, (R)0);
    std::vector<std::vector<R>> snaps; /* forward V snapshots for the reversibility check */
    if (o.recon_err) snaps.resize(nt + 1);
    Wave<R> Ssrc;
    if (rtm) Ssrc.init(n);

    for (int t = 0; t < nt; ++t) {
        stress_cpml(S, F, Sx, Sz, S.src[t]);
        velocity_cpml(S, F);
        for (int ix = 0; ix < nx; ++ix) {
            size_t k = (size_t)(ix + L) * Zn + Zr;
            rec[0][(size_t)ix * nt + t] = F.vx[k];
            rec[1][(size_t)ix * nt + t] = F.vz[k];
            if (o.seis) {
                o.seis[((size_t)0 * nx + ix) * nt + t] = (double)F.vx[k];
                o.seis[((size_t)1 * nx + ix) * nt + t] = (double)F.vz[k];
                o.seis[((size_t)2 * nx + ix) * nt + t] = (double)F.cur[0][k];
                o.seis[((size_t)3 * nx + ix) * nt + t] = (double)F.cur[1][k];
                o.seis[((size_t)4 * nx + ix) * nt + t] = (double)F.cur[2][k];
                o.seis[((size_t)5 * nx + ix) * nt + t] = (double)F.cur[3][k];
            }
        }
        if (o.recon_err) { /* px at step t+1 */
            snaps[t + 1] = F.cur[0];
        }
        if (o.fwd_v)
            for (int q = 0; q < 4; ++q)
                for (int ix = 0; ix < nx; ++ix)
                    for (int iz = 0; iz < nz; ++iz)
                        o.fwd_v[(((size_t)(t + 1) * 4 + q) * nx + ix) * nz + iz] =
                            (double)F.cur[q][(size_t)(ix + L) * Zn + iz + L];
        if (rtm) {
            if (t != nt - 1) strips.save(t, F);
            else copy_wave_stress(F, Ssrc); /* read_last_wavefiled2, P-S-seperating.cu:1318-1342 */
        }
    }

    const bool fwi = o.observed != nullptr;
    if (fwi) { /* adsource (P-S-seperating.cu:1484-1498): residual on every receiver and step */
        double m2 = 0;
        for (int q = 0; q < 2; ++q)
            for (size_t k = 0; k < (size_t)nx * nt; ++k) {
                rec[q][k] = rec[q][k] - (R)o.observed[(size_t)q * nx * nt + k];
                m2 += (double)rec[q][k] * (double)rec[q][k];
            }
        if (o.misfit) *o.misfit = 0.5 * m2;
    }
    if (s->mute && !fwi) {
        int t0 = ricker_t0(s->fm, p->dt);
        float vsrc = (float)S.vp_d[(size_t)Sx * Zn + Zr];
        for (int ix = 0; ix < nx; ++ix) {
            double thr = mute_threshold(Sx, Sz, Zr, ix + L, t0, p->dt, p->dx, vsrc);
            for (int t = 0; t < nt; ++t)
                if (t < thr) { rec[0][(size_t)ix * nt + t] = 0; rec[1][(size_t)ix * nt + t] = 0; }
        }
        if (o.seis)
            for (int q = 0; q < 6; ++q)
                for (int ix = 0; ix < nx; ++ix) {
                    double thr = mute_threshold(Sx, Sz, Zr, ix + L, t0, p->dt, p->dx, vsrc);
                    for (int t = 0; t < nt; ++t)
                        if (t < thr) o.seis[((size_t)q * nx + ix) * nt + t] = 0.0;
                }
    }
    if (rec_in) {
        for (int q = 0; q < 2; ++q)
            for (size_t k = 0; k < (size_t)nx * nt; ++k) rec[q][k] = (R)rec_in[(size_t)q * nx * nt + k];
    }
    if (o.rec_out)
        for (int q = 0; q < 2; ++q)
            for (size_t k = 0; k < (size_t)nx * nt; ++k) o.rec_out[(size_t)q * nx * nt + k] = (double)rec[q][k];
    if (!rtm) return OUC_OK;

    /* backward loop */
    Wave<R> Rw;
    Rw.init(n);
    std::vector<R> snew[4], now[4], past[4];
    for (int q = 0; q < 4; ++q) { snew[q] = F.cur[q]; now[q] = F.cur[q]; past[q] = F.prev[q]; }
    std::vector<R> num_pp(n, (R)0), num_ps(n, (R)0), den(n, (R)0);
    auto trace_src = [&](int lev, const std::vector<R> *v) {
        if (!o.src_v) return;
        for (int q = 0; q < 2; ++q)
            for (int ix = 0; ix < nx; ++ix)
                for (int iz = 0; iz < nz; ++iz)
                    o.src_v[(((size_t)lev * 2 + q) * nx + ix) * nz + iz] = (double)v[q][(size_t)(ix + L) * Zn + iz + L];
    };
    trace_src(nt, now);
    if (nt >= 1) trace_src(nt - 1, past);
    /* FWI: source totals at level t+1 for backward step t (V^nt first), and
     * the gradient accumulators */
    std::vector<R> u1[2], gvp, gvs, ill;
    if (fwi) {
        u1[0] = F.vx; u1[1] = F.vz;
        gvp.assign(n, (R)0); gvs.assign(n, (R)0); ill.assign(n, (R)0);
    }
    double vmax_snap = 0;
    if (o.recon_err) {
        for (int t = 1; t <= nt; ++t)
            for (int i = L; i < g.Xn - L; ++i)
                for (int j = L; j < Zn - L; ++j)
                    vmax_snap = std::max(vmax_snap, std::fabs((double)snaps[t][(size_t)i * Zn + j]));
        if (vmax_snap == 0) vmax_snap = 1;
    }
    for (int t = nt - 1; t >= 0; --t) {
        if (t < nt - 1) {
            recon_velocity(S, Ssrc, snew, now, past);
            trace_src(t, past);
            if (!exact) {
                std::vector<R> *dv[2] = {&Ssrc.vx, &Ssrc.vz};
                strips.restore(t, RS_VX, RS_VZ + 1, dv);
            } else if (t >= 1) {
                std::vector<R> *dv[2] = {&Ssrc.vx, &Ssrc.vz};
                strips.restore(t - 1, RS_VX, RS_VZ + 1, dv);
            } else {
                strips.zero_v(Ssrc.vx, Ssrc.vz);
            }
            if (exact) recon_stress(S, Ssrc, Sx, Sz, S.src[t]);
            else recon_stress(S, Ssrc, -1, -1, (R)0);
            std::vector<R> *dt6[6] = {&Ssrc.th, &Ssrc.om, &Ssrc.uzx, &Ssrc.uzz, &Ssrc.uxx, &Ssrc.uxz};
            strips.restore(t, RS_TH, RS_UXZ + 1, dt6);
            if (o.recon_err) { /* past holds V^t */
                double e = 0;
                for (int i = L; i < g.Xn - L; ++i)
                    for (int j = L; j < Zn - L; ++j) {
                        size_t k = (size_t)i * Zn + j;
                        double ref = t > 0 ? (double)snaps[t][k] : 0.0;
                        e = std::max(e, std::fabs((double)past[0][k] - ref));
                    }
                o.recon_err[t] = e / vmax_snap;
            }
        }
        /* load_record (P-S-seperating.cu:1471-1483): replacement at row Zr */
        for (int ix = 0; ix < nx; ++ix) {
            size_t k = (size_t)(ix + L) * Zn + Zr;
            Rw.vx[k] = rec[0][(size_t)ix * nt + t];
            Rw.vz[k] = rec[1][(size_t)ix * nt + t];
        }
        if (o.inj)
            for (int i = 0; i < g.Xn; ++i) {
                o.inj[((size_t)t * 2 + 0) * g.Xn + i] = (double)Rw.vx[(size_t)i * Zn + Zr];
                o.inj[((size_t)t * 2 + 1) * g.Xn + i] = (double)Rw.vz[(size_t)i * Zn + Zr];
            }
        stress_cpml(S, Rw, -1, -1, (R)0);
        if (fwi) {
            gradfun_step(S, u1[0].data(), u1[1].data(), Rw.th.data(), Rw.om.data(), Rw.uxx.data(), Rw.uxz.data(),
                         Rw.uzx.data(), Rw.uzz.data(), gvp.data(), gvs.data(), ill.data());
            if (t == nt - 1) { /* level nt-1: the forward's own totals */
                for (size_t k = 0; k < n; ++k) {
                    u1[0][k] = F.prev[0][k] + F.prev[2][k];
                    u1[1][k] = F.prev[1][k] + F.prev[3][k];
                }
            } else { /* level t: the reconstructed totals with the ring restored */
                u1[0] = Ssrc.vx;
                u1[1] = Ssrc.vz;
            }
        }
        velocity_cpml(S, Rw);
        /* corr_v (P-S-seperating.cu:1426-1439) */
        const std::vector<R> *sp = exact ? now : snew;
#pragma omp parallel for schedule(static)
        for (int i = L; i < g.Xn - L; ++i)
            for (int j = L; j < Zn - L; ++j) {
                size_t k = (size_t)i * Zn + j;
                R spx = sp[0][k], spz = sp[1][k];
                num_pp[k] += spx * Rw.cur[0][k] + spz * Rw.cur[1][k];
                den[k] += spx * spx + spz * spz;
                num_ps[k] += spx * Rw.cur[2][k] + spz * Rw.cur[3][k];
            }
        if (o.rcv)
            for (int q = 0; q < 4; ++q)
                for (int ix = 0; ix < nx; ++ix)
                    for (int iz = 0; iz < nz; ++iz)
                        o.rcv[(((size_t)t * 4 + q) * nx + ix) * nz + iz] = (double)Rw.cur[q][(size_t)(ix + L) * Zn + iz + L];
    }
    if (o.acc)
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz) {
                const size_t k = (size_t)(ix + L) * Zn + iz + L, oi = (size_t)ix * nz + iz;
                o.acc[oi] = (double)num_pp[k];
                o.acc[(size_t)nx * nz + oi] = (double)num_ps[k];
                o.acc[(size_t)2 * nx * nz + oi] = (double)den[k];
            }
    if (fwi)
        for (int i = L; i < g.Xn - L; ++i)
            for (int j = L; j < Zn - L; ++j) {
                size_t k = (size_t)i * Zn + j, oi = (size_t)(i - L) * nz + (j - L);
                if (o.grad_vp) o.grad_vp[oi] = (double)gvp[k];
                if (o.grad_vs) o.grad_vs[oi] = (double)gvs[k];
                if (o.illum) o.illum[oi] = (double)ill[k];
            }
    /* image_fun (P-S-seperating.cu:1441-1455) and Laplace (P-S-seperating.cu:1457-1469) */
    std::vector<R> ipp(n, (R)0), ips(n, (R)0);
    for (int i = L; i < g.Xn - L; ++i)
        for (int j = L; j < Zn - L; ++j) {
            size_t k = (size_t)i * Zn + j;
            ipp[k] = num_pp[k] / den[k];
            ips[k] = num_ps[k] / den[k];
        }
    const R dh = S.dh;
    std::vector<R> cpp((size_t)nx * nz), cps((size_t)nx * nz), lpp((size_t)nx * nz), lps((size_t)nx * nz);
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz) {
            const size_t k = (size_t)(ix + L) * Zn + iz + L, oi = (size_t)ix * nz + iz;
            cpp[oi] = ipp[k];
            cps[oi] = ips[k];
        }
    laplace_interior(nx, nz, dh, cpp.data(), lpp.data());
    laplace_interior(nx, nz, dh, cps.data(), lps.data());
    for (size_t oi = 0; oi < (size_t)nx * nz; ++oi) {
        if (o.img_pp) o.img_pp[oi] = (double)cpp[oi];
        if (o.img_ps) o.img_ps[oi] = (double)cps[oi];
        if (o.lap_pp) o.lap_pp[oi] = (double)lpp[oi];
        if (o.lap_ps) o.lap_ps[oi] = (double)lps[oi];
    }
    return OUC_OK;
}

} // namespace

extern "C" {

/* ---- small pieces, exposed for the closed-form and brute-force tests ---- */

int oracle_coeffs(int N, double *out) {
    if (N < 1 || N > 8 || !out) return OUC_INVALID_ARG;
    coeffs_float(N, out);
    return OUC_OK;
}

int oracle_ricker(int nt, float fm, float dt, double *out, int *t0) {
    if (nt < 1 || !out) return OUC_INVALID_ARG;
    std::vector<double> s;
    ricker(nt, fm, dt, s);
    std::copy(s.begin(), s.end(), out);
    if (t0) *t0 = ricker_t0(fm, dt);
    return OUC_OK;
}

/* Padded vp, vs, ax, az (Xn*Zn each, double). */
int oracle_setup(const ouc_params *p, const ouc_model *m, double *vp, double *vs, double *ax, double *az) {
    Setup<double> S;
    if (!make_setup(p, m, nullptr, S)) return OUC_INVALID_ARG;
    size_t n = (size_t)S.g.Xn * S.g.Zn;
    if (vp) std::copy(S.vp.begin(), S.vp.end(), vp);
    if (vs) std::copy(S.vs.begin(), S.vs.end(), vs);
    if (ax) std::copy(S.ax.begin(), S.ax.begin() + n, ax);
    if (az) std::copy(S.az.begin(), S.az.begin() + n, az);
    return OUC_OK;
}

/* Apply one staggered operator (0 Dm_x, 1 Dp_x, 2 Dm_z, 3 Dp_z) to f on an
 * Xn x Zn grid over [N, Xn-N) x [N, Zn-N). */
int oracle_dop(int N, int kind, const double *f, int Xn, int Zn, double h, double *out) {
    if (N < 1 || N > 8) return OUC_INVALID_ARG;
    double c[8];
    coeffs_float(N, c);
    for (int i = N; i < Xn - N; ++i)
        for (int j = N; j < Zn - N; ++j) {
            size_t k = (size_t)i * Zn + j;
            double v = 0;
            switch (kind) {
            case 0: v = dmx(f, k, Zn, c, N, h); break;
            case 1: v = dpx(f, k, Zn, c, N, h); break;
            case 2: v = dmz(f, k, c, N, h); break;
            default: v = dpz(f, k, c, N, h); break;
            }
            out[k] = v;
        }
    return OUC_OK;
}

int oracle_mute_thresholds(const ouc_params *p, const ouc_model *m, const ouc_shot *s, double *thr) {
    Setup<double> S;
    if (!make_setup(p, m, s, S)) return OUC_INVALID_ARG;
    int L = S.g.L, Zn = S.g.Zn;
    int Sx = s->sx + L, Sz = s->sz + L, Zr = s->rec_z + L;
    int t0 = ricker_t0(s->fm, p->dt);
    float vsrc = (float)S.vp_d[(size_t)Sx * Zn + Zr];
    for (int ix = 0; ix < p->nx; ++ix) thr[ix] = mute_threshold(Sx, Sz, Zr, ix + L, t0, p->dt, p->dx, vsrc);
    return OUC_OK;
}

/* Ring cell coordinates (padded i, j) in storage order. */
int oracle_ring(const ouc_params *p, int *ij, int64_t *ncells) {
    Geo g{p->nx, p->nz, p->npml, p->half_order, p->nx + 2 * p->npml, p->nz + 2 * p->npml, p->nt};
    Ring r;
    r.init(g);
    if (ncells) *ncells = (int64_t)r.ncells;
    if (ij)
        for (size_t k = 0; k < r.ncells; ++k) r.cell(k, ij[2 * k], ij[2 * k + 1]);
    return OUC_OK;
}

/* Strip codec: `steps` slots of one field, x as [steps][n], coded in order
 * with the predictive FP16 codec (the last slot is a keyframe).  h is
 * [steps][n] residual bits, e is [steps][ceil(n/64)] exponents; decoded
 * (optional) is [steps][n], decoded backwards as the reconstruction does. */
int oracle_fp16_encode(const float *x, int64_t n, int32_t steps, uint16_t *h, int8_t *e, float *decoded) {
    if (!x || n < 1 || steps < 1 || !h || !e) return OUC_INVALID_ARG;
    const int64_t nch = (n + 63) / 64;
    Codec16 c;
    c.init((size_t)n, steps, 1, std::vector<int>(1, steps - 1));
    for (int32_t t = 0; t < steps; ++t) c.encode(t, 0, x + (size_t)t * n);
    std::copy(c.h.begin(), c.h.end(), h);
    std::copy(c.e.begin(), c.e.begin() + (size_t)steps * nch, e);
    if (decoded)
        for (int32_t t = steps - 1; t >= 0; --t) {
            const float *d = c.decoded(t, 0);
            std::copy(d, d + n, decoded + (size_t)t * n);
        }
    return OUC_OK;
}
int oracle_fp16_decode(const uint16_t *h, const int8_t *e, int64_t n, float *x) {
    for (int64_t c0 = 0, ci = 0; c0 < n; c0 += 64, ++ci) {
        int k = (int)std::min<int64_t>(64, n - c0);
        decode_chunk(h + c0, k, e[ci], x + c0);
    }
    return OUC_OK;
}
int oracle_f32_to_f16(const float *x, int64_t n, uint16_t *h) {
    for (int64_t k = 0; k < n; ++k) h[k] = f32_to_f16_rne(x[k]);
    return OUC_OK;
}

/* ---- whole-path entry points, precision 64 (oracle) or 32 (baseline) ---- */

int oracle_forward_shot(const ouc_params *p, const ouc_model *m, const ouc_shot *s, int precision, double *seis) {
    if (!p || !m || !s) return OUC_INVALID_ARG;
    Outputs o;
    o.seis = seis;
    return precision == 32 ? run_shot<float>(p, m, s, nullptr, false, o) : run_shot<double>(p, m, s, nullptr, false, o);
}

int oracle_rtm_shot(const ouc_params *p, const ouc_model *m, const ouc_shot *s, int precision, const double *rec_in,
                    double *img_pp, double *img_ps, double *lap_pp, double *lap_ps, double *rec_out) {
    if (!p || !m || !s) return OUC_INVALID_ARG;
    Outputs o;
    o.img_pp = img_pp; o.img_ps = img_ps; o.lap_pp = lap_pp; o.lap_ps = lap_ps; o.rec_out = rec_out;
    return precision == 32 ? run_shot<float>(p, m, s, rec_in, true, o) : run_shot<double>(p, m, s, rec_in, true, o);
}

/* RTM of one shot with the backward-loop traces (see Outputs); any trace
 * pointer may be null.  Meant for tiny grids. */
int oracle_rtm_trace(const ouc_params *p, const ouc_model *m, const ouc_shot *s, int precision, const double *rec_in,
                     double *img_pp, double *img_ps, double *fwd_v, double *src_v, double *rcv, double *acc,
                     double *inj) {
    if (!p || !m || !s) return OUC_INVALID_ARG;
    Outputs o;
    o.img_pp = img_pp; o.img_ps = img_ps;
    o.fwd_v = fwd_v; o.src_v = src_v; o.rcv = rcv; o.acc = acc; o.inj = inj;
    if (fwd_v) std::fill(fwd_v, fwd_v + (size_t)4 * p->nx * p->nz, 0.0); /* level 0 */
    return precision == 32 ? run_shot<float>(p, m, s, rec_in, true, o) : run_shot<double>(p, m, s, rec_in, true, o);
}

/* Forward modelling only, with the interior split velocities of every level:
 * fwd_v [nt+1][4][nx][nz] (Vpx, Vpz, Vsx, Vsz; level 0 is zero). */
int oracle_forward_trace(const ouc_params *p, const ouc_model *m, const ouc_shot *s, int precision, double *fwd_v) {
    if (!p || !m || !s || !fwd_v) return OUC_INVALID_ARG;
    Outputs o;
    o.fwd_v = fwd_v;
    std::fill(fwd_v, fwd_v + (size_t)4 * p->nx * p->nz, 0.0);
    return precision == 32 ? run_shot<float>(p, m, s, nullptr, false, o) : run_shot<double>(p, m, s, nullptr, false, o);
}

/* Laplace (P-S-seperating.cu:1457-1469) of an interior image, FP64. */
int oracle_laplace(int nx, int nz, double dh, const double *img, double *lap) {
    if (nx <= 0 || nz <= 0 || !img || !lap || !(dh > 0)) return OUC_INVALID_ARG;
    laplace_interior(nx, nz, dh, img, lap);
    return OUC_OK;
}

/* Planted-mistake switch of the CPML memory-variable updates (tests only). */
int oracle_set_mutation(int u) {
    if (u < 0 || u > 14) return OUC_INVALID_ARG;
    g_mutation = u;
    return OUC_OK;
}

/* Split-PML first-order engine, forward seismograms [6][nx][nt]. */
int oracle_forward_splitpml(const ouc_params *p, const ouc_model *m, const ouc_shot *s, int precision, double *seis) {
    if (!p || !m || !s) return OUC_INVALID_ARG;
    return precision == 32 ? run_splitpml<float>(p, m, s, seis) : run_splitpml<double>(p, m, s, seis);
}

/* RTM with the first-order split-PML engine (run_splitpml): images, their
 * Laplacians and, optionally, the reconstruction's error per level. */
int oracle_rtm_splitpml(const ouc_params *p, const ouc_model *m, const ouc_shot *s, int precision, double *img_pp,
                        double *img_ps, double *lap_pp, double *lap_ps, double *recon_err) {
    if (!p || !m || !s) return OUC_INVALID_ARG;
    return precision == 32
               ? run_splitpml<float>(p, m, s, nullptr, true, img_pp, img_ps, lap_pp, lap_ps, recon_err)
               : run_splitpml<double>(p, m, s, nullptr, true, img_pp, img_ps, lap_pp, lap_ps, recon_err);
}

/* FWI gradient of one shot: observed records [2][nx][nt]; grad_vp, grad_vs,
 * illum are nx*nz (x-major); misfit = 0.5 * sum of squared residuals. */
int oracle_fwi_gradient(const ouc_params *p, const ouc_model *m, const ouc_shot *s, int precision,
                        const double *observed, double *grad_vp, double *grad_vs, double *illum, double *misfit) {
    if (!p || !m || !s || !observed) return OUC_INVALID_ARG;
    Outputs o;
    o.observed = observed;
    o.grad_vp = grad_vp; o.grad_vs = grad_vs; o.illum = illum; o.misfit = misfit;
    return precision == 32 ? run_shot<float>(p, m, s, nullptr, true, o) : run_shot<double>(p, m, s, nullptr, true, o);
}

/* One gradfun step on given padded Xn x Zn fields (FP64); the accumulators
 * are updated in place over the interior. */
int oracle_gradfun(const ouc_params *p, const ouc_model *m, const double *ux, const double *uz, const double *th,
                   const double *om, const double *uxx, const double *uxz, const double *uzx, const double *uzz,
                   double *gvp, double *gvs, double *ill) {
    Setup<double> S;
    if (!make_setup(p, m, nullptr, S)) return OUC_INVALID_ARG;
    gradfun_step(S, ux, uz, th, om, uxx, uxz, uzx, uzz, gvp, gvs, ill);
    return OUC_OK;
}

/* Reversibility: per backward step, max |reconstructed Vpx^t - forward Vpx^t|
 * over the interior divided by the forward maximum. */
int oracle_recon_error(const ouc_params *p, const ouc_model *m, const ouc_shot *s, int precision, double *err) {
    if (!p || !m || !s || !err) return OUC_INVALID_ARG;
    Outputs o;
    o.recon_err = err;
    for (int t = 0; t < p->nt; ++t) err[t] = 0;
    return precision == 32 ? run_shot<float>(p, m, s, nullptr, true, o) : run_shot<double>(p, m, s, nullptr, true, o);
}

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

} // extern "C"
