#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace ouc {

std::string make_grid(const ouc_params *p, Grid &g) {
    if (!p) return "params is NULL";
    if (p->half_order < 1 || p->half_order > 8) return "half_order must be in 1..8";
    if (p->nx < 2 * p->half_order || p->nz < 2 * p->half_order) return "interior must be at least 2*half_order cells";
    if (p->npml < p->half_order) return "npml must be >= half_order";
    if (p->nt < 2) return "nt must be >= 2";
    if (!(p->dx > 0) || !(p->dz > 0) || !(p->dt > 0)) return "dx, dz, dt must be positive";
    if (!(p->R > 0) || !(p->R < 1)) return "R must be in (0, 1)";
    if (p->semantics != OUC_SEM_FAITHFUL && p->semantics != OUC_SEM_EXACT) return "unknown semantics";
    if (p->strip_precision != OUC_STRIP_FP32 && p->strip_precision != OUC_STRIP_FP16) return "unknown strip precision";
    if (p->laplacian < OUC_LAP_STENCIL || p->laplacian > OUC_LAP_FFT_CONT) return "unknown laplacian";
    if (p->engine != OUC_ENGINE_PLAIN && p->engine != OUC_ENGINE_FUSED && p->engine != OUC_ENGINE_SPLIT_PML)
        return "unknown engine";
    if (p->boundary != OUC_BOUNDARY_CPML && p->boundary != OUC_BOUNDARY_LEARNED &&
        p->boundary != OUC_BOUNDARY_LEARNED_FALLBACK)
        return "unknown boundary";
    int pad = p->npml;
    if (p->boundary != OUC_BOUNDARY_CPML) {
        if (p->engine != OUC_ENGINE_FUSED) return "the learned boundary runs on the fused engine";
        pad = p->band > 0 ? p->band : 8;
        if (pad < p->half_order || pad < 4) return "band must be >= max(half_order, 4)";
    }
    int64_t xn = (int64_t)p->nx + 2 * pad, zn = (int64_t)p->nz + 2 * pad;
    if (xn * zn > (int64_t)1 << 31) return "grid too large";
    g.nx = p->nx; g.nz = p->nz; g.L = pad; g.N = p->half_order;
    g.learned = p->boundary != OUC_BOUNDARY_CPML;
    g.Xn = (int)xn; g.Zn = (int)zn;
    g.P = (g.Zn + 31) / 32 * 32;
    g.nt = p->nt;
    g.dx = p->dx; g.dz = p->dz; g.dt = p->dt;
    return "";
}

void stencil_coeffs(int N, float *c) {
    if (N == 6) {
        const float a[6] = {1.2213364f, -9.6931458e-2f, 1.7447662e-2f, -2.9672895e-3f, 3.5900540e-4f, -2.1847812e-5f};
        for (int k = 0; k < 6; ++k) c[k] = a[k];
        return;
    }
    // c_k = (-1)^{k+1}/(2k-1) prod_{m!=k} (2m-1)^2 / |(2k-1)^2 - (2m-1)^2|
    for (int k = 1; k <= N; ++k) {
        double num = 1.0, den = 2.0 * k - 1.0;
        for (int m = 1; m <= N; ++m) {
            if (m == k) continue;
            double a = (2.0 * m - 1) * (2.0 * m - 1), b = (2.0 * k - 1) * (2.0 * k - 1);
            num *= a;
            den *= std::fabs(b - a);
        }
        double v = num / den;
        c[k - 1] = (float)((k % 2) ? v : -v);
    }
}

void pad_model(const Grid &g, const float *interior, std::vector<float> &out) {
    out.assign(g.pitched(), 0.f);
    for (int i = 0; i < g.Xn; ++i) {
        int ii = std::min(std::max(i - g.L, 0), g.nx - 1);
        for (int j = 0; j < g.Zn; ++j) {
            int jj = std::min(std::max(j - g.L, 0), g.nz - 1);
            out[(size_t)i * g.P + j] = interior[(size_t)ii * g.nz + jj];
        }
    }
}

void cpml_profiles(const Grid &g, double R, const std::vector<float> &vp_pad, std::vector<float> &ax,
                   std::vector<float> &az) {
    ax.assign(g.pitched(), 0.f);
    az.assign(g.pitched(), 0.f);
    const int L = g.L;
    const double d0 = std::log10(1.0 / R) * 1.5;
    for (int i = 0; i < g.Xn; ++i)
        for (int j = 0; j < g.Zn; ++j) {
            size_t k = (size_t)i * g.P + j;
            double v = (double)vp_pad[k];
            double a = 0.0, b = 0.0;
            if (i < L) a = d0 * v / L * std::pow(1.0 * (L - i) / L, 4.0);
            else if (i > g.Xn - L) a = d0 * v / L * std::pow(1.0 * (i - g.Xn + L) / L, 4.0);
            if (j < L) b = d0 * v / L * std::pow(1.0 * (L - j) / L, 4.0);
            else if (j > g.Zn - L) b = d0 * v / L * std::pow(1.0 * (j - g.Zn + L) / L, 4.0);
            ax[k] = (float)a;
            az[k] = (float)b;
        }
}

int ricker_t0(float fm, float dt) { return (int)std::ceil(1.0 / ((double)fm * (double)dt)); }

void ricker(int nt, float fm, float dt, std::vector<float> &s) {
    const double pi = 3.1415926;
    double Nk = pi * pi * (double)fm * (double)fm * (double)dt * (double)dt;
    int t0 = ricker_t0(fm, dt);
    s.resize(nt);
    for (int t = 0; t < nt; ++t) {
        double u = (double)(t - t0) * (double)(t - t0);
        s[t] = (float)((1.0 - 2.0 * Nk * u) * std::exp(-Nk * u));
    }
}

void mute_counts(const Grid &g, const ouc_shot &s, float vp_src, std::vector<int32_t> &counts) {
    counts.assign(g.nx, 0);
    int t0 = ricker_t0(s.fm, g.dt);
    int Sx = s.sx + g.L, Sz = s.sz + g.L, Zr = s.rec_z + g.L;
    for (int ix = 0; ix < g.nx; ++ix) {
        int i = ix + g.L;
        int a = std::abs(Sx - i), b = std::abs(Zr - Sz);
        float dist = std::sqrt((float)(a * a + b * b));
        float dd = dist * g.dx;
        double thr = 2 * t0 + (double)dd * 1.0 / (double)(g.dt * vp_src);
        // integer t < thr  <=>  t < ceil(thr)
        double c = std::ceil(thr);
        counts[ix] = (int32_t)std::min<double>(std::max(c, 0.0), (double)g.nt);
    }
}

RingDesc make_ring(const Grid &g) {
    RingDesc r;
    r.L = g.L; r.N = g.N; r.Xn = g.Xn; r.Zn = g.Zn; r.nx = g.nx; r.nz = g.nz;
    r.ntop = (int64_t)(g.nx + 2 * g.N) * (2 * g.N);
    r.nside = (int64_t)(2 * g.N) * (g.nz - 2 * g.N);
    r.ncells = 2 * r.ntop + 2 * r.nside;
    r.nouter = (int64_t)(g.N - 1) * (2 * (g.nx + 2 * g.N) + 2 * (g.nz - 2 * g.N));
    return r;
}

} // namespace ouc
