// Host-side setup of one grid: padded model, CPML damping, Ricker wavelet,
// stencil coefficients, direct-wave mute thresholds and the strip ring.
// Everything here runs once per context or shot; the CPML setup stays in FP64.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ouc.h"

namespace ouc {

struct Grid {
    int nx = 0, nz = 0;  // interior
    int L = 0, N = 0;    // pad cells, stencil half-order
    int Xn = 0, Zn = 0;  // padded
    int P = 0;           // device row pitch (floats), Zn rounded up
    int nt = 0;
    float dx = 0, dz = 0, dt = 0;
    bool learned = false; // learned boundary: L is the low-order pad, no CPML damping
    size_t cells() const { return (size_t)Xn * Zn; }
    size_t pitched() const { return (size_t)Xn * P; }
};

// Validates p and fills g; returns an error message or "".
std::string make_grid(const ouc_params *p, Grid &g);

// Float stencil coefficients c_1..c_N (P-S-seperating.cu:221 for N = 6,
// the Taylor staggered closed form otherwise).
void stencil_coeffs(int N, float *c);

// Edge-replicated padding, xiangbian() (P-S-seperating.cu:37-119): every pad
// cell takes the nearest interior cell, corners the corner cell.  Output is
// pitched Xn x P (pitch columns zero).
void pad_model(const Grid &g, const float *interior, std::vector<float> &out);

// CPML damping in FP64 (P-S-seperating.cu:2067-2105), rounded to float,
// pitched.  Uses the padded vp.
void cpml_profiles(const Grid &g, double R, const std::vector<float> &vp_pad, std::vector<float> &ax,
                   std::vector<float> &az);

// Ricker wavelet (P-S-seperating.cu:1954-1958) with PI = 3.1415926
// (P-S-seperating.cu:14); t0 = ceil(1/(fm dt)).
int ricker_t0(float fm, float dt);
void ricker(int nt, float fm, float dt, std::vector<float> &s);

// Number of leading samples the mute of remove() (P-S-seperating.cu:1634-1657)
// zeroes in each interior receiver trace: t < 2 t0 + dist*dh/(dt*vp(Sx,Zr)).
void mute_counts(const Grid &g, const ouc_shot &s, float vp_src, std::vector<int32_t> &counts);

// The strip ring: cells of the frame [L-N, Xn-L+N)^2 within N of the interior
// edge, as four rectangles (top, bottom: (nx+2N) x 2N; left, right:
// 2N x (nz-2N)).  These are exactly the cells the interior-only source
// reconstruction reads outside its own update, so they replace the reference's
// full-length 2N-wide strips (P-S-seperating.cu:1230-1287).  nouter counts the
// cells of the N-1 outermost pad rows (columns) of each rectangle, stored first.
struct RingDesc {
    int L, N, Xn, Zn, nx, nz;
    int64_t ntop, nside, ncells, nouter;
};
RingDesc make_ring(const Grid &g);

} // namespace ouc
