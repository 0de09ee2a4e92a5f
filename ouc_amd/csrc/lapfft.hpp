// Laplacian image filter on rocFFT.
#pragma once

struct ouc_ctx;

// out = Laplacian of the x-major nx*nz interior image `img` (device pointers),
// applied as R2C -> multiply by the operator's symbol -> C2R on a zero-padded
// grid, then the reference's two-cell zero border (P-S-seperating.cu:1463).
void lap_fft(ouc_ctx *c, const float *img, float *out);
void lap_fft_destroy(ouc_ctx *c);
