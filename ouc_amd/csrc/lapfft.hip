#include "engine.hpp"
#include "lapfft.hpp"

void lap_fft(ouc_ctx *, const float *, float *) { throw ouc::Error(OUC_UNSUPPORTED, "FFT Laplacian not built yet"); }
void lap_fft_destroy(ouc_ctx *) {}
