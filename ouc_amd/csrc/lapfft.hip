// Laplacian image filter on rocFFT.
//
// The reference filters each image with a 5-point Laplacian stencil
// (Laplace, P-S-seperating.cu:1457-1469).  Here the image is zero-padded to an
// FFT-friendly Mx x Mz (at least one zero row/column past each edge, so the
// periodic convolution of the FFT equals the stencil on the interior),
// transformed R2C, multiplied by the operator's symbol and transformed back:
//   FFT5:     -(4/dh^2) [sin^2(pi kx/Mx) + sin^2(pi kz/Mz)]   (the 5-point
//             operator exactly: its eigenvalue on e^{2 pi i k n / M})
//   FFT_CONT: -|k|^2 = -[(2 pi kx/(Mx dh))^2 + (2 pi kz/(Mz dh))^2]
// The two outermost rows/columns are then zeroed as the stencil leaves them.
#include <rocfft/rocfft.h>

#include <cmath>
#include <mutex>

#include "engine.hpp"
#include "lapfft.hpp"

namespace {

struct LapFFT {
    int nx = 0, nz = 0, Mx = 0, Mz = 0;
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    ouc::DevBuf work, real, spec;
};

std::once_flag g_rocfft_once;

void rf_check(rocfft_status s, const char *what) {
    if (s != rocfft_status_success) throw ouc::Error(OUC_DEVICE, std::string(what) + " failed (rocfft status " +
                                                                     std::to_string((int)s) + ")");
}

// Smallest 2^a 3^b 5^c 7^d >= n.
int good_size(int n) {
    for (int m = n;; ++m) {
        int r = m;
        for (int f : {2, 3, 5, 7})
            while (r % f == 0) r /= f;
        if (r == 1) return m;
    }
}

__global__ void k_pad(int nx, int nz, int Mz, const float *__restrict__ img, float *__restrict__ buf, int Mx) {
    int z = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (z >= Mz || x >= Mx) return;
    buf[(size_t)x * Mz + z] = (x < nx && z < nz) ? img[(size_t)x * nz + z] : 0.f;
}

__global__ void k_symbol(int Mx, int Mz, float dh, int continuous, float2 *__restrict__ spec) {
    int kz = blockIdx.x * blockDim.x + threadIdx.x, kx = blockIdx.y;
    const int Hz = Mz / 2 + 1;
    if (kz >= Hz || kx >= Mx) return;
    const double pi = 3.14159265358979323846;
    double s;
    if (!continuous) {
        double a = sin(pi * kx / Mx), b = sin(pi * kz / Mz);
        s = -4.0 / ((double)dh * dh) * (a * a + b * b);
    } else {
        int kxs = kx <= Mx / 2 ? kx : kx - Mx;
        double wx = 2 * pi * kxs / (Mx * (double)dh), wz = 2 * pi * kz / (Mz * (double)dh);
        s = -(wx * wx + wz * wz);
    }
    s /= (double)Mx * Mz; // the inverse transform is unnormalised
    float2 v = spec[(size_t)kx * Hz + kz];
    v.x = (float)(v.x * s);
    v.y = (float)(v.y * s);
    spec[(size_t)kx * Hz + kz] = v;
}

__global__ void k_crop(int nx, int nz, int Mz, const float *__restrict__ buf, float *__restrict__ out) {
    int z = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (z >= nz || x >= nx) return;
    bool edge = x < 2 || x >= nx - 2 || z < 2 || z >= nz - 2;
    out[(size_t)x * nz + z] = edge ? 0.f : buf[(size_t)x * Mz + z];
}

LapFFT *state(ouc_ctx *c) {
    if (c->lapfft) return (LapFFT *)c->lapfft;
    std::call_once(g_rocfft_once, [] { rocfft_setup(); });
    auto *s = new LapFFT();
    c->lapfft = s;
    s->nx = c->g.nx;
    s->nz = c->g.nz;
    s->Mx = good_size(s->nx + 1);
    s->Mz = good_size(s->nz + 1);
    if (s->Mz % 2) s->Mz = good_size(s->Mz + 1); // even z length keeps the hermitian layout simple
    size_t lengths[2] = {(size_t)s->Mz, (size_t)s->Mx}; // fastest first
    rf_check(rocfft_plan_create(&s->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                rocfft_precision_single, 2, lengths, 1, nullptr),
             "rocfft_plan_create(forward)");
    rf_check(rocfft_plan_create(&s->inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                                rocfft_precision_single, 2, lengths, 1, nullptr),
             "rocfft_plan_create(inverse)");
    size_t w1 = 0, w2 = 0;
    rf_check(rocfft_plan_get_work_buffer_size(s->fwd, &w1), "work size");
    rf_check(rocfft_plan_get_work_buffer_size(s->inv, &w2), "work size");
    rf_check(rocfft_execution_info_create(&s->info), "execution info");
    size_t w = std::max(w1, w2);
    if (w) {
        s->work.alloc(w);
        rf_check(rocfft_execution_info_set_work_buffer(s->info, s->work.ptr, w), "work buffer");
    }
    rf_check(rocfft_execution_info_set_stream(s->info, c->stream), "set stream");
    s->real.alloc((size_t)s->Mx * s->Mz * sizeof(float));
    s->spec.alloc((size_t)s->Mx * (s->Mz / 2 + 1) * sizeof(float2));
    return s;
}

} // namespace

void lap_fft(ouc_ctx *c, const float *img, float *out) {
    LapFFT *s = state(c);
    dim3 gr((s->Mz + 255) / 256, s->Mx);
    k_pad<<<gr, 256, 0, c->stream>>>(s->nx, s->nz, s->Mz, img, s->real.f(), s->Mx);
    void *in[1] = {s->real.ptr}, *sp[1] = {s->spec.ptr};
    rf_check(rocfft_execute(s->fwd, in, sp, s->info), "rocfft_execute(forward)");
    dim3 gs((s->Mz / 2 + 1 + 255) / 256, s->Mx);
    k_symbol<<<gs, 256, 0, c->stream>>>(s->Mx, s->Mz, c->g.dx, c->p.laplacian == OUC_LAP_FFT_CONT,
                                        s->spec.as<float2>());
    rf_check(rocfft_execute(s->inv, sp, in, s->info), "rocfft_execute(inverse)");
    dim3 gc((s->nz + 255) / 256, s->nx);
    k_crop<<<gc, 256, 0, c->stream>>>(s->nx, s->nz, s->Mz, s->real.f(), out);
    OUC_HIP(hipGetLastError());
}

void lap_fft_destroy(ouc_ctx *c) {
    auto *s = (LapFFT *)c->lapfft;
    if (!s) return;
    if (s->fwd) rocfft_plan_destroy(s->fwd);
    if (s->inv) rocfft_plan_destroy(s->inv);
    if (s->info) rocfft_execution_info_destroy(s->info);
    delete s;
    c->lapfft = nullptr;
}
