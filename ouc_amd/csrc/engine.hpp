// Engine context shared by the plain and fused paths.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ouc.h"
#include "kernels_plain.hpp"
#include "setup.hpp"

namespace ouc {

struct Error : std::runtime_error {
    ouc_status code;
    Error(ouc_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what);
// Records the message ouc_last_error(NULL) returns and passes the status on.
ouc_status ouc_fail_global(ouc_status code, const std::string &msg);
#define OUC_HIP(x) ::ouc::hip_check((x), #x)

// Host <-> device copies on a context's own (non-blocking) stream, waited
// for.  The library issues nothing on the legacy null stream: a legacy-stream
// copy made while another thread's context captures a time loop into a graph
// would be ordered against that capture.
void copy_h2d(hipStream_t s, void *dst, const void *src, size_t bytes);
void copy_d2h(hipStream_t s, void *dst, const void *src, size_t bytes);

// Owning device allocation.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t n);
    void release();
    float *f() const { return (float *)ptr; }
    template <class T> T *as() const { return (T *)ptr; }
};

// One wavefield: 32 pitched arrays in a single allocation.
struct DevWave {
    DevBuf mem;
    KWave k{};
    void alloc(const Grid &g);
    void zero(hipStream_t s, const Grid &g);
};

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    void init();
    void destroy();
};

} // namespace ouc

struct ouc_ctx {
    ouc_params p{};
    ouc::Grid g;
    ouc::RingDesc ring{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool profiling = false;
    bool graphs = true; // replay captured time loops (fused engine, profiling off)
    int batch_cap = 0;  // shots per fused launch (0: as many as 85% of free HBM holds)

    std::vector<float> vp_pad; // host copy (mute velocity)
    std::vector<float> vp_in, vs_in; // interior model (the CPML fallback context is built from it)
    ouc_ctx *fallback = nullptr;     // CPML context of a learned-boundary context, made on first need
    ouc::DevBuf p2, q2, ax, az;
    ouc::DevBuf pq; // (vp^2, vs^2) pairs, interleaved
    ouc::DevWave F, R;
    ouc::DevBuf vp3[3][2];
    ouc::DevBuf npp, nps, den;
    ouc::DevBuf rec;    // [ncomp][nt][nx]
    ouc::DevBuf dX, dZ; // Xn
    ouc::DevBuf counts; // nx
    ouc::DevBuf img[4]; // pp, ps, pp_lap, ps_lap (nx*nz)
    ouc::DevBuf strips32, strips16, stripse8; // FP32 strips, or FP16 residuals and chunk exponents
    ouc::DevBuf strips_in32;                     // FP16: inner ring part of the FP32-kept fields
    ouc::DevBuf stripskey, stripserr, stripsdec; // FP16: keyframes, encoder state [2][RS_NF][ring], decoded [2][RS_NF][ring]
    int64_t strip_slots = 0, strip_chunks = 0;
    int strip_fp16 = -1;
    hipEvent_t ev[8] = {};
    hipEvent_t kev[2] = {};
    ouc_timings last{};
    void *lapfft = nullptr; // rocFFT plans and buffers (lapfft.hip)
    void *fused = nullptr;  // fused-engine state (fused.hip)

    ouc::KGrid kgrid() const;
    ouc::KModel kmodel() const;
    ouc::KRing kring() const;
    ouc::KStrips kstrips() const;
    void ensure_strips();
};
