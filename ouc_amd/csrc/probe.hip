// Streaming-rate probe: every thread reads `nin` arrays and writes `nout`
// arrays at its own index, one float each (a wavefront moves 256 contiguous
// bytes per array), which is the access pattern of the core kernel's state
// streams.  It measures what HBM delivers to that pattern on the device at
// hand, the practical ceiling beside the published peak.
#include "engine.hpp"

namespace ouc {
namespace {

constexpr int kProbeMax = 16;
struct ProbeArgs {
    const float *in[kProbeMax];
    float *out[kProbeMax];
};

// The last `nshared` inputs hold only m floats and are read at k mod m, as
// the core kernel reads the model arrays (one plane, shared by every shot).
__global__ __launch_bounds__(256) void k_stream_probe(ProbeArgs a, int nin, int nout, long n, int nshared, long m) {
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    float s = 0.f;
    const long km = k % m;
    for (int q = 0; q < nin; ++q) s += a.in[q][q < nin - nshared ? k : km];
    for (int q = 0; q < nout; ++q) a.out[q][k] = s + (float)q;
}

} // namespace
} // namespace ouc

extern "C" ouc_status ouc_stream_probe(int32_t device, int32_t nin, int32_t nout, int64_t n, int32_t nshared,
                                       int64_t m, int32_t reps, float *ms_per_launch) {
    using namespace ouc;
    if (nin < 1 || nout < 1 || nin > kProbeMax || nout > kProbeMax || n < 1 || reps < 1 || !ms_per_launch ||
        nshared < 0 || nshared >= nin || (nshared > 0 && (m < 1 || m > n)))
        return ouc_fail_global(OUC_INVALID_ARG, "bad arguments");
    if (nshared == 0) m = n;
    try {
        OUC_HIP(hipSetDevice(device));
        DevBuf mem;
        const int nfull = nin - nshared;
        mem.alloc(((size_t)(nfull + nout) * n + (size_t)nshared * m) * sizeof(float));
        OUC_HIP(hipMemset(mem.ptr, 0, mem.bytes));
        ProbeArgs a{};
        for (int q = 0; q < nfull; ++q) a.in[q] = mem.f() + (size_t)q * n;
        for (int q = 0; q < nout; ++q) a.out[q] = mem.f() + (size_t)(nfull + q) * n;
        for (int q = 0; q < nshared; ++q) a.in[nfull + q] = mem.f() + (size_t)(nfull + nout) * n + (size_t)q * m;
        const unsigned blocks = (unsigned)((n + 255) / 256);
        hipEvent_t e0, e1;
        OUC_HIP(hipEventCreate(&e0));
        OUC_HIP(hipEventCreate(&e1));
        k_stream_probe<<<blocks, 256>>>(a, nin, nout, (long)n, nshared, (long)m); // warm-up
        OUC_HIP(hipEventRecord(e0, nullptr));
        for (int r = 0; r < reps; ++r) k_stream_probe<<<blocks, 256>>>(a, nin, nout, (long)n, nshared, (long)m);
        OUC_HIP(hipEventRecord(e1, nullptr));
        OUC_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        OUC_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        OUC_HIP(hipGetLastError());
        *ms_per_launch = ms / reps;
        return OUC_OK;
    } catch (const Error &e) {
        return ouc_fail_global(e.code, e.what());
    }
}
