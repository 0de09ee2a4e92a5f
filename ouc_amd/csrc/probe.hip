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
// VW consecutive floats per thread (1: the core kernel's one cell per lane).
template <class T> struct Lanes;
template <> struct Lanes<float> { static constexpr int n = 1; };
template <> struct Lanes<float2> { static constexpr int n = 2; };
template <> struct Lanes<float4> { static constexpr int n = 4; };
__device__ __forceinline__ float2 operator+(float2 a, float2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ float4 operator+(float4 a, float4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }

template <class T>
__global__ __launch_bounds__(256) void k_stream_probe(ProbeArgs a, int nin, int nout, long n, int nshared, long m) {
    constexpr int VW = Lanes<T>::n;
    const long k = (long)blockIdx.x * 256 + threadIdx.x; // vector index
    if (k >= n / VW) return;
    T s{};
    const long km = k % (m / VW);
    for (int q = 0; q < nin; ++q) s = s + reinterpret_cast<const T *>(a.in[q])[q < nin - nshared ? k : km];
    for (int q = 0; q < nout; ++q) reinterpret_cast<T *>(a.out[q])[k] = s;
}

// The same streams with one value per thread for the arithmetic and two per
// thread for memory: a block moves 256-float chunks ("columns") of every array
// as float2 loads and stores, transposed through LDS, the staging the core
// kernel would need to issue 8-byte accesses while keeping one cell per lane.
__global__ __launch_bounds__(256) void k_stream_probe_lds(ProbeArgs a, int nin, int nout, long n) {
    __shared__ float lin[kProbeMax][256], lout[kProbeMax][256];
    const int t = threadIdx.x, half = t >> 7, pos = (t & 127) * 2;
    const long nchunks = n / 256;
    for (long c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const long base = c * 256;
        for (int q = half; q < nin; q += 2) {
            const float2 v = *reinterpret_cast<const float2 *>(a.in[q] + base + pos);
            lin[q][pos] = v.x;
            lin[q][pos + 1] = v.y;
        }
        __syncthreads();
        float s = 0.f;
        for (int q = 0; q < nin; ++q) s += lin[q][t];
        for (int q = 0; q < nout; ++q) lout[q][t] = s + (float)q;
        __syncthreads();
        for (int q = half; q < nout; q += 2)
            *reinterpret_cast<float2 *>(a.out[q] + base + pos) = make_float2(lout[q][pos], lout[q][pos + 1]);
        __syncthreads();
    }
}

// Two arrays per memory instruction inside one wave: lanes 0-31 move rows
// (2j, 2j+1) of array A and lanes 32-63 those of array B as float2, and
// cross-lane permutes (ds_bpermute, no LDS allocation or barrier) give every
// lane its own row of both, so the arithmetic keeps one row per lane.
__device__ __forceinline__ float lane_read(float v, int src) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__global__ __launch_bounds__(256) void k_stream_probe_perm(ProbeArgs a, int nin, int nout, long n) {
    const int lane = threadIdx.x & 63;
    const long row0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64; // this wave's 64 rows
    if (row0 >= n) return;
    const int hi = lane >> 5, j = lane & 31;        // memory role: array (pair member), row pair
    const int src = 32 * (lane & 1) + (lane >> 1);  // lane holding my row's pair... by half below
    float s = 0.f;
    for (int q = 0; q + 1 < nin; q += 2) {
        const float2 v = *reinterpret_cast<const float2 *>(a.in[q + hi] + row0 + 2 * j);
        // my row r = lane: pair member 0 holds it in lane r/2 (.x even, .y odd), member 1 in lane 32 + r/2
        const int from0 = lane >> 1, from1 = 32 + (lane >> 1);
        const float a0x = lane_read(v.x, from0), a0y = lane_read(v.y, from0);
        const float a1x = lane_read(v.x, from1), a1y = lane_read(v.y, from1);
        s += (lane & 1) ? a0y : a0x;
        s += (lane & 1) ? a1y : a1x;
    }
    if (nin & 1) s += a.in[nin - 1][row0 + lane];
    (void)src;
    for (int q = 0; q + 1 < nout; q += 2) {
        const float va = s + (float)q, vb = s + (float)(q + 1);
        // lane j of half hi stores rows 2j, 2j+1 of array q + hi
        const float x0 = lane_read(va, 2 * j), y0 = lane_read(va, 2 * j + 1);
        const float x1 = lane_read(vb, 2 * j), y1 = lane_read(vb, 2 * j + 1);
        *reinterpret_cast<float2 *>(a.out[q + hi] + row0 + 2 * j) = hi ? make_float2(x1, y1) : make_float2(x0, y0);
    }
    if (nout & 1) a.out[nout - 1][row0 + lane] = s + (float)(nout - 1);
}

} // namespace
} // namespace ouc

extern "C" ouc_status ouc_stream_probe(int32_t device, int32_t nin, int32_t nout, int64_t n, int32_t nshared,
                                       int64_t m, int32_t width, int32_t reps, float *ms_per_launch) {
    using namespace ouc;
    if (nin < 1 || nout < 1 || nin > kProbeMax || nout > kProbeMax || n < 1 || reps < 1 || !ms_per_launch ||
        nshared < 0 || nshared >= nin || (nshared > 0 && (m < 1 || m > n)) ||
        (width != 1 && width != 2 && width != 4 && width != -2 && width != -3) || n % 256 != 0 ||
        (nshared > 0 && m % 4 != 0) || (width < 0 && nshared > 0))
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
        const unsigned blocks = width == -2 ? 256u * 8u : width == -3 ? (unsigned)(n / 256) : (unsigned)((n / width + 255) / 256);
        auto launch = [&] {
            if (width == -2) k_stream_probe_lds<<<blocks, 256>>>(a, nin, nout, (long)n);
            else if (width == -3) k_stream_probe_perm<<<blocks, 256>>>(a, nin, nout, (long)n);
            else if (width == 1) k_stream_probe<float><<<blocks, 256>>>(a, nin, nout, (long)n, nshared, (long)m);
            else if (width == 2) k_stream_probe<float2><<<blocks, 256>>>(a, nin, nout, (long)n, nshared, (long)m);
            else k_stream_probe<float4><<<blocks, 256>>>(a, nin, nout, (long)n, nshared, (long)m);
        };
        hipEvent_t e0, e1;
        OUC_HIP(hipEventCreate(&e0));
        OUC_HIP(hipEventCreate(&e1));
        launch(); // warm-up
        OUC_HIP(hipEventRecord(e0, nullptr));
        for (int r = 0; r < reps; ++r) launch();
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
