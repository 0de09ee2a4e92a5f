// Engine runtime and C ABI: device memory, the forward and backward time loops
// of the plain path, imaging, timing.  The loops follow the reference's driver
// order (P-S-seperating.cu:2258-2373) with the kernels of kernels_plain.hpp.
#include "engine.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "fused.hpp"
#include "splitpml.hpp"
#include "lapfft.hpp"

namespace ouc {

void hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return;
    ouc_status code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? OUC_OOM : OUC_DEVICE;
    throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::alloc(size_t n) {
    release();
    if (n == 0) return;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        (void)hipGetLastError();
        throw Error(OUC_OOM, "hipMalloc of " + std::to_string(n) + " bytes failed: " + hipGetErrorString(e));
    }
    bytes = n;
}

void copy_h2d(hipStream_t s, void *dst, const void *src, size_t bytes) {
    OUC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    OUC_HIP(hipStreamSynchronize(s));
}

void copy_d2h(hipStream_t s, void *dst, const void *src, size_t bytes) {
    OUC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    OUC_HIP(hipStreamSynchronize(s));
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void DevWave::alloc(const Grid &g) {
    const size_t n = g.pitched();
    mem.alloc(32 * n * sizeof(float));
    float *b = mem.f();
    float **slots[32] = {&k.Phi, &k.DPhi, &k.Qx, &k.Qz, &k.DQx, &k.DQz, &k.Sx, &k.Sz, &k.DSx, &k.DSz, &k.Vx,
                         &k.Vz,  &k.Ux,   &k.Uz, &k.th, &k.om,  &k.gxx, &k.gxz, &k.gzx, &k.gzz, &k.F[0], &k.F[1],
                         &k.F[2], &k.F[3], &k.O[0], &k.O[1], &k.O[2], &k.O[3], &k.O[4], &k.O[5], &k.O[6], &k.O[7]};
    for (int a = 0; a < 32; ++a) *slots[a] = b + (size_t)a * n;
}

void DevWave::zero(hipStream_t s, const Grid &g) {
    (void)g;
    OUC_HIP(hipMemsetAsync(mem.ptr, 0, mem.bytes, s));
}

} // namespace ouc

using namespace ouc;

ouc::KGrid ouc_ctx::kgrid() const {
    KGrid k{};
    k.Xn = g.Xn; k.Zn = g.Zn; k.P = g.P; k.L = g.L; k.N = g.N; k.nx = g.nx; k.nz = g.nz;
    k.idx = 1.0f / g.dx; k.idz = 1.0f / g.dz; k.dt = g.dt; k.dt2 = g.dt * g.dt;
    float c[8] = {0};
    stencil_coeffs(g.N, c);
    for (int a = 0; a < 8; ++a) k.c[a] = c[a];
    return k;
}

ouc::KModel ouc_ctx::kmodel() const { return KModel{p2.f(), q2.f(), ax.f(), az.f()}; }

ouc::KRing ouc_ctx::kring() const {
    KRing r;
    r.L = ring.L; r.N = ring.N; r.Xn = ring.Xn; r.Zn = ring.Zn; r.nx = ring.nx; r.nz = ring.nz;
    r.ntop = ring.ntop; r.nside = ring.nside; r.ncells = ring.ncells; r.nouter = ring.nouter;
    return r;
}

ouc::KStrips ouc_ctx::kstrips() const {
    KStrips s{};
    s.f32 = strips32.f();
    s.f16 = strips16.as<__half>();
    s.in32 = strips_in32.f();
    s.e8 = stripse8.as<int8_t>();
    s.key = stripskey.f();
    s.err = stripserr.f();
    s.dec = stripsdec.f();
    s.ncells = ring.ncells;
    s.nouter = ring.nouter;
    s.nchunks = strip_chunks;
    s.fp16 = p.strip_precision == OUC_STRIP_FP16;
    s.last_v = s.last_th = (int)strip_slots - 1; // slots 0..nt-2 of both field groups
    return s;
}

void ouc_ctx::ensure_strips() {
    int want16 = p.strip_precision == OUC_STRIP_FP16;
    int64_t slots = g.nt - 1;
    if (strip_fp16 == want16 && strip_slots == slots) return;
    strips32.release(); strips16.release(); stripse8.release(); stripskey.release(); stripserr.release();
    stripsdec.release(); strips_in32.release();
    strip_chunks = (ring.ncells + 63) / 64;
    if (want16) {
        strips16.alloc((size_t)slots * strip16_offset(ring.ncells, ring.nouter, RS_NF) * sizeof(__half));
        strips_in32.alloc((size_t)slots * kNInner32 * (ring.ncells - ring.nouter) * sizeof(float));
        stripse8.alloc((size_t)slots * RS_NF * strip_chunks);
        stripskey.alloc((size_t)((slots + kStripKey - 1) / kStripKey) * RS_NF * ring.ncells * sizeof(float));
        stripserr.alloc((size_t)2 * RS_NF * ring.ncells * sizeof(float));
        stripsdec.alloc((size_t)2 * RS_NF * ring.ncells * sizeof(float));
    } else {
        strips32.alloc((size_t)slots * RS_NF * ring.ncells * sizeof(float));
    }
    strip_fp16 = want16;
    strip_slots = slots;
}

namespace {

#define OUC_DISPATCH_N(NV, ...)                                                  \
    switch (NV) {                                                                \
    case 1: { constexpr int NN = 1; __VA_ARGS__; } break;                        \
    case 2: { constexpr int NN = 2; __VA_ARGS__; } break;                        \
    case 3: { constexpr int NN = 3; __VA_ARGS__; } break;                        \
    case 4: { constexpr int NN = 4; __VA_ARGS__; } break;                        \
    case 5: { constexpr int NN = 5; __VA_ARGS__; } break;                        \
    case 6: { constexpr int NN = 6; __VA_ARGS__; } break;                        \
    case 7: { constexpr int NN = 7; __VA_ARGS__; } break;                        \
    default: { constexpr int NN = 8; __VA_ARGS__; } break;                       \
    }

inline dim3 tiles(int xcount, int zcount) { return dim3((zcount + 63) / 64, (xcount + 3) / 4); }
const dim3 kTile(64, 4);

thread_local std::string g_create_error;

ouc_status fail(ouc_ctx *c, ouc_status code, const std::string &msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

} // namespace

ouc_status ouc::ouc_fail_global(ouc_status code, const std::string &msg) { return fail(nullptr, code, msg); }

namespace {

template <class F>
ouc_status guarded(ouc_ctx *c, F &&f) {
    try {
        if (c) OUC_HIP(hipSetDevice(c->device));
        f();
        return OUC_OK;
    } catch (const Error &e) {
        return fail(c, e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(c, OUC_OOM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(c, OUC_DEVICE, e.what());
    }
}

float ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    OUC_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

void check_shot(const ouc_ctx *c, const ouc_shot *s) {
    if (!s) throw Error(OUC_INVALID_ARG, "shot is NULL");
    if (s->sx < 0 || s->sx >= c->g.nx || s->sz < 0 || s->sz >= c->g.nz)
        throw Error(OUC_INVALID_ARG, "source outside the interior");
    if (s->rec_z < 0 || s->rec_z >= c->g.nz) throw Error(OUC_INVALID_ARG, "receiver row outside the interior");
    if (!s->wavelet && !(s->fm > 0)) throw Error(OUC_INVALID_ARG, "fm must be positive when no wavelet is given");
    if (s->nt < 0 || s->nt > c->g.nt || (s->nt > 0 && s->nt < 2)) throw Error(OUC_INVALID_ARG, "shot nt out of range");
}

// Paths that run exactly the context's nt steps.
void check_full_length(const ouc_ctx *c, const ouc_shot *s) {
    if (s->nt != 0 && s->nt != c->g.nt)
        throw Error(OUC_UNSUPPORTED, "a shot nt other than the context's is supported by fused RTM only");
}

std::vector<float> shot_wavelet(const ouc_ctx *c, const ouc_shot *s) {
    std::vector<float> w;
    if (s->wavelet) w.assign(s->wavelet, s->wavelet + c->g.nt);
    else ricker(c->g.nt, s->fm, c->g.dt, w);
    return w;
}

// Sampled per-launch timing of the dominant step kernel (profiling mode).
struct StepProbe {
    std::vector<hipEvent_t> ev;
    int stride = 1, used = 0;
    void init(int nt, bool on) {
        if (!on) return;
        int slots = std::min(nt, 1024);
        stride = (nt + slots - 1) / slots;
        ev.resize(2 * slots);
        for (auto &e : ev) OUC_HIP(hipEventCreate(&e));
    }
    bool want(int t) const { return !ev.empty() && (t % stride) == 0 && used < (int)ev.size() / 2; }
    void begin(hipStream_t s) { OUC_HIP(hipEventRecord(ev[2 * used], s)); }
    void end(hipStream_t s) { OUC_HIP(hipEventRecord(ev[2 * used + 1], s)); ++used; }
    double total_ms() {
        double t = 0;
        for (int a = 0; a < used; ++a) t += ms_between(ev[2 * a], ev[2 * a + 1]);
        return t;
    }
    ~StepProbe() {
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
};

// Forward loop (P-S-seperating.cu:2258-2292).  Records ncomp components into
// c->rec as [ncomp][nt][nx]; with rtm the strips are saved.
void run_forward_plain(ouc_ctx *c, const ouc_shot *s, const std::vector<float> &wav, bool rtm, int ncomp,
                       StepProbe &probe) {
    const Grid &g = c->g;
    const KGrid kg = c->kgrid();
    const KModel km = c->kmodel();
    const KRing kr = c->kring();
    KStrips ks{};
    if (rtm) ks = c->kstrips();
    KWave w = c->F.k;
    const int N = g.N, Sx = s->sx + g.L, Sz = s->sz + g.L, Zr = s->rec_z + g.L;
    hipStream_t st = c->stream;
    const dim3 upd = tiles(g.Xn - 2 * N, g.Zn - 2 * N);
    const int rblocks = (int)((c->ring.ncells + 255) / 256);
    if (rtm && ks.err) OUC_HIP(hipMemsetAsync(ks.err, 0, c->stripserr.bytes, st));
    for (int t = 0; t < g.nt; ++t) {
        OUC_DISPATCH_N(N, k_assemble<NN><<<upd, kTile, 0, st>>>(kg, w));
        if (rtm && t >= 1) k_strip_save_v<<<rblocks, 256, 0, st>>>(kr, ks, t - 1, w.Vx, w.Vz, g.P);
        OUC_DISPATCH_N(N, k_stress<NN><<<upd, kTile, 0, st>>>(kg, w, km, Sx, Sz, wav[t], -1, nullptr, nullptr));
        if (rtm && t < g.nt - 1) k_strip_save_theta<<<rblocks, 256, 0, st>>>(kr, ks, t, w, g.P);
        bool pr = probe.want(t);
        if (pr) probe.begin(st);
        OUC_DISPATCH_N(N, k_velocity<NN><<<upd, kTile, 0, st>>>(kg, w, km));
        if (pr) probe.end(st);
        OUC_DISPATCH_N(N, k_record<NN><<<(g.nx + 255) / 256, 256, 0, st>>>(kg, w, Zr, t, g.nt, ncomp, c->rec.f()));
    }
    OUC_HIP(hipGetLastError());
}

void apply_mute(ouc_ctx *c, const ouc_shot *s, int ncomp) {
    const Grid &g = c->g;
    int Sx = s->sx + g.L, Zr = s->rec_z + g.L;
    float vsrc = c->vp_pad[(size_t)Sx * g.P + Zr];
    std::vector<int32_t> counts;
    mute_counts(g, *s, vsrc, counts);
    OUC_HIP(hipMemcpyAsync(c->counts.ptr, counts.data(), counts.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream));
    dim3 grid((g.nx + 255) / 256, g.nt);
    k_mute<<<grid, 256, 0, c->stream>>>(g.nx, g.nt, ncomp, c->counts.as<int32_t>(), c->rec.f());
    OUC_HIP(hipGetLastError());
}

// The caller's records [2][nx][nt] (vx, vz) in place of the forward run's,
// into c->rec's [component][nt][nx] layout.
void upload_records(ouc_ctx *c, const float *rec_in) {
    const Grid &g = c->g;
    std::vector<float> tr((size_t)2 * g.nt * g.nx);
    for (int q = 0; q < 2; ++q)
        for (int ix = 0; ix < g.nx; ++ix)
            for (int t = 0; t < g.nt; ++t)
                tr[((size_t)q * g.nt + t) * g.nx + ix] = rec_in[((size_t)q * g.nx + ix) * g.nt + t];
    copy_h2d(c->stream, c->rec.ptr, tr.data(), tr.size() * sizeof(float));
}

// Backward loop (P-S-seperating.cu:2313-2364): source reconstruction from the
// strips, receiver propagation with injection, imaging condition.
void run_backward_plain(ouc_ctx *c, const ouc_shot *s, const std::vector<float> &wav) {
    const Grid &g = c->g;
    const KGrid kg = c->kgrid();
    const KModel km = c->kmodel();
    const KRing kr = c->kring();
    const KStrips ks = c->kstrips();
    KWave fw = c->F.k, rw = c->R.k;
    const int N = g.N, L = g.L, nt = g.nt, Sx = s->sx + L, Sz = s->sz + L, Zr = s->rec_z + L;
    const bool exact = c->p.semantics == OUC_SEM_EXACT;
    hipStream_t st = c->stream;
    const dim3 upd = tiles(g.Xn - 2 * N, g.Zn - 2 * N);
    const dim3 inter = tiles(g.nx, g.nz);
    const dim3 frame = tiles(g.nx + 2 * N, g.nz + 2 * N);
    const dim3 full = tiles(g.Xn, g.Zn);
    const int rblocks = (int)((c->ring.ncells + 255) / 256);
    const size_t pn = g.pitched() * sizeof(float);

    c->R.zero(st, g);
    OUC_HIP(hipMemsetAsync(c->npp.ptr, 0, pn, st));
    OUC_HIP(hipMemsetAsync(c->nps.ptr, 0, pn, st));
    OUC_HIP(hipMemsetAsync(c->den.ptr, 0, pn, st));
    OUC_HIP(hipMemsetAsync(c->dX.ptr, 0, g.Xn * sizeof(float), st));
    OUC_HIP(hipMemsetAsync(c->dZ.ptr, 0, g.Xn * sizeof(float), st));
    // hand-off (read_last_wavefiled2, P-S-seperating.cu:2281): the forward state
    // becomes the source state; levels nt and nt-1 of the source P velocity
    OUC_DISPATCH_N(N, k_asm_vp<NN><<<inter, kTile, 0, st>>>(kg, fw, c->vp3[nt % 3][0].f(), c->vp3[nt % 3][1].f()));
    k_unstep<<<full, kTile, 0, st>>>(kg, fw);
    OUC_DISPATCH_N(N, k_asm_vp<NN><<<inter, kTile, 0, st>>>(kg, fw, c->vp3[(nt - 1) % 3][0].f(),
                                                            c->vp3[(nt - 1) % 3][1].f()));
    const float *recx = c->rec.f();
    const float *recz = c->rec.f() + (size_t)nt * g.nx;
    // FP16 strips decode backwards one slot per step: step t restores slots t
    // and (EXACT) t-1
    const dim3 dgrid(rblocks, RS_NF, 1);
    if (ks.fp16) k_strip_decode<<<dgrid, 256, 0, st>>>(ks, nt - 2, KStripStride{});
    for (int t = nt - 1; t >= 0; --t) {
        if (t < nt - 1) {
            if (ks.fp16 && t >= 1) k_strip_decode<<<dgrid, 256, 0, st>>>(ks, t - 1, KStripStride{});
            OUC_DISPATCH_N(N, k_recon_velocity<NN><<<frame, kTile, 0, st>>>(kg, fw, km));
            int sv = exact ? t - 1 : t;
            k_strip_restore_v<<<rblocks, 256, 0, st>>>(kr, ks, sv, fw.Vx, fw.Vz, g.P);
            OUC_DISPATCH_N(N, k_recon_assemble<NN><<<inter, kTile, 0, st>>>(kg, fw, c->vp3[t % 3][0].f(),
                                                                             c->vp3[t % 3][1].f()));
            OUC_DISPATCH_N(N, k_recon_stress<NN><<<inter, kTile, 0, st>>>(kg, fw, km, exact ? Sx : -1,
                                                                           exact ? Sz : -1, exact ? wav[t] : 0.f));
            k_strip_restore_theta<<<rblocks, 256, 0, st>>>(kr, ks, t, fw, g.P);
        }
        OUC_DISPATCH_N(N, k_assemble<NN><<<upd, kTile, 0, st>>>(kg, rw));
        k_inject<<<(g.nx + 255) / 256, 256, 0, st>>>(kg, rw, Zr, recx + (size_t)t * g.nx, recz + (size_t)t * g.nx,
                                                      c->dX.f(), c->dZ.f());
        OUC_DISPATCH_N(N, k_stress<NN><<<upd, kTile, 0, st>>>(kg, rw, km, -1, -1, 0.f, Zr, c->dX.f(), c->dZ.f()));
        OUC_DISPATCH_N(N, k_velocity<NN><<<upd, kTile, 0, st>>>(kg, rw, km));
        int lev = exact ? t + 1 : std::min(t + 2, nt);
        OUC_DISPATCH_N(N, k_corr<NN><<<inter, kTile, 0, st>>>(kg, rw, c->vp3[lev % 3][0].f(), c->vp3[lev % 3][1].f(),
                                                               c->npp.f(), c->nps.f(), c->den.f()));
    }
    OUC_HIP(hipGetLastError());
}

// image_fun + Laplace (P-S-seperating.cu:2367-2373) into c->img[0..3].
void make_images(ouc_ctx *c) {
    const Grid &g = c->g;
    const KGrid kg = c->kgrid();
    dim3 grid((g.nz + 255) / 256, g.nx);
    k_image<<<grid, 256, 0, c->stream>>>(kg, c->npp.f(), c->nps.f(), c->den.f(), c->img[0].f(), c->img[1].f());
    if (c->p.laplacian == OUC_LAP_STENCIL) {
        k_laplace5<<<grid, 256, 0, c->stream>>>(g.nx, g.nz, g.dx, c->img[0].f(), c->img[2].f());
        k_laplace5<<<grid, 256, 0, c->stream>>>(g.nx, g.nz, g.dx, c->img[1].f(), c->img[3].f());
    } else {
        lap_fft(c, c->img[0].f(), c->img[2].f());
        lap_fft(c, c->img[1].f(), c->img[3].f());
    }
    OUC_HIP(hipGetLastError());
}

void download(ouc_ctx *c, float *dst, const DevBuf &src, size_t n) {
    if (!dst) return;
    OUC_HIP(hipMemcpyAsync(dst, src.ptr, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
}

void fill_counts(ouc_ctx *c, double cell_updates, double canonical, double own) {
    c->last.cell_updates = cell_updates;
    c->last.canonical_bytes = canonical;
    c->last.own_bytes = own;
}

// Compulsory bytes per cell-update of the reference's two-pass form (canonical)
// and of the plain kernels (own): FP32 operands counted once.
struct ByteModel {
    double canonical_interior = 120, canonical_cpml = 232;
    double plain_interior = 4.0 * (2 + 2 + 4 /*assemble: Phi,Q,S -> V,U*/ + 4 + 2 + 6 /*stress*/ + 6 + 1 + 20 /*vel*/);
};

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
// The CPML context behind a learned-boundary context: same interior model,
// shot geometry and time stepping, CPML pad of p.npml cells.
ouc_ctx *fallback_ctx(ouc_ctx *c) {
    if (c->fallback) return c->fallback;
    ouc_params q = c->p;
    q.boundary = OUC_BOUNDARY_CPML;
    ouc_model m{c->vp_in.data(), c->vs_in.data(), nullptr};
    ouc_ctx *fb = nullptr;
    ouc_status st = ouc_create(&q, &m, c->device, &fb);
    if (st != OUC_OK) throw Error(st, std::string("CPML fallback context: ") + ouc_last_error(nullptr));
    fb->profiling = c->profiling;
    c->fallback = fb;
    return fb;
}

extern "C" {

ouc_status ouc_create(const ouc_params *p, const ouc_model *m, int32_t device, ouc_ctx **out) {
    if (!out) return fail(nullptr, OUC_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!m || !m->vp || !m->vs) return fail(nullptr, OUC_INVALID_ARG, "model arrays are NULL");
    Grid g;
    std::string e = make_grid(p, g);
    if (!e.empty()) return fail(nullptr, OUC_INVALID_ARG, e);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, OUC_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(nullptr, OUC_INVALID_ARG, "device ordinal out of range");
    for (size_t k = 0; k < (size_t)p->nx * p->nz; ++k)
        if (!(m->vp[k] > 0) || !(m->vs[k] >= 0) || !std::isfinite(m->vp[k]) || !std::isfinite(m->vs[k]))
            return fail(nullptr, OUC_INVALID_ARG, "model velocities must be finite, vp > 0, vs >= 0");
    ouc_ctx *c = new (std::nothrow) ouc_ctx();
    if (!c) return fail(nullptr, OUC_OOM, "host allocation failed");
    c->p = *p;
    c->g = g;
    c->ring = make_ring(g);
    c->device = device;
    ouc_status st = guarded(c, [&] {
        OUC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto &ev : c->ev) OUC_HIP(hipEventCreate(&ev));
        for (auto &ev : c->kev) OUC_HIP(hipEventCreate(&ev));
        std::vector<float> vp, vs, ax, az;
        pad_model(g, m->vp, vp);
        pad_model(g, m->vs, vs);
        cpml_profiles(g, p->R, vp, ax, az);
        if (g.learned) {
            std::fill(ax.begin(), ax.end(), 0.f);
            std::fill(az.begin(), az.end(), 0.f);
        }
        c->vp_in.assign(m->vp, m->vp + (size_t)p->nx * p->nz);
        c->vs_in.assign(m->vs, m->vs + (size_t)p->nx * p->nz);
        std::vector<float> p2(vp.size()), q2(vs.size());
        for (size_t k = 0; k < vp.size(); ++k) { p2[k] = vp[k] * vp[k]; q2[k] = vs[k] * vs[k]; }
        c->vp_pad = vp;
        const size_t pn = g.pitched() * sizeof(float);
        c->p2.alloc(pn); c->q2.alloc(pn); c->ax.alloc(pn); c->az.alloc(pn);
        copy_h2d(c->stream, c->p2.ptr, p2.data(), pn);
        copy_h2d(c->stream, c->q2.ptr, q2.data(), pn);
        copy_h2d(c->stream, c->ax.ptr, ax.data(), pn);
        copy_h2d(c->stream, c->az.ptr, az.data(), pn);
        {   // (vp^2, vs^2) interleaved: the fused core kernel reads both with one 8-byte load
            std::vector<float> pq(2 * p2.size());
            for (size_t k = 0; k < p2.size(); ++k) { pq[2 * k] = p2[k]; pq[2 * k + 1] = q2[k]; }
            c->pq.alloc(2 * pn);
            copy_h2d(c->stream, c->pq.ptr, pq.data(), 2 * pn);
        }
        c->F.alloc(g);
        c->R.alloc(g);
        for (auto &lv : c->vp3)
            for (auto &b : lv) b.alloc(pn);
        c->npp.alloc(pn); c->nps.alloc(pn); c->den.alloc(pn);
        c->dX.alloc(g.Xn * sizeof(float)); c->dZ.alloc(g.Xn * sizeof(float));
        c->counts.alloc(g.nx * sizeof(int32_t));
        for (auto &b : c->img) b.alloc((size_t)g.nx * g.nz * sizeof(float));
        c->rec.alloc((size_t)6 * g.nt * g.nx * sizeof(float));
        fused_init(c);
    });
    if (st != OUC_OK) {
        g_create_error = c->err;
        ouc_destroy(c);
        return st;
    }
    *out = c;
    return OUC_OK;
}

ouc_status ouc_set_profiling(ouc_ctx *c, int32_t on) {
    if (!c) return OUC_INVALID_ARG;
    c->profiling = on != 0;
    return OUC_OK;
}

ouc_status ouc_forward_shot(ouc_ctx *c, const ouc_shot *s, float *seis) {
    if (!c) return fail(nullptr, OUC_INVALID_ARG, "ctx is NULL");
    return guarded(c, [&] {
        check_shot(c, s);
        check_full_length(c, s);
        if (c->p.engine == OUC_ENGINE_FUSED) {
            fused_forward_one(c, s, seis);
            return;
        }
        if (c->p.engine == OUC_ENGINE_SPLIT_PML) {
            if (s->mute) throw Error(OUC_UNSUPPORTED, "the split-PML engine records unmuted seismograms");
            splitpml_forward(c, s, shot_wavelet(c, s), seis);
            return;
        }
        const Grid &g = c->g;
        std::vector<float> wav = shot_wavelet(c, s);
        c->last = ouc_timings{};
        StepProbe probe;
        probe.init(g.nt, c->profiling);
        c->F.zero(c->stream, g);
        OUC_HIP(hipEventRecord(c->ev[0], c->stream));
        run_forward_plain(c, s, wav, false, 6, probe);
        OUC_HIP(hipEventRecord(c->ev[1], c->stream));
        if (s->mute) apply_mute(c, s, 6);
        std::vector<float> tmp((size_t)6 * g.nt * g.nx);
        OUC_HIP(hipMemcpyAsync(tmp.data(), c->rec.ptr, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        OUC_HIP(hipEventRecord(c->ev[2], c->stream));
        OUC_HIP(hipStreamSynchronize(c->stream));
        c->last.forward_ms = ms_between(c->ev[0], c->ev[1]);
        c->last.total_ms = ms_between(c->ev[0], c->ev[2]);
        c->last.step_launches = g.nt;
        if (c->profiling) {
            c->last.step_kernel_ms = probe.total_ms();
            c->last.step_kernel_launches = probe.used;
        }
        double cells = (double)(g.Xn - 2 * g.N) * (g.Zn - 2 * g.N);
        fill_counts(c, cells * g.nt, 0, 0);
        if (seis)
            for (int q = 0; q < 6; ++q)
                for (int t = 0; t < g.nt; ++t)
                    for (int ix = 0; ix < g.nx; ++ix)
                        seis[((size_t)q * g.nx + ix) * g.nt + t] = tmp[((size_t)q * g.nt + t) * g.nx + ix];
    });
}

ouc_status ouc_rtm_shot(ouc_ctx *c, const ouc_shot *s, const float *rec_in, float *img_pp, float *img_ps,
                        float *img_pp_lap, float *img_ps_lap) {
    if (!c) return fail(nullptr, OUC_INVALID_ARG, "ctx is NULL");
    return guarded(c, [&] {
        check_shot(c, s);
        if (c->p.engine == OUC_ENGINE_SPLIT_PML) {
            check_full_length(c, s);
            const Grid &g = c->g;
            std::vector<float> wav = shot_wavelet(c, s);
            splitpml_rtm(c, s, wav, c->rec.f(), [&] {
                if (rec_in) upload_records(c, rec_in);
                if (s->mute) apply_mute(c, s, 2);
            }, c->npp.f(), c->nps.f(), c->den.f());
            make_images(c);
            OUC_HIP(hipEventRecord(c->ev[3], c->stream));
            const size_t ni = (size_t)g.nx * g.nz;
            download(c, img_pp, c->img[0], ni);
            download(c, img_ps, c->img[1], ni);
            download(c, img_pp_lap, c->img[2], ni);
            download(c, img_ps_lap, c->img[3], ni);
            OUC_HIP(hipStreamSynchronize(c->stream));
            c->last.forward_ms = ms_between(c->ev[0], c->ev[1]);
            c->last.backward_ms = ms_between(c->ev[1], c->ev[2]);
            c->last.image_ms = ms_between(c->ev[2], c->ev[3]);
            c->last.total_ms = ms_between(c->ev[0], c->ev[3]);
            return;
        }
        if (c->p.engine == OUC_ENGINE_FUSED) {
            if (rec_in) check_full_length(c, s);
            fused_rtm_one(c, s, rec_in, img_pp, img_ps, img_pp_lap, img_ps_lap);
            return;
        }
        check_full_length(c, s);
        const Grid &g = c->g;
        std::vector<float> wav = shot_wavelet(c, s);
        c->ensure_strips();
        c->last = ouc_timings{};
        StepProbe probe;
        probe.init(g.nt, c->profiling);
        c->F.zero(c->stream, g);
        OUC_HIP(hipEventRecord(c->ev[0], c->stream));
        run_forward_plain(c, s, wav, true, 2, probe);
        if (rec_in) upload_records(c, rec_in);
        if (s->mute) apply_mute(c, s, 2);
        OUC_HIP(hipEventRecord(c->ev[1], c->stream));
        run_backward_plain(c, s, wav);
        OUC_HIP(hipEventRecord(c->ev[2], c->stream));
        make_images(c);
        OUC_HIP(hipEventRecord(c->ev[3], c->stream));
        const size_t ni = (size_t)g.nx * g.nz;
        download(c, img_pp, c->img[0], ni);
        download(c, img_ps, c->img[1], ni);
        download(c, img_pp_lap, c->img[2], ni);
        download(c, img_ps_lap, c->img[3], ni);
        OUC_HIP(hipStreamSynchronize(c->stream));
        c->last.forward_ms = ms_between(c->ev[0], c->ev[1]);
        c->last.backward_ms = ms_between(c->ev[1], c->ev[2]);
        c->last.image_ms = ms_between(c->ev[2], c->ev[3]);
        c->last.total_ms = ms_between(c->ev[0], c->ev[3]);
        c->last.step_launches = 2 * (int64_t)g.nt;
        if (c->profiling) {
            c->last.step_kernel_ms = probe.total_ms();
            c->last.step_kernel_launches = probe.used;
        }
        double padded = (double)g.Xn * g.Zn, interior = (double)g.nx * g.nz;
        double cpml = padded - interior;
        double cu = (double)g.nt * (2 * padded + interior);
        double canon = (double)g.nt * (2 * (120 * interior + 232 * cpml) + 120 * interior + 48 * interior);
        fill_counts(c, cu, canon, 0);
    });
}

ouc_status ouc_rtm_grouped(ouc_ctx *const *ctxs, int32_t nctx, const ouc_shot *const *shots,
                           const int32_t *nshots, float *const *stack_pp_lap, float *const *stack_ps_lap) {
    if (!ctxs || nctx < 1 || !ctxs[0]) return fail(nullptr, OUC_INVALID_ARG, "need at least one context");
    if (!shots || !nshots) return fail(ctxs[0], OUC_INVALID_ARG, "shots and nshots must not be NULL");
    return guarded(ctxs[0], [&] { fused_rtm_group(ctxs, nctx, shots, nshots, stack_pp_lap, stack_ps_lap); });
}

ouc_status ouc_rtm_concurrent(ouc_ctx *const *ctxs, int32_t nctx, const ouc_shot *shots, int32_t nshots,
                              float *stack_pp_lap, float *stack_ps_lap, float *per_shot_ms) {
    if (!ctxs || nctx < 1 || !ctxs[0]) return fail(nullptr, OUC_INVALID_ARG, "need at least one context");
    if (nshots < 1 || !shots) return fail(ctxs[0], OUC_INVALID_ARG, "need at least one shot");
    return guarded(ctxs[0], [&] {
        for (int32_t k = 0; k < nshots; ++k) check_shot(ctxs[0], &shots[k]);
        fused_rtm_concurrent(ctxs, nctx, shots, nshots, stack_pp_lap, stack_ps_lap, per_shot_ms);
    });
}

ouc_status ouc_rtm_batch(ouc_ctx *c, const ouc_shot *shots, int32_t nshots, float *stack_pp_lap,
                         float *stack_ps_lap, float *per_shot_ms) {
    if (!c) return fail(nullptr, OUC_INVALID_ARG, "ctx is NULL");
    if (nshots < 1 || !shots) return fail(c, OUC_INVALID_ARG, "need at least one shot");
    if (c->p.engine == OUC_ENGINE_FUSED) {
        return guarded(c, [&] { fused_rtm_batch(c, shots, nshots, stack_pp_lap, stack_ps_lap, per_shot_ms); });
    }
    const Grid &g = c->g;
    const size_t ni = (size_t)g.nx * g.nz;
    std::vector<float> spp(ni, 0.f), sps(ni, 0.f), lpp(ni), lps(ni);
    ouc_timings agg{};
    double t_acc = 0;
    for (int32_t k = 0; k < nshots; ++k) {
        ouc_status st = ouc_rtm_shot(c, &shots[k], nullptr, nullptr, nullptr, lpp.data(), lps.data());
        if (st != OUC_OK) return st;
        for (size_t a = 0; a < ni; ++a) { spp[a] += lpp[a]; sps[a] += lps[a]; }
        t_acc += c->last.total_ms;
        if (per_shot_ms) per_shot_ms[k] = (float)t_acc;
        agg.total_ms += c->last.total_ms; agg.forward_ms += c->last.forward_ms;
        agg.backward_ms += c->last.backward_ms; agg.image_ms += c->last.image_ms;
        agg.step_launches += c->last.step_launches; agg.step_kernel_ms += c->last.step_kernel_ms;
        agg.step_kernel_launches += c->last.step_kernel_launches; agg.cell_updates += c->last.cell_updates;
        agg.canonical_bytes += c->last.canonical_bytes; agg.own_bytes += c->last.own_bytes;
    }
    c->last = agg;
    if (stack_pp_lap) std::memcpy(stack_pp_lap, spp.data(), ni * sizeof(float));
    if (stack_ps_lap) std::memcpy(stack_ps_lap, sps.data(), ni * sizeof(float));
    return OUC_OK;
}

ouc_status ouc_strip_codec_fp16(int32_t device, const float *x, int64_t n, int32_t steps, uint16_t *half_bits,
                                int8_t *exponents, float *decoded) {
    if (!x || n < 1 || steps < 1 || !half_bits || !exponents) return fail(nullptr, OUC_INVALID_ARG, "bad arguments");
    try {
        OUC_HIP(hipSetDevice(device));
        int64_t chunks = (n + 63) / 64;
        const size_t per = (size_t)RS_NF * n; // one slot of the store's layout; field 0 is used
        DevBuf din, dh, de, dkey, derr, ddec, dout;
        din.alloc((size_t)steps * n * sizeof(float));
        dh.alloc((size_t)steps * per * sizeof(__half));
        de.alloc((size_t)steps * RS_NF * chunks);
        dkey.alloc((size_t)((steps + kStripKey - 1) / kStripKey) * per * sizeof(float));
        derr.alloc(2 * per * sizeof(float));
        ddec.alloc(2 * per * sizeof(float));
        dout.alloc((size_t)steps * n * sizeof(float));
        OUC_HIP(hipMemcpy(din.ptr, x, (size_t)steps * n * sizeof(float), hipMemcpyHostToDevice));
        OUC_HIP(hipMemset(derr.ptr, 0, derr.bytes));
        KStrips s{};
        s.f16 = dh.as<__half>();
        s.e8 = de.as<int8_t>();
        s.key = dkey.f();
        s.err = derr.f();
        s.dec = ddec.f();
        s.ncells = n;
        s.nouter = n; // every cell residual-coded
        s.nchunks = chunks;
        s.fp16 = 1;
        s.last_v = s.last_th = steps - 1;
        const unsigned blocks = (unsigned)((n + 255) / 256);
        k_codec_fp16<<<blocks, 256>>>(s, din.f(), n, steps);
        for (int t = steps - 1; t >= 0; --t) { // backward decoding, as the reconstruction does
            k_strip_decode<<<dim3(blocks, 1, 1), 256>>>(s, t, KStripStride{});
            OUC_HIP(hipMemcpy(dout.f() + (size_t)t * n, ddec.f() + (size_t)(t & 1) * per, n * sizeof(float),
                              hipMemcpyDeviceToDevice));
        }
        OUC_HIP(hipGetLastError());
        OUC_HIP(hipMemcpy2D(half_bits, n * sizeof(uint16_t), dh.ptr, per * sizeof(__half), n * sizeof(uint16_t), steps,
                            hipMemcpyDeviceToHost));
        OUC_HIP(hipMemcpy2D(exponents, chunks, de.ptr, (size_t)RS_NF * chunks, chunks, steps, hipMemcpyDeviceToHost));
        if (decoded) OUC_HIP(hipMemcpy(decoded, dout.ptr, (size_t)steps * n * sizeof(float), hipMemcpyDeviceToHost));
        return OUC_OK;
    } catch (const Error &e) {
        return fail(nullptr, e.code, e.what());
    }
}

ouc_status ouc_fwi_gradient(ouc_ctx *c, const ouc_shot *shots, int32_t nshots, const float *observed,
                            float *grad_vp, float *grad_vs, float *illum, double *misfit) {
    if (!c) return fail(nullptr, OUC_INVALID_ARG, "ctx is NULL");
    if (nshots < 1 || !shots || !observed) return fail(c, OUC_INVALID_ARG, "need shots and their observed records");
    if (c->p.engine != OUC_ENGINE_FUSED) return fail(c, OUC_UNSUPPORTED, "the FWI gradient runs on the fused engine");
    return guarded(c, [&] { fused_fwi_gradient(c, shots, nshots, observed, grad_vp, grad_vs, illum, misfit); });
}

ouc_status ouc_model_update(int32_t device, int64_t n, const float *v, const float *grad, double step, int32_t iter,
                            float *out) {
    if (n < 1 || !v || !grad || !out || iter < 0) return fail(nullptr, OUC_INVALID_ARG, "bad arguments");
    try {
        fused_model_update(device, n, v, grad, step, iter, out);
        return OUC_OK;
    } catch (const Error &e) {
        return fail(nullptr, e.code, e.what());
    }
}

ouc_status ouc_get_timings(const ouc_ctx *c, ouc_timings *t) {
    if (!c || !t) return OUC_INVALID_ARG;
    *t = c->last;
    return OUC_OK;
}

ouc_status ouc_write_image_files(const char *dir, int32_t nx, int32_t nz, int32_t shot, const float *pp_lap,
                                 const float *ps_lap) {
    if (!dir || nx < 1 || nz < 1) return fail(nullptr, OUC_INVALID_ARG, "bad arguments");
    if (mkdir(dir, 0755) != 0 && errno != EEXIST) return fail(nullptr, OUC_IO, std::string("cannot create ") + dir);
    const float *imgs[2] = {pp_lap, ps_lap};
    const char *tags[2] = {"PP", "PS"};
    for (int a = 0; a < 2; ++a) {
        if (!imgs[a]) continue;
        char name[4096];
        if (shot >= 0) std::snprintf(name, sizeof name, "%s/image_%s_lap_%d_%d_%d.dat", dir, tags[a], nx, nz, shot);
        else std::snprintf(name, sizeof name, "%s/All_image_%s_lap.dat", dir, tags[a]);
        FILE *f = std::fopen(name, "wb");
        if (!f) return fail(nullptr, OUC_IO, std::string("cannot open ") + name);
        size_t n = (size_t)nx * nz;
        size_t w = std::fwrite(imgs[a], sizeof(float), n, f);
        if (std::fclose(f) != 0 || w != n) return fail(nullptr, OUC_IO, std::string("short write to ") + name);
    }
    return OUC_OK;
}

ouc_status ouc_set_graphs(ouc_ctx *c, int32_t on) {
    if (!c) return fail(nullptr, OUC_INVALID_ARG, "ctx is NULL");
    return guarded(c, [&] {
        c->graphs = on != 0;
        if (!c->graphs) fused_drop_graphs(c);
    });
}

ouc_status ouc_set_batch_capacity(ouc_ctx *c, int32_t shots, double *bytes_per_shot) {
    if (!c || shots < 0) return fail(c, OUC_INVALID_ARG, "bad arguments");
    return guarded(c, [&] {
        c->batch_cap = shots;
        if (bytes_per_shot) *bytes_per_shot = fused_shot_bytes(c);
    });
}

ouc_status ouc_set_tile(ouc_ctx *c, const ouc_tile *t) {
    if (!c || !t) return fail(c, OUC_INVALID_ARG, "bad arguments");
    return guarded(c, [&] { fused_set_tile(c, t->core_cols, t->prefetch); });
}

ouc_status ouc_get_tile(ouc_ctx *c, ouc_tile *t) {
    if (!c || !t) return fail(c, OUC_INVALID_ARG, "bad arguments");
    return guarded(c, [&] { fused_get_tile(c, &t->core_cols, &t->prefetch); });
}

ouc_status ouc_autotune(ouc_ctx *c, int32_t nshots, int32_t reps, ouc_tile *best, float *ms_per_launch) {
    if (!c || !best) return fail(c, OUC_INVALID_ARG, "bad arguments");
    return guarded(c, [&] {
        if (c->p.engine != OUC_ENGINE_FUSED) throw Error(OUC_INVALID_ARG, "the autotuner tunes the fused engine");
        fused_autotune(c, nshots, reps, &best->core_cols, &best->prefetch, ms_per_launch);
    });
}

ouc_status ouc_device_arch(int32_t device, char *buf, int32_t len) {
    if (!buf || len < 1) return fail(nullptr, OUC_INVALID_ARG, "bad arguments");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(nullptr, OUC_DEVICE, "no such HIP device");
    std::snprintf(buf, len, "%s", prop.gcnArchName);
    return OUC_OK;
}

ouc_status ouc_set_learned_boundary(ouc_ctx *c, const float *weights, int32_t n) {
    if (!c) return fail(nullptr, OUC_INVALID_ARG, "ctx is NULL");
    return guarded(c, [&] {
        if (!c->g.learned) throw Error(OUC_INVALID_ARG, "the context does not use the learned boundary");
        fused_set_learned(c, weights, n);
    });
}

ouc_status ouc_set_learned_damping(ouc_ctx *c, float beta) {
    if (!c) return fail(nullptr, OUC_INVALID_ARG, "ctx is NULL");
    return guarded(c, [&] {
        if (!c->g.learned) throw Error(OUC_INVALID_ARG, "the context does not use the learned boundary");
        fused_set_learned_damping(c, beta);
    });
}

ouc_status ouc_get_learned_filter(ouc_ctx *c, float *taps) {
    if (!c || !taps) return fail(c, OUC_INVALID_ARG, "bad arguments");
    return guarded(c, [&] { fused_get_learned_filter(c, taps); });
}

ouc_status ouc_forward_snapshots(ouc_ctx *c, const ouc_shot *s, int32_t n, const int32_t *steps, int32_t x0,
                                 int32_t z0, int32_t w, int32_t h, float *out) {
    if (!c) return fail(nullptr, OUC_INVALID_ARG, "ctx is NULL");
    return guarded(c, [&] {
        check_shot(c, s);
        check_full_length(c, s);
        if (c->p.engine != OUC_ENGINE_FUSED) throw Error(OUC_INVALID_ARG, "snapshots need the fused engine");
        fused_snapshots(c, s, n, steps, x0, z0, w, h, out);
    });
}

ouc_status ouc_recon_snapshots(ouc_ctx *c, const ouc_shot *s, int32_t n, const int32_t *steps, int32_t x0,
                               int32_t z0, int32_t w, int32_t h, float *out) {
    if (!c) return fail(nullptr, OUC_INVALID_ARG, "ctx is NULL");
    return guarded(c, [&] {
        check_shot(c, s);
        check_full_length(c, s);
        if (c->p.engine != OUC_ENGINE_FUSED) throw Error(OUC_INVALID_ARG, "snapshots need the fused engine");
        fused_recon_snapshots(c, s, n, steps, x0, z0, w, h, out);
    });
}

const char *ouc_last_error(const ouc_ctx *c) {
    if (c) return c->err.c_str();
    return g_create_error.c_str();
}

void ouc_destroy(ouc_ctx *c) {
    if (!c) return;
    if (c->fallback) ouc_destroy(c->fallback);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    fused_destroy(c);
    lap_fft_destroy(c);
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto &ev : c->kev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

} // extern "C"
