// On-disk formats either side of the path: the model dump the reference keeps
// commented out (P-S-seperating.cu:2042-2063) and the seismograms it never
// writes, as SEG-Y revision 1 files (one shot gather per file, one trace per
// receiver, IEEE float samples).  Host code only.
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "engine.hpp"

namespace {

// SEG-Y stores every header integer and sample big-endian.
void put16(uint8_t *p, int v) {
    const uint16_t u = (uint16_t)(int16_t)v;
    p[0] = (uint8_t)(u >> 8);
    p[1] = (uint8_t)u;
}
void put32(uint8_t *p, int64_t v) {
    const uint32_t u = (uint32_t)(int32_t)v;
    p[0] = (uint8_t)(u >> 24);
    p[1] = (uint8_t)(u >> 16);
    p[2] = (uint8_t)(u >> 8);
    p[3] = (uint8_t)u;
}
int get16(const uint8_t *p) { return (int16_t)(uint16_t)((p[0] << 8) | p[1]); }
int64_t get32(const uint8_t *p) {
    return (int32_t)(((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]);
}
void put_f32be(uint8_t *p, float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    put32(p, (int32_t)u);
}
float get_f32be(const uint8_t *p) {
    const uint32_t u = (uint32_t)get32(p);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// The textual header is EBCDIC (code page 037) in SEG-Y revision 1; the
// cards written here use upper-case letters, digits and a few punctuation marks.
char ebcdic(char c) {
    if (c >= '0' && c <= '9') return (char)(0xF0 + (c - '0'));
    if (c >= 'A' && c <= 'I') return (char)(0xC1 + (c - 'A'));
    if (c >= 'J' && c <= 'R') return (char)(0xD1 + (c - 'J'));
    if (c >= 'S' && c <= 'Z') return (char)(0xE2 + (c - 'S'));
    switch (c) {
    case '.': return (char)0x4B;
    case '(': return (char)0x4D;
    case '+': return (char)0x4E;
    case ')': return (char)0x5D;
    case '-': return (char)0x60;
    case '/': return (char)0x61;
    case ',': return (char)0x6B;
    case ':': return (char)0x7A;
    case '=': return (char)0x7E;
    default: return (char)0x40; // space (and anything else)
    }
}

constexpr int kText = 3200, kBin = 400, kTrace = 240;
constexpr int kCoordScale = 100; // coordinates stored in centimetres (scalar -100)

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

} // namespace

extern "C" {

ouc_status ouc_write_model_files(const char *dir, int32_t nx, int32_t nz, const float *vp, const float *vs) {
    if (!dir || nx < 1 || nz < 1 || (!vp && !vs)) return ouc::ouc_fail_global(OUC_INVALID_ARG, "bad arguments");
    if (mkdir(dir, 0755) != 0 && errno != EEXIST)
        return ouc::ouc_fail_global(OUC_IO, std::string("cannot create ") + dir);
    const float *arrs[2] = {vp, vs};
    const char *names[2] = {"Vp.dat", "Vs.dat"};
    for (int a = 0; a < 2; ++a) {
        if (!arrs[a]) continue;
        const std::string name = std::string(dir) + "/" + names[a];
        File out;
        out.f = std::fopen(name.c_str(), "wb");
        if (!out.f) return ouc::ouc_fail_global(OUC_IO, "cannot open " + name);
        const size_t n = (size_t)nx * nz;
        const size_t w = std::fwrite(arrs[a], sizeof(float), n, out.f);
        const int cl = std::fclose(out.f);
        out.f = nullptr;
        if (cl != 0 || w != n) return ouc::ouc_fail_global(OUC_IO, "short write to " + name);
    }
    return OUC_OK;
}

ouc_status ouc_write_segy(const char *path, const ouc_segy_gather *g, const float *traces) {
    if (!path || !g || !traces || g->ntraces < 1 || g->nt < 1)
        return ouc::ouc_fail_global(OUC_INVALID_ARG, "bad arguments");
    if (g->ntraces > 32767) return ouc::ouc_fail_global(OUC_UNSUPPORTED, "at most 32767 traces per gather");
    if (g->nt > 32767) return ouc::ouc_fail_global(OUC_UNSUPPORTED, "SEG-Y rev 1 holds at most 32767 samples per trace");
    const long us = std::lround((double)g->dt * 1e6);
    if (us < 1 || us > 32767) return ouc::ouc_fail_global(OUC_UNSUPPORTED, "sample interval outside 1..32767 us");
    File out;
    out.f = std::fopen(path, "wb");
    if (!out.f) return ouc::ouc_fail_global(OUC_IO, std::string("cannot open ") + path);
    // textual header: 40 card images of 80 EBCDIC characters
    std::vector<char> text(kText, ' ');
    const std::string cards[] = {
        "C 1 ELASTIC P/S RTM ENGINE SHOT GATHER, SEG-Y REV 1, IEEE FLOAT SAMPLES",
        "C 2 COMPONENT " + std::string(g->component[0] ? g->component : "VX"),
        "C 3 SHOT " + std::to_string(g->shot) + ", " + std::to_string(g->ntraces) + " RECEIVERS, " +
            std::to_string(g->nt) + " SAMPLES OF " + std::to_string(us) + " US",
        "C 4 COORDINATES IN METRES SCALED BY 1/100 (SCALAR -100), DEPTHS AS NEGATIVE ELEVATIONS",
        "C40 END TEXTUAL HEADER"};
    for (size_t k = 0; k < sizeof(cards) / sizeof(cards[0]); ++k) {
        const int line = k + 1 < sizeof(cards) / sizeof(cards[0]) ? (int)k : 39;
        std::memcpy(text.data() + 80 * line, cards[k].data(), std::min<size_t>(80, cards[k].size()));
    }
    for (char &ch : text) ch = ebcdic(ch);
    uint8_t bin[kBin] = {};
    put32(bin + 0, 1);               // 3201 job identification number
    put32(bin + 4, 1);               // 3205 line number
    put32(bin + 8, 1);               // 3209 reel number
    put16(bin + 12, g->ntraces);     // 3213 data traces per ensemble
    put16(bin + 16, (int)us);        // 3217 sample interval (us)
    put16(bin + 18, (int)us);        // 3219 sample interval of the original recording
    put16(bin + 20, g->nt);          // 3221 samples per trace
    put16(bin + 22, g->nt);          // 3223 samples per trace of the original recording
    put16(bin + 24, 5);              // 3225 sample format: 4-byte IEEE floating point
    put16(bin + 26, 1);              // 3227 ensemble fold
    put16(bin + 28, 1);              // 3229 trace sorting: as recorded
    put16(bin + 54, 1);              // 3255 measurement system: metres
    put16(bin + 300, 0x0100);        // 3501 SEG-Y format revision 1.0
    put16(bin + 302, 1);             // 3503 fixed-length traces
    put16(bin + 304, 0);             // 3505 no extended textual headers
    bool ok = std::fwrite(text.data(), 1, kText, out.f) == (size_t)kText && std::fwrite(bin, 1, kBin, out.f) == kBin;
    std::vector<uint8_t> tr(kTrace + (size_t)4 * g->nt);
    const double sx = g->sx_m, sz = g->sz_m, rz = g->rec_z_m;
    for (int k = 0; k < g->ntraces && ok; ++k) {
        std::memset(tr.data(), 0, kTrace);
        uint8_t *h = tr.data();
        const double rx = g->rec_x0_m + (double)k * g->rec_dx_m;
        put32(h + 0, k + 1);                                   // 1 trace sequence number within line
        put32(h + 4, k + 1);                                   // 5 trace sequence number within file
        put32(h + 8, g->shot);                                 // 9 original field record number
        put32(h + 12, k + 1);                                  // 13 trace number within the field record
        put32(h + 16, g->shot);                                // 17 energy source point number
        put16(h + 28, 1);                                      // 29 trace identification: seismic data
        put16(h + 30, 1);                                      // 31 vertically summed traces
        put16(h + 32, 1);                                      // 33 horizontally stacked traces
        put32(h + 36, std::lround((rx - sx) * kCoordScale));    // 37 source-receiver offset
        put32(h + 40, std::lround(-rz * kCoordScale));         // 41 receiver group elevation
        put32(h + 48, std::lround(sz * kCoordScale));          // 49 source depth below surface
        put16(h + 68, -kCoordScale);                           // 69 scalar for elevations and depths
        put16(h + 70, -kCoordScale);                           // 71 scalar for coordinates
        put32(h + 72, std::lround(sx * kCoordScale));          // 73 source X
        put32(h + 80, std::lround(rx * kCoordScale));          // 81 group X
        put16(h + 88, 1);                                      // 89 coordinate units: length
        put16(h + 114, g->nt);                                 // 115 samples in this trace
        put16(h + 116, (int)us);                               // 117 sample interval (us)
        const float *src = traces + (size_t)k * g->nt;
        for (int t = 0; t < g->nt; ++t) put_f32be(tr.data() + kTrace + 4 * (size_t)t, src[t]);
        ok = std::fwrite(tr.data(), 1, tr.size(), out.f) == tr.size();
    }
    const int cl = std::fclose(out.f);
    out.f = nullptr;
    if (!ok || cl != 0) return ouc::ouc_fail_global(OUC_IO, std::string("short write to ") + path);
    return OUC_OK;
}

ouc_status ouc_read_segy(const char *path, int32_t *ntraces, int32_t *nt, float *dt, float *traces, int64_t capacity) {
    if (!path || !ntraces || !nt) return ouc::ouc_fail_global(OUC_INVALID_ARG, "bad arguments");
    File in;
    in.f = std::fopen(path, "rb");
    if (!in.f) return ouc::ouc_fail_global(OUC_IO, std::string("cannot open ") + path);
    uint8_t bin[kBin];
    if (std::fseek(in.f, kText, SEEK_SET) != 0 || std::fread(bin, 1, kBin, in.f) != (size_t)kBin)
        return ouc::ouc_fail_global(OUC_IO, std::string("no SEG-Y binary header in ") + path);
    const int fmt = get16(bin + 24), ns = get16(bin + 20), us = get16(bin + 16), next = get16(bin + 304);
    if (fmt != 5) return ouc::ouc_fail_global(OUC_UNSUPPORTED, "only IEEE float (format 5) samples are read");
    if (ns < 1 || next != 0) return ouc::ouc_fail_global(OUC_UNSUPPORTED, "unsupported SEG-Y layout");
    if (std::fseek(in.f, 0, SEEK_END) != 0) return ouc::ouc_fail_global(OUC_IO, "seek failed");
    const long size = std::ftell(in.f);
    const long per = kTrace + 4L * ns;
    if (size < kText + kBin || (size - kText - kBin) % per != 0)
        return ouc::ouc_fail_global(OUC_IO, std::string("truncated SEG-Y file ") + path);
    const long n = (size - kText - kBin) / per;
    *ntraces = (int32_t)n;
    *nt = ns;
    if (dt) *dt = (float)(us * 1e-6);
    if (!traces) return OUC_OK; // size query
    if (capacity < (int64_t)n * ns) return ouc::ouc_fail_global(OUC_INVALID_ARG, "trace buffer too small");
    std::vector<uint8_t> tr(per);
    if (std::fseek(in.f, kText + kBin, SEEK_SET) != 0) return ouc::ouc_fail_global(OUC_IO, "seek failed");
    for (long k = 0; k < n; ++k) {
        if (std::fread(tr.data(), 1, per, in.f) != (size_t)per) return ouc::ouc_fail_global(OUC_IO, "short read");
        if (get16(tr.data() + 114) != ns) return ouc::ouc_fail_global(OUC_UNSUPPORTED, "variable trace length");
        for (int t = 0; t < ns; ++t) traces[k * ns + t] = get_f32be(tr.data() + kTrace + 4 * (size_t)t);
    }
    return OUC_OK;
}

} // extern "C"
