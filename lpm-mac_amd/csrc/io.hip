// Asynchronous Assembly::writeDump (assembly.h:309-335).
//
// The reference formats one text line per particle with fprintf inside the load-step loop
// (solver_static.h:89, about 50 MB per step at 262,144 particles), which would hold the GPU idle
// for longer than a step takes. Here the step's fields are copied on the device in stream order
// (so later steps may overwrite the live arrays at once), copied to pinned host memory on a side
// stream, and a writer thread formats the lines in parallel chunks and appends them to the file
// in particle order. The text is byte-identical to the reference's fprintf.
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "lpm_internal.h"

struct DumpState {
    int64_t n = 0;
    double *dsnap = nullptr;   // device: xyz, xyz0, stress, damage, damage_visual, state_var[0] (15 per particle)
    double *hsnap = nullptr;   // pinned host copy of dsnap
    int32_t *dtype = nullptr;  // device copy of type
    int32_t *htype = nullptr;  // pinned host copy
    hipStream_t side = nullptr;
    hipEvent_t snapped = nullptr, copied = nullptr;
    std::thread writer;
    bool busy = false;
    int status = LPM_OK;       // of the last completed write
    std::string error;
};

namespace {

#define RC(x)                \
    do {                     \
        const int rc_ = (x); \
        if (rc_) return rc_; \
    } while (0)

constexpr int DUMP_FIELDS = 15;

// printf("%.4e") for the dump. Fast path: with |E - 4| <= 22 the power of ten is exact, so the
// scaled mantissa a / 10^(E-4) (or a * 10^(4-E)) is within half an ulp of the exact value, an
// absolute error below 2e-11 at five digits; when its fractional part is farther than 1e-9 from
// one half, rounding to five digits is decided and the text equals printf's. Zero is written
// directly; near-ties, very large or small magnitudes, infinities and NaN go to snprintf.
const double POW10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                          1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

char *fmt_e4(char *p, double x) {
    if (x == 0.0) {
        const char *z = std::signbit(x) ? "-0.0000e+00" : "0.0000e+00";
        while (*z) *p++ = *z++;
        return p;
    }
    const double a = std::fabs(x);
    if (std::isfinite(a)) {
        int E = (int)std::floor(std::log10(a));
        for (int pass = 0; pass < 3; ++pass) {
            const int s = E - 4;
            if (s > 22 || s < -22) break;
            const double m = s >= 0 ? a / POW10[s] : a * POW10[-s];
            if (m < 10000.0) { --E; continue; }     // log10 rounded up across a power of ten
            if (m >= 100000.0) { ++E; continue; }   // ... or down
            const double fl = std::floor(m), fr = m - fl;
            if (std::fabs(fr - 0.5) < 1e-9) break;  // too close to a tie: let printf decide
            long q = (long)fl + (fr > 0.5 ? 1 : 0);
            if (q == 100000) {                      // rounded up to the next decade: 9.99995.. -> 1.0000
                q = 10000;
                ++E;
            }
            if (x < 0) *p++ = '-';
            *p++ = (char)('0' + q / 10000);
            *p++ = '.';
            const long r = q % 10000;
            *p++ = (char)('0' + r / 1000);
            *p++ = (char)('0' + r / 100 % 10);
            *p++ = (char)('0' + r / 10 % 10);
            *p++ = (char)('0' + r % 10);
            *p++ = 'e';
            *p++ = E < 0 ? '-' : '+';
            int ae = E < 0 ? -E : E;
            if (ae >= 100) {
                *p++ = (char)('0' + ae / 100);
                ae %= 100;
            }
            *p++ = (char)('0' + ae / 10);
            *p++ = (char)('0' + ae % 10);
            return p;
        }
    }
    char buf[48];
    const int len = snprintf(buf, sizeof buf, "%.4e", x);
    for (int k = 0; k < len && k < (int)sizeof buf - 1; ++k) *p++ = buf[k];
    return p;
}

char *fmt_int(char *p, long v) {
    char buf[24];
    int k = 0;
    const bool neg = v < 0;
    unsigned long u = neg ? 0UL - (unsigned long)v : (unsigned long)v;
    do {
        buf[k++] = (char)('0' + u % 10);
        u /= 10;
    } while (u);
    if (neg) *p++ = '-';
    while (k) *p++ = buf[--k];
    return p;
}

int dump_setup(lpm_ctx *c) {
    if (c->dump) return LPM_OK;
    DumpState *d = new DumpState();
    c->dump = d;
    d->n = c->n;
    const size_t bytes = sizeof(double) * DUMP_FIELDS * (size_t)c->n;
    if (hipMalloc(&d->dsnap, bytes) != hipSuccess || hipMalloc(&d->dtype, sizeof(int32_t) * c->n) != hipSuccess ||
        hipHostMalloc((void **)&d->hsnap, bytes) != hipSuccess ||
        hipHostMalloc((void **)&d->htype, sizeof(int32_t) * c->n) != hipSuccess)
        return lpm_fail(c, LPM_E_NOMEM, "dump buffers");
    if (hipStreamCreateWithFlags(&d->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&d->snapped, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&d->copied, hipEventDisableTiming) != hipSuccess)
        return lpm_fail(c, LPM_E_HIP, "dump stream / events");
    return LPM_OK;
}

// The writer thread: wait for the host copy, format in parallel chunks, append in order.
void write_dump(DumpState *d, int device, std::string path, int step, std::array<double, 6> box) {
    if (hipSetDevice(device) != hipSuccess || hipEventSynchronize(d->copied) != hipSuccess) {
        d->status = LPM_E_HIP;
        d->error = "dump copy failed";
        return;
    }
    const int64_t n = d->n;
    const double *xyz = d->hsnap, *x0 = xyz + 3 * n, *st = x0 + 3 * n, *dm = st + 6 * n, *dv = dm + n, *s0 = dv + n;
    constexpr int64_t CH = 8192;
    const int64_t nch = (n + CH - 1) / CH;
    std::vector<std::string> text(nch);
    // leave two cores to the thread that keeps launching the solver's work
    const int nt = std::max(1, omp_get_max_threads() - 2);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t k = 0; k < nch; ++k) {
        std::string &t = text[k];
        const int64_t i0 = k * CH, i1 = std::min(n, (k + 1) * CH);
        t.resize((size_t)(i1 - i0) * 17 * 20);  // at most 17 fields of 20 bytes per line
        char *p = &t[0];
        // "%d %d %.4e x15\n", the reference's line
        for (int64_t i = i0; i < i1; ++i) {
            const double v[15] = {xyz[3 * i],          xyz[3 * i + 1],          xyz[3 * i + 2],
                                  xyz[3 * i] - x0[3 * i], xyz[3 * i + 1] - x0[3 * i + 1], xyz[3 * i + 2] - x0[3 * i + 2],
                                  st[6 * i],           st[6 * i + 1],           st[6 * i + 2],
                                  st[6 * i + 3],       st[6 * i + 4],           st[6 * i + 5],
                                  dm[i],               dv[i],                   s0[i]};
            p = fmt_int(p, (long)i);
            *p++ = ' ';
            p = fmt_int(p, d->htype[i]);
            for (int f = 0; f < 15; ++f) {
                *p++ = ' ';
                p = fmt_e4(p, v[f]);
            }
            *p++ = '\n';
        }
        t.resize((size_t)(p - &t[0]));
    }
    FILE *f = fopen(path.c_str(), "a+");
    if (!f) {
        d->status = LPM_E_ARG;
        d->error = "cannot open " + path;
        return;
    }
    fprintf(f, "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n", step, (int)n);
    fprintf(f, "%8.8f %8.8f\n%8.8f %8.8f\n%8.8f %8.8f\n", box[0], box[1], box[2], box[3], box[4], box[5]);
    fprintf(f, "ITEM: ATOMS id type x y z dx dy dz s11 s22 s33 s23 s13 s12 damage damage_visual state_var[0] \n");
    bool ok = true;
    for (const std::string &t : text) ok = ok && fwrite(t.data(), 1, t.size(), f) == t.size();
    ok = (fclose(f) == 0) && ok;
    d->status = ok ? LPM_OK : LPM_E_ARG;
    if (!ok) d->error = "write to " + path + " failed";
}

}  // namespace

void dump_free(lpm_ctx *c) {
    DumpState *d = c->dump;
    if (!d) return;
    if (d->writer.joinable()) d->writer.join();
    if (d->side) hipStreamSynchronize(d->side);
    if (d->snapped) hipEventDestroy(d->snapped);
    if (d->copied) hipEventDestroy(d->copied);
    if (d->side) hipStreamDestroy(d->side);
    if (d->dsnap) hipFree(d->dsnap);
    if (d->dtype) hipFree(d->dtype);
    if (d->hsnap) hipHostFree(d->hsnap);
    if (d->htype) hipHostFree(d->htype);
    delete d;
    c->dump = nullptr;
}

extern "C" {

int lpm_format_e4(const double *x, int64_t n, char *out, int64_t cap) {
    if ((!x && n > 0) || !out || cap < 1) return LPM_E_ARG;
    char *p = out;
    for (int64_t i = 0; i < n; ++i) {
        if (p - out + 48 >= cap) return LPM_E_LIMIT;
        p = fmt_e4(p, x[i]);
        *p++ = '\n';
    }
    *p = 0;
    return LPM_OK;
}

int lpm_dump_wait(lpm_ctx *c) {
    if (!c) return LPM_E_ARG;
    DumpState *d = c->dump;
    if (!d || !d->busy) return LPM_OK;
    d->writer.join();
    d->busy = false;
    if (d->status != LPM_OK) return lpm_fail(c, d->status, d->error);
    return LPM_OK;
}

int lpm_dump_async(lpm_ctx *c, const char *path, int32_t step, const double *box) {
    if (!c || !path || !box) return LPM_E_ARG;
    RC(lpm_dump_wait(c));  // one write in flight: the previous one must have left the buffers
    RC(dump_setup(c));
    DumpState *d = c->dump;
    const size_t n = c->n, D = sizeof(double);
    double *s = d->dsnap;
    // snapshot in stream order, after everything queued so far
    LPM_HIP(c, hipMemcpyAsync(s, c->xyz, 3 * n * D, hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(s + 3 * n, c->xyz0, 3 * n * D, hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(s + 6 * n, c->stress, 6 * n * D, hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(s + 12 * n, c->damage, n * D, hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(s + 13 * n, c->damage_visual, n * D, hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipMemcpy2DAsync(s + 14 * n, D, c->sv, LPM_NSV * D, D, n, hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(d->dtype, c->type, n * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipEventRecord(d->snapped, c->stream));
    // host copy on the side stream, overlapping whatever the context stream runs next
    LPM_HIP(c, hipStreamWaitEvent(d->side, d->snapped, 0));
    LPM_HIP(c, hipMemcpyAsync(d->hsnap, s, DUMP_FIELDS * n * D, hipMemcpyDeviceToHost, d->side));
    LPM_HIP(c, hipMemcpyAsync(d->htype, d->dtype, n * sizeof(int32_t), hipMemcpyDeviceToHost, d->side));
    LPM_HIP(c, hipEventRecord(d->copied, d->side));
    std::array<double, 6> b;  // copied by value into the thread: the caller's array may not outlive it
    for (int k = 0; k < 6; ++k) b[k] = box[k];
    d->status = LPM_OK;
    d->busy = true;
    d->writer = std::thread(write_dump, d, c->device, std::string(path), (int)step, b);
    return LPM_OK;
}

}  // extern "C"
