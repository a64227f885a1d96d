// HILA gauge-configuration files on the device: GaugeField<SU<3,double>>::config_read /
// config_write (libraries/plumbing/gaugefield.h:194-272) and the gauge download.
//
// File format (gaugefield.h:200-212, Field<T>::write field_io.h:125-160): 7 native-endian
// int64 {394824242, NDIM = 4, sizeof(T) = 144, Lx, Ly, Lz, Lt}, then for mu = 0..3 the link
// U_mu(x) of every site in lexicographic order, x fastest (lattice.h:389-396), 18 doubles each
// (3x3 complex, row-major).  A file whose lattice divides the context's in every direction is
// replicated periodically to fill it (Field<T>::read(stream, insize), field_io.h:245-312).
//
// The file is streamed through two pinned chunk buffers into the context's device staging
// buffer (one direction at a time), and one kernel per direction scatters the lexicographic
// links into the gauge blocks, tiling a smaller file lattice as it goes; writing runs the
// same pipeline backwards.  Host code only reads and writes bytes: every reordering is done
// on the GPU.

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <memory>
#include <sys/stat.h>

#include "wl_host.h"
#include "wl_store.h"

namespace wl {

namespace {

constexpr int64_t kConfigFlag = 394824242;   // gaugefield.h:28
constexpr int64_t kLinkBytes = 144;          // sizeof(SU<3,double>)
constexpr int kHeaderWords = 3 + 4;
constexpr size_t kChunk = (size_t)16 << 20;  // bytes per pinned chunk
const char *const kDirName[4] = {"x", "y", "z", "t"};

struct d2 {
    double x, y;
};

// Lexicographic links of a file lattice `in` (lex index x + in0 (y + in1 (z + in2 t)), 18
// doubles each) -> direction mu of the gauge blocks; checkerboard site (p, k) takes the link
// of its coordinates reduced modulo the file lattice.
__global__ __launch_bounds__(kBlock) void k_gauge_from_lex(const d2 *__restrict__ lex, cx<double> *__restrict__ gauge,
                                                           int mu, Geom g, int in0, int in1, int in2, int in3) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= 2 * g.vh) return;
    const int p = s >= g.vh;
    const int k = s - p * g.vh;
    const unsigned q = 2u * (unsigned)k;
    const unsigned x0 = q % (unsigned)g.L[0];
    unsigned r = q / (unsigned)g.L[0];
    const unsigned y = r % (unsigned)g.L[1];
    r /= (unsigned)g.L[1];
    const unsigned z = r % (unsigned)g.L[2];
    const unsigned t = r / (unsigned)g.L[2];
    const unsigned x = x0 + (((unsigned)p + y + z + t) & 1u);
    const size_t src = (size_t)(x % (unsigned)in0) +
                       (size_t)in0 * ((y % (unsigned)in1) + (size_t)in1 * ((z % (unsigned)in2) +
                                                                           (size_t)in2 * (t % (unsigned)in3)));
    const d2 *a = lex + 9 * src;
    cx<double> *dst = gauge + (size_t)((p * 4 + mu) * 9) * g.stride;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        const d2 v = a[e];
        dst[link_entry<double>(e, g.stride, k)] = {v.x, v.y};
    }
}

// Direction mu of the gauge blocks -> 18 doubles per site, sites in lexicographic order
// (lexicographic != 0, the file's order) or in HILA's EVEN_SITES_FIRST order.
__global__ __launch_bounds__(kBlock) void k_gauge_to_host(const cx<double> *__restrict__ gauge, d2 *__restrict__ out,
                                                          int mu, Geom g, int lexicographic) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= 2 * g.vh) return;
    int p, k;
    if (lexicographic) {
        unsigned r = (unsigned)i;
        unsigned sum = r % (unsigned)g.L[0];
        r /= (unsigned)g.L[0];
        sum += r % (unsigned)g.L[1];
        r /= (unsigned)g.L[1];
        sum += r % (unsigned)g.L[2] + r / (unsigned)g.L[2];
        p = (int)(sum & 1u);
        k = i >> 1;
    } else {
        p = i >= g.vh;
        k = i - p * g.vh;
    }
    const cx<double> *src = gauge + (size_t)((p * 4 + mu) * 9) * g.stride;
    d2 *o = out + (size_t)9 * i;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        const cx<double> v = src[link_entry<double>(e, g.stride, k)];
        o[e] = {v.x, v.y};
    }
}

// Two pinned host buffers, each with the event of the last copy that used it.
struct PinnedPair {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    size_t used[2] = {0, 0};
    size_t bytes = 0;

    int init(size_t n) {
        bytes = n;
        for (int b = 0; b < 2; ++b) {
            HIPCHK(hipHostMalloc(&buf[b], n, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
        }
        return WL_OK;
    }
    int wait(int b) {
        if (pending[b]) {
            HIPCHK(hipEventSynchronize(ev[b]));
            pending[b] = false;
        }
        return WL_OK;
    }
    ~PinnedPair() {
        for (int b = 0; b < 2; ++b) {
            if (ev[b]) {
                (void)hipEventSynchronize(ev[b]);
                (void)hipEventDestroy(ev[b]);
            }
            if (buf[b]) (void)hipHostFree(buf[b]);
        }
    }
};

struct FileCloser {
    void operator()(FILE *f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

dim3 site_grid(const Geom &g) { return dim3((unsigned)((2 * (size_t)g.vh + kBlock - 1) / kBlock)); }

// Header checks of config_read (gaugefield.h:219-262), with its messages.  `f` is positioned
// after the header on success.
int read_header(FILE *f, const char *path, int insize[4]) {
    int64_t h[kHeaderWords];
    if (std::fread(h, sizeof(int64_t), kHeaderWords, f) != (size_t)kHeaderWords)
        return fail(WL_E_ARG, "CONFIG ERROR in file %s: header too short", path);
    if (h[0] != kConfigFlag)
        return fail(WL_E_ARG, "CONFIG ERROR in file %s: wrong id, should be %lld is %lld", path,
                    (long long)kConfigFlag, (long long)h[0]);
    if (h[1] != 4)
        return fail(WL_E_ARG, "CONFIG ERROR in file %s: wrong dimensionality, should be 4 is %lld", path,
                    (long long)h[1]);
    if (h[2] != kLinkBytes)
        return fail(WL_E_ARG, "CONFIG ERROR in file %s: wrong size of field element, should be %lld is %lld",
                    path, (long long)kLinkBytes, (long long)h[2]);
    for (int d = 0; d < 4; ++d) {
        if (h[3 + d] < 1 || h[3 + d] > (1 << 20))
            return fail(WL_E_ARG, "CONFIG ERROR in file %s: incorrect lattice dimension %s is %lld", path,
                        kDirName[d], (long long)h[3 + d]);
        insize[d] = (int)h[3 + d];
    }
    return WL_OK;
}

int open_config(const char *path, File &f, int insize[4], size_t *dir_bytes) {
    if (!path) return fail(WL_E_ARG, "null path");
    f.reset(std::fopen(path, "rb"));
    if (!f) return fail(WL_E_ARG, "cannot open %s: %s", path, std::strerror(errno));
    if (int rc = read_header(f.get(), path, insize)) return rc;
    if ((double)insize[0] * insize[1] * insize[2] * insize[3] > 2147483648.0)
        return fail(WL_E_ARG, "CONFIG ERROR in file %s: lattice %d x %d x %d x %d too large", path, insize[0],
                    insize[1], insize[2], insize[3]);
    const size_t vin = (size_t)insize[0] * insize[1] * insize[2] * insize[3];
    *dir_bytes = vin * kLinkBytes;
    struct stat sb;
    if (fstat(fileno(f.get()), &sb) != 0) return fail(WL_E_ARG, "cannot stat %s: %s", path, std::strerror(errno));
    const size_t need = sizeof(int64_t) * kHeaderWords + 4 * *dir_bytes;
    if ((size_t)sb.st_size < need)
        return fail(WL_E_ARG, "CONFIG ERROR in file %s: %lld bytes, the header's lattice needs %zu", path,
                    (long long)sb.st_size, need);
    return WL_OK;
}

int ensure_gauge64(wl_ctx *c) {
    if (c->gauge64) return WL_OK;
    if (int rc = dev_alloc(c, &c->gauge64, gauge_elems(c->g) * 16, "gauge field")) return rc;
    HIPCHK(hipMemsetAsync(c->gauge64, 0, gauge_elems(c->g) * 16, c->stream));
    return WL_OK;
}

}  // namespace

hipError_t launch_links_to_host(const Geom &g, const void *planes, int mu, double *out, int lexicographic,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_gauge_to_host, site_grid(g), dim3(kBlock), 0, s, (const cx<double> *)planes, (d2 *)out, mu,
                       g, lexicographic);
    return hipGetLastError();
}

int config_upload(wl_ctx *c, const char *path) {
    File f;
    int insize[4];
    size_t dir_bytes = 0;
    if (int rc = open_config(path, f, insize, &dir_bytes)) return rc;
    for (int d = 0; d < 4; ++d)
        if (c->g.L[d] % insize[d] != 0)
            return fail(WL_E_ARG, "CONFIG ERROR in file %s: incorrect lattice dimension %s is %d should be (or divide)%d",
                        path, kDirName[d], insize[d], c->g.L[d]);
    if (int rc = ensure_gauge64(c)) return rc;
    if (int rc = ensure_staging(c, dir_bytes)) return rc;
    PinnedPair pp;
    if (int rc = pp.init(std::min(kChunk, dir_bytes))) return rc;
    // From here on a failure leaves the gauge partly overwritten: mark it unusable.
    c->gauge_ok = c->gauge32_ok = c->gauge16_ok = false;
    links_changed(c);
    size_t n_chunk = 0;
    for (int mu = 0; mu < 4; ++mu) {
        for (size_t off = 0; off < dir_bytes; off += pp.bytes, ++n_chunk) {
            const int b = (int)(n_chunk & 1);
            const size_t n = std::min(pp.bytes, dir_bytes - off);
            if (int rc = pp.wait(b)) return rc;
            if (std::fread(pp.buf[b], 1, n, f.get()) != n)
                return fail(WL_E_ARG, "CONFIG ERROR in file %s: read failed at direction %d", path, mu);
            HIPCHK(hipMemcpyAsync((char *)c->staging + off, pp.buf[b], n, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipEventRecord(pp.ev[b], c->stream));
            pp.pending[b] = true;
        }
        hipLaunchKernelGGL(k_gauge_from_lex, site_grid(c->g), dim3(kBlock), 0, c->stream, (const d2 *)c->staging,
                           (cx<double> *)c->gauge64, mu, c->g, insize[0], insize[1], insize[2], insize[3]);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->gauge_ok = true;
    return WL_OK;
}

namespace {

int config_write(wl_ctx *c, const char *path) {
    if (!c->gauge_ok) return fail(WL_E_STATE, "no gauge field uploaded");
    File f(std::fopen(path, "wb"));
    if (!f) return fail(WL_E_ARG, "cannot open %s for writing: %s", path, std::strerror(errno));
    const int64_t h[kHeaderWords] = {kConfigFlag, 4, kLinkBytes, c->g.L[0], c->g.L[1], c->g.L[2], c->g.L[3]};
    if (std::fwrite(h, sizeof(int64_t), kHeaderWords, f.get()) != (size_t)kHeaderWords)
        return fail(WL_E_ARG, "write to %s failed: %s", path, std::strerror(errno));
    const size_t dir_bytes = (size_t)c->V * kLinkBytes;
    if (int rc = ensure_staging(c, dir_bytes)) return rc;
    PinnedPair pp;
    if (int rc = pp.init(std::min(kChunk, dir_bytes))) return rc;
    // Chunk i goes to buffer i & 1; before buffer b is refilled, the chunk it holds (i - 2)
    // is written out, so the file receives the chunks in order.
    auto flush = [&](int b) -> int {
        if (!pp.pending[b]) return WL_OK;
        if (int rc = pp.wait(b)) return rc;
        if (std::fwrite(pp.buf[b], 1, pp.used[b], f.get()) != pp.used[b])
            return fail(WL_E_ARG, "write to %s failed: %s", path, std::strerror(errno));
        return WL_OK;
    };
    size_t n_chunk = 0;
    for (int mu = 0; mu < 4; ++mu) {
        hipLaunchKernelGGL(k_gauge_to_host, site_grid(c->g), dim3(kBlock), 0, c->stream, (const cx<double> *)c->gauge64,
                           (d2 *)c->staging, mu, c->g, 1);
        HIPCHK(hipGetLastError());
        for (size_t off = 0; off < dir_bytes; off += pp.bytes, ++n_chunk) {
            const int b = (int)(n_chunk & 1);
            const size_t n = std::min(pp.bytes, dir_bytes - off);
            if (int rc = flush(b)) return rc;
            HIPCHK(hipMemcpyAsync(pp.buf[b], (const char *)c->staging + off, n, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipEventRecord(pp.ev[b], c->stream));
            pp.pending[b] = true;
            pp.used[b] = n;
        }
    }
    if (int rc = flush((int)(n_chunk & 1))) return rc;
    if (int rc = flush((int)((n_chunk + 1) & 1))) return rc;
    FILE *raw = f.release();
    if (std::fclose(raw) != 0) return fail(WL_E_ARG, "closing %s failed: %s", path, std::strerror(errno));
    return WL_OK;
}

}  // namespace

}  // namespace wl

using namespace wl;

extern "C" int wl_config_header(const char *path, int dims[4]) {
    if (!dims) return fail(WL_E_ARG, "null argument");
    File f;
    size_t dir_bytes;
    return open_config(path, f, dims, &dir_bytes);
}

extern "C" int wl_config_read(wl_ctx *c, const char *path) {
    if (!c || !path) return fail(WL_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    return config_upload(c, path);
}

extern "C" int wl_config_write(wl_ctx *c, const char *path) {
    if (!c || !path) return fail(WL_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    return config_write(c, path);
}

extern "C" int wl_gauge_download(wl_ctx *c, double *const U[4]) {
    if (!c || !U) return fail(WL_E_ARG, "null argument");
    for (int mu = 0; mu < 4; ++mu)
        if (!U[mu]) return fail(WL_E_ARG, "null gauge direction %d", mu);
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    if (!c->gauge_ok) return fail(WL_E_STATE, "no gauge field uploaded");
    const size_t bytes = (size_t)c->V * kLinkBytes;
    if (int rc = ensure_staging(c, bytes)) return rc;
    for (int mu = 0; mu < 4; ++mu) {
        hipLaunchKernelGGL(k_gauge_to_host, site_grid(c->g), dim3(kBlock), 0, c->stream, (const cx<double> *)c->gauge64,
                           (d2 *)c->staging, mu, c->g, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(U[mu], c->staging, bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return WL_OK;
}
