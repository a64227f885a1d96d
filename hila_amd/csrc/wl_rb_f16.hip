// Multi-rhs hopping kernel instantiated for F16 storage (see wl_rb.h), and the diagnostic that
// runs the packed FP16 projection's operand forms (pk_cadd, wl_store.h) on given inputs.
#include "wl_host.h"
#include "wl_rb.h"

namespace wl {

hipError_t launch_dslash_mr_f16(const Geom &g, const MrLaunch &L, hipStream_t s) {
    return L.sign > 0 ? rb_nr<f16s, 1>(g, L, s) : rb_nr<f16s, -1>(g, L, s);
}

namespace {

// out[f * n + i] = pk_cadd<C, F>(a[i], b[i]) for the 16 forms f = 4 (C index) + F, C index
// 0..3 = C +1, -1, +i, -i.
__global__ void k_fp16_forms(const h2 *a, const h2 *b, int n, h2 *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const h2 x = a[i], y = b[i];
    h2 r[16];
    r[0] = pk_cadd<1, 0>(x, y);
    r[1] = pk_cadd<1, 1>(x, y);
    r[2] = pk_cadd<1, 2>(x, y);
    r[3] = pk_cadd<1, 3>(x, y);
    r[4] = pk_cadd<-1, 0>(x, y);
    r[5] = pk_cadd<-1, 1>(x, y);
    r[6] = pk_cadd<-1, 2>(x, y);
    r[7] = pk_cadd<-1, 3>(x, y);
    r[8] = pk_cadd<2, 0>(x, y);
    r[9] = pk_cadd<2, 1>(x, y);
    r[10] = pk_cadd<2, 2>(x, y);
    r[11] = pk_cadd<2, 3>(x, y);
    r[12] = pk_cadd<-2, 0>(x, y);
    r[13] = pk_cadd<-2, 1>(x, y);
    r[14] = pk_cadd<-2, 2>(x, y);
    r[15] = pk_cadd<-2, 3>(x, y);
#pragma unroll
    for (int f = 0; f < 16; ++f) out[(size_t)f * n + i] = r[f];
}

}  // namespace

}  // namespace wl

using namespace wl;

extern "C" int wl_fp16_forms(int device, const uint32_t *a, const uint32_t *b, int n, uint32_t *out) {
    if (!a || !b || !out || n < 1) return fail(WL_E_ARG, "null argument or n < 1");
    DeviceGuard dg(device);
    void *d = nullptr;
    const size_t in_bytes = (size_t)n * 4, out_bytes = (size_t)16 * n * 4;
    if (hipMalloc(&d, 2 * in_bytes + out_bytes) != hipSuccess) return fail(WL_E_OOM, "device buffer of %d pairs", n);
    char *p = (char *)d;
    hipError_t e = hipMemcpy(p, a, in_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p + in_bytes, b, in_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fp16_forms, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, (const h2 *)p,
                           (const h2 *)(p + in_bytes), n, (h2 *)(p + 2 * in_bytes));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, p + 2 * in_bytes, out_bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(WL_E_DEVICE, "FP16 operand forms: %s", hipGetErrorString(e));
    return WL_OK;
}
