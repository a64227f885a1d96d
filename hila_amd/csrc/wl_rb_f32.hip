// Multi-rhs hopping kernel instantiated for F32 storage (see wl_rb.h).
#include "wl_rb.h"

namespace wl {

hipError_t launch_dslash_mr_f32(const Geom &g, const MrLaunch &L, hipStream_t s) {
    return L.sign > 0 ? rb_nr<float, 1>(g, L, s) : rb_nr<float, -1>(g, L, s);
}

}  // namespace wl
