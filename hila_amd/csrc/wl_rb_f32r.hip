// Multi-rhs hopping kernel instantiated for F32 storage with two-row links (see wl_rb.h and
// ld_link_raw in wl_store.h); a unit of its own so that it compiles in parallel with the
// full-link instantiations.
#include "wl_rb.h"

namespace wl {

hipError_t launch_dslash_mr_f32_rows(const Geom &g, const MrLaunch &L, hipStream_t s) {
    return L.sign > 0 ? rb_nr<float, 1, 6>(g, L, s) : rb_nr<float, -1, 6>(g, L, s);
}

}  // namespace wl
