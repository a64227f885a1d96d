// Host check of the site-tiled device layout (tile_elem, hila_amd/csrc/wl_store.h): for every
// pack size and block shape the kernels use, the map (entry e, site k) -> element is a
// bijection of the block [0, ne * S) onto itself, a 16-byte vector (P consecutive entries of
// one site) is P consecutive elements starting at a multiple of P, and the same vector of the
// 64 sites of a tile is one contiguous run (what a wave's coalesced load reads).
#include <cstdio>
#include <vector>

#include "wl_store.h"

template <int P>
static int check(int ne, int S) {
    std::vector<unsigned char> seen((size_t)ne * S, 0);
    const int npk = ne - ne % P;
    for (int k = 0; k < S; ++k)
        for (int e = 0; e < ne; ++e) {
            const size_t i = wl::tile_elem<P>(e, ne, k);
            if (i >= seen.size() || seen[i]) {
                std::printf("FAIL P=%d ne=%d S=%d e=%d k=%d -> %zu\n", P, ne, S, e, k, i);
                return 1;
            }
            seen[i] = 1;
            const int lane = k % wl::kTile;
            const size_t first = wl::tile_elem<P>(e, ne, k - lane);
            if (e < npk) {
                const int q0 = e - e % P;  // first entry of e's vector
                if (wl::tile_elem<P>(q0, ne, k) % P != 0 || i != wl::tile_elem<P>(q0, ne, k) + e % P ||
                    i != first + (size_t)P * lane) {
                    std::printf("FAIL packing P=%d ne=%d S=%d e=%d k=%d\n", P, ne, S, e, k);
                    return 1;
                }
            } else if (i != first + lane) {
                std::printf("FAIL run P=%d ne=%d S=%d e=%d k=%d\n", P, ne, S, e, k);
                return 1;
            }
        }
    return 0;
}

int main() {
    int bad = 0;
    for (int S : {64, 256, 320, 2176, 5184}) {
        bad += check<1>(12, S) + check<2>(12, S) + check<4>(12, S);  // spinors
        bad += check<1>(9, S) + check<2>(9, S) + check<4>(9, S);     // links
        bad += check<1>(6, S) + check<2>(6, S);                      // two-row links
        bad += check<4>(6, S);  // unused shape with two unpacked entries: the general rule
    }
    std::printf(bad ? "layout FAILED\n" : "layout ok\n");
    return bad ? 1 : 0;
}
