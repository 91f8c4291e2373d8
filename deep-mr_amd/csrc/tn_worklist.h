// Work items of the gridding and interpolation launches, and the host-side
// construction of the tile-matrix gridding items. Plain C++ (no HIP types),
// so the construction can be checked by a host-only program.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace tn {

// Work item. Gathering interpolation: unit = sample block. Staged
// interpolation: unit = tile, [e0, e1) = slice of its entries. Gridding:
// unit = tile, [e0, e1) = slice of its entries (tile matrix: of its batches);
// part / slot >= 0 when the tile is summed in parts (partial `part` of fold
// slot `slot`), slot = -1 otherwise. Fold items: unit = tile, e0 = slot,
// e1 = number of parts.
struct alignas(16) Item {
    int32_t job, unit, c0, c1;
    int32_t e0, e1, part, slot;
};

// Tile-matrix gridding items of one job (frame) with `nbatch` coils:
// for every tile t (batch_off[t + 1] - batch_off[t] batches of its bin) and
// coil record [c0, c0 + rc): one item over all batches when there are at
// most kb, else ceil(nb / kb) part items of kb consecutive batches, each
// leaving a [coil][cell] partial of the tile (2 * coils * cells floats) in
// the scratch region of its fold slot, and one fold item that sums the parts
// in order. Appends to grid / fold / slot_base and advances scratch (floats).
inline void tile_matrix_items(int job, const int32_t *batch_off, int64_t ntiles, int nbatch, int rc, int kb,
                              int64_t cells, std::vector<Item> &grid, std::vector<Item> &fold,
                              std::vector<int64_t> &slot_base, int64_t &scratch) {
    for (int64_t tl = 0; tl < ntiles; ++tl) {
        const int nb = batch_off[tl + 1] - batch_off[tl];
        const int nparts = nb <= kb ? 1 : (nb + kb - 1) / kb;
        for (int c0 = 0; c0 < nbatch; c0 += rc) {
            const int c1 = std::min(nbatch, c0 + rc);
            if (nparts == 1) {
                grid.push_back(Item{job, (int32_t)tl, c0, c1, 0, nb, 0, -1});
                continue;
            }
            const int slot = (int)slot_base.size();
            slot_base.push_back(scratch);
            scratch += (int64_t)nparts * 2 * (c1 - c0) * cells;
            for (int q = 0; q < nparts; ++q)
                grid.push_back(Item{job, (int32_t)tl, c0, c1, q * kb, std::min(nb, (q + 1) * kb), q, slot});
            fold.push_back(Item{job, (int32_t)tl, c0, c1, slot, nparts, 0, 0});
        }
    }
}

}  // namespace tn
