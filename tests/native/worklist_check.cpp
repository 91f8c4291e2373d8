// Host-only check of the tile-matrix gridding work list (tn_worklist.h):
// every (tile, coil record) gets items covering its batches exactly once, in
// part order, with fold slots whose scratch regions hold every part's
// partial tile and do not overlap. Prints "ok <cases>" or the first failure.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "tn_worklist.h"

static int fail(const char *what, int cs, long t) {
    std::printf("FAIL case %d tile %ld: %s\n", cs, t, what);
    return 1;
}

int main() {
    std::mt19937 rng(7);
    int cases = 0;
    for (int cs = 0; cs < 400; ++cs) {
        const int64_t ntiles = 1 + rng() % 300;
        const int nbatch = 1 + rng() % 40;
        const int rc = 1 << (rng() % 5);
        const int kb = cs % 7 == 0 ? 1 << 30 : 1 + rng() % 6;
        const int64_t cells = 16 + rng() % 1009;
        std::vector<int32_t> boff(ntiles + 1, 0);
        for (int64_t t = 0; t < ntiles; ++t) {
            const int r = rng() % 10;
            const int nb = r < 3 ? 0 : r < 8 ? 1 + rng() % 3 : rng() % 60;
            boff[t + 1] = boff[t] + nb;
        }
        std::vector<tn::Item> grid, fold;
        std::vector<int64_t> slot_base;
        int64_t scratch = 0;
        const int job = cs % 5;
        tn::tile_matrix_items(job, boff.data(), ntiles, nbatch, rc, kb, cells, grid, fold, slot_base, scratch);
        // per (tile, chunk): the batches each item covers
        const int nchunk = (nbatch + rc - 1) / rc;
        std::vector<std::vector<const tn::Item *>> by(ntiles * nchunk);
        for (const tn::Item &it : grid) {
            if (it.job != job || it.unit < 0 || it.unit >= ntiles) return fail("item job / tile", cs, it.unit);
            if (it.c0 < 0 || it.c0 % rc || it.c1 <= it.c0 || it.c1 > nbatch || it.c1 - it.c0 > rc)
                return fail("coil range", cs, it.unit);
            const int nb = boff[it.unit + 1] - boff[it.unit];
            if (it.e0 < 0 || it.e1 < it.e0 || it.e1 > nb || it.e1 - it.e0 > kb) return fail("batch range", cs, it.unit);
            by[(int64_t)it.unit * nchunk + it.c0 / rc].push_back(&it);
        }
        std::vector<int> parts_of_slot(slot_base.size(), 0);
        for (int64_t t = 0; t < ntiles; ++t) {
            const int nb = boff[t + 1] - boff[t];
            for (int c = 0; c < nchunk; ++c) {
                const auto &v = by[t * nchunk + c];
                if (v.empty()) return fail("tile without an item", cs, t);
                if (v.size() == 1 && v[0]->slot < 0) {
                    if (v[0]->e0 != 0 || v[0]->e1 != nb || v[0]->part != 0) return fail("whole item", cs, t);
                    continue;
                }
                const int slot = v[0]->slot;
                if (slot < 0 || slot >= (int)slot_base.size()) return fail("fold slot", cs, t);
                int next = 0;
                for (size_t q = 0; q < v.size(); ++q) {
                    if (v[q]->slot != slot || v[q]->part != (int)q) return fail("part order", cs, t);
                    if (v[q]->e0 != next) return fail("batches not contiguous", cs, t);
                    next = v[q]->e1;
                }
                if (next != nb) return fail("batches not covered", cs, t);
                if ((int)v.size() < 2) return fail("split into one part", cs, t);
                parts_of_slot[slot] = (int)v.size();
            }
        }
        // fold items: one per slot, with its part count; scratch regions
        if (fold.size() != slot_base.size()) return fail("fold items != slots", cs, -1);
        int64_t end = 0;
        for (const tn::Item &f : fold) {
            if (f.e0 < 0 || f.e0 >= (int)slot_base.size() || f.e1 != parts_of_slot[f.e0])
                return fail("fold item", cs, f.unit);
            const int64_t need = (int64_t)f.e1 * 2 * (f.c1 - f.c0) * cells;
            if (slot_base[f.e0] != end) return fail("scratch regions not packed", cs, f.unit);
            end += need;
        }
        if (end != scratch) return fail("scratch size", cs, -1);
        ++cases;
    }
    std::printf("ok %d\n", cases);
    return 0;
}
