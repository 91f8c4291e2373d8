// Host-only check of the matrix-gridding work list (tn_worklist.h): for
// random CSR row offsets, the runs cover every cell exactly once (long rows
// first), short runs stay within the kernel's cell and LDS-stage limits and
// carry the entry range their cells' rows span, and the items cover every
// (run, coil record) with long rows split into in-order parts whose partial
// records are disjoint and inside the scratch size; and that the fold items
// of split gridding tiles cover each partial tile's elements once in
// contiguous ranges. Prints "ok <cases>" or the first failure.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "tn_worklist.h"

static int fail(const char *what, int cs, long c) {
    std::printf("FAIL case %d cell %ld: %s\n", cs, c, what);
    return 1;
}

int main() {
    std::mt19937 rng(7);
    const int cpis[] = {16, 32, 64, 128, 256};
    int cases = 0;
    for (int cs = 0; cs < 400; ++cs) {
        const int64_t ncell = cs % 50 == 0 ? 0 : 1 + rng() % 5000;
        const int cpi = cpis[rng() % 5];
        const int stage = cs % 2 ? tn::kMatrixStage : tn::kMatrixStage2d;
        const int mlong = 16 + rng() % (stage - 15);
        const int nbatch = 1 + rng() % 40;
        const int mc = 1 + rng() % tn::kMatrixCoils;
        // row lengths: mostly empty or short, a few long (a dense centre);
        // every seventh case has no entries at all (a plan without samples)
        const bool empty = cs % 7 == 3;
        std::vector<int32_t> row(ncell + 1, 0);
        for (int64_t c = 0; c < ncell; ++c) {
            const int r = rng() % 100;
            const int len = empty ? 0 : r < 30 ? 0 : r < 95 ? rng() % 40 : r < 98 ? rng() % (2 * mlong) : rng() % 9000;
            row[c + 1] = row[c] + len;
        }
        std::vector<tn::MatrixRun> runs;
        tn::matrix_runs_of(row.data(), ncell, cpi, mlong, runs, stage);
        std::vector<int> seen(ncell, 0);
        bool in_shorts = false;
        for (const tn::MatrixRun &r : runs) {
            if (r.ncells == 0) return fail("empty run", cs, r.cell);
            if (r.ncells < 0) {
                if (in_shorts) return fail("long row after a short run", cs, r.cell);
                if (r.cell < 0 || r.cell >= ncell) return fail("long row cell", cs, r.cell);
                const int len = row[r.cell + 1] - row[r.cell];
                if (len <= mlong) return fail("long row not above the threshold", cs, r.cell);
                if (r.e0 != row[r.cell] || r.ne != len) return fail("long row entries", cs, r.cell);
                if (-r.ncells != (len + tn::kMatrixPart - 1) / tn::kMatrixPart) return fail("part count", cs, r.cell);
                ++seen[r.cell];
                continue;
            }
            in_shorts = true;
            if (r.cell < 0 || r.cell + r.ncells > ncell) return fail("run past the grid", cs, r.cell);
            if (r.ncells > cpi) return fail("run longer than cells per item", cs, r.cell);
            if (r.ne > stage || r.ne < 0) return fail("run entries exceed the LDS stage", cs, r.cell);
            if (r.e0 != row[r.cell] || r.ne != row[r.cell + r.ncells] - row[r.cell]) return fail("run entries", cs, r.cell);
            for (int i = 0; i < r.ncells; ++i) {
                if (row[r.cell + i + 1] - row[r.cell + i] > mlong) return fail("long row inside a run", cs, r.cell + i);
                ++seen[r.cell + i];
            }
        }
        for (int64_t c = 0; c < ncell; ++c)
            if (seen[c] != 1) return fail("cell not covered exactly once", cs, c);

        std::vector<tn::Item> grid, fold;
        int64_t scratch = 0;
        const int job = cs % 5;
        tn::matrix_items(job, runs, nbatch, mc, grid, fold, scratch);
        // per (run, chunk): the items, in order
        const int nchunk = (nbatch + mc - 1) / mc;
        size_t gi = 0, fi = 0;
        std::vector<char> used(scratch / (2 * tn::kMatrixCoils), 0);
        if (scratch % (2 * tn::kMatrixCoils)) return fail("scratch not whole records", cs, -1);
        for (const tn::MatrixRun &r : runs) {
            for (int q = 0; q < nchunk; ++q) {
                const int c0 = q * mc, c1 = std::min(nbatch, c0 + mc);
                auto same = [&](const tn::Item &it) {
                    return it.job == job && it.unit == r.cell && it.c0 == c0 && it.c1 == c1 && it.e0 == q;
                };
                if (r.ncells > 0) {
                    if (gi >= grid.size() || !same(grid[gi])) return fail("run item", cs, r.cell);
                    const tn::Item &it = grid[gi++];
                    if (it.part != 0 || (it.e1 & 0xffff) != r.ncells || (it.e1 >> 16) != r.ne || it.slot != r.e0)
                        return fail("run item fields", cs, r.cell);
                    continue;
                }
                const int nparts = -r.ncells;
                int rec = -1;
                for (int k = 0; k < nparts; ++k) {
                    if (gi >= grid.size() || !same(grid[gi])) return fail("part item", cs, r.cell);
                    const tn::Item &it = grid[gi++];
                    if (it.part != 1 || it.e1 != k) return fail("part order", cs, r.cell);
                    // the kernel sums entries [row + k * part, min(row end, + part))
                    if (row[r.cell] + (int64_t)k * tn::kMatrixPart >= row[r.cell + 1]) return fail("empty part", cs, r.cell);
                    if (k == 0) rec = it.slot;
                    if (it.slot != rec) return fail("parts of one row in different records", cs, r.cell);
                }
                if (nparts == 1) {
                    if (rec != -1) return fail("single part with a record", cs, r.cell);
                    continue;
                }
                if (rec < 0 || (int64_t)(rec + nparts) * 2 * tn::kMatrixCoils > scratch)
                    return fail("partial records past the scratch", cs, r.cell);
                for (int k = 0; k < nparts; ++k) {
                    if (used[rec + k]) return fail("partial records overlap", cs, r.cell);
                    used[rec + k] = 1;
                }
                if (fi >= fold.size() || !same(fold[fi])) return fail("fold item", cs, r.cell);
                const tn::Item &f = fold[fi++];
                if (f.e1 != nparts || f.slot != rec || f.part != 0) return fail("fold item fields", cs, r.cell);
            }
        }
        if (gi != grid.size()) return fail("extra grid items", cs, -1);
        if (empty && (!fold.empty() || scratch != 0)) return fail("an empty matrix needs no parts", cs, -1);
        if (fi != fold.size()) return fail("extra fold items", cs, -1);
        // tile folds: the element ranges of a split tile are contiguous,
        // cover [0, elems) once, carry the tile's fields, number at most
        // kFoldChunks (one for a tile of at most kFoldChainParts parts), and
        // hold at most one element per lane when the chain is long enough
        {
            const int block = 256;
            const int64_t elems = 1 + rng() % (16 * 4096);
            const int nparts = 2 + (int)(rng() % (cs % 3 ? 40 : 3000));
            std::vector<tn::Item> tf;
            tn::tile_fold_items(3, 17, 2, 9, 5, nparts, elems, block, tf);
            if (tf.empty() || (int)tf.size() > tn::kFoldChunks) return fail("tile fold item count", cs, elems);
            if (nparts <= tn::kFoldChainParts && tf.size() != 1) return fail("light tile split", cs, elems);
            int64_t next = 0;
            for (const tn::Item &f : tf) {
                if (f.job != 3 || f.unit != 17 || f.c0 != 2 || f.c1 != 9 || f.e0 != 5 || f.e1 != nparts)
                    return fail("tile fold item fields", cs, elems);
                if (f.part != next || f.slot <= f.part) return fail("tile fold ranges", cs, elems);
                const int64_t lanes_needed = (elems + block - 1) / block;
                if (lanes_needed <= tn::kFoldChunks &&
                    (nparts + tn::kFoldChainParts - 1) / tn::kFoldChainParts >= lanes_needed &&
                    f.slot - f.part > block)
                    return fail("tile fold range wider than a workgroup", cs, elems);
                next = f.slot;
            }
            if (next != elems) return fail("tile fold ranges do not cover the tile", cs, elems);
        }
        ++cases;
    }
    std::printf("ok %d\n", cases);
    return 0;
}
