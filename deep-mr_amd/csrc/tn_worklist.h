// Work items of the gridding and interpolation launches, and the host-side
// construction of the sparse-matrix gridding items. Plain C++ (no HIP
// types), so the construction can be checked by a host-only program.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace tn {

constexpr int kMatrixCoils = 16;   // widest record of the matrix gridding (coils per item)
constexpr int kMatrixStage = 1024; // entries of a run of cells staged in LDS (also the longest short row), 3D
constexpr int kMatrixStage2d = 512; // the same in 2D: a smaller stage leaves LDS for more workgroups per CU
                                    // (spiral 8x16: 0.175 against 0.181 ms; in 3D 512 measured slower)
constexpr int kMatrixPart = 2048;  // entries per workgroup of a long row; longer rows are summed
                                   // in parts (kMatrixCoils-wide partial records, folded in order)

// LDS-staged entries per matrix-gridding item for plans of `ndim` dimensions.
inline int matrix_stage(int ndim) { return ndim == 3 ? kMatrixStage : kMatrixStage2d; }

// Work item. Gathering interpolation: unit = sample block. Staged
// interpolation: unit = tile, [e0, e1) = slice of its entries. Gridding:
// unit = tile, [e0, e1) = slice of its entries; part / slot >= 0 when the
// tile is split (partial `part` of fold slot `slot`), slot = -1 otherwise.
// Tile fold items: unit = tile, e0 = slot, e1 = number of parts, elements
// [part, slot) of the partial tile. Matrix gridding items: see matrix_items.
struct alignas(16) Item {
    int32_t job, unit, c0, c1;
    int32_t e0, e1, part, slot;
};

constexpr int kFoldChunks = 32;      // most workgroups per tile fold item
constexpr int kFoldChainParts = 16;  // parts per element chain above which a tile fold is spread wider

// Fold items of one split gridding tile (coils c0 .. c1 - 1, `nparts`
// partial tiles of `elems` elements in fold slot `slot`): one workgroup of
// `block` lanes walks a tile of at most kFoldChainParts parts; a tile of more
// parts (a longer dependent chain per element) spreads its elements over
// ceil(parts / kFoldChainParts) contiguous ranges, at most kFoldChunks and at
// most one per `block` elements (one element per lane at the limit).
inline void tile_fold_items(int job, int tile, int c0, int c1, int slot, int nparts, int64_t elems, int block,
                            std::vector<Item> &fold) {
    const int64_t nsub = std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)kFoldChunks, (elems + block - 1) / block,
                              ((int64_t)nparts + kFoldChainParts - 1) / kFoldChainParts}));
    const int64_t per = (elems + nsub - 1) / nsub;
    for (int64_t k = 0; k < nsub && k * per < elems; ++k)
        fold.push_back(Item{job, tile, c0, c1, slot, nparts, (int32_t)(k * per),
                            (int32_t)std::min(elems, (k + 1) * per)});
}

// Matrix-gridding work of one frame, independent of the coils: a run of
// ncells consecutive cells from `cell` whose entries are [e0, e0 + ne), or,
// with ncells < 0, a long row `cell` summed in -ncells parts.
struct MatrixRun {
    int32_t cell, ncells, e0, ne;
};

// Runs of one frame from its CSR row offsets row[0 .. ncell]: consecutive
// cells with at most cpi cells and at most `stage` entries per run, so that
// a run's entries (contiguous in the CSR order) fit the kernel's LDS stage;
// a row longer than mlong (<= stage) ends the run and becomes a long-row run
// of ceil(len / kMatrixPart) parts. Long rows are listed first.
inline void matrix_runs_of(const int32_t *row, int64_t ncell, int cpi, int mlong, std::vector<MatrixRun> &out,
                           int stage = kMatrixStage) {
    std::vector<MatrixRun> shorts;
    int64_t start = 0;
    int ncells = 0, nent = 0;
    auto close = [&]() {
        if (ncells > 0) shorts.push_back(MatrixRun{(int32_t)start, ncells, row[start], row[start + ncells] - row[start]});
        ncells = 0;
        nent = 0;
    };
    for (int64_t c = 0; c < ncell; ++c) {
        const int len = row[c + 1] - row[c];
        if (len > mlong) {
            close();
            out.push_back(MatrixRun{(int32_t)c, -((len + kMatrixPart - 1) / kMatrixPart), row[c], len});
            continue;
        }
        if (ncells == cpi || nent + len > stage) close();
        if (ncells == 0) start = c;
        ++ncells;
        nent += len;
    }
    close();
    out.insert(out.end(), shorts.begin(), shorts.end());
}

// Gridding items of one job (frame, `nbatch` coils in records of mc coils)
// from its runs. A run of cells: unit = first cell, e1 = cell count | entry
// count << 16, slot = first entry, e0 = coil record, part = 0. A long row:
// one item per part (unit = cell, e1 = part, part = 1, e0 = coil record);
// with several parts, slot = first partial record (kMatrixCoils complex
// values each, 2 * kMatrixCoils floats of scratch) and one fold item (e1 =
// number of parts, part = 0, slot = first record) sums them in part order;
// slot = -1 for a single part. Appends to grid / fold, advances scratch
// (floats).
inline void matrix_items(int job, const std::vector<MatrixRun> &runs, int nbatch, int mc, std::vector<Item> &grid,
                         std::vector<Item> &fold, int64_t &scratch) {
    for (const MatrixRun &r : runs) {
        for (int c0 = 0; c0 < nbatch; c0 += mc) {
            const int c1 = std::min(nbatch, c0 + mc);
            if (r.ncells > 0) {
                grid.push_back(Item{job, r.cell, c0, c1, c0 / mc, r.ncells | (r.ne << 16), 0, r.e0});
                continue;
            }
            const int nparts = -r.ncells;
            if (nparts == 1) {
                grid.push_back(Item{job, r.cell, c0, c1, c0 / mc, 0, 1, -1});
                continue;
            }
            const int32_t rec = (int32_t)(scratch / (2 * kMatrixCoils));
            scratch += (int64_t)nparts * 2 * kMatrixCoils;
            for (int k = 0; k < nparts; ++k) grid.push_back(Item{job, r.cell, c0, c1, c0 / mc, k, 1, rec});
            fold.push_back(Item{job, r.cell, c0, c1, c0 / mc, nparts, 0, rec});
        }
    }
}

}  // namespace tn
