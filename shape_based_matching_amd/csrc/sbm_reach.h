// Address arithmetic of the matching kernels: where each load from a level's linear memories and from the
// nonzero-response list starts, and how many bytes it takes.  The kernels compute their load operands with
// these functions; tests/cpp/geometry_bounds.cpp runs the same functions on the host over every workload's
// geometry and shows that no load leaves its buffer (the zeroed guard after the planes, the sentinel
// padding after the list) and that no 32-bit offset wraps.
#pragma once
#include "sbm_internal.h"

namespace sbm {

// ---------------------------------------------------------------------------------------------------------
// Linear memories: [planes][T*T phases][H*W], then the zeroed guard (lm_guard_bytes) at zero_off.
// ---------------------------------------------------------------------------------------------------------

// Byte offset of the response to orientation `lab` at level pixel (fx, fy): plane lab, phase
// (fy % T, fx % T), grid cell (fy / T, fx / T) (line2Dup.cpp:897-957).  Pixels outside the level read the
// guard.
__host__ __device__ inline int feature_offset(const LevelGeom& g, int fx, int fy, int lab, int zero_off) {
    if (fx < 0 || fx >= g.cols || fy < 0 || fy >= g.rows) return zero_off;
    const int T = g.T, hw = (int)g.hw;
    const int lxq = fx / T, lyq = fy / T;
    return lab * (T * T * hw) + ((fy - lyq * T) * T + (fx - lxq * T)) * hw + lyq * g.W + lxq;
}

// Grid positions j = r * W + c at which a width x height template is scored (line2Dup.cpp:786-798); the
// sums at later positions are computed and dropped.
__host__ __device__ inline int template_positions(const LevelGeom& g, int width, int height) {
    const int wf = (width - 1) / g.T + 1, hf = (height - 1) / g.T + 1;
    return (g.H - hf) * g.W + (g.W - wf) + 1;
}

// One lane's load for one feature: buffer offset voff (vector, per lane) + soff (scalar, per wave) and its
// width in bytes.
struct LaneLoad {
    int voff, soff, bytes;
};

// ---------------------------------------------------------------------------------------------------------
// Coarse gather (k_coarse): a wave scores one tile of positions j0 .. j0 + tile - 1 of one template.
// ---------------------------------------------------------------------------------------------------------
// Lanes that own positions: all 64, or 63 when the last lane only feeds its neighbour (gather modes 2, 3).
__host__ __device__ constexpr int coarse_position_lanes(int gather) { return gather >= 2 ? 63 : 64; }
__host__ __device__ constexpr int coarse_tile_positions(int ppl, int gather) {
    return coarse_position_lanes(gather) * ppl;
}
// The load of lane `lane` for a feature at offset o, tile start j0 (ppl positions per lane):
//   gather 0  ppl/4 + 1 dwords from the dword boundary at or below o + j0 + lane * ppl (plain pointer loads)
//   gather 1  ppl bytes at o + j0 + lane * ppl (buffer load)
//   gather 2  ppl bytes at lane * ppl past the ppl-aligned boundary at or below o + j0 (buffer load)
//   gather 3  16 bytes of the nibble-packed copy at lane * 16 past byte (o + j0) / 2 rounded down to 16
//             (buffer load; positions o + j0 rounded down to 32)
__host__ __device__ inline LaneLoad coarse_lane_load(int o, int j0, int lane, int ppl, int gather) {
    const int a = o + j0;
    if (gather == 0) return LaneLoad{0, (a + lane * ppl) & ~3, ppl + 4};
    if (gather == 1) return LaneLoad{lane * ppl, a, ppl};
    if (gather == 2) return LaneLoad{lane * ppl, a & ~(ppl - 1), ppl};
    return LaneLoad{lane * 16, (a & ~31) >> 1, 16};
}

// ---------------------------------------------------------------------------------------------------------
// Refinement (k_refine): a 16 x 16 window of grid positions around each coarse candidate, at the level
// below.
// ---------------------------------------------------------------------------------------------------------
constexpr int kRefineWindow = 16;

struct RefineWindow {
    int x, y;        // the candidate's centre in level pixels, clamped
    int offx, offy;  // pixel offset of the window's first grid position
};
// Candidate (cx, cy) of the coarser level: its centre one level down, kept 8T from the borders for a
// width x height template (line2Dup.cpp:1207-1216, 1226-1229).
__host__ __device__ inline RefineWindow refine_window(int cx, int cy, int width, int height, int rows, int cols,
                                                      int T) {
    const int border = 8 * T;
    const int max_x = cols - width - border, max_y = rows - height - border;
    int x = cx * 2 + 1, y = cy * 2 + 1;
    x = x < border ? border : x;
    y = y < border ? border : y;
    x = x > max_x ? max_x : x;
    y = y > max_y ? max_y : y;
    return RefineWindow{x, y, (x / T - 8) * T, (y / T - 8) * T};
}
// Lane l holds window row l / 4, positions 4 (l % 4) .. 4 (l % 4) + 3: its bytes sit at this offset past the
// feature's; they are loaded as 8 bytes from the dword boundary at or below.
__host__ __device__ inline int refine_lane_offset(const LevelGeom& g, int lane) { return (lane >> 2) * g.W + (lane & 3) * 4; }
__host__ __device__ inline LaneLoad refine_lane_load(int feature_off, int lane_off) {
    const int ad = feature_off + lane_off;
    return LaneLoad{ad & ~3, 0, 8};
}

// Wide loads (refine layout 2): lanes 0..31 take one feature, lanes 32..63 the next; lane l holds window row
// (l % 32) / 2, positions 8 (l % 2) .. + 7, loaded as 16 bytes from the dword boundary at or below.
__host__ __device__ inline int wide_lane_offset(const LevelGeom& g, int lane) {
    return ((lane & 31) >> 1) * g.W + (lane & 1) * 8;
}
__host__ __device__ inline LaneLoad refine_wide_load(int feature_off, int lane_off) {
    const int ad = feature_off + lane_off;
    return LaneLoad{ad & ~3, 0, 16};
}
// The window position (row-major index) of position p (0..7) of lane l < 32 after the two halves are added.
__host__ __device__ inline int wide_window_index(int lane, int p) { return ((lane & 31) >> 1) * 16 + (lane & 1) * 8 + p; }

// ---------------------------------------------------------------------------------------------------------
// Banded copy of a refined level (k_band, read by k_refine when the tuning asks for it): every 16 x 16
// window of the linear memories is one contiguous run of bytes.  For each plane-phase pp and band b of 16
// grid rows, column c (0 .. W + 15) holds a 32-byte slot: the linear bytes at pp * H * W + (16 b + s) * W + c
// for s = 0..31 -- the rows 16 b .. 16 b + 31 of that column, with the linear layout's reads past a row's end
// (columns W.., the next row) and past the plane's last row (the next plane, then the zeroed guard) kept as
// they are.  A window at (row r, column c) is column c .. c + 15 of band r / 16, slots r % 16 .. r % 16 + 15:
// 512 bytes in a row instead of sixteen rows W bytes apart.  After the slots, a zeroed block that features
// outside the level read.
// ---------------------------------------------------------------------------------------------------------
constexpr int kBandRows = 16, kBandSlot = 32, kBandZeroBytes = 1024;

__host__ __device__ inline BandGeom band_geom(int planes, int T, int W, int H) {
    BandGeom b;
    b.nb = (H + kBandRows - 1) / kBandRows;
    b.wp = W + kBandRows;
    b.zero_off = (long long)planes * T * T * b.nb * b.wp * kBandSlot;
    b.bytes = b.zero_off + kBandZeroBytes;
    return b;
}
// Byte offset in the banded copy of the response to orientation `lab` at level pixel (fx, fy); pixels outside
// the level read the zeroed block.
__host__ __device__ inline int band_feature_offset(const LevelGeom& g, const BandGeom& b, int fx, int fy, int lab) {
    if (fx < 0 || fx >= g.cols || fy < 0 || fy >= g.rows) return (int)b.zero_off;
    const int T = g.T;
    const int lxq = fx / T, lyq = fy / T;
    const int pp = lab * T * T + (fy - lyq * T) * T + (fx - lxq * T);
    return ((pp * b.nb + (lyq >> 4)) * b.wp + lxq) * kBandSlot + (lyq & 15);
}
// Lane l holds window column l / 4, rows 4 (l % 4) .. 4 (l % 4) + 3: four bytes of one slot, at this offset
// past the feature's (a multiple of 4; the feature's own offset sets the misalignment, the same for every
// lane), loaded as 8 bytes from the dword boundary at or below -- inside the slot, whose 32 bytes start on
// a 32-byte boundary.
__host__ __device__ inline int band_lane_offset(int lane) { return (lane >> 2) * kBandSlot + (lane & 3) * 4; }
// The window position (row, column) of byte k (0..3) of lane l: banded (row-major index row * 16 + column)
__host__ __device__ inline int band_window_index(int lane, int k) { return ((lane & 3) * 4 + k) * 16 + (lane >> 2); }
// Column loads (refine layout 3): lanes 16 g .. 16 g + 15 take feature g of every four; lane l holds window
// column l % 16, its sixteen rows, as one 16-byte load at the feature's offset (slot row r % 16 of the first
// column) plus 32 (l % 16): bytes (r % 16) .. (r % 16) + 15 of that column's slot, unaligned, inside the slot.
__host__ __device__ inline int band_column_offset(int lane) { return (lane & 15) * kBandSlot; }
__host__ __device__ inline LaneLoad band_column_load(int feature_off, int lane_off) {
    return LaneLoad{feature_off + lane_off, 0, 16};
}
// The window position (row-major index) of row p (0..15) of lane l < 16 after the four groups are added.
__host__ __device__ inline int band_column_window_index(int lane, int p) { return p * 16 + (lane & 15); }
// The linear byte (in the level's [planes][T*T][H*W] block) that slot byte s of column c, band b, plane-phase
// pp copies.
__host__ __device__ inline long long band_source(int W, int H, int pp, int b, int c, int s) {
    return (long long)pp * H * W + (long long)(kBandRows * b + s) * W + c;
}

// ---------------------------------------------------------------------------------------------------------
// Nibble-banded copy of a refined level (k_band_nib, refine layout 4): the banded copy with two rows per byte.
// Responses are 0..4, so a 16-byte slot holds the 32 rows of one column of one band: row s (0..31) in byte
// s / 2, the low nibble for even s; the rows are band_source's, as in the banded copy.  A window at (row r,
// column c) is then sixteen consecutive slots, 256 bytes in a row, and row r % 16 of each slot its first row.
// After the slots, the same zeroed block.  About 1.06x the level's linear bytes (W + 16 columns of W).
// ---------------------------------------------------------------------------------------------------------
constexpr int kNibSlot = 16;

__host__ __device__ inline BandGeom nib_band_geom(int planes, int T, int W, int H) {
    BandGeom b;
    b.nb = (H + kBandRows - 1) / kBandRows;
    b.wp = W + kBandRows;
    b.zero_off = (long long)planes * T * T * b.nb * b.wp * kNibSlot;
    b.bytes = b.zero_off + kBandZeroBytes;
    return b;
}
// The slot of the response to orientation `lab` at level pixel (fx, fy) (a multiple of 16) plus its row in
// the slot (0..15) in the low four bits; pixels outside the level read row 0 of the zeroed block.
__host__ __device__ inline int nib_band_feature_offset(const LevelGeom& g, const BandGeom& b, int fx, int fy, int lab) {
    if (fx < 0 || fx >= g.cols || fy < 0 || fy >= g.rows) return (int)b.zero_off;
    const int T = g.T;
    const int lxq = fx / T, lyq = fy / T;
    const int pp = lab * T * T + (fy - lyq * T) * T + (fx - lxq * T);
    return ((pp * b.nb + (lyq >> 4)) * b.wp + lxq) * kNibSlot + (lyq & 15);
}
// Lanes 16 g .. 16 g + 15 take feature g of every four; lane l loads window column l % 16: 16 bytes at the
// feature's slot plus 16 (l % 16), dword-aligned -- 4 bytes further when the window's first row r0 is 8 or more
// (bit 3 of the offset), so that its rows start in the first dword loaded, at bit 4 (r0 % 8), either way.  The
// load then ends at most 4 bytes into the next column's slot of the same band (window columns reach W + 14, a
// band has W + 16), or inside the zeroed block for a feature outside the level (r0 = 0).
__host__ __device__ inline int nib_column_offset(int lane) { return (lane & 15) * kNibSlot; }
__host__ __device__ inline LaneLoad nib_column_load(int feature_off, int lane_off) {
    return LaneLoad{(feature_off & ~(kNibSlot - 1)) + ((feature_off >> 1) & 4) + lane_off, 0, 16};
}
// The bit in the first loaded dword where the window's first row starts, as v_alignbit takes it (bits 4:0):
// 4 (r0 % 8).
__host__ __device__ inline uint32_t nib_column_shift(int feature_off) { return ((uint32_t)feature_off << 2) & 31u; }
// Where window row p (0..15) of lane l < 16 sits after the four feature groups are added: 16-bit half `half`
// of accumulator word `word` (the kernel's nibble sums are widened into even-row and odd-row bytes, those into
// 16-bit pairs of bytes 0, 2 and 1, 3); its window position is row-major p * 16 + l % 16.
__host__ __device__ inline int nib_row_word(int p) { return 2 * ((p >> 3) * 2 + (p & 1)) + (((p & 7) >> 1) & 1); }
__host__ __device__ inline int nib_row_half(int p) { return ((p & 7) >> 1) >> 1; }

// ---------------------------------------------------------------------------------------------------------
// Nonzero-response list (sbm_sparse.hip): entry = flat index << 3 | 4 (response == 4), in flat order, then
// kSparseListPad sentinel entries 0xFFFFFFFF.
// ---------------------------------------------------------------------------------------------------------
constexpr int kSparseChunk = 1024;       // flat positions per list chunk
constexpr int kSparseRowEntries = 256;   // entries per scatter row load: four per lane, 16 bytes
// Whether a coarse level may take the list path: a feature's range starts below cells + hw, and the
// scatter's decode (sparse_map_byte) needs 8 (cells + hw) below 2^31.
__host__ __device__ inline bool sparse_list_fits(const LevelGeom& g) {
    return (long long)g.planes * g.T * g.T * g.hw + g.hw < kern::kSparseMaxCells;
}
// Entries the list of a rows x cols coarse level may hold (one response in eight: denser levels take the
// gather path); its buffer holds kSparseListPad more.
inline long long sparse_list_capacity(int rows, int cols) {
    const long long px = (long long)rows * cols;
    return px > 1024 ? px : 1024;
}
inline size_t sparse_entries_bytes(int rows, int cols) {
    return (size_t)(sparse_list_capacity(rows, cols) + kern::kSparseListPad) * sizeof(uint32_t);
}
// Chunk offsets (one per 1024 flat positions, plus the end) and the count kernel's per-block sums.
inline size_t sparse_rows_bytes(int rows, int cols, int T, int planes) {
    const size_t chunks = ((size_t)planes * T * T * (rows / T) * (cols / T) + kSparseChunk - 1) / kSparseChunk;
    return (chunks + 1 + (chunks + 3) / 4) * sizeof(int);
}
// The score-map byte that entry x (flat << 3 | 4 (response == 4)) adds 1 to for a feature at offset o and a
// segment starting at s0 (o8 = 8 (o + s0)): 8 j for the entry's position j = flat - o - s0 when the response is
// the low one (3, or 1 with 16 orientations), 8 j + 4 when it is 4 -- position j's two counters; entries before
// the segment wrap and, like the sentinel, land on the lane's spare dword.
__host__ __device__ inline uint32_t sparse_map_byte(uint32_t x, uint32_t o8, uint32_t spare) {
    const uint32_t d = x - o8;
    return d < spare ? d : spare;
}

}  // namespace sbm
