// Host check of the matching kernels' reach: for every workload's geometry, the furthest byte each kernel
// loads from a level's linear memories stays inside the zeroed guard after the planes, the furthest list
// entry the nonzero-response scatter loads stays inside the sentinel padding after the list, the scatter's
// decode sends the sentinel and out-of-segment entries to the spare word, and no 32-bit offset wraps.  The
// kernels compute their load operands with the functions of sbm_reach.h; this program calls the same
// functions.
//
//   geometry_bounds               every geometry; prints one line per level, exit 1 on any violation
//   geometry_bounds --plant guard  the same with a guard that leaves out the coarse tiles' overshoot (the
//                                 W * H term): must fail, which shows the check sees a short guard
//   geometry_bounds --plant pad    the same with 128 sentinel entries instead of kSparseListPad: must fail
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sbm_reach.h"

using namespace sbm;

namespace {

bool g_short_guard = false;
long long g_list_pad = kern::kSparseListPad;
long long g_violations = 0;
long long g_checks = 0;

void fail(const std::string& what) {
    if (g_violations < 20) std::printf("VIOLATION %s\n", what.c_str());
    ++g_violations;
}

void expect(bool ok, const char* what, long long a, long long b, const std::string& where) {
    ++g_checks;
    if (!ok) fail(where + ": " + what + " (" + std::to_string(a) + " vs " + std::to_string(b) + ")");
}

// The zeroed bytes of one image's linear memories as the pipeline sizes them (or the shortened guard).
long long zeroed_lm_bytes(int planes, int rows, int cols, int T) {
    if (!g_short_guard) return lm_level_bytes(planes, rows, cols, T);
    const long long W = cols / T;
    return (long long)planes * rows * cols + ((17 * W + 4096 + 15) & ~15LL);
}

LevelGeom geometry(int planes, int rows, int cols, int T) {
    LevelGeom g{};
    g.rows = rows;
    g.cols = cols;
    g.T = T;
    g.W = cols / T;
    g.H = rows / T;
    g.hw = (long long)g.W * g.H;
    g.planes = planes;
    g.lm = nullptr;
    g.lm_bytes = (int)zeroed_lm_bytes(planes, rows, cols, T);
    g.lm_nib = nullptr;
    g.nib_bytes = (int)lm_nibble_bytes(g.lm_bytes);
    return g;
}

// The furthest in-image feature offset and the guard offset that every pixel outside the level reads.
int max_feature_offset(const LevelGeom& g) {
    return feature_offset(g, g.cols - 1, g.rows - 1, g.planes - 1, g.zero_off());
}

struct Shape {
    int ppl, gather;
};
const Shape kCoarseShapes[] = {{4, 0}, {8, 0}, {16, 0}, {4, 1}, {8, 1}, {16, 1}, {4, 2}, {8, 2}, {16, 2}, {32, 3}};

// Coarse gather: the last tile a wave scores starts below the position limit (hw when every position is
// reported, the template's position count otherwise); every lane loads for every feature.
long long check_coarse(const LevelGeom& g, const std::string& where) {
    long long reach = 0;
    const int offsets[] = {0, max_feature_offset(g), g.zero_off()};
    expect(max_feature_offset(g) < g.zero_off(), "in-image offsets below the guard", max_feature_offset(g),
           g.zero_off(), where);
    for (const Shape& s : kCoarseShapes) {
        const long long tile = coarse_tile_positions(s.ppl, s.gather);
        const long long j0max = ((g.hw - 1) / tile) * tile;
        const long long limit_bytes = s.gather == 3 ? lm_nibble_bytes(g.lm_bytes) : (long long)g.lm_bytes;
        for (int o : offsets)
            for (int lane = 0; lane < 64; ++lane) {
                const long long a = (long long)o + j0max + (long long)lane * s.ppl;
                expect(a < (1LL << 31), "coarse offset fits 32 bits", a, 1LL << 31, where);
                const LaneLoad ld = coarse_lane_load(o, (int)j0max, lane, s.ppl, s.gather);
                const long long end = (long long)ld.voff + ld.soff + ld.bytes;
                expect(ld.voff >= 0 && ld.soff >= 0, "coarse load starts inside", ld.soff, 0,
                       where + " gather " + std::to_string(s.gather) + " ppl " + std::to_string(s.ppl));
                expect(end <= limit_bytes, "coarse load ends inside the zeroed guard", end, limit_bytes,
                       where + " gather " + std::to_string(s.gather) + " ppl " + std::to_string(s.ppl));
                if (s.gather != 3 && end > reach) reach = end;
            }
    }
    return reach;
}

// Refinement at a level below the coarse one: features of any template, at any window, read at their offset
// (at most the guard's) plus the lane's part of the 16 x 16 window.
long long check_refine(const LevelGeom& g, const std::string& where) {
    long long reach = 0;
    const int offsets[] = {0, max_feature_offset(g), g.zero_off()};
    for (int o : offsets)
        for (int lane = 0; lane < 64; ++lane) {
            const int lane_off = refine_lane_offset(g, lane);
            const LaneLoad ld = refine_lane_load(o, lane_off);
            const long long end = (long long)ld.voff + ld.soff + ld.bytes;
            expect((long long)o + lane_off < (1LL << 31), "refine offset fits 32 bits", o + (long long)lane_off,
                   1LL << 31, where);
            expect(end <= g.lm_bytes, "refine load ends inside the zeroed guard", end, g.lm_bytes, where);
            if (end > reach) reach = end;
            const LaneLoad wd = refine_wide_load(o, wide_lane_offset(g, lane));  // refine layout 2
            const long long wend = (long long)wd.voff + wd.soff + wd.bytes;
            expect(wend <= g.lm_bytes, "wide refine load ends inside the zeroed guard", wend, g.lm_bytes, where);
            if (wend > reach) reach = wend;
        }
    // the banded copy (refine_layout 1), when the pipeline makes one (below 2^31 bytes): every lane's 8-byte
    // load of an in-level feature stays inside its 32-byte slot, and the zeroed block covers the lanes of a
    // feature outside the level
    const BandGeom bg = band_geom(g.planes, g.T, g.W, g.H);
    if (bg.bytes < (1LL << 31)) {
        const int band_offsets[] = {0, band_feature_offset(g, bg, g.cols - 1, g.rows - 1, g.planes - 1),
                                    band_feature_offset(g, bg, g.cols - 1, 0, 0),
                                    band_feature_offset(g, bg, 0, g.rows - 1, g.planes - 1)};
        for (int o : band_offsets)
            for (int lane = 0; lane < 64; ++lane) {
                const LaneLoad ld = refine_lane_load(o, band_lane_offset(lane));
                const long long slot = (long long)(o + band_lane_offset(lane)) / kBandSlot;
                expect(ld.voff / kBandSlot == slot && (ld.voff + ld.bytes - 1) / kBandSlot == slot,
                       "banded load inside its slot", ld.voff, slot * kBandSlot, where);
                expect((long long)ld.voff + ld.bytes <= bg.zero_off, "banded load before the zeroed block",
                       ld.voff + ld.bytes, bg.zero_off, where);
            }
        for (int lane = 0; lane < 64; ++lane) {
            const LaneLoad ld = refine_lane_load((int)bg.zero_off, band_lane_offset(lane));
            expect((long long)ld.voff + ld.bytes <= bg.bytes, "banded load of an outside feature in the zeroed block",
                   ld.voff + ld.bytes, bg.bytes, where);
            const LaneLoad cl = band_column_load((int)bg.zero_off, band_column_offset(lane));  // layout 3
            expect((long long)cl.voff + cl.bytes <= bg.bytes, "column load of an outside feature in the zeroed block",
                   cl.voff + cl.bytes, bg.bytes, where);
        }
        // layout 3: a column load of an in-level feature (slot row r % 16 <= 15) stays inside its slot
        for (int o : band_offsets)
            for (int lane = 0; lane < 64; ++lane) {
                const LaneLoad cl = band_column_load(o, band_column_offset(lane));
                const long long slot = (long long)cl.voff / kBandSlot;
                expect((cl.voff + cl.bytes - 1) / kBandSlot == slot, "column load inside its slot", cl.voff,
                       slot * kBandSlot, where);
            }
        // windows reach columns up to W + 14 and slot rows up to 30
        expect(g.W - 1 + kRefineWindow - 1 < bg.wp, "window columns inside the band", g.W + 14, bg.wp, where);
    }
    // the nibble-banded copy (refine_layout 4): a feature's offset is its slot plus its row (0..15); lane l loads
    // the whole slot of window column l % 16, which for an in-level feature is a slot of the same band, before
    // the zeroed block, and for a feature outside the level lies inside the zeroed block
    const BandGeom nb = nib_band_geom(g.planes, g.T, g.W, g.H);
    if (nb.bytes < (1LL << 31)) {
        std::vector<int> nib_offsets = {nib_band_feature_offset(g, nb, 0, 0, 0),
                                        nib_band_feature_offset(g, nb, g.cols - 1, g.rows - 1, g.planes - 1),
                                        nib_band_feature_offset(g, nb, g.cols - 1, 0, 0),
                                        nib_band_feature_offset(g, nb, 0, g.rows - 1, g.planes - 1)};
        if (g.rows - 1 - 16 * g.T >= 0)  // a band above the last
            nib_offsets.push_back(nib_band_feature_offset(g, nb, g.cols - 1, g.rows - 1 - 16 * g.T, g.planes - 1));
        // every first row r0 = 0..15 of a band, at the last column and the last plane
        for (int r0 = 0; r0 < 16 && r0 * g.T < g.rows; ++r0)
            nib_offsets.push_back(nib_band_feature_offset(g, nb, g.cols - 1, r0 * g.T, g.planes - 1));
        for (int o : nib_offsets) {
            expect(o >= 0 && (long long)o < nb.zero_off, "nibble feature offset before the zeroed block", o,
                   nb.zero_off, where);
            const long long slot0 = (long long)(o & ~(kNibSlot - 1)) / kNibSlot;
            const long long band_row = slot0 / nb.wp;  // pp * nb + b
            const int r0 = o & (kNibSlot - 1);
            for (int lane = 0; lane < 64; ++lane) {
                const LaneLoad ld = nib_column_load(o, nib_column_offset(lane));
                // the window's rows start at bit 4 (r0 % 8) of the first dword loaded
                const long long first_bit = 8LL * (ld.voff - (slot0 + (lane & 15)) * kNibSlot);
                expect(ld.bytes == kNibSlot && ld.voff % 4 == 0 && first_bit + nib_column_shift(o) == 4LL * r0,
                       "nibble load starts at the window's first row", first_bit + nib_column_shift(o), 4LL * r0, where);
                expect((long long)(ld.voff + ld.bytes - 1) / kNibSlot / nb.wp == band_row,
                       "nibble slot load ends in the feature's band", (long long)(ld.voff + ld.bytes - 1) / kNibSlot / nb.wp,
                       band_row, where);
                expect((long long)ld.voff + ld.bytes <= nb.zero_off, "nibble slot load before the zeroed block",
                       ld.voff + ld.bytes, nb.zero_off, where);
                // the 64 bits taken (rows r0 .. r0 + 15) lie inside the feature's column slot
                expect(first_bit + nib_column_shift(o) + 64 <= 8LL * kNibSlot, "the window's rows inside the column's slot",
                       first_bit + nib_column_shift(o) + 64, 8LL * kNibSlot, where);
            }
        }
        const int outside = nib_band_feature_offset(g, nb, -1, 0, 0);
        expect(outside == nb.zero_off, "outside feature reads the zeroed block", outside, nb.zero_off, where);
        for (int lane = 0; lane < 64; ++lane) {
            const LaneLoad ld = nib_column_load(outside, nib_column_offset(lane));
            expect((long long)ld.voff + ld.bytes <= nb.bytes, "nibble load of an outside feature in the zeroed block",
                   ld.voff + ld.bytes, nb.bytes, where);
        }
        // the rows of every slot row r0 (0..15) and window row p map onto distinct accumulator halves
        int seen[8][2] = {};
        for (int p = 0; p < kRefineWindow; ++p) ++seen[nib_row_word(p)][nib_row_half(p)];
        for (int w = 0; w < 8; ++w)
            for (int h = 0; h < 2; ++h) expect(seen[w][h] == 1, "window rows to accumulator halves one to one",
                                               seen[w][h], 1, where);
    }
    // windows of candidates anywhere on the coarser level, templates from one pixel to larger than the
    // level: feature offsets stay at or below the guard's
    const int T = g.T;
    const int widths[] = {1, T, 16 * T, g.cols / 2, g.cols - 1, g.cols, g.cols + 40};
    const int cxs[] = {-1, 0, 1, g.cols / 4, g.cols / 2 - 1, g.cols / 2, g.cols / 2 + 16 * T};
    for (int w : widths)
        for (int cx : cxs) {
            const RefineWindow win = refine_window(cx, cx * g.rows / (g.cols > 0 ? g.cols : 1), w, w, g.rows, g.cols, T);
            const int fxs[] = {0, w / 2, w};
            for (int fx : fxs)
                for (int lab : {0, g.planes - 1}) {
                    const int o = feature_offset(g, fx + win.offx, fx + win.offy, lab, g.zero_off());
                    expect(o >= 0 && o <= g.zero_off(), "window feature offset inside the level or the guard", o,
                           g.zero_off(), where);
                }
        }
    return reach;
}

// Nonzero-response list of the coarse level, as the pipeline sizes it: capacity + pad entries per image,
// a list of at most `cap` entries (the scan's condition for the path), then the sentinels.
void check_list(const LevelGeom& g, const std::string& where) {
    if (!sparse_list_fits(g)) {
        std::printf("  %s: list path off (cells + hw >= 2^29)\n", where.c_str());
        return;
    }
    const long long cells = (long long)g.planes * g.T * g.T * g.hw;
    const long long ent_b = (((long long)sparse_entries_bytes(g.rows, g.cols) + 255) & ~255LL) / 4;  // per image
    const long long cap = ent_b - g_list_pad;
    expect(cap >= sparse_list_capacity(g.rows, g.cols), "list capacity", cap, sparse_list_capacity(g.rows, g.cols),
           where);
    const long long nnzs[] = {0, 1, 255, 256, cap - 1, cap};
    for (long long nnz : nnzs) {
        // sentinels written at nnz .. nnz + pad - 1
        expect(nnz + g_list_pad <= ent_b, "sentinels inside the list buffer", nnz + g_list_pad,
               ent_b, where);
        // a slice [ea, eb) with eb <= nnz is read in rows of 256 from ea rounded down to 4; the prefetch of
        // the next slice starts at its start (<= nnz) rounded down
        // the prefetch of a slice starting at nnz (an empty one) reads a row from nnz rounded down to 4; a
        // slice's further rows start below its end (<= nnz)
        const long long starts[] = {nnz & ~3LL, nnz > 0 ? nnz - 1 : 0};
        for (long long e : starts) {
            const long long end = e + kSparseRowEntries;
            expect(end <= nnz + g_list_pad, "row read inside the sentinel padding", end,
                   nnz + g_list_pad, where);
        }
    }
    // the count and fill kernels read 16 bytes at every 16-byte offset below the cells
    const long long last = ((cells - 1) & ~15LL) + 16;
    expect(last <= g.lm_bytes, "list build reads inside the zeroed guard", last, g.lm_bytes, where);
    // chunk offsets: chunk c1 <= nchunks is read, the array holds nchunks + 1 and the block sums after it
    const long long nchunks = (cells + kSparseChunk - 1) / kSparseChunk;
    expect((long long)sparse_rows_bytes(g.rows, g.cols, g.T, g.planes) >= (nchunks + 1 + (nchunks + 3) / 4) * 4,
           "chunk offsets and block sums", (long long)sparse_rows_bytes(g.rows, g.cols, g.T, g.planes),
           (nchunks + 1) * 4, where);
    // decode: entries are flat << 3 | 4 (response == 4) < the sentinel; for any feature offset (<= the guard's)
    // and segment start (< hw) the sentinel and the entries before the segment land on the spare word
    const uint32_t max_entry = (uint32_t)(((cells - 1) << 3) | 4);
    expect(((cells - 1) << 3 | 4) < 0xFFFFFFFFLL, "largest entry below the sentinel", max_entry, 0xFFFFFFFFLL, where);
    for (int seg : {4096, 8192, kern::kSparseMaxSeg}) {
        for (int lane : {0, 63}) {
            const uint32_t spare = 4u * (uint32_t)(2 * seg + lane);
            const long long os[] = {0, max_feature_offset(g), g.zero_off()};
            const long long s0s[] = {0, seg, g.hw - 1};
            for (long long o : os)
                for (long long s0 : s0s) {
                    const long long o8 = 8 * (o + s0);
                    expect(o8 < (1LL << 31), "8 (offset + segment start) below 2^31", o8, 1LL << 31, where);
                    const uint32_t o8u = (uint32_t)o8;
                    expect(sparse_map_byte(0xFFFFFFFFu, o8u, spare) == spare, "sentinel decodes to the spare",
                           sparse_map_byte(0xFFFFFFFFu, o8u, spare), spare, where);
                    if (o + s0 > 0) {  // the entry just before the segment, either response
                        for (uint32_t b : {0u, 4u}) {
                            const uint32_t x = (uint32_t)((o + s0 - 1) << 3) | b;
                            expect(sparse_map_byte(x, o8u, spare) == spare, "entry before the segment to the spare",
                                   sparse_map_byte(x, o8u, spare), spare, where);
                        }
                    }
                    if (o + s0 + seg - 1 < cells) {  // the segment's own first and last positions, both counters
                        const uint32_t x0 = (uint32_t)((o + s0) << 3), x1 = (uint32_t)((o + s0 + seg - 1) << 3) | 4u;
                        expect(sparse_map_byte(x0, o8u, spare) == 0u, "first position", sparse_map_byte(x0, o8u, spare),
                               0, where);
                        expect(sparse_map_byte(x1, o8u, spare) == 8u * (uint32_t)(seg - 1) + 4u, "last position",
                               sparse_map_byte(x1, o8u, spare), 8LL * (seg - 1) + 4, where);
                    }
                }
            // every word the decode can name lies in the map and the spares, below the item slots
            expect(spare / 4 < (uint32_t)(2 * seg + 64), "decoded word inside the map and spares", spare / 4, 2 * seg + 64,
                   where);
        }
    }
}

// Exhaustive check on small levels: every feature offset the level can produce (every pixel and plane, and
// the guard), every tile and every lane of every coarse shape, every refinement lane.
void check_exhaustive(int planes, int rows, int cols, int T) {
    const LevelGeom g = geometry(planes, rows, cols, T);
    const std::string where = "exhaustive " + std::to_string(rows) + "x" + std::to_string(cols) + " T" +
                              std::to_string(T) + " planes " + std::to_string(planes);
    long long max_coarse = 0, max_nib = 0, max_refine = 0;
    for (int lab = 0; lab < planes; ++lab)
        for (int y = -1; y <= rows; ++y)
            for (int x = -1; x <= cols; ++x) {
                const int o = feature_offset(g, x, y, lab, g.zero_off());
                if (o < 0 || o > g.zero_off()) fail(where + ": feature offset outside the level");
                for (const Shape& s : kCoarseShapes) {
                    const int tile = coarse_tile_positions(s.ppl, s.gather);
                    for (long long j0 = 0; j0 < g.hw; j0 += tile)
                        for (int lane = 0; lane < 64; lane += 9) {
                            const LaneLoad ld = coarse_lane_load(o, (int)j0, lane, s.ppl, s.gather);
                            const long long end = (long long)ld.voff + ld.soff + ld.bytes;
                            if (s.gather == 3) max_nib = end > max_nib ? end : max_nib;
                            else max_coarse = end > max_coarse ? end : max_coarse;
                        }
                }
                for (int lane = 0; lane < 64; ++lane) {
                    const LaneLoad ld = refine_lane_load(o, refine_lane_offset(g, lane));
                    const long long end = (long long)ld.voff + ld.soff + ld.bytes;
                    max_refine = end > max_refine ? end : max_refine;
                    const LaneLoad wd = refine_wide_load(o, wide_lane_offset(g, lane));
                    const long long wend = (long long)wd.voff + wd.soff + wd.bytes;
                    max_refine = wend > max_refine ? wend : max_refine;
                }
            }
    expect(max_coarse <= g.lm_bytes, "coarse loads (all offsets)", max_coarse, g.lm_bytes, where);
    expect(max_nib <= lm_nibble_bytes(g.lm_bytes), "nibble loads (all offsets)", max_nib, lm_nibble_bytes(g.lm_bytes),
           where);
    expect(max_refine <= g.lm_bytes, "refine loads (all offsets)", max_refine, g.lm_bytes, where);
    std::printf("  %s: coarse %lld / nibble %lld / refine %lld of %d zeroed bytes\n", where.c_str(), max_coarse,
                max_nib, max_refine, g.lm_bytes);
}

struct Pyramid {
    int planes;
    std::vector<int> T;
};

// Every level of one image size and pyramid, as Detector::match runs it.
void check_image(int rows, int cols, const Pyramid& p, const char* name) {
    int r = rows, c = cols;
    const int L = (int)p.T.size();
    for (int l = 0; l < L; ++l) {
        if (l > 0) {
            r /= 2;
            c /= 2;
        }
        const int T = p.T[l];
        if (r % T || c % T || ((long long)r * c) % 16) {
            std::printf("  %s %dx%d: level %d not divisible, size refused\n", name, rows, cols, l);
            return;
        }
        if (lm_level_bytes(p.planes, r, c, T) >= (1LL << 31)) {
            std::printf("  %s %dx%d: level %d too large, size refused\n", name, rows, cols, l);
            return;
        }
    }
    r = rows;
    c = cols;
    for (int l = 0; l < L; ++l) {
        if (l > 0) {
            r /= 2;
            c /= 2;
        }
        const LevelGeom g = geometry(p.planes, r, c, p.T[l]);
        const std::string where = std::string(name) + " " + std::to_string(rows) + "x" + std::to_string(cols) +
                                  " level " + std::to_string(l) + " T" + std::to_string(p.T[l]) + " planes " +
                                  std::to_string(p.planes);
        long long reach;
        if (l == L - 1) {
            reach = check_coarse(g, where);
            check_list(g, where);
        } else {
            reach = check_refine(g, where);
        }
        std::printf("  %s: %s loads reach byte %lld of %d zeroed (guard from %d)\n", where.c_str(),
                    l == L - 1 ? "coarse" : "refine", reach, g.lm_bytes, g.zero_off());
    }
}

// The largest rows x rows image a pyramid accepts (levels below 2^31 bytes).
int largest_square(const Pyramid& p) {
    int step = 16;
    for (int T : p.T) step = step % T ? step * T : step;
    step <<= (int)p.T.size();
    int best = 0;
    for (int n = step; n < 65536; n += step) {
        int r = n;
        bool ok = true;
        for (size_t l = 0; l < p.T.size(); ++l) {
            if (l > 0) r /= 2;
            if (r % p.T[l] || lm_level_bytes(p.planes, r, r, p.T[l]) >= (1LL << 31)) ok = false;
        }
        if (ok) best = n;
    }
    return best;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc > 2 && std::strcmp(argv[1], "--plant") == 0) {
        g_short_guard = std::strcmp(argv[2], "guard") == 0;
        if (std::strcmp(argv[2], "pad") == 0) g_list_pad = 128;
    }
    const Pyramid pyramids[] = {{8, {4, 8}}, {8, {5, 8}}, {8, {2, 4, 8}}, {8, {8}}, {8, {16}}, {16, {4, 8}}};
    const int sizes[][2] = {{480, 640},   {1024, 1024}, {2048, 2048}, {960, 1280}, {1936, 2592},
                            {400, 640},   {256, 384},   {16, 16},     {64, 80},    {4096, 4096}};
    for (const Pyramid& p : pyramids) {
        std::string name = "T{";
        for (size_t i = 0; i < p.T.size(); ++i) name += (i ? "," : "") + std::to_string(p.T[i]);
        name += "}";
        for (const auto& s : sizes) check_image(s[0], s[1], p, name.c_str());
        const int big = largest_square(p);
        if (big > 0) check_image(big, big, p, (name + " largest").c_str());
    }
    check_exhaustive(8, 32, 48, 4);
    check_exhaustive(8, 40, 40, 5);
    check_exhaustive(16, 32, 32, 8);
    check_exhaustive(8, 16, 16, 16);
    std::printf("%lld checks, %lld violations\n", g_checks, g_violations);
    return g_violations ? 1 : 0;
}
