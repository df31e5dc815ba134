// Internal declarations shared by the host C++ and the HIP kernels of libsbm_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

namespace sbm {

// Per-template, per-level record in HBM: width, height, feature count, first feature index.
struct TemplLevelDev {
    int width, height, nfeat, off;
};

// Geometry of one pyramid level's linear memories (layout [ori planes][phase T*T][H*W] + zero guard).
struct LevelGeom {
    int rows, cols, T, W, H;
    long long hw;         // H*W
    const uint8_t* lm;    // device pointer
    int lm_bytes;         // planes*T*T*hw plus the zeroed guard (< 2^31)
    const uint8_t* lm_nib;  // the same positions two per byte (low nibble = even position), or null
    int nib_bytes;
    int planes = 8;       // orientations: 8, or 16 for the 16-orientation variant
    // batched launches: image z's linear memories (and nibble copy) start z * img_stride (nib_stride)
    // bytes after lm (lm_nib); lm_bytes / nib_bytes stay the size of one image's block
    long long img_stride = 0, nib_stride = 0;
    // byte offset of the zeroed guard, where features outside the image read
    __host__ __device__ int zero_off() const { return planes * T * T * (int)hw; }
    // the smaller nonzero response: 3 (neighbouring orientation) with 8 planes, 1 with 16 (a response is 0,
    // this value, or 4)
    __host__ __device__ int low_response() const { return planes == 16 ? 1 : 3; }
};

// Geometry of a level's banded copy (the refinement's layout; band_geom in sbm_reach.h).
struct BandGeom {
    int nb = 0;               // bands per plane-phase: ceil(H / 16)
    int wp = 0;               // columns per band: W + 16
    long long zero_off = 0;   // byte offset of the zeroed block
    long long bytes = 0;      // the whole copy, zeroed block included
};

// Bytes of the nibble-packed copy of a level's linear memories (lm_bytes is a multiple of 16).
inline long long lm_nibble_bytes(long long lm_bytes) { return ((lm_bytes / 2) + 15) & ~15LL; }

// Device-side candidate: global template index (with the image of a batched launch in bits 24..30),
// x, y (level coordinates), raw score.
struct Cand {
    int g, x, y, raw;
};
constexpr int kCandImageShift = 24;
constexpr int kCandTemplMask = (1 << kCandImageShift) - 1;  // banks hold fewer than 2^24 templates

// Images of one batched launch (blockIdx.z of the preprocessing kernels, the image of a coarse item, the
// image bits of a candidate).
constexpr int kMaxBatch = 16;
// Level-0 inputs of a batched launch, per image (device memory): the caller's pixels, row stride, mask.
struct SrcTable {
    const uint8_t* img[kMaxBatch];
    const uint8_t* mask[kMaxBatch];
    int stride[kMaxBatch];
};
// Where image z of a batched preprocessing launch finds its buffers: inputs at in + z * in_b (or, when
// `src` is set, the level-0 pixels, stride and mask of src entry z), outputs at out + z * out_b.
struct BatchStrides {
    const SrcTable* src = nullptr;
    bool src_aligned = false;  // every src image starts on a 4-byte boundary with a row stride divisible by 4
    long long in_b = 0, in_mask_b = 0;
    long long out_b = 0, out_mask_b = 0, nib_b = 0;
    long long down_b = 0, down_mask_b = 0;  // the gradient kernel's next-level image and mask (pyrDown fused)
};

// Longest match list the device finalizes itself (longer lists are finished on the host).
constexpr int kFinalMax = 2048;

// Bytes of zeroed guard placed after each level's linear memories; covers every read the
// coarse (tile overshoot) and refine (16-row window) kernels can make past the last phase row.
// A multiple of 16 bytes; the spread kernel re-zeroes it on every pass.
__host__ __device__ inline long long lm_guard_bytes(int W, int H) { return (((long long)W * H + 17LL * W + 4096) + 15) & ~15LL; }
// Bytes of one image's linear memories at a rows x cols level (rows, cols divisible by T), guard included;
// offsets into them are 32-bit, so a match refuses levels where this reaches 2^31.
__host__ __device__ inline long long lm_level_bytes(int planes, int rows, int cols, int T) {
    return (long long)planes * rows * cols + lm_guard_bytes(cols / T, rows / T);
}

// Kernel launchers (sbm_kernels.hip).  All are asynchronous on `stream`.
namespace kern {
// zero_ptr (optional): zero_n words cleared by the launch (the match counters)
// bins: 8, or 16 for the 16-orientation variant (3 channels only; q then holds 2 bytes per pixel)
// nimg > 1 (or bs.src): one launch over the images of a batch (blockIdx.z = image), see BatchStrides;
// the counters are cleared once
// down (optional): also write the next pyramid level's image there (pyr_down's result; image z at
// + z * bs.down_b), and down_mask (with mask) its mask
hipError_t gradient_quantize(const uint8_t* img, int rows, int cols, int stride, int channels,
                             const uint8_t* mask, float weak_threshold, uint8_t* q, float* mag,
                             float* angle, hipStream_t stream, uint32_t* zero_ptr = nullptr, int zero_n = 0,
                             int bins = 8, int nimg = 1, const BatchStrides& bs = BatchStrides{},
                             uint8_t* down = nullptr, uint8_t* down_mask = nullptr);
hipError_t pyr_down(const uint8_t* src, int rows, int cols, int stride, int channels, uint8_t* dst,
                    const uint8_t* mask, uint8_t* mask_dst, hipStream_t stream, int nimg = 1,
                    const BatchStrides& bs = BatchStrides{});
// nib (optional): also write the two-positions-per-byte copy (see pack_nibbles) when the banded kernel
// applies; *nib_written says whether it did (else pack_nibbles must follow)
hipError_t spread_linearize(const uint8_t* q, int rows, int cols, int T, uint8_t* lm, hipStream_t stream,
                            uint8_t* nib = nullptr, bool* nib_written = nullptr, int nimg = 1,
                            const BatchStrides& bs = BatchStrides{});
// The 16-orientation variant: 16-bit OR-spread, the compiled 1024-entry table's responses, 16 planes of
// linear memories and the zeroed guard after them (line2Dup_16bit_ori.cpp:542-730).
hipError_t spread_linearize16(const uint16_t* q, int rows, int cols, int T, uint8_t* lm, hipStream_t stream,
                              int nimg = 1, const BatchStrides& bs = BatchStrides{});
// nib[i] = lm[2i] | lm[2i+1] << 4 for the first lm_bytes / 32 * 16 bytes of nib (image z: lm + z * bs.in_b,
// nib + z * bs.out_b)
hipError_t pack_nibbles(const uint8_t* lm, long long lm_bytes, uint8_t* nib, hipStream_t stream, int nimg = 1,
                        const BatchStrides& bs = BatchStrides{});
// n warps of one source by the inverse maps inv[6*i..6*i+5] into dst (image i at i*rows*cols*channels);
// binary_mask writes 255 where the warped value is > 0, else 0 (sbm_warp.hip)
hipError_t warp_affine(const uint8_t* src, int rows, int cols, int stride, int channels, const double* inv, int n,
                       uint8_t* dst, bool binary_mask, hipStream_t stream);

struct CoarseArgs {
    LevelGeom geom;
    const TemplLevelDev* tl;   // [g * levels + level]
    int levels, level;
    const uint2* feats;        // {x | y << 16 (int16 each), label}
    const int* order;          // templates in processing order
    int ntempl;                // entries of order
    int ntiles;                // tiles per template (tile = 64 * PPL positions)
    const int* raw_min;        // [g] smallest passing raw score (coarse, strict >)
    int* work_counter;         // persistent work queue head (nullptr: one item per wave)
    Cand* cand;
    int* cand_count;
    int cand_cap;
    int nimg = 1;                     // images of the launch (their LMs at geom.img_stride apart)
    const int* dense_skip = nullptr;  // per image: when set and nonzero, the sparse path has that image's work
    // Work distribution of the persistent grid: a device-scope atomic on one address costs several
    // nanoseconds per update, so few items per worker go by fixed stride (static_order), many in pairs.
    int static_order = 0;             // items worker, worker + workers, ... instead of the atomic queue
    int grab = 1;                     // consecutive items taken per queue update
};
// gather: 0 = aligned dword loads + v_alignbyte, 1 = unaligned buffer loads of the lane's PPL bytes,
// 2 = aligned buffer vector loads + the next lane's dwords by DPP (63 position lanes per wave),
// 3 = as 2 on the nibble-packed linear memories (ppl 32)
int coarse_tiles(long long hw, int ppl, int gather);  // tiles per template for this shape
// split > 1: that many wavefronts (one block) share each work item's features
hipError_t coarse(const CoarseArgs& a, int ppl, int waves_per_block, int persistent_blocks, int gather, int split,
                  hipStream_t stream);

// Coarse stage over the nonzero responses (sbm_sparse.hip).
constexpr int kSparseMaxSeg = 16384;  // largest score map of one work item (two dwords per position, in LDS)
// entries hold flat index << 3, and the end-of-list sentinel must decode past every segment: the list path
// takes coarse levels whose cells plus one plane of positions stay below 2^28
constexpr long long kSparseMaxCells = 1LL << 28;
// room past the list's end for the scatter's reads (a row read from a slice's start rounded down to a
// multiple of four ends fewer than 256 entries past the slice); the list buffer holds this many entries
// beyond its capacity, filled with the sentinel 0xFFFFFFFF
constexpr int kSparseListPad = 1024;
// pruned scatter: the positions a work item may keep after the first part of its features (LDS, two dwords each)
constexpr int kSparseAliveCap = 256;
// LDS of one block of the scatter with a score map of seg positions: the map, a spare per lane, two item slots,
// the kept-position list, its gathered sums and its count
inline long long sparse_block_lds(int seg) { return 8LL * seg + 4 * 66 + 8 * kSparseAliveCap + 16; }
struct SparseCoarseArgs {
    LevelGeom geom;
    const TemplLevelDev* tl;
    int levels, level;
    const uint2* feats;
    const int* order;
    int ntempl;
    int seg, nseg;                 // positions per item, items per template
    const int* raw_min;
    int* work_counter;
    int static_order = 0;          // nonzero: fixed XCD shares of the items, no queue (needs >= 8 blocks)
    int split_pct = 100;           // < 100: pruned scatter after this percent of each template's features
    const int* chunkstart;         // [nchunks + 1] entry offset of every 1024-position chunk of the flat level
    int nchunks;
    const uint32_t* entries;       // flat index << 3 | 4 (response == 4)
    int entries_bytes;             // of one image's list
    unsigned long long* entries_read = nullptr;  // optional: list entries read, summed
    const int* flag;               // per image, nonzero: this path runs for that image
    int nimg = 1;                  // images of the launch: image z's chunk offsets at chunkstart + z * cs_b,
    long long cs_b = 0, ent_b = 0; // its entries at entries + z * ent_b
    Cand* cand;
    int* cand_count;
    int cand_cap;
};
// Lists the coarse level's nonzero responses and decides the path: *flag_out = 1 when mode is 1 and
// at most max_density of the responses are nonzero, or mode is 2, and the list fits `cap` entries.
int sparse_chunks(const LevelGeom& g);  // 1024-position chunks of the flat coarse level
// chunkstart: nchunks + 1 ints; block_sum: ceil(nchunks / 4) ints of scratch.  nimg images: image z's level
// at g.lm + z * g.img_stride, its chunk offsets and block sums at + z * rows_b ints, its entries at
// + z * ent_b entries, its count and flag at nnz_out[z], flag_out[z]
hipError_t sparse_lists(const LevelGeom& g, int* chunkstart, int* block_sum, uint32_t* entries, long long cap,
                        double max_density, int mode, int* nnz_out, int* flag_out, hipStream_t stream, int nimg = 1,
                        long long rows_b = 0, long long ent_b = 0);
// threads: 512 or 1024 per block; seg_max: the score map's positions, 4096, 8192 or 16384 (a.seg <= seg_max)
hipError_t coarse_sparse(const SparseCoarseArgs& a, int blocks, int threads, int seg_max, hipStream_t stream);
// out = the first min(*count_in, cap) candidates of `in`, grouped by image and coarse-level row (y / T, H rows
// per image, nimg images); cursor: scratch of min(nimg * H, 1024) ints
hipError_t sort_cands_by_row(const Cand* in, const int* count_in, int cap, int T, int H, int* cursor, Cand* out,
                             int blocks, hipStream_t stream, int nimg = 1);

struct RefineArgs {
    LevelGeom geom;
    const TemplLevelDev* tl;
    int levels, level;
    const uint2* feats;
    const Cand* cand_in;
    const int* count_in;
    int cap_in;
    const int* raw_min;        // [g * levels + level] smallest kept raw (refine, >=)
    Cand* cand_out;
    int* count_out;
    int cap_out;
    unsigned long long* feat_count = nullptr;  // optional: features of the candidates refined, summed
    // optional: eight zeroed counters, one per XCD, from which the blocks take their candidates (else each
    // block takes a fixed share)
    int* queue = nullptr;
    // 0: the linear memories, 8-byte loads; 1: the level's banded copy (band_copy; image z at band +
    // z * band_stride), 8-byte loads; 2: the linear memories, 16-byte loads, two features per load instruction;
    // 3: the banded copy, 16-byte column loads, four features per load instruction; 4: the nibble-banded copy
    // (nib_band_copy), 16-byte slot loads, four features per load instruction
    int layout = 0;
    const uint8_t* band = nullptr;
    long long band_stride = 0;
    BandGeom bgeom{};
};
// waves (1, 2, 4, 8, 16) per candidate, group (8, 16) features' loads in flight per wave
hipError_t refine(const RefineArgs& a, int blocks, hipStream_t stream, int waves = 8, int group = 16);
// The banded copy of a level's linear memories (lm_bytes of them, guard included) for the refinement: image z
// from lm + z * lm_b to out + z * out_b; refuses copies of 2^31 bytes or more
hipError_t band_copy(const uint8_t* lm, long long lm_bytes, int planes, int T, int W, int H, uint8_t* out,
                     hipStream_t stream, int nimg = 1, long long lm_b = 0, long long out_b = 0);
// The nibble-banded copy (refine layout 4; nib_band_geom in sbm_reach.h), with band_copy's arguments and limits
hipError_t nib_band_copy(const uint8_t* lm, long long lm_bytes, int planes, int T, int W, int H, uint8_t* out,
                         hipStream_t stream, int nimg = 1, long long lm_b = 0, long long out_b = 0);

// Scores, canonical order + unique, and NMS of the final candidates in one block (sbm_finalize.hip).
constexpr int kFinalCountSlot = 60, kKeptCountSlot = 61, kFinalFlagSlot = 62;  // counter slots written
struct FinalArgs {
    const Cand* cand;          // refined candidates
    const int* count_in;
    int cap_in;
    const TemplLevelDev* tl;   // [g * levels + level]; level 0 gives n and the box size
    int levels;
    const int* gclass;         // class rank of each template
    const int* gtid;           // template id within its class
    float score_threshold, nms_threshold;
    int top_k;
    int* out;                  // 5 ints per match: x, y, similarity bits, class, template id
    int* kept;                 // list positions kept by NMS, in order
    uint64_t* mask;            // kFinalMax * kFinalMax / 64 words of scratch
    int* counters;             // match count, kept count, "too many for the device" flag
};
hipError_t finalize(const FinalArgs& a, hipStream_t stream);

hipError_t fill_u32(uint32_t* p, uint32_t v, long long n, hipStream_t stream);
hipError_t copy_probe(const uint4* src, uint4* dst, long long n16, hipStream_t stream);

// Debug / parity helpers
hipError_t similarity_map(const LevelGeom& g, const TemplLevelDev* tl_host_dev, const uint2* feats, int width,
                          int height, int nfeat, int off, uint16_t* out, hipStream_t stream);
hipError_t similarity_local(const LevelGeom& g, const uint2* feats, int nfeat, int off, int cx, int cy,
                            uint16_t* out, hipStream_t stream);
hipError_t bin_table(int lo, int hi, uint8_t* out, hipStream_t stream);  // dx, dy in [-1020, 1020]
// *mismatches (zeroed by the caller) += the pairs 0 <= num <= den <= 1020 whose gradient-kernel quotient differs
// from the IEEE one
hipError_t division_check(int* mismatches, hipStream_t stream);
}  // namespace kern

}  // namespace sbm
