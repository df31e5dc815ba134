// The coarse similarity stage of Detector::match over the nonzero responses only
// (line2Dup.cpp:780-831, 897-957, 1149-1188: the same sums, organised the other way round).
//
// The coarse kernel in sbm_kernels.hip gathers, for every template and every position j, the bytes
// LM[o_f + j] of its features f (o_f: the feature's offset in the flat [ori][T*T][H*W] block) and adds
// them.  On edge maps most responses are 0.  Here the nonzero responses of the coarse level are listed
// once per image, in flat order, and a block scatters, for each feature, the entries q of its flat range
// [o_f, o_f + P) into a score map in LDS at j = q - o_f.  The sums are the same integers, so the
// candidates are the same.  A map dword counts, for its position, the addends in its high half and the
// sum of (addend - 1) in its low half; the raw score is the sum of the halves.  With fewer than 16384
// features per template neither half can carry (3 * 16383 < 2^16).
//
//   k_sparse_count   nonzero bytes of every 1024-position chunk of the flat coarse level, summed per block
//   k_sparse_scan    one block: exclusive block offsets, the list length, and the path flag
//   k_sparse_fill    chunk offsets and the list: entry = flat index << 3 | 4 (response == 4), chunk by chunk,
//                    grouped by LDS bank class inside a chunk
//   k_coarse_sparse  persistent blocks over (template, position segment) items from an atomic queue
//
// Which path runs is decided on the device, so a captured graph holds both: the dense kernel returns
// at once when the flag says sparse, and this one when it says dense.
#include "sbm_internal.h"
#include "sbm_reach.h"

namespace sbm {
namespace {

constexpr int kWave = 64;
typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));
typedef unsigned int v2u32 __attribute__((ext_vector_type(2)));
constexpr int kBufferWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle

__device__ __forceinline__ int nonzero_bytes(uint32_t x) {
    const uint32_t t = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
    return __popc(t & 0x80808080u);
}

constexpr int kChunk = kSparseChunk;  // positions (bytes) of the flat coarse level per list chunk: one wave, 16 each
static_assert(kChunk == 16 * kWave, "a chunk is one 16-byte load per lane");
constexpr int kListThreads = 256;   // four chunks per block of the count and fill kernels
constexpr int kChunksPerBlock = kListThreads / kWave;
constexpr int kClasses = 64;        // LDS bank classes: position & 63 of a one-dword-per-position map

__device__ __forceinline__ int uint4_nonzero(const uint4 v) {
    return nonzero_bytes(v.x) + nonzero_bytes(v.y) + nonzero_bytes(v.z) + nonzero_bytes(v.w);
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// one wave per chunk of 1024 positions: nonzero count; one sum per block (four chunks)
// Batched launches: image blockIdx.z's level at lm + z * lm_b, its block sums at block_sum + z * rows_b.
__global__ __launch_bounds__(kListThreads) void k_sparse_count(const uint8_t* __restrict__ lm, long long cells,
                                                               int* __restrict__ block_sum, long long lm_b,
                                                               long long rows_b) {
    __shared__ int s_chunk[kChunksPerBlock];
    lm += blockIdx.z * lm_b;
    block_sum += blockIdx.z * rows_b;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const long long c = (long long)blockIdx.x * kChunksPerBlock + w;
    const long long off = c * kChunk + 16LL * lane;
    // the zeroed guard follows the planes, so 16 bytes reaching past them read zeros
    const int n = off < cells ? uint4_nonzero(*reinterpret_cast<const uint4*>(lm + off)) : 0;
    const int tot = wave_sum(n);
    if (lane == 0) s_chunk[w] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
#pragma unroll
        for (int k = 0; k < kChunksPerBlock; ++k) s += s_chunk[k];
        block_sum[blockIdx.x] = s;
    }
}

// one block of 1024 threads: block bases (exclusive, in place), the list length and the path flag
__global__ __launch_bounds__(1024) void k_sparse_scan(int* __restrict__ block_sum, int nblocks, long long cap,
                                                      long long dense_limit, int* __restrict__ nnz_out,
                                                      int* __restrict__ flag_out, long long rows_b) {
    // the list path takes levels of fewer than 2^28 cells, so every count and offset fits an int
    __shared__ int s_wave[1024 / kWave];
    block_sum += blockIdx.z * rows_b;
    nnz_out += blockIdx.z;
    flag_out += blockIdx.z;
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    const int per = (nblocks + 1023) / 1024;
    const int b = min(nblocks, t * per), e = min(nblocks, b + per);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += block_sum[i];
    // exclusive offsets of the threads' partial sums: inclusive scans inside each wavefront, then of the
    // sixteen wave totals
    int incl = sum;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int o = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += o;
    }
    if (lane == kWave - 1) s_wave[w] = incl;
    __syncthreads();
    int base = 0, total = 0;
    for (int k = 0; k < 1024 / kWave; ++k) {
        const int v = s_wave[k];
        base += k < w ? v : 0;
        total += v;
    }
    int run = base + incl - sum;
    for (int i = b; i < e; ++i) {
        const int c = block_sum[i];
        block_sum[i] = run;
        run += c;
    }
    if (t == 0) {
        *nnz_out = total;
        *flag_out = ((long long)total <= cap && (long long)total <= dense_limit) ? 1 : 0;
    }
}

// one wave per chunk: its entry offset (chunkstart) and its entries.  Inside a chunk the entries are
// grouped by q & 63, which spreads a chunk's consecutive entries -- the lanes of one scatter
// instruction -- over the banks of the score map.
__global__ __launch_bounds__(kListThreads) void k_sparse_fill(const uint8_t* __restrict__ lm, long long cells, int nchunks,
                                                              const int* __restrict__ block_base,
                                                              const int* __restrict__ flag, const int* __restrict__ nnz,
                                                              int* __restrict__ chunkstart, uint32_t* __restrict__ entries,
                                                              long long lm_b, long long rows_b, long long ent_b) {
    __shared__ int s_chunk[kChunksPerBlock];
    __shared__ int s_cls[kChunksPerBlock][kClasses];
    {  // image blockIdx.z of a batched launch
        const int z = blockIdx.z;
        lm += z * lm_b;
        block_base += z * rows_b;
        chunkstart += z * rows_b;
        entries += z * ent_b;
        flag += z;
        nnz += z;
    }
    if (*flag == 0) return;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const long long c = (long long)blockIdx.x * kChunksPerBlock + w;
    const long long off = c * kChunk + 16LL * lane;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (off < cells) v = *reinterpret_cast<const uint4*>(lm + off);
    const int n = uint4_nonzero(v);
    const int tot = wave_sum(n);
    s_cls[w][lane] = 0;
    if (lane == 0) s_chunk[w] = tot;
    __syncthreads();
    int start = block_base[blockIdx.x];
    for (int k = 0; k < w; ++k) start += s_chunk[k];
    if (c < nchunks && lane == 0) chunkstart[c] = start;
    if (c == 0 && lane == 0) chunkstart[nchunks] = *nnz;
    // ranks inside each class
    const uint32_t words[4] = {v.x, v.y, v.z, v.w};
    int rank[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t r = (words[i >> 2] >> (8 * (i & 3))) & 0xFFu;
        rank[i] = r ? atomicAdd(&s_cls[w][(16 * lane + i) & (kClasses - 1)], 1) : -1;
    }
    __syncthreads();
    // class bases: exclusive scan of the 64 class counts, one per lane
    const int cnt = s_cls[w][lane];
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int o = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += o;
    }
    __syncthreads();
    s_cls[w][lane] = start + incl - cnt;
    __syncthreads();
    const uint32_t base = (uint32_t)off;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        if (rank[i] < 0) continue;
        const uint32_t r = (words[i >> 2] >> (8 * (i & 3))) & 0xFFu;
        entries[s_cls[w][(16 * lane + i) & (kClasses - 1)] + rank[i]] = ((base + i) << 3) | (r == 4u ? 4u : 0u);
    }
    // the sentinel after the list: the scatter reads up to kern::kSparseListPad entries past a slice's end
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < kern::kSparseListPad; i += kListThreads) entries[*nnz + i] = 0xFFFFFFFFu;
}

constexpr int kRowEntries = kSparseRowEntries;  // list entries per row: one 16-byte load per lane
static_assert(kRowEntries == 4 * kWave, "a row is one 16-byte load per lane");

// Row of 256 entries from entry e (a multiple of four): lane l gets entries e + 4 l .. e + 4 l + 3.
// Reads stay below the list's length plus kern::kSparseListPad, inside its buffer.  The lane's part of
// the offset is a fixed vector offset and 4 e the scalar offset, so no vector register but the loaded
// ones is written (a write to a register with a load in flight would wait for every load).
__device__ __forceinline__ uint4 load_row(__amdgpu_buffer_rsrc_t r, int e, int lane) {
    const v4u32 v = __builtin_amdgcn_raw_buffer_load_b128(r, 16 * lane, 4 * e, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
}

// Half row of 128 entries from entry e (a multiple of four): lane l gets entries e + 2 l, e + 2 l + 1.
__device__ __forceinline__ uint2 load_half_row(__amdgpu_buffer_rsrc_t r, int e, int lane) {
    const v2u32 v = __builtin_amdgcn_raw_buffer_load_b64(r, 8 * lane, 4 * e, 0);
    return make_uint2(v[0], v[1]);
}

__device__ __forceinline__ void add_entry(uint32_t x, uint32_t o8, uint32_t spare, uint8_t* map_bytes) {
    atomicAdd(reinterpret_cast<uint32_t*>(map_bytes + sparse_map_byte(x, o8, spare)), 1u);
}

// The adds of one row: entry x adds 1 at byte min(x - o8, spare).
__device__ __forceinline__ void add_row(const uint4 x, uint32_t o8, uint32_t spare, uint8_t* map_bytes) {
    add_entry(x.x, o8, spare, map_bytes);
    add_entry(x.y, o8, spare, map_bytes);
    add_entry(x.z, o8, spare, map_bytes);
    add_entry(x.w, o8, spare, map_bytes);
}

// Waits for a row's load without using the values.
__device__ __forceinline__ void retire(const uint4 x) {
    __asm__ volatile("" : : "v"(x.x), "v"(x.y), "v"(x.z), "v"(x.w));
}

// Feature u of a wave's batch (cnt features; per-lane of, ea, eb: offset and entry slice of each): xc holds
// the first row of its slice, from the slice's start rounded down to a multiple of four; xn receives the
// next feature's first row (a row at entry 0 when there is none).  The feature's own further loads are
// issued first and the prefetch after them, so that the adds wait only for loads older than xn (the
// wait counter retires loads in order).  Rows of 256 entries, four per lane; a last part of at most 128
// entries takes a half row, or lanes 0..31 of the first row.  The entries read outside the slice need no
// mask.
__device__ __forceinline__ void feature_step(int u, int cnt, const uint4& xc, uint4& xn, int of, int ea, int eb,
                                             int s0, __amdgpu_buffer_rsrc_t rsrc, int lane, uint32_t spare,
                                             uint8_t* map_bytes) {
    const int next = u + 1 < cnt ? __builtin_amdgcn_readlane(ea, u + 1) & ~3 : 0;
    // of + s0 < 2^28 (the list path's cell bound), so o8 does not wrap
    const uint32_t o8 = 8u * (uint32_t)(__builtin_amdgcn_readlane(of, u) + s0);
    const int e0 = __builtin_amdgcn_readlane(ea, u), e1 = __builtin_amdgcn_readlane(eb, u);
    const int e = e0 & ~3;
    const int left = e1 - e;  // wave-uniform
    if (e0 >= e1) {
        xn = load_row(rsrc, next, lane);
        retire(xc);
    } else if (left > kRowEntries + kRowEntries / 2) {
        const uint4 x1 = load_row(rsrc, e + kRowEntries, lane);
        xn = load_row(rsrc, next, lane);
        add_row(xc, o8, spare, map_bytes);
        add_row(x1, o8, spare, map_bytes);
        for (int f = e + 2 * kRowEntries; f < e1; f += kRowEntries) {  // long slices: one row at a time
            if (e1 - f > kRowEntries / 2) {
                add_row(load_row(rsrc, f, lane), o8, spare, map_bytes);
            } else {
                const uint2 h = load_half_row(rsrc, f, lane);
                add_entry(h.x, o8, spare, map_bytes);
                add_entry(h.y, o8, spare, map_bytes);
            }
        }
    } else if (left > kRowEntries) {
        const uint2 h = load_half_row(rsrc, e + kRowEntries, lane);
        xn = load_row(rsrc, next, lane);
        add_row(xc, o8, spare, map_bytes);
        add_entry(h.x, o8, spare, map_bytes);
        add_entry(h.y, o8, spare, map_bytes);
    } else if (left > kRowEntries / 2) {
        xn = load_row(rsrc, next, lane);
        add_row(xc, o8, spare, map_bytes);
    } else {
        xn = load_row(rsrc, next, lane);
        if (lane < kWave / 2) add_row(xc, o8, spare, map_bytes);  // entries e .. e + 127
        else retire(xc);
    }
}

// Work item = (template k of the order, segment s of SEG positions).  The score map of the segment
// lives in LDS, two counters per position (the low responses and the 4s), followed by one spare dword per lane.  Wave w of the block's
// sixteen takes features w, w + 16, ...: 64 at a time, one per lane, it finds their flat ranges' entry
// bounds; then, feature by feature, the lanes stride over the entries in rows of 256 (four entries per
// lane, one 16-byte load) and add 1 to position j's counter for its response, at byte
// d = x - 8 (o_f + s0) = 8 j + 4 (response == 4).  Entries read outside the slice need no mask: the up to
// three before it (rounding its start down to a multiple of four) and its boundary chunks' entries
// before the segment lie at j < 0, wrap, and go to the lane's spare dword; the entries after it lie at
// j >= s1 - s0 and go to the spare or to the word of a position past the segment's end, which the
// threshold scan ignores and clears (or never reads); past the list, the sentinel 0xFFFFFFFF decodes to
// j >= 2^28 (o_f + s0 < 2^28).  Two VALU instructions per entry, no branch: sub and min, then one LDS add of
// a constant (with one dword per position holding count << 16 | sum of (response - 1) it took four: the
// eight-image 36k-bank launch 11.75 -> 11.41 ms).  Not bound by the VALU work or the LDS bank conflicts (an XOR swizzle of the map halved them
// and was slower).  On large banks (long slices) the list-load instructions bind it, hence four entries per
// load.  On small scenes (slices shorter than a row) the vector L1's misses to L2 bind it: a row load makes
// about five L2 requests (the list stays in L2), and more wavefronts per CU only queue behind the L1's
// misses (at eight blocks of 256 threads per CU the L2 read latency quadruples and the kernel slows by half);
// loading rows further ahead (as a list of row jobs with seven in flight) changed nothing.
// THREADS (256, 512 or 1024) threads per block and a score map of SEG positions (4096, 8192 or 16384: 16, 32 or 64 KB
// of LDS) are the autotuner's to choose, with the blocks per CU.
// Positions a work item may keep after the first part of its features (the pruned scatter): their segment
// index and partial raw score, one dword each, and their gathered sums, in LDS beside the score map.
constexpr int kAliveCap = kern::kSparseAliveCap;

// Wavefronts per SIMD that the block size and the score map's LDS allow (a CU: 2048 threads, 160 KB of LDS, four
// SIMDs), at most eight; the register budget is set to fit them (the kernel's scalar state would otherwise hold
// a SIMD to seven, and a second 16-wave block would never start beside the first).
template <int THREADS, int SEG>
constexpr int sparse_waves_per_simd() {
    constexpr int by_lds = 160 * 1024 / (8 * SEG + 4 * 66 + 8 * kAliveCap + 16);
    constexpr int by_threads = 2048 / THREADS;
    constexpr int blocks = by_lds < by_threads ? by_lds : by_threads;
    constexpr int w = blocks * (THREADS / 64) / 4;
    return w < 8 ? (w < 1 ? 1 : w) : 8;
}

// A passing position of a lane (segment index i, grid position j, raw score) appended to the candidate list: one
// atomic per wave.
__device__ __forceinline__ void emit_candidate(const kern::SparseCoarseArgs& a, bool pass, int j, int raw, int gidx,
                                               int img, int W, int T, int offs, int lane) {
    const uint64_t ballot = __ballot(pass);
    if (!ballot) return;
    const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ballot >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)ballot, 0u));
    int slot0 = 0;
    if (lane == 0) slot0 = atomicAdd(a.cand_count, (int)__popcll(ballot));
    slot0 = __shfl(slot0, 0, kWave);
    const int slot = slot0 + below;
    if (pass && slot < a.cand_cap) {
        const int r = j / W, c = j - r * W;
        Cand cd;
        cd.g = gidx | (img << kCandImageShift);
        cd.x = c * T + offs;
        cd.y = r * T + offs;
        cd.raw = raw;
        a.cand[slot] = cd;
    }
}

// Work item = (template k of the order, segment s of SEG positions).  The score map of the segment
// lives in LDS, two counters per position (the low responses and the 4s), followed by one spare dword per lane.  Wave w of the block's
// waves takes features w, w + waves, ...: 64 at a time, one per lane, it finds their flat ranges' entry
// bounds; then, feature by feature, the lanes stride over the entries in rows of 256 (four entries per
// lane, one 16-byte load) and add 1 to position j's counter for its response, at byte
// d = x - 8 (o_f + s0) = 8 j + 4 (response == 4).  Entries read outside the slice need no mask: the up to
// three before it (rounding its start down to a multiple of four) and its boundary chunks' entries
// before the segment lie at j < 0, wrap, and go to the lane's spare dword; the entries after it lie at
// j >= s1 - s0 and go to the spare or to the word of a position past the segment's end, which the
// threshold scan ignores and clears (or never reads); past the list, the sentinel 0xFFFFFFFF decodes to
// j >= 2^28 (o_f + s0 < 2^28).  Two VALU instructions per entry, no branch: sub and min, then one LDS add of
// a constant (with one dword per position holding count << 16 | sum of (response - 1) it took four: the
// eight-image 36k-bank launch 11.75 -> 11.41 ms).  Not bound by the VALU work or the LDS bank conflicts (an XOR swizzle of the map halved them
// and was slower).  On large banks (long slices) the list-load instructions bind it, hence four entries per
// load.  On small scenes (slices shorter than a row) the vector L1's misses to L2 bind it: a row load makes
// about five L2 requests (the list stays in L2), and more wavefronts per CU only queue behind the L1's
// misses (at eight blocks of 256 threads per CU the L2 read latency quadruples and the kernel slows by half);
// loading rows further ahead (as a list of row jobs with seven in flight) changed nothing.
//
// Pruned scatter (a.split_pct < 100): only the first split_pct percent of the features are scattered; then a
// position whose partial raw score plus 4 for each remaining feature stays below rmin cannot pass and is
// dropped (split_pct is used only where that drops every position with no partial score: 4 * remaining <
// rmin).  The kept positions (a few dozen of a 16384-position segment at threshold 80) go to a list in LDS; the
// block's waves split the remaining features among them, and for each kept position a lane adds its wave's
// features' responses from the level's linear memories directly, LM[o_f + j] -- one byte load per feature, the
// feature offset in the scalar operand, the position in the vector one -- the same addends the scatter would
// have added, so the candidates are identical.  A segment keeping more than kAliveCap positions scatters its
// remaining features instead.  Each such byte load reads a cache line of its own, so the gather pays only while
// the kept positions number well below a feature's slice length / 32: on the 36,000-template bank the kernel
// gains about 4% at 60 percent and loses at 40 (a few hundred kept positions), hence the autotuner's choice.
// THREADS (256, 512 or 1024) threads per block and a score map of SEG positions (4096, 8192 or 16384: 16, 32 or
// 64 KB of LDS) are the autotuner's to choose, with the blocks per CU and the split.
template <int THREADS, int SEG>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(sparse_waves_per_simd<THREADS, SEG>())))
void k_coarse_sparse(kern::SparseCoarseArgs a) {
    constexpr int kSparseThreads = THREADS;
    constexpr int kSparseWaves = THREADS / kWave;
    // score map (two counters per position: words 2 i and 2 i + 1 count the low and the 4 responses at position
    // i), one spare dword per lane, then two item-index slots used in turn
    __shared__ __align__(16) uint32_t s_map[2 * SEG + kWave + 2];
    __shared__ uint32_t s_alive[kAliveCap];  // pruned scatter: (partial raw << 16) | segment index
    __shared__ int s_gsum[kAliveCap];        // ... the gathered responses of each kept position
    __shared__ int s_nalive;
    const int seg = a.seg;
    int* const s_items = reinterpret_cast<int*>(s_map + 2 * SEG + kWave);
    {  // nothing to do when the gather kernel has every image
        bool none = true;
        for (int z = 0; z < a.nimg; ++z) none &= a.flag[z] == 0;
        if (none) return;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wave-uniform: the feature loop stays scalar
    const LevelGeom& g = a.geom;
    const int hw = (int)g.hw, W = g.W;
    const int T = g.T;
    const int offs = T / 2 + (T % 2 - 1);  // line2Dup.cpp:1181
    const int zero_off = g.zero_off();
    const int low = g.low_response();  // raw score of position i: low * map[2 i] + 4 * map[2 i + 1]
    const long long per_image = (long long)a.ntempl * a.nseg;
    const long long total = per_image * a.nimg;
    uint8_t* const map_bytes = reinterpret_cast<uint8_t*>(s_map);
    const uint32_t spare = 4u * (uint32_t)(2 * SEG + lane);  // byte offset of this lane's spare
    for (int i = threadIdx.x; i < 2 * seg; i += kSparseThreads) s_map[i] = 0u;
    // With static_order (gridDim.x >= 8) no queue: blocks go to the XCDs in turn by index, XCD x takes the
    // x-th eighth of the items (one image's items when a launch holds eight) and its blocks take them in
    // turn.  Otherwise blocks take items from the work counter.
    const int xcd = blockIdx.x & 7, nbx = (int)(gridDim.x >> 3) + (xcd < (int)(gridDim.x & 7) ? 1 : 0);
    const long long sbeg = xcd * (total >> 3) + min((long long)xcd, total & 7);
    const long long send = sbeg + (total >> 3) + (xcd < (total & 7) ? 1 : 0);
    if (threadIdx.x == 0 && !a.static_order) s_items[0] = atomicAdd(a.work_counter, 1);
    unsigned long long read = 0;  // list entries this lane's features read, over all items
    for (int it = 0;; ++it) {
        __syncthreads();  // the map is clear, and this item's index is in its slot
        long long item;
        if (a.static_order) {
            item = sbeg + (blockIdx.x >> 3) + (long long)it * nbx;
            if (item >= send) break;
        } else {
            item = __builtin_amdgcn_readfirstlane(s_items[it & 1]);  // the same for the whole block
            if (item >= total) break;
            // the next item's index, taken while this one runs (every thread read the other slot before the
            // barrier above)
            if (threadIdx.x == 0) s_items[(it + 1) & 1] = atomicAdd(a.work_counter, 1);
        }
        if (threadIdx.x == 0) s_nalive = 0;  // read only after the barrier that follows the first scatter
        const int img = (int)(item / per_image);
        const long long rem = item - (long long)img * per_image;
        const int k = (int)(rem / a.nseg), sgi = (int)(rem - (long long)k * a.nseg);
        const int gidx = a.order[k];
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(a.entries + img * a.ent_b), (short)0, a.entries_bytes, kBufferWord3);
        const int* const chunkstart = a.chunkstart + img * a.cs_b;
        const TemplLevelDev t = a.tl[(long long)gidx * a.levels + a.level];
        const int nfeat = t.nfeat;
        const int rmin = a.raw_min[gidx];
        const int P = template_positions(g, t.width, t.height);
        const int limit = rmin == 0 ? hw : min(P, hw);
        const int s0 = sgi * seg;
        const int s1 = min(s0 + seg, limit);
        const int len = s1 - s0;
        const bool live = rmin <= 4 * nfeat && s0 < s1 && a.flag[img] != 0;  // block-uniform
        const uint2* F = a.feats + t.off;
        // features [0, ksplit) are scattered first; the rest are gathered at the kept positions
        int ksplit = nfeat;
        if (a.split_pct < 100) {
            const int ks = (nfeat * a.split_pct + 99) / 100;
            if (4 * (nfeat - ks) < rmin) ksplit = ks;
        }
        // scatter of features [f0, f1) of this item into the map
        auto scatter = [&](int f0, int f1) {
            for (int fb = f0 + w; fb < f1; fb += kSparseWaves * kWave) {
                // this wave's next 64 features, one per lane: offset and entry range
                const int f = fb + kSparseWaves * lane;
                int of = zero_off, ea = 0, eb = 0;
                if (f < f1) {
                    const uint2 fr = F[f];
                    of = feature_offset(g, (int)(int16_t)(fr.x & 0xFFFFu), (int)(int16_t)(fr.x >> 16), (int)fr.y,
                                        zero_off);
                    const long long lo = (long long)of + s0, hi = (long long)of + s1 - 1;  // flat range
                    const int c0 = (int)min(lo / kChunk, (long long)a.nchunks);
                    const int c1 = (int)min(hi / kChunk + 1, (long long)a.nchunks);
                    ea = chunkstart[c0];
                    eb = chunkstart[c1];
                }
                const int cnt = min(kWave, (f1 - fb + kSparseWaves - 1) / kSparseWaves);
                read += (unsigned long long)(eb - ea);
                // Feature by feature, each one's first row loaded during the previous feature's adds (two
                // register sets in turn, so no copy waits on a load in flight).
                uint4 xa = load_row(rsrc, cnt > 0 ? __builtin_amdgcn_readlane(ea, 0) & ~3 : 0, lane), xb;
                for (int u = 0; u < cnt; u += 2) {
                    feature_step(u, cnt, xa, xb, of, ea, eb, s0, rsrc, lane, spare, map_bytes);
                    if (u + 1 < cnt) feature_step(u + 1, cnt, xb, xa, of, ea, eb, s0, rsrc, lane, spare, map_bytes);
                }
                retire(xa);  // the last prefetch (of no feature) retires here
                retire(xb);
            }
        };
        if (live) scatter(0, ksplit);
        __syncthreads();
        bool done = false;  // block-uniform: the pruned path finished this item
        if (live && ksplit < nfeat) {
            // the positions that can still reach rmin: partial + 4 (nfeat - ksplit) >= rmin, i.e. partial >= keep
            // (> 0); two positions (four counters) per lane, one 16-byte LDS read; one LDS atomic per wave that
            // keeps any
            const int keep = rmin - 4 * (nfeat - ksplit);
            for (int base = w * 2 * kWave; base < seg; base += 2 * kSparseThreads) {
                const int i0 = base + 2 * lane;  // even: its counters start on a 16-byte boundary
                uint4 m4 = make_uint4(0u, 0u, 0u, 0u);
                if (i0 < seg) m4 = *reinterpret_cast<const uint4*>(s_map + 2 * i0);
                const int parts[2] = {low * (int)m4.x + 4 * (int)m4.y, low * (int)m4.z + 4 * (int)m4.w};
                if (!__ballot(parts[0] >= keep || parts[1] >= keep)) continue;
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int i = i0 + q, j = s0 + i;
                    const int part = parts[q];
                    const bool kept = i < len && j < P && part >= keep;
                    const uint64_t ballot = __ballot(kept);
                    if (!ballot) continue;
                    const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ballot >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)ballot, 0u));
                    int slot0 = 0;
                    if (lane == 0) slot0 = atomicAdd(&s_nalive, (int)__popcll(ballot));
                    slot0 = __shfl(slot0, 0, kWave);
                    if (kept && slot0 + below < kAliveCap) {
                        s_alive[slot0 + below] = (uint32_t)part << 16 | (uint32_t)i;
                        s_gsum[slot0 + below] = 0;
                    }
                }
            }
            __syncthreads();
            const int nalive = __builtin_amdgcn_readfirstlane(s_nalive);
            if (nalive <= kAliveCap) {
                done = true;
                // The remaining features at the kept positions: wave w takes features ksplit + w, + waves, ...
                // and a lane per kept position adds LM[o_f + j] (the feature offset, the scatter's o_f, in the
                // scalar operand, j in the vector one; eight loads in flight); one LDS add per lane and wave
                // meets the waves' sums.
                const __amdgpu_buffer_rsrc_t lrs = __builtin_amdgcn_make_buffer_rsrc(
                    (void*)(g.lm + img * g.img_stride), (short)0, g.lm_bytes, kBufferWord3);
                const int mine = (nfeat - ksplit - w + kSparseWaves - 1) / kSparseWaves;  // this wave's features
                for (int e0 = 0; e0 < nalive && mine > 0; e0 += kWave) {
                    const int e = e0 + lane;
                    const int j = s0 + (e < nalive ? (int)(s_alive[e] & 0xFFFFu) : 0);
                    int acc = 0;
                    for (int mb = 0; mb < mine; mb += kWave) {  // 64 of the wave's features, one per lane
                        const int f = ksplit + w + kSparseWaves * (mb + lane);
                        int of = zero_off;
                        if (mb + lane < mine) {
                            const uint2 fr = F[f];
                            of = feature_offset(g, (int)(int16_t)(fr.x & 0xFFFFu), (int)(int16_t)(fr.x >> 16),
                                                (int)fr.y, zero_off);
                        }
                        const int cnt = min(kWave, mine - mb);
                        for (int u = 0; u < cnt; u += 8) {
                            uint32_t v[8];
#pragma unroll
                            for (int q = 0; q < 8; ++q) {
                                // past the wave's features: the zeroed guard (j < hw: inside it)
                                const int o = u + q < cnt ? __builtin_amdgcn_readlane(of, u + q) : zero_off;
                                v[q] = __builtin_amdgcn_raw_buffer_load_b8(lrs, j, o, 0);
                            }
#pragma unroll
                            for (int q = 0; q < 8; ++q) acc += (int)v[q];
                        }
                    }
                    if (e < nalive) atomicAdd(&s_gsum[e], acc);
                }
                // clear the map for the next item (the words up to i0 + 3 < 2 seg + 3 lie inside its allocation)
                for (int i0 = 4 * threadIdx.x; i0 < 2 * seg; i0 += 4 * kSparseThreads)
                    *reinterpret_cast<uint4*>(s_map + i0) = make_uint4(0u, 0u, 0u, 0u);
                __syncthreads();
                for (int e0 = w * kWave; e0 < nalive; e0 += kSparseThreads) {
                    const int e = e0 + lane;
                    const uint32_t ent = e < nalive ? s_alive[e] : 0u;
                    const int raw = (int)(ent >> 16) + (e < nalive ? s_gsum[e] : 0);
                    emit_candidate(a, e < nalive && raw >= rmin, s0 + (int)(ent & 0xFFFFu), raw, gidx, img, W, T,
                                   offs, lane);
                }
            } else {
                scatter(ksplit, nfeat);  // too many kept positions: the rest of the features as well
                __syncthreads();
            }
        }
        if (live && !done) {
            // threshold scan of the segment (and clear it for the next item): two positions (four counters) per
            // lane, one 16-byte LDS read and write; a wave with no candidate costs a few VALU instructions per
            // position, otherwise each position is checked and appended, one atomic per wave
            for (int base = w * 2 * kWave; base < seg; base += 2 * kSparseThreads) {
                const int i0 = base + 2 * lane;  // even: its counters start on a 16-byte boundary
                uint4 m4 = make_uint4(0u, 0u, 0u, 0u);
                if (i0 < seg) {  // words up to 2 i0 + 3 < 2 seg + 2 lie inside the map's allocation
                    m4 = *reinterpret_cast<const uint4*>(s_map + 2 * i0);
                    *reinterpret_cast<uint4*>(s_map + 2 * i0) = make_uint4(0u, 0u, 0u, 0u);
                }
                const int raws[2] = {low * (int)m4.x + 4 * (int)m4.y, low * (int)m4.z + 4 * (int)m4.w};
                if (!__ballot(raws[0] >= rmin || raws[1] >= rmin)) continue;
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int i = i0 + q, j = s0 + i;
                    const int raw = j < P ? raws[q] : 0;
                    emit_candidate(a, i < len && raw >= rmin, j, raw, gidx, img, W, T, offs, lane);
                }
            }
        }
    }
    if (a.entries_read) {  // one atomic per wave at the end: a shared counter hit per item would serialize
        for (int d = 32; d >= 1; d >>= 1) read += __shfl_xor(read, d, kWave);
        if (lane == 0 && read) atomicAdd(a.entries_read, read);
    }
}

// Coarse candidates reordered by their coarse-level row, so that the refinement's XCD shares (each XCD
// takes a contiguous part of the list) hold spatially close candidates, whichever coarse path ran.
// One block builds the histogram of rows (LDS atomics) and the bin offsets; then every candidate takes a
// slot in its bin.  The order inside a bin is arbitrary; the final list is sorted canonically anyway.
constexpr int kMaxBins = 1024;

// Row of a candidate over the images of a launch: image * H + its coarse-level row.
__device__ __forceinline__ int cand_row(const Cand& c, int T, int H) {
    return (c.g >> kCandImageShift) * H + min(max(c.y / T, 0), H - 1);
}

__global__ __launch_bounds__(1024) void k_cand_bins(const Cand* __restrict__ cand, const int* __restrict__ count_in,
                                                    int cap, int T, int H, int rows_all, int nbins,
                                                    int* __restrict__ cursor) {
    __shared__ int s_hist[kMaxBins];
    for (int i = threadIdx.x; i < nbins; i += 1024) s_hist[i] = 0;
    __syncthreads();
    const int n = min(*count_in, cap);
    for (int i = threadIdx.x; i < n; i += 1024) {
        const int ly = cand_row(cand[i], T, H);
        atomicAdd(&s_hist[(int)((long long)ly * nbins / rows_all)], 1);
    }
    __syncthreads();
    // exclusive scan of the nbins <= 1024 counts, one per thread: inclusive scans inside each wavefront, then
    // of the sixteen wave totals
    __shared__ int s_wave[1024 / kWave];
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    const int c = t < nbins ? s_hist[t] : 0;
    int incl = c;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int o = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += o;
    }
    if (lane == kWave - 1) s_wave[w] = incl;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; ++k) base += s_wave[k];
    if (t < nbins) cursor[t] = base + incl - c;
}

// 1024 candidates per block step: slots inside the block from LDS atomics, one global reservation per
// (block step, bin)
__global__ __launch_bounds__(256) void k_cand_scatter(const Cand* __restrict__ cand, const int* __restrict__ count_in,
                                                      int cap, int T, int H, int rows_all, int nbins,
                                                      int* __restrict__ cursor, Cand* __restrict__ out) {
    __shared__ int s_cnt[kMaxBins];
    const int n = min(*count_in, cap);
    for (int c0 = blockIdx.x * 1024; c0 < n; c0 += gridDim.x * 1024) {
        for (int b = threadIdx.x; b < nbins; b += 256) s_cnt[b] = 0;
        __syncthreads();
        Cand cd[4];
        int bin[4], loc[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = c0 + threadIdx.x + 256 * k;
            bin[k] = -1;
            if (i < n) {
                cd[k] = cand[i];
                const int ly = cand_row(cd[k], T, H);
                bin[k] = (int)((long long)ly * nbins / rows_all);
                loc[k] = atomicAdd(&s_cnt[bin[k]], 1);
            }
        }
        __syncthreads();
        for (int b = threadIdx.x; b < nbins; b += 256) {
            const int c = s_cnt[b];
            s_cnt[b] = c ? atomicAdd(&cursor[b], c) : 0;  // now the bin's base for this step
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (bin[k] >= 0) out[s_cnt[bin[k]] + loc[k]] = cd[k];
        __syncthreads();  // s_cnt is reset by the next step
    }
}

}  // namespace

namespace kern {

hipError_t sort_cands_by_row(const Cand* in, const int* count_in, int cap, int T, int H, int* cursor, Cand* out,
                             int blocks, hipStream_t stream, int nimg) {
    if (H <= 0 || T <= 0 || nimg < 1 || nimg > kMaxBatch) return hipErrorInvalidValue;
    const int rows_all = H * nimg;
    const int nbins = std::min(rows_all, kMaxBins);
    hipLaunchKernelGGL(k_cand_bins, dim3(1), dim3(1024), 0, stream, in, count_in, cap, T, H, rows_all, nbins, cursor);
    hipLaunchKernelGGL(k_cand_scatter, dim3(blocks), dim3(256), 0, stream, in, count_in, cap, T, H, rows_all, nbins,
                       cursor, out);
    return hipGetLastError();
}


int sparse_chunks(const LevelGeom& g) {
    const long long cells = (long long)g.planes * g.T * g.T * g.hw;
    return (int)((cells + kChunk - 1) / kChunk);
}

hipError_t sparse_lists(const LevelGeom& g, int* chunkstart, int* block_sum, uint32_t* entries, long long cap,
                        double max_density, int mode, int* nnz_out, int* flag_out, hipStream_t stream, int nimg,
                        long long rows_b, long long ent_b) {
    const long long cells = (long long)g.planes * g.T * g.T * g.hw;
    const int nchunks = sparse_chunks(g);
    const int nblocks = (nchunks + kChunksPerBlock - 1) / kChunksPerBlock;
    if (nchunks <= 0 || nimg < 1 || nimg > kMaxBatch) return hipErrorInvalidValue;
    const long long dense_limit = mode == 2 ? cells : (long long)(max_density * (double)cells);
    hipLaunchKernelGGL(k_sparse_count, dim3(nblocks, 1, nimg), dim3(kListThreads), 0, stream, g.lm, cells, block_sum,
                       g.img_stride, rows_b);
    hipLaunchKernelGGL(k_sparse_scan, dim3(1, 1, nimg), dim3(1024), 0, stream, block_sum, nblocks, mode == 0 ? -1LL : cap,
                       dense_limit, nnz_out, flag_out, rows_b);
    hipLaunchKernelGGL(k_sparse_fill, dim3(nblocks, 1, nimg), dim3(kListThreads), 0, stream, g.lm, cells, nchunks,
                       block_sum, flag_out, nnz_out, chunkstart, entries, g.img_stride, rows_b, ent_b);
    return hipGetLastError();
}

hipError_t coarse_sparse(const SparseCoarseArgs& a, int blocks, int threads, int seg_max, hipStream_t stream) {
    if (a.seg <= 0 || a.seg > seg_max || a.nimg < 1 || a.nimg > kMaxBatch) return hipErrorInvalidValue;
#define SBM_SPARSE(TH_, SEG_) hipLaunchKernelGGL((k_coarse_sparse<TH_, SEG_>), dim3(blocks), dim3(TH_), 0, stream, a)
    if (threads == 1024 && seg_max == 16384) SBM_SPARSE(1024, 16384);
    else if (threads == 1024 && seg_max == 8192) SBM_SPARSE(1024, 8192);
    else if (threads == 1024 && seg_max == 4096) SBM_SPARSE(1024, 4096);
    else if (threads == 512 && seg_max == 16384) SBM_SPARSE(512, 16384);
    else if (threads == 512 && seg_max == 8192) SBM_SPARSE(512, 8192);
    else if (threads == 512 && seg_max == 4096) SBM_SPARSE(512, 4096);
    else if (threads == 256 && seg_max == 16384) SBM_SPARSE(256, 16384);
    else if (threads == 256 && seg_max == 8192) SBM_SPARSE(256, 8192);
    else if (threads == 256 && seg_max == 4096) SBM_SPARSE(256, 4096);
    else return hipErrorInvalidValue;
#undef SBM_SPARSE
    return hipGetLastError();
}

}  // namespace kern
}  // namespace sbm
