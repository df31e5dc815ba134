// gfx950 kernels for line2Dup's matching path.
//
//   k_gradient      GaussianBlur 7x7 + Sobel + fastAtan2 phase + 16->8 bin + 3x3 vote + mask, fused,
//                   LDS-tiled (line2Dup.cpp:206-302, 419-423)
//   k_pyr_down      cv::pyrDown for the image, nearest 2x decimation for the mask (line2Dup.cpp:397-417)
//   k_spread_lm     OR-spread over TxT + LUT response + linearize, fused: one pass from the quantized
//                   image to the [ori][T*T][H*W] linear memories (line2Dup.cpp:556-750)
//   k_coarse        similarity / similarity_64 gather-accumulate over a template bank with in-kernel
//                   threshold compaction; plain (one template per launch) or persistent over a
//                   feature-count-sorted work list (line2Dup.cpp:780-831, 897-957, 1149-1188)
//   k_refine        similarityLocal / similarityLocal_64 + first-strict-argmax + threshold filter, eight
//                   wavefronts per candidate (line2Dup.cpp:833-895, 959-1021, 1192-1264)
//
// All integer work is exact.  The one float step (the phase angle) is evaluated operation by
// operation in round-to-nearest single precision with contraction disabled, exactly as OpenCV 3.x's
// fastAtan2 does it, so the 8-bin orientation equals the CPU's for every (dx, dy).
#include <type_traits>

#include "sbm_internal.h"
#include "sbm_reach.h"

#pragma clang fp contract(off)

namespace sbm {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// two 16-bit lanes of a dword (packed 16-bit integer arithmetic)
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// BORDER_REFLECT_101 index (pyrDown's border, line2Dup.cpp:397-417)
__device__ __forceinline__ int reflect101(int p, int n) {
    if (n == 1) return 0;
    p = p < 0 ? -p : p;
    p = p >= n ? 2 * n - 2 - p : p;
    return p;  // |overshoot| <= 2 < n for every image the kernels see (n >= 4)
}

// Workgroups go to the eight XCDs in turn by flat index (x fastest, then y, then z).  The tile a block
// takes: flat block lb of nb goes to tile x q + min(x, r) + lb / 8 (x = lb % 8, q = nb / 8, r = nb % 8), so
// each XCD takes a contiguous run of tiles in (image, row, column) order and neighbouring tiles -- the halo
// rows and columns they share -- meet in one L2.  A bijection for any nb.
struct GridTile {
    int x, y, z;
};
__device__ __forceinline__ GridTile xcd_tile() {
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int nb = gx * gy * (int)gridDim.z;
    const int lb = ((int)blockIdx.z * gy + (int)blockIdx.y) * gx + (int)blockIdx.x;
    const int x = lb & 7;
    const int t = x * (nb >> 3) + min(x, nb & 7) + (lb >> 3);
    const int z = t / (gx * gy), rem = t - z * gx * gy;
    return GridTile{rem % gx, rem / gx, z};
}

// RN(num / (den + DBL_EPSILON)) in single precision for integers 0 <= num <= den <= 1020, the magnitudes of the
// Sobel responses of 8-bit images: for den >= 1 the epsilon vanishes in den's rounding, and den = 0 has num = 0
// and the quotient 0.  One reciprocal (v_rcp_f32, within an ulp), the product, its exact residual by fma and one
// correction (Markstein's step) instead of the IEEE division sequence; that this equals the IEEE quotient is not
// a general property of the step but is checked on the device for every pair of the domain (k_div_check).
__device__ __forceinline__ float div_sobel(float num, float den) {
    const float d = fmaxf(den, 1.f);
    const float r = __builtin_amdgcn_rcpf(d);
    const float q = __fmul_rn(num, r);
    return __fmaf_rn(__fmaf_rn(-q, d, num), r, q);
}

// OpenCV 3.x fastAtan2 in degrees (cv::phase(dx, dy, angle, true), line2Dup.cpp:300), for dx, dy in
// [-1020, 1020] (div_sobel's domain).
__device__ __forceinline__ float fast_atan2_deg(float y, float x) {
    const float s = (float)(180.0 / 3.1415926535897932384626433832795);
    const float p1 = 0.9997878412794807f * s;
    const float p3 = -0.3258083974640975f * s;
    const float p5 = 0.1555786518463281f * s;
    const float p7 = -0.04432655554792128f * s;
    const float eps = (float)2.220446049250313080847e-16;
    const float ax = fabsf(x), ay = fabsf(y);
    const bool x_major = ax >= ay;
    const float num = x_major ? ay : ax;
    const float den = x_major ? ax : ay;
    (void)eps;  // c = num / (den + eps), the epsilon vanishing for den >= 1 (div_sobel)
    const float c = div_sobel(num, den);
    const float c2 = __fmul_rn(c, c);
    float a = __fmul_rn(p7, c2);
    a = __fmul_rn(__fadd_rn(a, p5), c2);
    a = __fmul_rn(__fadd_rn(a, p3), c2);
    a = __fmul_rn(__fadd_rn(a, p1), c);
    if (!x_major) a = __fsub_rn(90.f, a);
    if (x < 0.f) a = __fsub_rn(180.f, a);
    if (y < 0.f) a = __fsub_rn(360.f, a);
    return a;
}

// angle.convertTo(CV_8U, 16/360): saturate_cast<uchar>(RN(a * (float)(16/360))), round half to even.
__device__ __forceinline__ int angle_bin(float a) {
    const float scale = (float)(16.0 / 360.0);
    const int q = __float2int_rn(__fmul_rn(a, scale));
    return clampi(q, 0, 255);
}

// The 16-orientation variant: angle.convertTo(CV_16U, 32/360) (line2Dup_16bit_ori.cpp:222-223).
__device__ __forceinline__ int angle_bin32(float a) {
    const float scale = (float)(32.0 / 360.0);
    const int q = __float2int_rn(__fmul_rn(a, scale));
    return clampi(q, 0, 65535);
}

// ------------------------------------------------------------------------------------------------
// k_gradient: one block of 256 threads per 60 x TH output tile; halo 5 (blur 3 + Sobel 1 + vote 1).
//
// The reference blurs with BORDER_REPLICATE and then applies Sobel to the blurred image, again with
// BORDER_REPLICATE (line2Dup.cpp:290-297).  So every Sobel tap is the blur of the replicate-padded
// image at a clamped, in-image position.  The blur passes below run unclamped over the padded input
// tile (their values at out-of-image positions are never read), and the clamp is applied where the
// Sobel stage reads the blurred tile.  That keeps the blur passes plain stencils:
//   1. input tile (TH + 10) x 70 with replicated borders, byte loads, all in flight at once
//   2. horizontal [4 14 28 36 28 14 4]: four adjacent outputs per thread from three LDS dwords
//   3. vertical pass + (sum + 8192) >> 14: one column and a strip of rows per thread, register window
//   4. Sobel (clamped reads), magnitude flag, fastAtan2 phase, 16-bucket bin, "& 7", zeroed ring:
//      one column and a strip of rows per thread
//   5. 3x3 vote over eight 4-bit counters (row sums reused down the strip), mask, byte stores
// ------------------------------------------------------------------------------------------------
constexpr int GQ_TW = 60;                 // output columns per tile
constexpr int GB_W = GQ_TW + 4;           // blurred columns (x0 - 2 .. x0 + 61)
constexpr int GI_W = GQ_TW + 10;          // input columns (x0 - 5 .. x0 + 64)
constexpr int GI_P = 72;                  // input row pitch in LDS (dword reads up to byte 71)

// BINS == 16 is the 16-orientation variant (line2Dup_16bit_ori.cpp:216-296): the phase is binned by
// 32/360, "& 15", the vote counts sixteen buckets, and q holds 16-bit values 1 << bin.
template <int CH, bool FLOATS, int TH, int BINS = 8>
__global__ __launch_bounds__(256) void k_gradient(const uint8_t* __restrict__ img, int rows, int cols, int stride,
                                                  const uint8_t* __restrict__ mask, float thr,
                                                  uint8_t* __restrict__ q, float* __restrict__ mag_out,
                                                  float* __restrict__ ang_out, uint32_t* __restrict__ zero_ptr,
                                                  int zero_n, uint8_t* __restrict__ down,
                                                  uint8_t* __restrict__ down_mask, BatchStrides bs) {
    static_assert(BINS == 8 || BINS == 16, "8 or 16 orientations");
    using Hist = typename std::conditional<BINS == 16, unsigned long long, uint32_t>::type;  // 4-bit counters
    constexpr int IN_H = TH + 10;         // input rows y0 - 5 .. y0 + TH + 4
    constexpr int B_H = TH + 4;           // blurred rows y0 - 2 .. y0 + TH + 1
    constexpr int Q_H = TH + 2;           // quantized rows y0 - 1 .. y0 + TH
    constexpr int SB = B_H / 4;           // blurred rows per strip
    constexpr int SQ = (Q_H + 3) / 4;     // longest quantized strip
    constexpr int SO = TH / 4;            // output rows per strip
    static_assert(TH % 4 == 0, "strips of four");
    __shared__ __align__(16) uint8_t s_in[CH][IN_H][GI_P];
    __shared__ __align__(16) uint16_t s_h[CH][IN_H][GB_W];
    __shared__ uint8_t s_b[CH][B_H][GB_W];
    __shared__ uint8_t s_q[Q_H][GB_W];
    __shared__ Hist s_vote[Q_H][GB_W];  // 1 << 4 bin: the pixel's vote as a 4-bit counter of its bin
    if (zero_ptr && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && (int)threadIdx.x < zero_n)
        zero_ptr[threadIdx.x] = 0u;
    const GridTile tile = xcd_tile();
    {  // image tile.z of a batched launch
        const int z = tile.z;
        if (bs.src) {
            img = bs.src->img[z];
            stride = bs.src->stride[z];
            mask = bs.src->mask[z];
        } else {
            img += z * bs.in_b;
            if (mask) mask += z * bs.in_mask_b;
        }
        q += z * bs.out_b;
        if (down) down += z * bs.down_b;
        if (down_mask) down_mask += z * bs.down_mask_b;
    }
    const int x0 = tile.x * GQ_TW, y0 = tile.y * TH;
    const int tid = threadIdx.x;
    const int lane = tid & 63, strip = tid >> 6;

    // 1. input tile with replicated borders: every load of the thread in flight before the LDS stores.
    // A gray tile away from the borders of a dword-aligned image (x0 is a multiple of 4) takes dword loads:
    // LDS dword m of a row holds image columns x0 - 5 + 4m .. x0 - 2 + 4m, bytes 3..6 of the aligned dwords
    // at x0 - 8 + 4m and x0 - 4 + 4m; the dwords reach column x0 + 67 < cols.
    const bool interior = CH == 1 && (((uintptr_t)img | (uintptr_t)stride) & 3) == 0 && x0 >= 8 &&
                          x0 + 68 <= cols && y0 >= 5 && y0 + TH + 5 <= rows;
    if (interior) {
        constexpr int ND = IN_H * (GI_P / 4);  // LDS dwords of the tile
        constexpr int NL = (ND + 255) / 256;
        uint32_t lo[NL], hi[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int i = min(tid + 256 * k, ND - 1);
            const int a = i / (GI_P / 4), m = i - a * (GI_P / 4);
            const uint32_t* p = reinterpret_cast<const uint32_t*>(img + (size_t)(y0 - 5 + a) * stride + (x0 - 8 + 4 * m));
            lo[k] = p[0];
            hi[k] = p[1];
        }
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int i = tid + 256 * k;
            if (i < ND) {
                const int a = i / (GI_P / 4), m = i - a * (GI_P / 4);
                *reinterpret_cast<uint32_t*>(&s_in[0][a][4 * m]) = __builtin_amdgcn_alignbyte(hi[k], lo[k], 3u);
            }
        }
    } else {
        constexpr int NLOAD = (IN_H * GI_W + 255) / 256;
        uint8_t v[NLOAD][CH];
#pragma unroll
        for (int k = 0; k < NLOAD; ++k) {
            const int i = min(tid + 256 * k, IN_H * GI_W - 1);  // unconditional: the loads issue back to back
            const int a = i / GI_W, b = i - a * GI_W;
            const int ry = clampi(y0 - 5 + a, 0, rows - 1), rx = clampi(x0 - 5 + b, 0, cols - 1);
            const uint8_t* p = img + (size_t)ry * stride + (size_t)rx * CH;
#pragma unroll
            for (int c = 0; c < CH; ++c) v[k][c] = p[c];
        }
#pragma unroll
        for (int k = 0; k < NLOAD; ++k) {
            const int i = tid + 256 * k;
            if (i < IN_H * GI_W) {
                const int a = i / GI_W, b = i - a * GI_W;
#pragma unroll
                for (int c = 0; c < CH; ++c) s_in[c][a][b] = v[k][c];
            }
        }
    }
    __syncthreads();
    // 2. horizontal taps, four outputs (blurred columns 4m .. 4m + 3) per item
    constexpr int NH = IN_H * (GB_W / 4);
#pragma unroll
    for (int k = 0; k < (NH + 255) / 256; ++k) {
        const int i = tid + 256 * k;
        if (i < NH) {
            const int a = i / (GB_W / 4), m = i - a * (GB_W / 4);
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const uint32_t* w = reinterpret_cast<const uint32_t*>(&s_in[c][a][4 * m]);
                const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
                // outputs in 16-bit pairs: x[k] = (byte k, byte k + 1) of the row, and output pair (e, e + 1) is
                // 4 (x[e] + x[e+6]) + 14 (x[e+1] + x[e+5]) + 28 (x[e+2] + x[e+4]) + 36 x[e+3], each half at most
                // 255 * 128 < 2^16
                u16x2 x[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    const uint32_t sel = k < 7 ? (uint32_t)k | (uint32_t)(k + 1) << 16 | 0x0C000C00u
                                               : (uint32_t)(k - 4) | (uint32_t)(k - 3) << 16 | 0x0C000C00u;
                    x[k] = as_u16x2(k < 7 ? __builtin_amdgcn_perm(w1, w0, sel) : __builtin_amdgcn_perm(w2, w1, sel));
                }
                u16x2 s01 = (u16x2)(x[0] + x[6]) * (u16x2)4 + (u16x2)(x[1] + x[5]) * (u16x2)14 +
                            (u16x2)(x[2] + x[4]) * (u16x2)28 + x[3] * (u16x2)36;
                u16x2 s23 = (u16x2)(x[2] + x[8]) * (u16x2)4 + (u16x2)(x[3] + x[7]) * (u16x2)14 +
                            (u16x2)(x[4] + x[6]) * (u16x2)28 + x[5] * (u16x2)36;
                *reinterpret_cast<uint2*>(&s_h[c][a][4 * m]) = make_uint2(as_u32(s01), as_u32(s23));
            }
        }
    }
    // 2b. the next pyramid level of this tile when `down` is set (pyrDown, line2Dup.cpp:397-417: 5x5
    // [1 4 6 4 1]^2 / 256 rounded half up, BORDER_REFLECT_101), 30 x TH/2 outputs from the input tile:
    // horizontal taps around the even columns of input rows y0 - 2 .. y0 + TH; taps past the image border
    // reflect to in-image pixels, which the tile holds.  Rows and columns past the image are not needed.
    constexpr int DR = TH + 3, DW = GQ_TW / 2;
    __shared__ int s_d[CH][DR][DW];
    if (down) {
        // thread: output column X = tid % 32 (X < 30), rows a = tid / 32, + 8, + 16; the five reflected taps of
        // its column once
        const int X = tid & 31, xc = x0 + 2 * X;
        if (X < DW && xc < cols) {
            int cx[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) cx[j] = reflect101(xc - 2 + j, cols) - (x0 - 5);
#pragma unroll
            for (int a = tid >> 5; a < DR; a += 8) {
                const int y = y0 - 2 + a;
                if (y > rows) break;  // and every later row
                const int ry = reflect101(y, rows) - (y0 - 5);
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    const uint8_t* r = s_in[c][ry];
                    s_d[c][a][X] = r[cx[0]] + 4 * r[cx[1]] + 6 * r[cx[2]] + 4 * r[cx[3]] + r[cx[4]];
                }
            }
        }
    }
    __syncthreads();
    // 3. vertical taps: column `lane`, blurred rows strip*SB .. strip*SB + SB - 1
    {
        const int j = lane, b0 = strip * SB;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            int h[SB + 6];
#pragma unroll
            for (int u = 0; u < SB + 6; ++u) h[u] = s_h[c][b0 + u][j];
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int s = 4 * (h[u] + h[u + 6]) + 14 * (h[u + 1] + h[u + 5]) + 28 * (h[u + 2] + h[u + 4]) +
                              36 * h[u + 3];
                s_b[c][b0 + u][j] = (uint8_t)((s + 8192) >> 14);
            }
        }
    }
    // 3b. the pyrDown's vertical taps and its stores (and the mask's, src[2y][2x])
    if (down) {
        const int dr = rows >> 1, dc = cols >> 1;
        for (int i = tid; i < (TH / 2) * DW; i += 256) {
            const int yl = i / DW, X = i - yl * DW;
            const int Y = (y0 >> 1) + yl, Xg = (x0 >> 1) + X;
            if (Y >= dr || Xg >= dc) continue;
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int v = s_d[c][2 * yl][X] + 4 * s_d[c][2 * yl + 1][X] + 6 * s_d[c][2 * yl + 2][X] +
                              4 * s_d[c][2 * yl + 3][X] + s_d[c][2 * yl + 4][X];
                down[((size_t)Y * dc + Xg) * CH + c] = (uint8_t)((v + 128) >> 8);
            }
            if (down_mask) down_mask[(size_t)Y * dc + Xg] = mask[(size_t)(2 * Y) * cols + 2 * Xg];
        }
    }
    __syncthreads();
    // Tiles whose Sobel taps, border ring and vote all stay inside the image (nearly all of them) take stages 4 and
    // 5 without clamps or bounds tests: their LDS indices are then fixed offsets from one base per thread.
    const bool inner = y0 >= 2 && y0 + TH + 2 <= rows && x0 >= 2 && x0 + GQ_TW + 2 <= cols;
    // 4. Sobel with clamped taps, phase, bin: column qj = lane (< 62), rows lo .. hi - 1
    auto sobel_phase = [&](auto inner_tag) {
        constexpr bool IN = decltype(inner_tag)::value;
        const int qj = lane;
        const int lo = (strip * Q_H) / 4, hi = ((strip + 1) * Q_H) / 4;
        const int c = x0 - 1 + qj;
        const int jl = IN ? qj : clampi(c - 1, 0, cols - 1) - (x0 - 2);
        const int jm = IN ? qj + 1 : clampi(c, 0, cols - 1) - (x0 - 2);
        const int jr = IN ? qj + 2 : clampi(c + 1, 0, cols - 1) - (x0 - 2);
        int bl[CH][SQ + 2], bm[CH][SQ + 2], br[CH][SQ + 2];
#pragma unroll
        for (int u = 0; u < SQ + 2; ++u) {
            const int R = IN ? lo + u : clampi(y0 - 2 + lo + u, 0, rows - 1) - (y0 - 2);
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) {
                bl[ch][u] = s_b[ch][R][jl];
                bm[ch][u] = s_b[ch][R][jm];
                br[ch][u] = s_b[ch][R][jr];
            }
        }
#pragma unroll
        for (int u = 0; u < SQ; ++u) {
            const int qi = lo + u;
            if (qi >= hi) break;
            const int r = y0 - 1 + qi;
            uint8_t val = 0;
            if (IN || (r >= 0 && r < rows && c >= 0 && c < cols)) {
                int dx = 0, dy = 0, m = 0;
#pragma unroll
                for (int ch = 0; ch < CH; ++ch) {
                    const int b00 = bl[ch][u], b01 = bm[ch][u], b02 = br[ch][u];
                    const int b10 = bl[ch][u + 1], b12 = br[ch][u + 1];
                    const int b20 = bl[ch][u + 2], b21 = bm[ch][u + 2], b22 = br[ch][u + 2];
                    const int gx = (b02 - b00) + 2 * (b12 - b10) + (b22 - b20);
                    const int gy = (b20 - b00) + 2 * (b21 - b01) + (b22 - b02);
                    const int mm = gx * gx + gy * gy;
                    // colour branch: keep the first channel reaching the maximum (line2Dup.cpp:343-360)
                    if (ch == 0 || mm > m) {
                        dx = gx;
                        dy = gy;
                        m = mm;
                    }
                }
                const float fm = (float)m;
                const float ang = fast_atan2_deg((float)dy, (float)dx);
                const int bin = BINS == 16 ? angle_bin32(ang) : angle_bin(ang);
                const bool ring = !IN && ((r == 0) | (r == rows - 1) | (c == 0) | (c == cols - 1));
                val = ring ? 0 : (uint8_t)(bin & (BINS - 1));
                if (fm > thr) val |= 0x80;
                if (FLOATS && qi >= 1 && qi <= TH && qj >= 1 && qj <= GQ_TW) {
                    mag_out[(size_t)r * cols + c] = fm;
                    ang_out[(size_t)r * cols + c] = ang;
                }
            }
            s_q[qi][qj] = val;
            s_vote[qi][qj] = (Hist)1 << (4 * (val & (BINS - 1)));
        }
    };
    if (lane < GQ_TW + 2) {
        if (inner) sobel_phase(std::true_type{});
        else sobel_phase(std::false_type{});
    }
    __syncthreads();
    // 5. 3x3 vote: column oj = lane (< 60), output rows strip*SO .. strip*SO + SO - 1
    auto vote = [&](auto inner_tag) {
        constexpr bool IN = decltype(inner_tag)::value;
        const int oj = lane, o0 = strip * SO;
        const int c = x0 + oj;
        Hist hrow[SO + 2];
        uint8_t centre[SO];
#pragma unroll
        for (int u = 0; u < SO + 2; ++u) {
            hrow[u] = s_vote[o0 + u][oj] + s_vote[o0 + u][oj + 1] + s_vote[o0 + u][oj + 2];
            if (u >= 1 && u <= SO) centre[u - 1] = s_q[o0 + u][oj + 1];
        }
#pragma unroll
        for (int u = 0; u < SO; ++u) {
            const int r = y0 + o0 + u;
            if (!IN && (r >= rows || c >= cols)) continue;
            uint32_t out = 0;
            if ((IN || (r >= 1 && r <= rows - 2 && c >= 1 && c <= cols - 2)) && (centre[u] & 0x80)) {
                const Hist hist = hrow[u] + hrow[u + 1] + hrow[u + 2];  // BINS 4-bit counters
                // The nine votes leave at most one counter at 5 or more, so the first maximum reaches 5
                // exactly when some counter does, and it is that counter: c + 3 (<= 12, no carry) has bit 3
                // set iff c >= 5.
                constexpr Hist kThree = (Hist)0x3333333333333333ULL, kEight = (Hist)0x8888888888888888ULL;
                const Hist ge5 = (hist + kThree) & kEight;
                if (ge5) {
                    const int idx = BINS == 16 ? (int)__builtin_ctzll((unsigned long long)ge5) >> 2
                                               : (int)__builtin_ctz((uint32_t)ge5) >> 2;
                    out = 1u << idx;
                }
            }
            if (mask && !mask[(size_t)r * cols + c]) out = 0;
            if constexpr (BINS == 16) reinterpret_cast<uint16_t*>(q)[(size_t)r * cols + c] = (uint16_t)out;
            else q[(size_t)r * cols + c] = (uint8_t)out;
        }
    };
    if (lane < GQ_TW) {
        if (inner) vote(std::true_type{});
        else vote(std::false_type{});
    }
}

// ------------------------------------------------------------------------------------------------
// k_pyr_down: 5x5 [1,4,6,4,1]^2 / 256, round half up, BORDER_REFLECT_101; mask: src[2y][2x].
// ------------------------------------------------------------------------------------------------

// Image blockIdx.z of a batched launch: sources from bs.src or at + z * in_b, outputs at + z * out_b.
__device__ __forceinline__ void batch_image(const BatchStrides& bs, const uint8_t* __restrict__& src, int& stride,
                                            const uint8_t* __restrict__& mask, uint8_t* __restrict__& dst,
                                            uint8_t* __restrict__& mask_dst) {
    const int z = blockIdx.z;
    if (bs.src) {
        src = bs.src->img[z];
        stride = bs.src->stride[z];
        mask = bs.src->mask[z];
    } else {
        src += z * bs.in_b;
        if (mask) mask += z * bs.in_mask_b;
    }
    dst += z * bs.out_b;
    if (mask_dst) mask_dst += z * bs.out_mask_b;
}

template <int CH>
__global__ __launch_bounds__(256) void k_pyr_down(const uint8_t* __restrict__ src, int rows, int cols, int stride,
                                                  uint8_t* __restrict__ dst, const uint8_t* __restrict__ mask,
                                                  uint8_t* __restrict__ mask_dst, BatchStrides bs) {
    batch_image(bs, src, stride, mask, dst, mask_dst);
    const int dc = cols / 2, dr = rows / 2;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= dc || y >= dr) return;
    const int w[5] = {1, 4, 6, 4, 1};
    int sr[5], sc[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        sr[i] = reflect101(2 * y + i - 2, rows);
        sc[i] = reflect101(2 * x + i - 2, cols);
    }
#pragma unroll
    for (int ch = 0; ch < CH; ++ch) {
        int s = 0;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const uint8_t* rp = src + (size_t)sr[i] * stride + ch;
            int h = 0;
#pragma unroll
            for (int j = 0; j < 5; ++j) h += w[j] * rp[(size_t)sc[j] * CH];
            s += w[i] * h;
        }
        dst[((size_t)y * dc + x) * CH + ch] = (uint8_t)((s + 128) >> 8);
    }
    if (mask) mask_dst[(size_t)y * dc + x] = mask[(size_t)(2 * y) * cols + 2 * x];
}

// k_pyr_down_w<OUT>: the same for gray images, OUT (2 or 4) adjacent outputs per thread.  The taps of
// the OUT outputs lie in OUT + 1 (3 or 5 ... rounded to whole dwords: 3 or 4) aligned dwords per input
// row starting at byte 2x - 4, so a thread makes 15 or 20 dword loads instead of 50 or 100 byte
// loads; threads whose taps reach the left or right border use reflected columns byte by byte.
template <int OUT>
__global__ __launch_bounds__(256) void k_pyr_down_w(const uint8_t* __restrict__ src, int rows, int cols, int stride,
                                                    uint8_t* __restrict__ dst, const uint8_t* __restrict__ mask,
                                                    uint8_t* __restrict__ mask_dst, BatchStrides bs) {
    batch_image(bs, src, stride, mask, dst, mask_dst);
    constexpr int NDW = OUT == 4 ? 4 : 3;  // dwords per input row
    const int dc = cols / 2, dr = rows / 2;
    const int x = OUT * (blockIdx.x * 64 + (threadIdx.x & 63));
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= dc || y >= dr) return;
    const int w[5] = {1, 4, 6, 4, 1};
    int sr[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) sr[i] = reflect101(2 * y + i - 2, rows);
    uint8_t* out = dst + (size_t)y * dc + x;
    if (x >= 2 && 2 * x - 4 + 4 * NDW <= cols && x + OUT <= dc) {
        uint32_t d[5][NDW];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const uint32_t* p = reinterpret_cast<const uint32_t*>(src + (size_t)sr[i] * stride + 2 * x - 4);
#pragma unroll
            for (int k = 0; k < NDW; ++k) d[i][k] = p[k];
        }
        int acc[OUT];
#pragma unroll
        for (int e = 0; e < OUT; ++e) acc[e] = 0;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            int b[4 * NDW];
#pragma unroll
            for (int k = 0; k < 4 * NDW; ++k) b[k] = (d[i][k >> 2] >> (8 * (k & 3))) & 0xFF;
#pragma unroll
            for (int e = 0; e < OUT; ++e) {  // centre column 2(x + e) is byte 2e + 4
                const int h = b[2 * e + 2] + 4 * b[2 * e + 3] + 6 * b[2 * e + 4] + 4 * b[2 * e + 5] + b[2 * e + 6];
                acc[e] += w[i] * h;
            }
        }
        uint32_t o = 0;
#pragma unroll
        for (int e = 0; e < OUT; ++e) o |= (uint32_t)((acc[e] + 128) >> 8) << (8 * e);
        if (OUT == 4) *reinterpret_cast<uint32_t*>(out) = o;
        else *reinterpret_cast<uint16_t*>(out) = (uint16_t)o;
    } else {
        for (int e = 0; e < OUT && x + e < dc; ++e) {
            int sc[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) sc[j] = reflect101(2 * (x + e) + j - 2, cols);
            int s = 0;
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                const uint8_t* rp = src + (size_t)sr[i] * stride;
                int h = 0;
#pragma unroll
                for (int j = 0; j < 5; ++j) h += w[j] * rp[sc[j]];
                s += w[i] * h;
            }
            out[e] = (uint8_t)((s + 128) >> 8);
        }
    }
    if (mask) {
        for (int e = 0; e < OUT && x + e < dc; ++e)
            mask_dst[(size_t)y * dc + x + e] = mask[(size_t)(2 * y) * cols + 2 * (x + e)];
    }
}

// ------------------------------------------------------------------------------------------------
// k_spread_lm: quantized -> spread (forward TxT OR) -> 8 responses -> linear memories, one pass.
// One block per (image row r, 64 LM columns): the vertical OR of rows r..r+T-1 is staged in LDS,
// then each thread ORs T neighbours of it and writes the 8 responses of that pixel to phase row
// (r % T, c % T); consecutive lanes write consecutive bytes of each phase row.
// ------------------------------------------------------------------------------------------------
constexpr int SL_BX = 64;

// Response of orientation `ori` to a spread bitmask: 4 if bit ori is set, 3 if a circular neighbour
// is, else 0 -- the closed form of SIMILARITY_LUT (line2Dup.cpp:605-608, 654-716).
__device__ __forceinline__ uint32_t response(uint32_t s, int ori) {
    const uint32_t rot = ((s >> ori) | (s << (8 - ori))) & 0xFFu;
    return (rot & 1u) ? 4u : ((rot & 0x82u) ? 3u : 0u);
}

__global__ __launch_bounds__(256) void k_spread_lm(const uint8_t* __restrict__ q, int rows, int cols, int T,
                                                   uint8_t* __restrict__ lm, int hw, int guard16, BatchStrides bs) {
    extern __shared__ uint8_t s_v[];  // SL_BX * T + T - 1 bytes
    q += blockIdx.z * bs.in_b;
    lm += blockIdx.z * bs.out_b;
    const int W = cols / T;
    const int r = blockIdx.y;
    if (r == rows) {
        // the extra row of blocks re-zeroes the guard after the 8 orientations (reads of features
        // outside the image and of tile overshoot land there), whatever an earlier, larger image left
        uint4* g = reinterpret_cast<uint4*>(lm + (size_t)8 * T * T * hw);
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < guard16; i += gridDim.x * blockDim.x)
            g[i] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const int lx0 = blockIdx.x * SL_BX;
    const int c0 = lx0 * T;
    const int vw = SL_BX * T + T - 1;
    const int rmax = min(rows, r + T);
    // vertical OR of rows r .. rmax-1 (at most 16), every load in flight before the OR
    for (int i = threadIdx.x; i < vw; i += blockDim.x) {
        const int c = c0 + i;
        uint32_t v = 0;
        if (c < cols) {
            const uint8_t* col = q + (size_t)r * cols + c;
            uint8_t t[16];
#pragma unroll
            for (int d = 0; d < 16; ++d) t[d] = (r + d < rmax) ? col[(size_t)d * cols] : (uint8_t)0;
#pragma unroll
            for (int d = 0; d < 16; ++d) v |= t[d];
        }
        s_v[i] = (uint8_t)v;
    }
    __syncthreads();
    const int r0 = r % T, ly = r / T;
    const int plane = T * T * hw;
    for (int k = threadIdx.x; k < SL_BX * T; k += blockDim.x) {
        const int lx = k & (SL_BX - 1), coff = k / SL_BX;
        if (lx0 + lx >= W) continue;
        const int pc = lx * T + coff;
        uint32_t sp = 0;
        for (int d = 0; d < T; ++d) sp |= s_v[pc + d];
        uint8_t* out = lm + (size_t)(r0 * T + coff) * hw + (size_t)ly * W + lx0 + lx;
#pragma unroll
        for (int ori = 0; ori < 8; ++ori) out[(size_t)ori * plane] = (uint8_t)response(sp, ori);
    }
}

// k_pack_nibbles: the coarse level's linear memories two positions per byte (responses are 0, 3 or 4,
// so a nibble holds one), halving what the coarse gather moves.  Position order is unchanged: byte i
// holds positions 2i (low nibble) and 2i+1 (high nibble) of the whole [ori][T*T][H*W] + guard block.
__device__ __forceinline__ uint32_t pack4(uint32_t x) {  // bytes b0..b3 (each < 16) -> b0 | b1<<4 | b2<<8 | b3<<12
    x = (x | (x >> 4)) & 0x00FF00FFu;
    return (x | (x >> 8)) & 0x0000FFFFu;
}

__global__ __launch_bounds__(256) void k_pack_nibbles(const uint4* __restrict__ in, uint4* __restrict__ out,
                                                      long long n16, BatchStrides bs) {
    in += blockIdx.z * (bs.in_b / 16);
    out += blockIdx.z * (bs.out_b / 16);
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n16) return;
    const uint4 a = in[2 * i], b = in[2 * i + 1];
    out[i] = make_uint4(pack4(a.x) | (pack4(a.y) << 16), pack4(a.z) | (pack4(a.w) << 16),
                        pack4(b.x) | (pack4(b.y) << 16), pack4(b.z) | (pack4(b.w) << 16));
}

// ------------------------------------------------------------------------------------------------
// k_spread_band<T>: the same result as k_spread_lm, one block per (LM row ly, 64 LM columns) covering
// all T image rows of that LM row, so each quantized byte is loaded once per block instead of once per
// image row, and every store is a dword (four consecutive LM columns).  Needs T % 4 == 0, cols % 4 == 0
// and W % 4 == 0 (the launcher falls back to k_spread_lm otherwise).
//   1. V[r0][c] = OR of q rows ly*T + r0 .. ly*T + r0 + T - 1 (rows past the image read as 0) for the
//      block's columns c0 .. c0 + 64T + T - 2 (columns past the image read as 0).  OR is bitwise, so a
//      dword holds four columns; the T windows over 2T - 1 rows come from suffix/prefix ORs (van Herk).
//   2. a thread takes phase (r0, coff) and LM columns lx .. lx + 3: their spread masks are the ORs of
//      the four T-byte groups of V[r0][lx*T + coff ...], 4T consecutive bytes (dwords realigned).
//   3. eight responses per mask, four masks per dword (the closed form of `response` on each byte),
//      dword stores; with `nib`, the two-positions-per-byte copy of the same positions as well.
// The extra row of blocks (ly == H) re-zeroes both guards.
// ------------------------------------------------------------------------------------------------
constexpr int SB_BX = 64;  // LM columns per block

__device__ __forceinline__ uint32_t fold_or_bytes(uint32_t g) {  // OR of the four bytes, in the low byte
    g |= g >> 16;
    g |= g >> 8;
    return g & 0xFFu;
}

template <int T>
__global__ __launch_bounds__(256) void k_spread_band(const uint8_t* __restrict__ q, int rows, int cols, int W, int H,
                                                     uint8_t* __restrict__ lm, uint8_t* __restrict__ nib,
                                                     int guard16, int nib_guard16, BatchStrides bs) {
    const GridTile tile = xcd_tile();
    q += tile.z * bs.in_b;
    lm += tile.z * bs.out_b;
    if (nib) nib += tile.z * bs.nib_b;
    constexpr int SPAN = SB_BX * T + T - 1;  // image columns a block reads
    constexpr int NW = (SPAN + 3) / 4 + 1;   // dwords per V row; realignment reads one dword past SPAN
    constexpr int NR = 2 * T - 1;            // image rows a block reads
    __shared__ uint32_t s_v[T][NW];
    const long long hw = (long long)W * H;
    const long long plane = (long long)T * T * hw;
    const int ly = tile.y;
    if (ly == H) {
        uint4* g = reinterpret_cast<uint4*>(lm + 8 * plane);
        for (int i = tile.x * blockDim.x + threadIdx.x; i < guard16; i += gridDim.x * blockDim.x)
            g[i] = make_uint4(0u, 0u, 0u, 0u);
        if (nib) {
            uint4* gn = reinterpret_cast<uint4*>(nib + 4 * plane);
            for (int i = tile.x * blockDim.x + threadIdx.x; i < nib_guard16; i += gridDim.x * blockDim.x)
                gn[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        return;
    }
    const int lx0 = tile.x * SB_BX;
    const int c0 = lx0 * T;
    const int rb = ly * T;
    // 1. vertical window ORs, four columns per dword; all loads of a thread issue before the ORs
    for (int w = threadIdx.x; w < NW; w += blockDim.x) {
        const int c = c0 + 4 * w;
        const bool cin = c < cols;  // cols % 4 == 0: a dword is wholly inside or wholly outside
        const int cc = cin ? c : 0;
        uint32_t x[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const int r = min(rb + k, rows - 1);
            x[k] = *reinterpret_cast<const uint32_t*>(q + (size_t)r * cols + cc);
        }
#pragma unroll
        for (int k = 0; k < NR; ++k)
            if (!cin || rb + k >= rows) x[k] = 0u;
        uint32_t suf[T];
        suf[T - 1] = x[T - 1];
#pragma unroll
        for (int i = T - 2; i >= 0; --i) suf[i] = x[i] | suf[i + 1];
        s_v[0][w] = suf[0];
        uint32_t pre = 0u;
#pragma unroll
        for (int r0 = 1; r0 < T; ++r0) {
            pre |= x[T + r0 - 1];
            s_v[r0][w] = suf[r0] | pre;
        }
    }
    __syncthreads();
    // 2-3. spread masks of four LM columns, responses, dword (and nibble) stores
    for (int it = threadIdx.x; it < T * T * (SB_BX / 4); it += blockDim.x) {
        const int phase = it / (SB_BX / 4), quad = it - phase * (SB_BX / 4);
        const int lx = lx0 + 4 * quad;
        if (lx >= W) continue;
        const int r0 = phase / T, coff = phase - r0 * T;
        const int c = 4 * quad * T + coff;  // first byte of the 4T-byte run, relative to c0
        const int w0 = c >> 2, sh = c & 3;
        uint32_t u[T];
#pragma unroll
        for (int k = 0; k < T; ++k) u[k] = __builtin_amdgcn_alignbyte(s_v[r0][w0 + k + 1], s_v[r0][w0 + k], sh);
        uint32_t s = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t g = 0u;
#pragma unroll
            for (int k = 0; k < T / 4; ++k) g |= u[j * (T / 4) + k];
            s |= fold_or_bytes(g) << (8 * j);
        }
        const long long pos = (long long)phase * hw + (long long)ly * W + lx;
#pragma unroll
        for (int ori = 0; ori < 8; ++ori) {
            // per byte: 4 if bit ori is set, else 3 if bit ori+1 or ori-1 (mod 8) is, else 0
            const uint32_t b0 = (s >> ori) & 0x01010101u;
            const uint32_t b1 = (s >> ((ori + 1) & 7)) & 0x01010101u;
            const uint32_t b7 = (s >> ((ori + 7) & 7)) & 0x01010101u;
            const uint32_t r = (b0 << 2) | (((b1 | b7) & ~b0) * 3u);
            *reinterpret_cast<uint32_t*>(lm + ori * plane + pos) = r;
            if (nib) *reinterpret_cast<uint16_t*>(nib + ((ori * plane + pos) >> 1)) = (uint16_t)pack4(r);
        }
    }
}

// ------------------------------------------------------------------------------------------------
// k_spread16: the 16-orientation variant's spread + response + linearize (line2Dup_16bit_ori.cpp:542-730)
// for any T in [1, 16], one block per (LM row ly, 32 LM columns).
//   1. V[r0][c] = OR of q rows ly*T + r0 .. ly*T + r0 + T - 1 (rows and columns past the image read 0),
//      16-bit, for the block's columns; one column per thread item, its 2T - 1 loads in flight
//   2. a thread takes phase (r0, coff) and four LM columns: each spread value is the OR of T
//      consecutive V values
//   3. sixteen responses per value from the compiled table's closed form, dword stores of four columns
// The response of orientation o to a spread value s: the table is indexed by the nibbles of s, the fourth
// being (s & (15 << 16)) >> 16 == 0 (line2Dup_16bit_ori.cpp:636-639), so only bits 0..11 count.  Rotating
// them so that bit o lands on bit 0, the value is 4 when a set bit lies within 2 orientations of o
// (bits 0, 1, 2, 14, 15), else 1 within 4 (bits 3, 4, 12, 13), else 0.
// The extra row of blocks (ly == H) re-zeroes the guard after the 16 planes.
// ------------------------------------------------------------------------------------------------
constexpr int S16_BX = 32;  // LM columns per block

__device__ __forceinline__ uint32_t response16(uint32_t s, int ori) {
    const uint32_t m = s & 0x0FFFu;
    const uint32_t rot = ((m >> ori) | (m << (16 - ori))) & 0xFFFFu;
    return (rot & 0xC007u) ? 4u : ((rot & 0x3018u) ? 1u : 0u);
}

__global__ __launch_bounds__(256) void k_spread16(const uint16_t* __restrict__ q, int rows, int cols, int T, int W, int H,
                                                  uint8_t* __restrict__ lm, int guard16, BatchStrides bs) {
    extern __shared__ uint16_t s_v16[];  // [T][span]
    q = reinterpret_cast<const uint16_t*>(reinterpret_cast<const uint8_t*>(q) + blockIdx.z * bs.in_b);
    lm += blockIdx.z * bs.out_b;
    const long long hw = (long long)W * H;
    const long long plane = (long long)T * T * hw;
    const int ly = blockIdx.y;
    if (ly == H) {
        uint4* g = reinterpret_cast<uint4*>(lm + 16 * plane);
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < guard16; i += gridDim.x * blockDim.x)
            g[i] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const int lx0 = blockIdx.x * S16_BX;
    const int c0 = lx0 * T;
    const int span = S16_BX * T + T - 1;
    const int rb = ly * T;
    // 1. vertical ORs: rows rb .. rb + 2T - 2 of column c, then the T windows (suffix / prefix ORs)
    for (int i = threadIdx.x; i < span; i += blockDim.x) {
        const int c = c0 + i;
        uint32_t x[31];
#pragma unroll
        for (int k = 0; k < 31; ++k) {
            const int r = rb + k;
            x[k] = (k < 2 * T - 1 && c < cols && r < rows) ? q[(size_t)r * cols + c] : 0u;
        }
        uint32_t suf = 0u;
#pragma unroll
        for (int k = 15; k >= 0; --k) {  // suf = OR of x[k .. T-1]; V[k] = suf | OR of x[T .. T+k-1]
            if (k < T) suf |= x[k];
            if (k < T) {
                uint32_t pre = 0u;
#pragma unroll
                for (int e = 0; e < 15; ++e)
                    if (e < k) pre |= x[T + e];
                s_v16[k * span + i] = (uint16_t)(suf | pre);
            }
        }
    }
    __syncthreads();
    // 2-3. four LM columns of one phase per item
    const bool words = (W % 4) == 0;  // then every dword store below is aligned
    for (int it = threadIdx.x; it < T * T * (S16_BX / 4); it += blockDim.x) {
        const int phase = it / (S16_BX / 4), quad = it - phase * (S16_BX / 4);
        const int lx = lx0 + 4 * quad;
        if (lx >= W) continue;
        const int r0 = phase / T, coff = phase - r0 * T;
        const uint16_t* vrow = s_v16 + r0 * span + 4 * quad * T + coff;
        uint32_t sp[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t v = 0u;
            for (int d = 0; d < T; ++d) v |= vrow[j * T + d];
            sp[j] = v;
        }
        const long long pos = (long long)phase * hw + (long long)ly * W + lx;
        const int nx = min(4, W - lx);
#pragma unroll
        for (int ori = 0; ori < 16; ++ori) {
            uint32_t r = 0u;
#pragma unroll
            for (int j = 0; j < 4; ++j) r |= response16(sp[j], ori) << (8 * j);
            uint8_t* dst = lm + ori * plane + pos;
            if (words && nx == 4) {
                *reinterpret_cast<uint32_t*>(dst) = r;
            } else {
                for (int j = 0; j < nx; ++j) dst[j] = (uint8_t)(r >> (8 * j));
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Gather-accumulate core shared by the coarse kernel and its debug twin.
//
// A wavefront owns PPL consecutive positions per lane (64*PPL per tile).  Each feature adds the
// bytes lm_row[off + j] for its positions j.  Feature records are read 64 at a time, one per lane;
// each lane turns its record into a 32-bit byte offset into the linear memories (features outside
// the image point at the zeroed guard after the last orientation, so no branch is needed), and the
// loop broadcasts the offsets with readlane.  Four positions are summed per 32-bit register
// (bytewise: at most 32 features of value <= 4 per chunk, so no byte carries), then widened into
// 16-bit pairs, which hold up to 4 * 8191 = 32764.  A feature's byte offset is uniform across the
// wave, so a lane's PPL bytes are either one unaligned buffer vector load (the feature's offset in
// the scalar offset register) or aligned dword loads realigned with v_alignbyte.
// ------------------------------------------------------------------------------------------------
typedef const uint32_t __attribute__((address_space(1)))* gptr_u32;

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Byte offset of feature (fx, fy, label) in the [ori][T*T][H*W] block, or `zero_off` (the zeroed
// guard) when the feature lies outside the image (line2Dup.cpp:755-778, 809-810, 851-852).
// NW + 1 dwords from the dword-aligned address p, realigned by sh bytes (uniform: a lane's positions start
// at a multiple of 4 past the tile's).
template <int NW>
__device__ __forceinline__ void gather_add(const uint8_t* p, uint32_t sh, uint32_t (&acc8)[NW]) {
    const gptr_u32 wp = (gptr_u32)p;
    uint32_t v[NW + 1];
#pragma unroll
    for (int k = 0; k <= NW; ++k) v[k] = wp[k];
#pragma unroll
    for (int k = 0; k < NW; ++k) acc8[k] += __builtin_amdgcn_alignbyte(v[k + 1], v[k], sh);
}

// Unaligned vector load of a lane's NW dwords at byte offset voff + soff of the linear memories,
// through a buffer resource: soff (the feature's offset plus the tile start) is wave-uniform and rides
// in an SGPR, voff (the lane's position in the tile) is fixed per lane, so a feature costs one load
// instruction and no address arithmetic.  Loads past lm_bytes return zero.
typedef int32_t i32x2 __attribute__((ext_vector_type(2)));
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
constexpr int kBufferWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle

template <int NW>
__device__ __forceinline__ void gather_add_buf(__amdgpu_buffer_rsrc_t rsrc, int voff, int soff, uint32_t (&acc8)[NW]) {
    if constexpr (NW == 4) {
        const i32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff, soff, 0);
        acc8[0] += (uint32_t)v.x;
        acc8[1] += (uint32_t)v.y;
        acc8[2] += (uint32_t)v.z;
        acc8[3] += (uint32_t)v.w;
    } else if constexpr (NW == 2) {
        const i32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, voff, soff, 0);
        acc8[0] += (uint32_t)v.x;
        acc8[1] += (uint32_t)v.y;
    } else {
        acc8[0] += (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, soff, 0);
    }
}

// Position k of a lane's PPL positions inside the interleaved 16-bit accumulators.
template <int PPL>
__device__ __forceinline__ int lane_raw(const uint32_t (&acc16)[PPL / 2], int k) {
    const uint32_t w = acc16[(k >> 2) * 2 + (k & 1)];
    return (k & 2) ? (int)(w >> 16) : (int)(w & 0xFFFFu);
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = max(v, __shfl_xor(v, s, kWave));
    return v;
}

// Gather modes of the coarse kernel (a lane's PPL consecutive bytes at a wave-uniform byte offset):
//   0  aligned dword loads per lane (NW + 1 of them) realigned with v_alignbyte
//   1  one unaligned buffer vector load per lane; the hardware splits every lane's access
//   2  one aligned buffer vector load per lane, the next lane's dwords fetched with a DPP wave shift
//      and realigned with v_alignbyte; lane 63 only supplies data, so a tile is 63 * PPL positions
template <int GATHER>
__device__ __forceinline__ constexpr int position_lanes() { return coarse_position_lanes(GATHER); }

__device__ __forceinline__ uint32_t from_next_lane(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, false);  // wave_shl:1
}

template <int NW>
__device__ __forceinline__ void load_aligned(__amdgpu_buffer_rsrc_t rsrc, int voff, int soff, uint32_t (&c)[NW]) {
    if constexpr (NW == 4) {
        const i32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff, soff, 0);
        c[0] = (uint32_t)v.x;
        c[1] = (uint32_t)v.y;
        c[2] = (uint32_t)v.z;
        c[3] = (uint32_t)v.w;
    } else if constexpr (NW == 2) {
        const i32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, voff, soff, 0);
        c[0] = (uint32_t)v.x;
        c[1] = (uint32_t)v.y;
    } else {
        c[0] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, soff, 0);
    }
}

// acc8[k] += bytes [4(Q+k) + s3, +4) of (own chunk ++ next lane's chunk); Q = dword skip (uniform).
template <int NW, int Q>
__device__ __forceinline__ void combine_next(const uint32_t (&c)[NW], uint32_t s3, uint32_t (&acc8)[NW]) {
    uint32_t w[2 * NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) w[k] = c[k];
#pragma unroll
    for (int k = 0; k <= Q; ++k) w[NW + k] = from_next_lane(c[k]);
#pragma unroll
    for (int k = 0; k < NW; ++k) acc8[k] += __builtin_amdgcn_alignbyte(w[Q + k + 1], w[Q + k], s3);
}

template <int NW>
__device__ __forceinline__ void add_realigned(const uint32_t (&c)[NW], int sh, uint32_t (&acc8)[NW]) {
    const uint32_t s3 = (uint32_t)(sh & 3);
    switch (sh >> 2) {  // wave-uniform
        case 0: combine_next<NW, 0>(c, s3, acc8); break;
        case 1: if constexpr (NW > 1) combine_next<NW, 1>(c, s3, acc8); break;
        case 2: if constexpr (NW > 2) combine_next<NW, 2>(c, s3, acc8); break;
        default: if constexpr (NW > 3) combine_next<NW, 3>(c, s3, acc8); break;
    }
}

// One feature at byte offset `o` (uniform) for every lane of the tile starting at j0.
template <int NW, int GATHER>
__device__ __forceinline__ void gather_feature(const LevelGeom& g, __amdgpu_buffer_rsrc_t rsrc, int o, int j0,
                                               int lane, uint32_t (&acc8)[NW]) {
    constexpr int PPL = 4 * NW;
    const LaneLoad ld = coarse_lane_load(o, j0, lane, PPL, GATHER);
    if constexpr (GATHER == 0) {
        gather_add<NW>(g.lm + ld.soff, (uint32_t)((o + j0) & 3), acc8);
    } else if constexpr (GATHER == 1) {
        gather_add_buf<NW>(rsrc, ld.voff, ld.soff, acc8);
    } else {
        uint32_t c[NW];
        load_aligned<NW>(rsrc, ld.voff, ld.soff, c);
        add_realigned<NW>(c, (o + j0) & (PPL - 1), acc8);
    }
}

// Sum over the template's features of the LM bytes at positions jlane .. jlane+PPL-1 (jlane = the
// wave's tile start j0 plus lane * PPL), as 16-bit pairs.  Features are taken kChunk at a time so the
// byte lanes of the SWAR accumulators (<= kChunk * 4) never carry.
//
// Early exit (rmin > 0): every response is at most 4, so after `done` features no position can end
// above partial + 4 * (nfeat - done).  When that bound is below rmin for every position j < limit of
// the tile, no candidate can come out of it and the rest of the features are skipped; the function
// returns false.  Positions at or past `limit` are zero or unreported, so they do not hold the tile.
//
// first / step: the chunks this wave sums (first, first + step, ...), so that `step` wavefronts can
// share one template's features; the early exit needs every feature, so it applies to step == 1 only.
constexpr int kChunk = 32;
template <int PPL, int GATHER>
__device__ __forceinline__ bool accumulate_tile(const LevelGeom& g, const uint2* __restrict__ F, int nfeat,
                                                int j0, int lane, int rmin, int limit, uint32_t (&acc16)[PPL / 2],
                                                int first = 0, int step = 1) {
    constexpr int NW = PPL / 4;
#pragma unroll
    for (int k = 0; k < PPL / 2; ++k) acc16[k] = 0;
    const int zero_off = g.zero_off();
    const int jlane = j0 + lane * PPL;
    const bool positions = lane < position_lanes<GATHER>();
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)g.lm, (short)0, g.lm_bytes, kBufferWord3);
    int my_off = zero_off, loaded = -1;
    for (int base = first * kChunk; base < nfeat; base += step * kChunk) {
        const int cnt = min(kChunk, nfeat - base);
        if ((base >> 6) != loaded) {  // offsets of the 64 features of this block, one per lane
            loaded = base >> 6;
            my_off = zero_off;
            if ((base & ~63) + lane < nfeat) {
                const uint2 f = F[(base & ~63) + lane];
                my_off = feature_offset(g, (int)(int16_t)(f.x & 0xFFFFu), (int)(int16_t)(f.x >> 16), (int)f.y, zero_off);
            }
        }
        const int sl = base & 63;
        uint32_t acc8[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) acc8[k] = 0;
        int i = 0;
        if constexpr (GATHER == 2) {  // all eight loads in flight before the first realignment
            for (; i + 8 <= cnt; i += 8) {
                uint32_t c[8][NW];
                int sh[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int o = __builtin_amdgcn_readlane(my_off, sl + i + u);
                    const LaneLoad ld = coarse_lane_load(o, j0, lane, PPL, GATHER);
                    sh[u] = (o + j0) & (PPL - 1);
                    load_aligned<NW>(rsrc, ld.voff, ld.soff, c[u]);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) add_realigned<NW>(c[u], sh[u], acc8);
            }
            for (; i + 4 <= cnt; i += 4) {
                uint32_t c[4][NW];
                int sh[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int o = __builtin_amdgcn_readlane(my_off, sl + i + u);
                    const LaneLoad ld = coarse_lane_load(o, j0, lane, PPL, GATHER);
                    sh[u] = (o + j0) & (PPL - 1);
                    load_aligned<NW>(rsrc, ld.voff, ld.soff, c[u]);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) add_realigned<NW>(c[u], sh[u], acc8);
            }
        }
        for (; i + 4 <= cnt; i += 4) {
            const int o0 = __builtin_amdgcn_readlane(my_off, sl + i);
            const int o1 = __builtin_amdgcn_readlane(my_off, sl + i + 1);
            const int o2 = __builtin_amdgcn_readlane(my_off, sl + i + 2);
            const int o3 = __builtin_amdgcn_readlane(my_off, sl + i + 3);
            gather_feature<NW, GATHER>(g, rsrc, o0, j0, lane, acc8);
            gather_feature<NW, GATHER>(g, rsrc, o1, j0, lane, acc8);
            gather_feature<NW, GATHER>(g, rsrc, o2, j0, lane, acc8);
            gather_feature<NW, GATHER>(g, rsrc, o3, j0, lane, acc8);
        }
        for (; i < cnt; ++i)
            gather_feature<NW, GATHER>(g, rsrc, __builtin_amdgcn_readlane(my_off, sl + i), j0, lane, acc8);
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            acc16[2 * k] += acc8[k] & 0x00FF00FFu;
            acc16[2 * k + 1] += (acc8[k] >> 8) & 0x00FF00FFu;
        }
        const int left = nfeat - base - cnt;
        if (step == 1 && rmin > 0 && left > 0) {
            int m = 0;
            if (positions) {
#pragma unroll
                for (int k = 0; k < PPL; ++k)
                    if (jlane + k < limit) m = max(m, lane_raw<PPL>(acc16, k));
            }
            if (wave_max(m) + 4 * left < rmin) return false;
        }
    }
    return true;
}

// Nibble-packed gather (mode 3): a lane owns 32 positions = 16 packed bytes.  The aligned chunk and
// the next lane's (DPP) are realigned at nibble granularity with v_alignbit; three features are summed
// per nibble (<= 12, no carry) before the nibbles are split into byte accumulators lo (even positions)
// and hi (odd positions).  acc16 layout: position 8k + m sits in acc16[4k + (m&1)*2 + ((m>>1)&1)],
// half (m >> 2) & 1.
__device__ __forceinline__ int lane_raw_nib(const uint32_t (&acc16)[16], int p) {
    const int k = p >> 3, m = p & 7;
    const uint32_t w = acc16[4 * k + (m & 1) * 2 + ((m >> 1) & 1)];
    return ((m >> 2) & 1) ? (int)(w >> 16) : (int)(w & 0xFFFFu);
}

template <int Q>
__device__ __forceinline__ void combine_next_nib(const uint32_t (&c)[4], uint32_t bits, uint32_t (&s)[4]) {
    uint32_t w[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = c[k];
#pragma unroll
    for (int k = 0; k <= Q; ++k) w[4 + k] = from_next_lane(c[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) s[k] += __builtin_amdgcn_alignbit(w[Q + k + 1], w[Q + k], bits);
}

__device__ __forceinline__ void add_realigned_nib(const uint32_t (&c)[4], int sh, uint32_t (&s)[4]) {
    const uint32_t bits = (uint32_t)(sh & 7) * 4u;
    switch (sh >> 3) {  // wave-uniform
        case 0: combine_next_nib<0>(c, bits, s); break;
        case 1: combine_next_nib<1>(c, bits, s); break;
        case 2: combine_next_nib<2>(c, bits, s); break;
        default: combine_next_nib<3>(c, bits, s); break;
    }
}

__device__ __forceinline__ void split_nibbles(const uint32_t (&s)[4], uint32_t (&lo)[4], uint32_t (&hi)[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        lo[k] += s[k] & 0x0F0F0F0Fu;
        hi[k] += (s[k] >> 4) & 0x0F0F0F0Fu;
    }
}

constexpr int kNibChunk = 30;  // byte accumulators: <= 30 * 4 = 120
__device__ __forceinline__ bool accumulate_tile_nib(const LevelGeom& g, const uint2* __restrict__ F, int nfeat, int j0,
                                                    int lane, int rmin, int limit, uint32_t (&acc16)[16],
                                                    int first = 0, int step = 1) {
#pragma unroll
    for (int k = 0; k < 16; ++k) acc16[k] = 0;
    const int zero_off = g.zero_off();
    const int jlane = j0 + lane * 32;
    const bool positions = lane < kWave - 1;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void*)g.lm_nib, (short)0, g.nib_bytes, kBufferWord3);
    int my_off = zero_off, loaded = -1;
    for (int base = first * kNibChunk; base < nfeat; base += step * kNibChunk) {
        const int cnt = min(kNibChunk, nfeat - base);
        const int blk = base / (2 * kNibChunk), sl = base - blk * (2 * kNibChunk);
        if (blk != loaded) {  // offsets of the 60 features of this block, one per lane
            loaded = blk;
            my_off = zero_off;
            const int fb = blk * (2 * kNibChunk);
            if (lane < 2 * kNibChunk && fb + lane < nfeat) {
                const uint2 f = F[fb + lane];
                my_off = feature_offset(g, (int)(int16_t)(f.x & 0xFFFFu), (int)(int16_t)(f.x >> 16), (int)f.y, zero_off);
            }
        }
        uint32_t lo[4] = {0u, 0u, 0u, 0u}, hi[4] = {0u, 0u, 0u, 0u};
        int i = 0;
        for (; i + 6 <= cnt; i += 6) {  // six loads in flight, summed three at a time
            uint32_t c[6][4];
            int sh[6];
#pragma unroll
            for (int u = 0; u < 6; ++u) {
                const int o = __builtin_amdgcn_readlane(my_off, sl + i + u);
                const LaneLoad ld = coarse_lane_load(o, j0, lane, 32, 3);
                sh[u] = (o + j0) & 31;
                load_aligned<4>(rsrc, ld.voff, ld.soff, c[u]);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                uint32_t s[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int u = 0; u < 3; ++u) add_realigned_nib(c[3 * h + u], sh[3 * h + u], s);
                split_nibbles(s, lo, hi);
            }
        }
        for (; i < cnt; i += 3) {
            uint32_t s[4] = {0u, 0u, 0u, 0u};
            const int n3 = min(3, cnt - i);
            for (int u = 0; u < n3; ++u) {
                const int o = __builtin_amdgcn_readlane(my_off, sl + i + u);
                const LaneLoad ld = coarse_lane_load(o, j0, lane, 32, 3);
                uint32_t c[4];
                load_aligned<4>(rsrc, ld.voff, ld.soff, c);
                add_realigned_nib(c, (o + j0) & 31, s);
            }
            split_nibbles(s, lo, hi);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            acc16[4 * k + 0] += lo[k] & 0x00FF00FFu;
            acc16[4 * k + 1] += (lo[k] >> 8) & 0x00FF00FFu;
            acc16[4 * k + 2] += hi[k] & 0x00FF00FFu;
            acc16[4 * k + 3] += (hi[k] >> 8) & 0x00FF00FFu;
        }
        const int left = nfeat - base - cnt;
        if (step == 1 && rmin > 0 && left > 0) {
            int m = 0;
            if (positions) {
#pragma unroll
                for (int k = 0; k < 32; ++k)
                    if (jlane + k < limit) m = max(m, lane_raw_nib(acc16, k));
            }
            if (wave_max(m) + 4 * left < rmin) return false;
        }
    }
    return true;
}

template <int PPL, int GATHER>
__device__ __forceinline__ int raw_at(const uint32_t (&acc16)[PPL / 2], int k) {
    if constexpr (GATHER == 3) return lane_raw_nib(acc16, k);
    else return lane_raw<PPL>(acc16, k);
}

// Threshold compaction of one tile: positions j < hw of this lane whose raw score reaches rmin (and
// j < P; later positions score 0) are appended to the candidate list with one atomic per wave.
template <int PPL, int GATHER>
__device__ __forceinline__ void emit_tile(const kern::CoarseArgs& a, const uint32_t (&acc16)[PPL / 2], bool keep, int jl,
                                          int P, int rmin, int gidx) {
    const int hw = (int)a.geom.hw;
    int pass = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int j = jl + k;
        const int raw = j < P ? raw_at<PPL, GATHER>(acc16, k) : 0;
        pass += (keep && j < hw && raw >= rmin) ? 1 : 0;
    }
    if (!pass) return;
    int slot = atomicAdd(a.cand_count, pass);
    const int T = a.geom.T, W = a.geom.W;
    const int offs = T / 2 + (T % 2 - 1);  // line2Dup.cpp:1181
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int j = jl + k;
        const int raw = j < P ? raw_at<PPL, GATHER>(acc16, k) : 0;
        if (j < hw && raw >= rmin) {
            if (slot < a.cand_cap) {
                const int r = j / W, c = j - r * W;
                Cand cd;
                cd.g = gidx;
                cd.x = c * T + offs;
                cd.y = r * T + offs;
                cd.raw = raw;
                a.cand[slot] = cd;
            }
            ++slot;
        }
    }
}

template <int PPL, int GATHER>
__device__ __forceinline__ bool accumulate_any(const LevelGeom& g, const uint2* F, int nfeat, int j0, int lane, int rmin,
                                               int limit, uint32_t (&acc16)[PPL / 2], int first, int step) {
    if constexpr (GATHER == 3) return accumulate_tile_nib(g, F, nfeat, j0, lane, rmin, limit, acc16, first, step);
    else return accumulate_tile<PPL, GATHER>(g, F, nfeat, j0, lane, rmin, limit, acc16, first, step);
}

// Work item = (tile, template).  SPLIT == 1: one wavefront per item, items from the queue per wave.
// SPLIT > 1: the block's SPLIT wavefronts share one item's features (chunks interleaved) and add their
// partial sums through LDS -- a shorter critical path per item when there are few items.
template <int PPL, int GATHER, int SPLIT>
__global__ __launch_bounds__(SPLIT == 8 ? 512 : 256) void k_coarse(kern::CoarseArgs a) {
    if (a.dense_skip) {  // nothing to do when the nonzero-response path has every image
        bool all = true;
        for (int z = 0; z < a.nimg; ++z) all &= a.dense_skip[z] != 0;
        if (all) return;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const long long per_image = (long long)a.ntempl * a.ntiles;
    const long long total = per_image * a.nimg;
    const int tile_pos = position_lanes<GATHER>() * PPL;
    const bool positions = lane < position_lanes<GATHER>();
    const int hw = (int)a.geom.hw;
    const int w = (int)(threadIdx.x / kWave);
    long long item = SPLIT > 1 ? (long long)blockIdx.x
                               : (long long)uniform(blockIdx.x * (blockDim.x / kWave) + w);
    long long workers = SPLIT > 1 ? (long long)gridDim.x : (long long)gridDim.x * (blockDim.x / kWave);
    long long end = total;
    if (a.work_counter && a.static_order && gridDim.x % 8 == 0) {
        // Workgroups go to the eight XCDs in turn (block b on XCD b % 8): XCD x takes the x-th eighth of the
        // image-major, tile-major item list, so its L2 holds only that share's linear-memory windows.
        const int xcd = (int)(blockIdx.x % 8);
        const long long per_block = SPLIT > 1 ? 1 : (long long)(blockDim.x / kWave);
        workers /= 8;
        item = total * xcd / 8 + (long long)(blockIdx.x / 8) * per_block + (SPLIT > 1 ? 0 : w);
        end = total * (xcd + 1) / 8;
    }
    const int grab = a.grab > 1 ? a.grab : 1;
    bool first_item = true;
    int left = 0;  // items left in the current grab
    for (;;) {
        if (a.work_counter && a.static_order) {  // item = worker, worker + workers, ... (of its XCD's share)
            if (!first_item) item += workers;
            first_item = false;
        } else if (a.work_counter) {
            if (left > 0) {
                ++item;
                --left;
            } else {
                if constexpr (SPLIT > 1) {
                    __shared__ int s_item;
                    if (threadIdx.x == 0) s_item = atomicAdd(a.work_counter, 1);
                    __syncthreads();
                    item = (long long)s_item * grab;
                    __syncthreads();  // every wave has read it before the next grab overwrites it
                } else {
                    int v = 0;
                    if (lane == 0) v = atomicAdd(a.work_counter, 1);
                    item = (long long)uniform(v) * grab;
                }
                left = grab - 1;
            }
        }
        if (item >= end) break;
        // image-major, then tile-major: the waves taking neighbouring items work on the same LM window
        const int img = uniform((int)(item / per_image));
        const long long rem = item - (long long)img * per_image;
        const int tile = uniform((int)(rem / a.ntempl));
        const int gidx = uniform(a.order[rem - (long long)tile * a.ntempl]);
        LevelGeom geo = a.geom;
        geo.lm += img * a.geom.img_stride;
        if (geo.lm_nib) geo.lm_nib += img * a.geom.nib_stride;
        const bool skip_img = a.dense_skip && a.dense_skip[img] != 0;
        const TemplLevelDev t = a.tl[(long long)gidx * a.levels + a.level];
        const int nfeat = uniform(t.nfeat), foff = uniform(t.off);
        const int rmin = uniform(a.raw_min[gidx]);
        const int P = template_positions(a.geom, uniform(t.width), uniform(t.height));
        const int limit = rmin == 0 ? hw : min(P, hw);
        const int j0 = tile * tile_pos;
        const int jl = j0 + lane * PPL;
        const bool valid = rmin <= 4 * nfeat && j0 < limit && !skip_img;
        const int gtag = gidx | (img << kCandImageShift);
        uint32_t acc16[PPL / 2];
        if constexpr (SPLIT == 1) {
            if (valid) {
                const bool live = accumulate_any<PPL, GATHER>(geo, a.feats + foff, nfeat, j0, lane, rmin, limit, acc16, 0, 1);
                emit_tile<PPL, GATHER>(a, acc16, live && positions, jl, P, rmin, gtag);
            }
        } else {
            __shared__ uint32_t s_part[SPLIT - 1][PPL / 2][kWave];
            if (valid) {
                accumulate_any<PPL, GATHER>(geo, a.feats + foff, nfeat, j0, lane, 0, limit, acc16, w, SPLIT);
                if (w > 0) {
#pragma unroll
                    for (int k = 0; k < PPL / 2; ++k) s_part[w - 1][k][lane] = acc16[k];
                }
            }
            __syncthreads();
            if (valid && w == 0) {
#pragma unroll
                for (int s = 0; s < SPLIT - 1; ++s)
#pragma unroll
                    for (int k = 0; k < PPL / 2; ++k) acc16[k] += s_part[s][k][lane];
                emit_tile<PPL, GATHER>(a, acc16, positions, jl, P, rmin, gtag);
            }
            __syncthreads();  // s_part is reused by the next item
        }
        if (!a.work_counter) break;
    }
}

// Full similarity map of one template (zeros at positions >= template_positions), for parity tests.
template <int PPL>
__global__ __launch_bounds__(64) void k_simmap(LevelGeom g, const uint2* feats, int width, int height, int nfeat,
                                               int off, uint16_t* out) {
    const int lane = threadIdx.x;
    const int j0 = blockIdx.x * kWave * PPL;
    const int jl = j0 + lane * PPL;
    const int P = template_positions(g, width, height);
    uint32_t acc16[PPL / 2];
    accumulate_tile<PPL, 1>(g, feats + off, nfeat, j0, lane, 0, 0, acc16);
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int j = jl + k;
        if (j < g.hw) out[j] = (uint16_t)(j < P ? lane_raw<PPL>(acc16, k) : 0);
    }
}

// ------------------------------------------------------------------------------------------------
// Refinement: a 16x16 window per candidate; lane (r, c0..c0+3) of the window, features split over
// the block's wavefronts and summed through LDS.
// ------------------------------------------------------------------------------------------------
// Sums the features of chunks first_chunk, first_chunk + chunk_step, ... (63 features per chunk, so
// the bytewise sums stay below 256).  A lane's four window bytes come from the two aligned dwords
// around them -- one 8-byte load at a 4-byte aligned address, so that each feature costs the address
// unit one instruction over the window's 16 rows -- realigned with v_alignbyte; GROUP features' loads
// are in flight before they are added.
// Feature slots past the chunk read the zeroed guard.
// BANDED: the features read the level's banded copy (`band`, see band_geom): the same bytes, a window's
// sixteen rows 32 bytes apart instead of W; lane l then holds window column l / 4, rows 4 (l % 4) .. + 3.
template <int GROUP, bool BANDED = false>
__device__ __forceinline__ void accumulate_local(const LevelGeom& g, const uint2* __restrict__ F, int nfeat, int offx,
                                                 int offy, int lane, int first_chunk, int chunk_step, uint32_t& acc_lo,
                                                 uint32_t& acc_hi, const uint8_t* band = nullptr,
                                                 const BandGeom& bg = BandGeom{}) {
    const int lane_off = BANDED ? band_lane_offset(lane) : refine_lane_offset(g, lane);
    const int zero_off = BANDED ? (int)bg.zero_off : g.zero_off();
    const __amdgpu_buffer_rsrc_t rsrc =
        BANDED ? __builtin_amdgcn_make_buffer_rsrc((void*)band, (short)0, (int)bg.bytes, kBufferWord3)
               : __builtin_amdgcn_make_buffer_rsrc((void*)g.lm, (short)0, g.lm_bytes, kBufferWord3);
    acc_lo = 0;
    acc_hi = 0;
    for (int base = first_chunk * 63; base < nfeat; base += 63 * chunk_step) {
        const int cnt = min(63, nfeat - base);
        int my_off = zero_off;
        if (lane < cnt) {
            const uint2 f = F[base + lane];
            const int fx = (int)(int16_t)(f.x & 0xFFFFu) + offx, fy = (int)(int16_t)(f.x >> 16) + offy;
            my_off = BANDED ? band_feature_offset(g, bg, fx, fy, (int)f.y) : feature_offset(g, fx, fy, (int)f.y, zero_off);
        }
        uint32_t acc8 = 0;
#pragma unroll
        for (int grp = 0; grp < 64 / GROUP; ++grp) {
            if (grp * GROUP >= cnt) break;
            i32x2 vw[GROUP];
            int sh[GROUP];
#pragma unroll
            for (int k = 0; k < GROUP; ++k) {
                const int o = __builtin_amdgcn_readlane(my_off, grp * GROUP + k);
                const LaneLoad ld = refine_lane_load(o, lane_off);
                sh[k] = (o + lane_off) & 3;
                vw[k] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, ld.voff, ld.soff, 0);
            }
#pragma unroll
            for (int k = 0; k < GROUP; ++k)
                acc8 += __builtin_amdgcn_alignbyte((uint32_t)vw[k][1], (uint32_t)vw[k][0], (uint32_t)sh[k]);
        }
        acc_lo += acc8 & 0x00FF00FFu;         // the lane's positions 0 and 2
        acc_hi += (acc8 >> 8) & 0x00FF00FFu;  // positions 1 and 3
    }
}

// Wide loads (refine layout 2) on the linear memories: a 16-byte load gives a lane eight positions of a window
// row (misaligned by up to three bytes), so 32 lanes cover a feature's window and one load instruction serves
// two features (the even features of a chunk in lanes 0..31, the odd ones in lanes 32..63): half the
// address-unit cycles per feature of the 8-byte loads.  acc: 16-bit pairs of positions (0, 2), (1, 3), (4, 6),
// (5, 7) of the lane's eight; the two halves are added by the caller.
template <int GROUP>
__device__ __forceinline__ void accumulate_local_wide(const LevelGeom& g, const uint2* __restrict__ F, int nfeat,
                                                      int offx, int offy, int lane, int first_chunk, int chunk_step,
                                                      uint32_t (&acc)[4]) {
    constexpr int P = GROUP / 2;  // feature pairs in flight
    const bool odd = lane >= 32;
    const int lane_off = wide_lane_offset(g, lane);
    const int zero_off = g.zero_off();
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)g.lm, (short)0, g.lm_bytes, kBufferWord3);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] = 0;
    for (int base = first_chunk * 63; base < nfeat; base += 63 * chunk_step) {
        const int cnt = min(63, nfeat - base);
        int my_off = zero_off;  // lane 63 (and every lane past the chunk) reads the guard
        if (lane < cnt) {
            const uint2 f = F[base + lane];
            my_off = feature_offset(g, (int)(int16_t)(f.x & 0xFFFFu) + offx, (int)(int16_t)(f.x >> 16) + offy, (int)f.y,
                                    zero_off);
        }
        uint32_t a8[2] = {0u, 0u};  // <= 32 features of the chunk per lane: byte sums <= 128
#pragma unroll
        for (int grp = 0; grp < 32 / P; ++grp) {
            if (grp * 2 * P >= cnt) break;
            i32x4 vw[P];
            int sh[P];
#pragma unroll
            for (int k = 0; k < P; ++k) {
                const int j = grp * P + k;
                const int o0 = __builtin_amdgcn_readlane(my_off, 2 * j), o1 = __builtin_amdgcn_readlane(my_off, 2 * j + 1);
                const LaneLoad ld = refine_wide_load(odd ? o1 : o0, lane_off);
                sh[k] = ((odd ? o1 : o0) + lane_off) & 3;
                vw[k] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, ld.voff, ld.soff, 0);
            }
#pragma unroll
            for (int k = 0; k < P; ++k) {
                a8[0] += __builtin_amdgcn_alignbyte((uint32_t)vw[k][1], (uint32_t)vw[k][0], (uint32_t)sh[k]);
                a8[1] += __builtin_amdgcn_alignbyte((uint32_t)vw[k][2], (uint32_t)vw[k][1], (uint32_t)sh[k]);
            }
        }
        acc[0] += a8[0] & 0x00FF00FFu;
        acc[1] += (a8[0] >> 8) & 0x00FF00FFu;
        acc[2] += a8[1] & 0x00FF00FFu;
        acc[3] += (a8[1] >> 8) & 0x00FF00FFu;
    }
}

// Column loads (refine layout 3) on the banded copy: a window column's sixteen rows lie inside one 32-byte
// slot, so one unaligned 16-byte load gives a lane a whole column; 16 lanes cover a feature's window and one
// load instruction serves four features (feature g of every four in lanes 16 g .. 16 g + 15).  acc: 16-bit
// pairs of rows (4i, 4i + 2) in acc[2i], (4i + 1, 4i + 3) in acc[2i + 1]; the four groups are added by the
// caller.
template <int GROUP>
__device__ __forceinline__ void accumulate_local_columns(const LevelGeom& g, const uint2* __restrict__ F, int nfeat,
                                                         int offx, int offy, int lane, int first_chunk, int chunk_step,
                                                         uint32_t (&acc)[8], const uint8_t* band, const BandGeom& bg) {
    constexpr int P = GROUP / 2;  // loads in flight, four features each
    const int grp4 = lane >> 4;
    const int lane_off = band_column_offset(lane);
    const int zero_off = (int)bg.zero_off;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)band, (short)0, (int)bg.bytes, kBufferWord3);
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0;
    for (int base = first_chunk * 63; base < nfeat; base += 63 * chunk_step) {
        const int cnt = min(63, nfeat - base);
        int my_off = zero_off;  // lane 63 (and every lane past the chunk) reads the zeroed block
        if (lane < cnt) {
            const uint2 f = F[base + lane];
            my_off = band_feature_offset(g, bg, (int)(int16_t)(f.x & 0xFFFFu) + offx, (int)(int16_t)(f.x >> 16) + offy,
                                         (int)f.y);
        }
        uint32_t a8[4] = {0u, 0u, 0u, 0u};  // <= 16 features of the chunk per lane: byte sums <= 64
#pragma unroll
        for (int grp = 0; grp < 16 / P; ++grp) {
            if (grp * 4 * P >= cnt) break;
            i32x4 vw[P];
#pragma unroll
            for (int k = 0; k < P; ++k) {
                const int j = 4 * (grp * P + k);
                const int o0 = __builtin_amdgcn_readlane(my_off, j), o1 = __builtin_amdgcn_readlane(my_off, j + 1);
                const int o2 = __builtin_amdgcn_readlane(my_off, j + 2), o3 = __builtin_amdgcn_readlane(my_off, j + 3);
                const int o = grp4 == 0 ? o0 : grp4 == 1 ? o1 : grp4 == 2 ? o2 : o3;
                const LaneLoad ld = band_column_load(o, lane_off);
                vw[k] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, ld.voff, ld.soff, 0);
            }
#pragma unroll
            for (int k = 0; k < P; ++k)
#pragma unroll
                for (int i = 0; i < 4; ++i) a8[i] += (uint32_t)vw[k][i];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[2 * i] += a8[i] & 0x00FF00FFu;
            acc[2 * i + 1] += (a8[i] >> 8) & 0x00FF00FFu;
        }
    }
}

// Nibble column loads (refine layout 4) on the nibble-banded copy: a lane loads one window column's whole
// 16-byte slot (32 rows, two per byte), 16 lanes read a feature's window as 256 consecutive bytes and one load
// instruction serves four features (feature g of every four in lanes 16 g .. 16 g + 15), each lane's offset
// fetched from the lane that decoded the feature by one ds_bpermute.  The window's first row r0 (the offset's
// low four bits) picks bits 4 r0 .. 4 r0 + 63 of the slot -- the load starts 4 bytes in when r0 >= 8, so the
// rows start at bit 4 (r0 % 8) of the first dword either way: rows 0..7 of the window in one dword, 8..15 in
// the other, a nibble each.  Nibble sums hold three features (3 * 4 < 16) and are then
// widened into bytes (even rows, odd rows), the bytes into 16-bit pairs at the end of each 63-feature chunk;
// acc[w] half h holds window row p where nib_row_word(p) == w, nib_row_half(p) == h.
template <int GROUP>
__device__ __forceinline__ void accumulate_local_nibcols(const LevelGeom& g, const uint2* __restrict__ F, int nfeat,
                                                         int offx, int offy, int lane, int first_chunk, int chunk_step,
                                                         uint32_t (&acc)[8], const uint8_t* band, const BandGeom& bg) {
    constexpr int P = GROUP / 2;  // loads in flight, four features each
    const int src4 = 4 * (lane >> 4);  // ds_bpermute byte address of lane lane / 16
    const int lane_off = nib_column_offset(lane);
    const int zero_off = (int)bg.zero_off;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)band, (short)0, (int)bg.bytes, kBufferWord3);
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0;
    for (int base = first_chunk * 63; base < nfeat; base += 63 * chunk_step) {
        const int cnt = min(63, nfeat - base);
        int my_off = zero_off;  // lane 63 (and every lane past the chunk) reads the zeroed block
        if (lane < cnt) {
            const uint2 f = F[base + lane];
            my_off = nib_band_feature_offset(g, bg, (int)(int16_t)(f.x & 0xFFFFu) + offx,
                                             (int)(int16_t)(f.x >> 16) + offy, (int)f.y);
        }
        // byte sums of rows (0, 2, 4, 6), (1, 3, 5, 7), (8, 10, 12, 14), (9, 11, 13, 15): <= 16 features of the
        // chunk per lane, <= 64
        uint32_t b8[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int grp = 0; grp < 16 / P; ++grp) {
            if (grp * 4 * P >= cnt) break;
            i32x4 vw[P];
            uint32_t sh[P];
#pragma unroll
            for (int k = 0; k < P; ++k) {
                // feature 4 (grp P + k) + lane / 16 of the chunk, decoded by that lane
                const int o = __builtin_amdgcn_ds_bpermute(src4 + 16 * (grp * P + k), my_off);
                const LaneLoad ld = nib_column_load(o, lane_off);
                sh[k] = nib_column_shift(o);
                vw[k] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, ld.voff, ld.soff, 0);
            }
            uint32_t n4[2] = {0u, 0u};  // nibble sums of at most three features
#pragma unroll
            for (int k = 0; k < P; ++k) {
                const uint32_t a = (uint32_t)vw[k][0], b = (uint32_t)vw[k][1], c = (uint32_t)vw[k][2];
                n4[0] += __builtin_amdgcn_alignbit(b, a, sh[k]);  // rows 0..7 of the window
                n4[1] += __builtin_amdgcn_alignbit(c, b, sh[k]);  // rows 8..15
                if (k % 3 == 2 || k == P - 1) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        b8[2 * h] += n4[h] & 0x0F0F0F0Fu;
                        b8[2 * h + 1] += (n4[h] >> 4) & 0x0F0F0F0Fu;
                        n4[h] = 0u;
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[2 * i] += b8[i] & 0x00FF00FFu;
            acc[2 * i + 1] += (b8[i] >> 8) & 0x00FF00FFu;
        }
    }
}

// LAYOUT: 0 the linear memories, 8-byte loads (lane l: window row l / 4, positions 4 (l % 4) .. + 3); 1 the
// banded copy (see accumulate_local); 2 the linear memories with 16-byte loads (accumulate_local_wide); 3 the
// banded copy with 16-byte column loads (accumulate_local_columns); 4 the nibble-banded copy with 16-byte slot
// loads (accumulate_local_nibcols).
template <int WAVES, int GROUP, int LAYOUT>
__global__ __launch_bounds__(64 * WAVES) void k_refine(kern::RefineArgs a) {
    constexpr bool BANDED = LAYOUT == 1, WIDE = LAYOUT == 2, COLUMNS = LAYOUT == 3, NIBCOLS = LAYOUT == 4;
    constexpr int NACC = COLUMNS || NIBCOLS ? 8 : WIDE ? 4 : 2;
    __shared__ uint32_t s_part[WAVES][kWave][NACC];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = uniform((int)(threadIdx.x / kWave));
    const int count = min(*a.count_in, a.cap_in);
    const int T = a.geom.T;
    const int offs = T / 2 + (T % 2 - 1);
    // Workgroups go to the eight XCDs in turn (block b on XCD b % 8).  Consecutive candidates come from
    // the same coarse tile and read the same level-0 windows, so each XCD takes a contiguous eighth of
    // the list and its L2 keeps those windows.  Inside its eighth a block takes a fixed share, or, with many
    // candidates per block (their feature counts differ up to 8191 / 32, so fixed shares end unevenly), the
    // next candidate from the XCD's counter.
    const int nx = (gridDim.x % 8 == 0) ? 8 : 1;
    const int xcd = blockIdx.x % nx, per = gridDim.x / nx;
    const int seg0 = (int)((long long)count * xcd / nx), seg1 = (int)((long long)count * (xcd + 1) / nx);
    int* const queue = a.queue && nx == 8 && count >= 8 * (int)gridDim.x ? a.queue + xcd : nullptr;
    __shared__ int s_next;
    for (int k = 0;; ++k) {
        int ci;
        if (queue) {
            if (threadIdx.x == 0) s_next = seg0 + atomicAdd(queue, 1);
            __syncthreads();  // read by every thread before the barriers of the body let it change
            ci = uniform(s_next);
        } else {
            ci = seg0 + (int)blockIdx.x / nx + k * per;
        }
        if (ci >= seg1) break;
        const Cand cd = a.cand_in[ci];
        const int gtag = uniform(cd.g);  // template, with the image of a batched launch in the high bits
        const int g = gtag & kCandTemplMask, img = gtag >> kCandImageShift;
        LevelGeom geo = a.geom;
        geo.lm += img * a.geom.img_stride;
        const uint8_t* band = BANDED || COLUMNS || NIBCOLS ? a.band + img * a.band_stride : nullptr;
        const TemplLevelDev t = a.tl[(long long)g * a.levels + a.level];
        const int rmin = a.raw_min[(long long)g * a.levels + a.level];  // loaded beside the record
        const int nfeat = uniform(t.nfeat), foff = uniform(t.off);
        if (a.feat_count && threadIdx.x == 0) atomicAdd(a.feat_count, (unsigned long long)nfeat);
        const RefineWindow win = refine_window(uniform(cd.x), uniform(cd.y), uniform(t.width), uniform(t.height),
                                               a.geom.rows, a.geom.cols, T);  // line2Dup.cpp:1207-1216
        const int x = win.x, y = win.y, offx = win.offx, offy = win.offy;
        uint32_t acc[NACC];
        if constexpr (NIBCOLS) {
            accumulate_local_nibcols<GROUP>(geo, a.feats + foff, nfeat, offx, offy, lane, wave, WAVES, acc, band,
                                            a.bgeom);
        } else if constexpr (COLUMNS) {
            accumulate_local_columns<GROUP>(geo, a.feats + foff, nfeat, offx, offy, lane, wave, WAVES, acc, band,
                                            a.bgeom);
        } else if constexpr (WIDE) {
            accumulate_local_wide<GROUP>(geo, a.feats + foff, nfeat, offx, offy, lane, wave, WAVES, acc);
        } else {
            accumulate_local<GROUP, BANDED>(geo, a.feats + foff, nfeat, offx, offy, lane, wave, WAVES, acc[0], acc[1],
                                            band, a.bgeom);
        }
#pragma unroll
        for (int k = 0; k < NACC; ++k) s_part[wave][lane][k] = acc[k];
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int w = 1; w < WAVES; ++w)
#pragma unroll
                for (int k = 0; k < NACC; ++k) acc[k] += s_part[w][lane][k];
            // first strict maximum in row-major order == max of (raw, -index)
            uint32_t key = 0;
            if constexpr (NIBCOLS) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {  // the four feature groups
                    acc[k] += (uint32_t)__shfl_down((int)acc[k], 32);
                    acc[k] += (uint32_t)__shfl_down((int)acc[k], 16);
                }
                if (lane < 16) {
#pragma unroll
                    for (int p = 0; p < 16; ++p) {
                        const uint32_t w = acc[nib_row_word(p)];
                        const int raw = nib_row_half(p) ? (int)(w >> 16) : (int)(w & 0xFFFFu);
                        key = max(key, ((uint32_t)raw << 8) | (uint32_t)(255 - band_column_window_index(lane, p)));
                    }
                }
            } else if constexpr (COLUMNS) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {  // the four feature groups
                    acc[k] += (uint32_t)__shfl_down((int)acc[k], 32);
                    acc[k] += (uint32_t)__shfl_down((int)acc[k], 16);
                }
                if (lane < 16) {
#pragma unroll
                    for (int p = 0; p < 16; ++p) {
                        const uint32_t w = acc[(p >> 2) * 2 + (p & 1)];
                        const int raw = (p & 2) ? (int)(w >> 16) : (int)(w & 0xFFFFu);
                        key = max(key, ((uint32_t)raw << 8) | (uint32_t)(255 - band_column_window_index(lane, p)));
                    }
                }
            } else if constexpr (WIDE) {
#pragma unroll
                for (int k = 0; k < 4; ++k) acc[k] += (uint32_t)__shfl_down((int)acc[k], 32);  // the odd features
                if (lane < 32) {
#pragma unroll
                    for (int p = 0; p < 8; ++p) {
                        const uint32_t w = acc[(p >> 2) * 2 + (p & 1)];
                        const int raw = (p & 2) ? (int)(w >> 16) : (int)(w & 0xFFFFu);
                        key = max(key, ((uint32_t)raw << 8) | (uint32_t)(255 - wide_window_index(lane, p)));
                    }
                }
            } else {
                const uint32_t lo = acc[0], hi = acc[1];
                const int raws[4] = {(int)(lo & 0xFFFFu), (int)(hi & 0xFFFFu), (int)(lo >> 16), (int)(hi >> 16)};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int idx = BANDED ? band_window_index(lane, k) : (lane >> 2) * 16 + (lane & 3) * 4 + k;
                    key = max(key, ((uint32_t)raws[k] << 8) | (uint32_t)(255 - idx));
                }
            }
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) key = max(key, (uint32_t)__shfl_xor((int)key, m));
            const int best = (int)(key >> 8);
            int br = -1, bc = -1;
            if (best > 0) {
                const int idx = 255 - (int)(key & 255u);
                br = idx >> 4;
                bc = idx & 15;
            }
            if (lane == 0 && best >= rmin) {
                const int slot = atomicAdd(a.count_out, 1);
                if (slot < a.cap_out) {
                    Cand o;
                    o.g = gtag;
                    o.x = (x / T - 8 + bc) * T + offs;  // line2Dup.cpp:1256-1257
                    o.y = (y / T - 8 + br) * T + offs;
                    o.raw = best;
                    a.cand_out[slot] = o;
                }
            }
        }
        __syncthreads();
    }
}

// Banded copy of a level's linear memories (see band_geom): one thread per four columns of one band of one
// plane-phase, 32 rows of four bytes each (one dword load per row when W is a multiple of four, byte loads
// otherwise; past the linear block's zeroed guard it reads zeros), transposed into four 32-byte slots, eight
// 16-byte stores; block 0 also clears the zeroed block after the slots.  Image blockIdx.y of a batched launch:
// linear memories at + z * lm_b, the copy at + z * out_b.
__global__ __launch_bounds__(256) void k_band(const uint8_t* __restrict__ lm, long long lm_bytes, int W, int H,
                                              int rows_all, BandGeom bg, uint8_t* __restrict__ out, long long lm_b,
                                              long long out_b) {
    const int z = blockIdx.y;
    lm += z * lm_b;
    out += z * out_b;
    // the zeroed block that features outside the level read, rewritten on every pass: the buffer may have
    // held a larger level's copy, whose slots lie where this one's block is
    if (blockIdx.x == 0 && threadIdx.x < kBandZeroBytes / 16)
        reinterpret_cast<uint4*>(out + bg.zero_off)[threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
    const int quads = (bg.wp + 3) / 4;  // wp = W + 16 columns; the last quad may hold fewer
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)rows_all * quads) return;
    const int row = (int)(i / quads), q = (int)(i - (long long)row * quads);  // row = pp * nb + b
    const int pp = row / bg.nb, b = row - pp * bg.nb;
    uint32_t d[kBandSlot];
    const long long src0 = band_source(W, H, pp, b, 4 * q, 0);
    if ((W & 3) == 0) {
#pragma unroll
        for (int s = 0; s < kBandSlot; ++s) {
            const long long src = src0 + (long long)s * W;  // a multiple of four, as lm_bytes is
            d[s] = src < lm_bytes ? *reinterpret_cast<const uint32_t*>(lm + src) : 0u;
        }
    } else {
#pragma unroll
        for (int s = 0; s < kBandSlot; ++s) {
            uint32_t v = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long src = src0 + (long long)s * W + j;
                v |= (src < lm_bytes ? (uint32_t)lm[src] : 0u) << (8 * j);
            }
            d[s] = v;
        }
    }
    uint4* o = reinterpret_cast<uint4*>(out + ((long long)row * bg.wp + 4 * q) * kBandSlot);
#pragma unroll
    for (int j = 0; j < 4; ++j) {  // column 4 q + j: byte j of every row
        if (4 * q + j >= bg.wp) break;
        uint32_t w[kBandSlot / 4];
#pragma unroll
        for (int k = 0; k < kBandSlot / 4; ++k)
            w[k] = ((d[4 * k] >> (8 * j)) & 0xFFu) | (((d[4 * k + 1] >> (8 * j)) & 0xFFu) << 8) |
                   (((d[4 * k + 2] >> (8 * j)) & 0xFFu) << 16) | (((d[4 * k + 3] >> (8 * j)) & 0xFFu) << 24);
        o[2 * j] = make_uint4(w[0], w[1], w[2], w[3]);
        o[2 * j + 1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
}

// Nibble-banded copy of a level's linear memories (refine layout 4, see nib_band_geom): one thread per four
// columns of one band of one plane-phase, 32 rows of four bytes each (as k_band), rows 2k and 2k + 1 packed into
// one byte per column (responses are below 16), four 16-byte stores; block 0 also clears the zeroed block.  (Reading
// each linear row once, with 8-byte half-slot stores, or one thread per slot with byte loads, measured no faster.)
// Image blockIdx.y of a batched launch: linear memories at + z * lm_b, the copy at + z * out_b.
__global__ __launch_bounds__(256) void k_band_nib(const uint8_t* __restrict__ lm, long long lm_bytes, int W, int H,
                                                  int rows_all, BandGeom bg, uint8_t* __restrict__ out, long long lm_b,
                                                  long long out_b) {
    const int z = blockIdx.y;
    lm += z * lm_b;
    out += z * out_b;
    if (blockIdx.x == 0 && threadIdx.x < kBandZeroBytes / 16)
        reinterpret_cast<uint4*>(out + bg.zero_off)[threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
    const int quads = (bg.wp + 3) / 4;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)rows_all * quads) return;
    const int row = (int)(i / quads), q = (int)(i - (long long)row * quads);  // row = pp * nb + b
    const int pp = row / bg.nb, b = row - pp * bg.nb;
    uint32_t e[kNibSlot];  // byte j of e[k]: rows 2k (low nibble) and 2k + 1 (high) of column 4 q + j
    const long long src0 = band_source(W, H, pp, b, 4 * q, 0);
    if ((W & 3) == 0) {
#pragma unroll
        for (int k = 0; k < kNibSlot; ++k) {
            const long long s0 = src0 + (long long)(2 * k) * W, s1 = s0 + W;  // multiples of four, as lm_bytes is
            const uint32_t d0 = s0 < lm_bytes ? *reinterpret_cast<const uint32_t*>(lm + s0) : 0u;
            const uint32_t d1 = s1 < lm_bytes ? *reinterpret_cast<const uint32_t*>(lm + s1) : 0u;
            e[k] = d0 | (d1 << 4);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kNibSlot; ++k) {
            uint32_t v = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long s0 = src0 + (long long)(2 * k) * W + j, s1 = s0 + W;
                v |= ((s0 < lm_bytes ? (uint32_t)lm[s0] : 0u) | ((s1 < lm_bytes ? (uint32_t)lm[s1] : 0u) << 4)) << (8 * j);
            }
            e[k] = v;
        }
    }
    uint4* o = reinterpret_cast<uint4*>(out + ((long long)row * bg.wp + 4 * q) * kNibSlot);
#pragma unroll
    for (int j = 0; j < 4; ++j) {  // column 4 q + j: byte j of every e
        if (4 * q + j >= bg.wp) break;
        uint32_t w[4];
#pragma unroll
        for (int m = 0; m < 4; ++m)
            w[m] = ((e[4 * m] >> (8 * j)) & 0xFFu) | (((e[4 * m + 1] >> (8 * j)) & 0xFFu) << 8) |
                   (((e[4 * m + 2] >> (8 * j)) & 0xFFu) << 16) | (((e[4 * m + 3] >> (8 * j)) & 0xFFu) << 24);
        o[j] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

__global__ void k_simlocal(LevelGeom g, const uint2* feats, int nfeat, int off, int cx, int cy, uint16_t* out) {
    const int lane = threadIdx.x;
    const int T = g.T;
    uint32_t lo, hi;
    accumulate_local<16>(g, feats + off, nfeat, (cx / T - 8) * T, (cy / T - 8) * T, lane, 0, 1, lo, hi);
    const int base = (lane >> 2) * 16 + (lane & 3) * 4;
    out[base + 0] = (uint16_t)(lo & 0xFFFFu);
    out[base + 1] = (uint16_t)(hi & 0xFFFFu);
    out[base + 2] = (uint16_t)(lo >> 16);
    out[base + 3] = (uint16_t)(hi >> 16);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Mismatches of div_sobel against the IEEE quotient num / (den + DBL_EPSILON) over 0 <= num <= den <= 1020.
__global__ void k_div_check(int* mismatches) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 1021 * 1021) return;
    const int num = i / 1021, den = i - num * 1021;
    if (num > den) return;
    const float eps = (float)2.220446049250313080847e-16;
    const float a = div_sobel((float)num, (float)den), b = __fdiv_rn((float)num, __fadd_rn((float)den, eps));
    if (__float_as_uint(a) != __float_as_uint(b)) atomicAdd(mismatches, 1);
}

__global__ void k_bin_table(int lo, int hi, uint8_t* out) {
    const int n = hi - lo + 1;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * n) return;
    const int dy = (int)(i / n) + lo, dx = (int)(i % n) + lo;
    out[i] = (uint8_t)(angle_bin(fast_atan2_deg((float)dy, (float)dx)) & 7);
}

}  // namespace

namespace kern {

hipError_t gradient_quantize(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask,
                             float weak, uint8_t* q, float* mag, float* angle, hipStream_t stream, uint32_t* zero_ptr,
                             int zero_n, int bins, int nimg, const BatchStrides& bs, uint8_t* down,
                             uint8_t* down_mask) {
    if (zero_n > 256) return hipErrorInvalidValue;
    if ((down_mask && !mask) || (down && (rows % 2 || cols % 2 || rows < 4 || cols < 4))) return hipErrorInvalidValue;
    if (nimg < 1 || nimg > kMaxBatch || (mag && nimg > 1)) return hipErrorInvalidValue;
    if (channels != 1 && channels != 3) return hipErrorInvalidValue;
    if (bins != 8 && (bins != 16 || channels != 3)) return hipErrorInvalidValue;  // the 16-bin variant: colour only
    const float thr = weak * weak;  // hysteresisGradient(..., threshold * threshold), line2Dup.cpp:301
    const bool floats = mag != nullptr;
    // 16-row tiles when the launch's images give the chip at least four blocks per CU, else 8-row tiles
    const long long tiles16 = (long long)((cols + GQ_TW - 1) / GQ_TW) * ((rows + 15) / 16) * nimg;
    const bool tall = tiles16 >= 1024;
    const dim3 grid((cols + GQ_TW - 1) / GQ_TW, tall ? (rows + 15) / 16 : (rows + 7) / 8, nimg);
#define SBM_GRAD(CH_, FL_, TH_, BINS_) \
    hipLaunchKernelGGL((k_gradient<CH_, FL_, TH_, BINS_>), grid, dim3(256), 0, stream, img, rows, cols, stride, mask, \
                       thr, q, mag, angle, zero_ptr, zero_n, down, down_mask, bs)
    if (bins == 16) {
        if (tall) { if (floats) SBM_GRAD(3, true, 16, 16); else SBM_GRAD(3, false, 16, 16); }
        else { if (floats) SBM_GRAD(3, true, 8, 16); else SBM_GRAD(3, false, 8, 16); }
    } else if (channels == 1) {
        if (tall) { if (floats) SBM_GRAD(1, true, 16, 8); else SBM_GRAD(1, false, 16, 8); }
        else { if (floats) SBM_GRAD(1, true, 8, 8); else SBM_GRAD(1, false, 8, 8); }
    } else {
        if (tall) { if (floats) SBM_GRAD(3, true, 16, 8); else SBM_GRAD(3, false, 16, 8); }
        else { if (floats) SBM_GRAD(3, true, 8, 8); else SBM_GRAD(3, false, 8, 8); }
    }
#undef SBM_GRAD
    return hipGetLastError();
}

hipError_t pyr_down(const uint8_t* src, int rows, int cols, int stride, int channels, uint8_t* dst, const uint8_t* mask,
                    uint8_t* mask_dst, hipStream_t stream, int nimg, const BatchStrides& bs) {
    if (nimg < 1 || nimg > kMaxBatch) return hipErrorInvalidValue;
    const dim3 grid((cols / 2 + 63) / 64, (rows / 2 + 3) / 4, nimg);
    // aligned dword rows for every image of the launch (a source table's rows are the caller's)
    bool aligned = ((uintptr_t)src % 4 == 0) && stride % 4 == 0 && bs.in_b % 4 == 0;
    if (bs.src) aligned = bs.src_aligned;
    const bool words = aligned && ((uintptr_t)dst % 4 == 0) && bs.out_b % 4 == 0 && (cols / 2) % 4 == 0;
    if (channels == 1 && words) {
        // four outputs per thread when that still gives the chip four wavefronts per SIMD, else two
        const int out = (long long)(cols / 2) * (rows / 2) * nimg / 4 >= 4096LL * 64 ? 4 : 2;
        const dim3 gw((cols / 2 / out + 63) / 64, (rows / 2 + 3) / 4, nimg);
        if (out == 4)
            hipLaunchKernelGGL(k_pyr_down_w<4>, gw, dim3(256), 0, stream, src, rows, cols, stride, dst, mask, mask_dst, bs);
        else
            hipLaunchKernelGGL(k_pyr_down_w<2>, gw, dim3(256), 0, stream, src, rows, cols, stride, dst, mask, mask_dst, bs);
    } else if (channels == 1)
        hipLaunchKernelGGL((k_pyr_down<1>), grid, dim3(256), 0, stream, src, rows, cols, stride, dst, mask, mask_dst, bs);
    else if (channels == 3)
        hipLaunchKernelGGL((k_pyr_down<3>), grid, dim3(256), 0, stream, src, rows, cols, stride, dst, mask, mask_dst, bs);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t pack_nibbles(const uint8_t* lm, long long lm_bytes, uint8_t* nib, hipStream_t stream, int nimg,
                        const BatchStrides& bs) {
    const long long n16 = lm_bytes / 32;
    if (n16 <= 0) return hipSuccess;
    if (nimg < 1 || nimg > kMaxBatch || bs.in_b % 16 || bs.out_b % 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pack_nibbles, dim3((unsigned)((n16 + 255) / 256), 1, nimg), dim3(256), 0, stream,
                       reinterpret_cast<const uint4*>(lm), reinterpret_cast<uint4*>(nib), n16, bs);
    return hipGetLastError();
}

hipError_t spread_linearize(const uint8_t* q, int rows, int cols, int T, uint8_t* lm, hipStream_t stream,
                            uint8_t* nib, bool* nib_written, int nimg, const BatchStrides& bs) {
    if (T < 1 || T > 16 || nimg < 1 || nimg > kMaxBatch) return hipErrorInvalidValue;
    const int W = cols / T, H = rows / T;
    if (nib_written) *nib_written = false;
    if ((T == 4 || T == 8 || T == 16) && cols % 4 == 0 && W % 4 == 0 && W > 0 && H > 0) {
        const long long lmb = 8LL * T * T * W * H + lm_guard_bytes(W, H);
        const int guard16 = (int)(lm_guard_bytes(W, H) / 16);
        const int nib_guard16 = (int)((lm_nibble_bytes(lmb) - 4LL * T * T * W * H) / 16);
        const dim3 grid((W + SB_BX - 1) / SB_BX, H + 1, nimg);
        switch (T) {
            case 4: hipLaunchKernelGGL(k_spread_band<4>, grid, dim3(256), 0, stream, q, rows, cols, W, H, lm, nib, guard16, nib_guard16, bs); break;
            case 8: hipLaunchKernelGGL(k_spread_band<8>, grid, dim3(256), 0, stream, q, rows, cols, W, H, lm, nib, guard16, nib_guard16, bs); break;
            default: hipLaunchKernelGGL(k_spread_band<16>, grid, dim3(256), 0, stream, q, rows, cols, W, H, lm, nib, guard16, nib_guard16, bs); break;
        }
        if (nib_written) *nib_written = nib != nullptr;
        return hipGetLastError();
    }
    const size_t shmem = (size_t)SL_BX * T + T;
    const dim3 grid((W + SL_BX - 1) / SL_BX, rows + 1, nimg);
    const int guard16 = (int)(lm_guard_bytes(W, H) / 16);
    hipLaunchKernelGGL(k_spread_lm, grid, dim3(256), shmem, stream, q, rows, cols, T, lm, W * H, guard16, bs);
    return hipGetLastError();
}

hipError_t spread_linearize16(const uint16_t* q, int rows, int cols, int T, uint8_t* lm, hipStream_t stream, int nimg,
                              const BatchStrides& bs) {
    if (T < 1 || T > 16 || nimg < 1 || nimg > kMaxBatch) return hipErrorInvalidValue;
    const int W = cols / T, H = rows / T;
    if (W <= 0 || H <= 0) return hipErrorInvalidValue;
    const int guard16 = (int)(lm_guard_bytes(W, H) / 16);
    const dim3 grid((W + S16_BX - 1) / S16_BX, H + 1, nimg);
    const size_t shmem = (size_t)T * (S16_BX * T + T - 1) * sizeof(uint16_t);
    hipLaunchKernelGGL(k_spread16, grid, dim3(256), shmem, stream, q, rows, cols, T, W, H, lm, guard16, bs);
    return hipGetLastError();
}

int coarse_tiles(long long hw, int ppl, int gather) {
    const long long tp = coarse_tile_positions(ppl, gather);
    return (int)((hw + tp - 1) / tp);
}

template <int PPL, int GATHER>
hipError_t launch_coarse(const CoarseArgs& a, dim3 grid, int wpb, int split, hipStream_t stream) {
    switch (split) {
        case 1: hipLaunchKernelGGL((k_coarse<PPL, GATHER, 1>), grid, dim3(64 * wpb), 0, stream, a); break;
        case 2: hipLaunchKernelGGL((k_coarse<PPL, GATHER, 2>), grid, dim3(128), 0, stream, a); break;
        case 4: hipLaunchKernelGGL((k_coarse<PPL, GATHER, 4>), grid, dim3(256), 0, stream, a); break;
        case 8: hipLaunchKernelGGL((k_coarse<PPL, GATHER, 8>), grid, dim3(512), 0, stream, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t coarse(const CoarseArgs& a, int ppl, int wpb, int persistent_blocks, int gather, int split,
                  hipStream_t stream) {
    if (a.nimg < 1 || a.nimg > kMaxBatch) return hipErrorInvalidValue;
    const long long total = (long long)a.ntempl * a.ntiles * a.nimg;
    if (total <= 0) return hipSuccess;
    if (gather < 0 || gather > 3 || (gather == 3) != (ppl == 32)) return hipErrorInvalidValue;
    if (gather == 3 && a.geom.lm_nib == nullptr) return hipErrorInvalidValue;
    if (a.ntiles != coarse_tiles(a.geom.hw, ppl, gather)) return hipErrorInvalidValue;
    if (split != 1 && split != 2 && split != 4 && split != 8) return hipErrorInvalidValue;
    if (wpb < 1 || wpb > 4) return hipErrorInvalidValue;  // (split > 1 launches `split` wavefronts per block)
    // one item per wave (split 1) or per block (split > 1) when not persistent
    const int per_block = split == 1 ? wpb : 1;
    long long blocks = persistent_blocks > 0 ? persistent_blocks : (total + per_block - 1) / per_block;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)blocks);
    switch (ppl * 4 + gather) {
        case 16: return launch_coarse<4, 0>(a, grid, wpb, split, stream);
        case 17: return launch_coarse<4, 1>(a, grid, wpb, split, stream);
        case 18: return launch_coarse<4, 2>(a, grid, wpb, split, stream);
        case 32: return launch_coarse<8, 0>(a, grid, wpb, split, stream);
        case 33: return launch_coarse<8, 1>(a, grid, wpb, split, stream);
        case 34: return launch_coarse<8, 2>(a, grid, wpb, split, stream);
        case 64: return launch_coarse<16, 0>(a, grid, wpb, split, stream);
        case 65: return launch_coarse<16, 1>(a, grid, wpb, split, stream);
        case 66: return launch_coarse<16, 2>(a, grid, wpb, split, stream);
        case 131: return launch_coarse<32, 3>(a, grid, wpb, split, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t refine(const RefineArgs& a, int blocks, hipStream_t stream, int waves, int group) {
    // `waves` wavefronts share a candidate's features; eight XCDs each take a contiguous share of the list
    const int layout = a.layout;
    if (layout < 0 || layout > 4 || ((layout == 1 || layout == 3 || layout == 4) && !a.band)) return hipErrorInvalidValue;
#define SBM_REFINE(W_, G_)                                                                                         \
    do {                                                                                                           \
        if (layout == 1) hipLaunchKernelGGL((k_refine<W_, G_, 1>), dim3(blocks), dim3(64 * W_), 0, stream, a);      \
        else if (layout == 2) hipLaunchKernelGGL((k_refine<W_, G_, 2>), dim3(blocks), dim3(64 * W_), 0, stream, a); \
        else if (layout == 3) hipLaunchKernelGGL((k_refine<W_, G_, 3>), dim3(blocks), dim3(64 * W_), 0, stream, a); \
        else if (layout == 4) hipLaunchKernelGGL((k_refine<W_, G_, 4>), dim3(blocks), dim3(64 * W_), 0, stream, a); \
        else hipLaunchKernelGGL((k_refine<W_, G_, 0>), dim3(blocks), dim3(64 * W_), 0, stream, a);                  \
    } while (0)
    if (waves == 8 && group == 16) SBM_REFINE(8, 16);
    else if (waves == 8 && group == 8) SBM_REFINE(8, 8);
    else if (waves == 4 && group == 16) SBM_REFINE(4, 16);
    else if (waves == 4 && group == 8) SBM_REFINE(4, 8);
    else if (waves == 16 && group == 16) SBM_REFINE(16, 16);
    else if (waves == 16 && group == 8) SBM_REFINE(16, 8);
    else if (waves == 2 && group == 16) SBM_REFINE(2, 16);
    else if (waves == 2 && group == 8) SBM_REFINE(2, 8);
    else if (waves == 1 && group == 16) SBM_REFINE(1, 16);
    else if (waves == 1 && group == 8) SBM_REFINE(1, 8);
    else return hipErrorInvalidValue;
#undef SBM_REFINE
    return hipGetLastError();
}

hipError_t band_copy(const uint8_t* lm, long long lm_bytes, int planes, int T, int W, int H, uint8_t* out,
                     hipStream_t stream, int nimg, long long lm_b, long long out_b) {
    if (W <= 0 || H <= 0 || nimg < 1 || nimg > kMaxBatch) return hipErrorInvalidValue;
    const BandGeom bg = band_geom(planes, T, W, H);
    if (bg.bytes >= (1LL << 31) || (nimg > 1 && (out_b < bg.bytes || lm_b < lm_bytes))) return hipErrorInvalidValue;
    const int rows_all = planes * T * T * bg.nb;
    const long long threads = (long long)rows_all * ((bg.wp + 3) / 4);
    hipLaunchKernelGGL(k_band, dim3((unsigned)((threads + 255) / 256), nimg), dim3(256), 0, stream, lm, lm_bytes, W, H,
                       rows_all, bg, out, lm_b, out_b);
    return hipGetLastError();
}

hipError_t nib_band_copy(const uint8_t* lm, long long lm_bytes, int planes, int T, int W, int H, uint8_t* out,
                         hipStream_t stream, int nimg, long long lm_b, long long out_b) {
    if (W <= 0 || H <= 0 || nimg < 1 || nimg > kMaxBatch) return hipErrorInvalidValue;
    const BandGeom bg = nib_band_geom(planes, T, W, H);
    if (bg.bytes >= (1LL << 31) || (nimg > 1 && (out_b < bg.bytes || lm_b < lm_bytes))) return hipErrorInvalidValue;
    const int rows_all = planes * T * T * bg.nb;
    const long long threads = (long long)rows_all * ((bg.wp + 3) / 4);
    hipLaunchKernelGGL(k_band_nib, dim3((unsigned)((threads + 255) / 256), nimg), dim3(256), 0, stream, lm, lm_bytes,
                       W, H, rows_all, bg, out, lm_b, out_b);
    return hipGetLastError();
}

hipError_t fill_u32(uint32_t* p, uint32_t v, long long n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, v, n);
    return hipGetLastError();
}

hipError_t similarity_map(const LevelGeom& g, const TemplLevelDev*, const uint2* feats, int width, int height,
                          int nfeat, int off, uint16_t* out, hipStream_t stream) {
    const long long tiles = (g.hw + 64 * 8 - 1) / (64 * 8);
    hipLaunchKernelGGL((k_simmap<8>), dim3((unsigned)tiles), dim3(64), 0, stream, g, feats, width, height, nfeat, off, out);
    return hipGetLastError();
}

hipError_t similarity_local(const LevelGeom& g, const uint2* feats, int nfeat, int off, int cx, int cy, uint16_t* out,
                            hipStream_t stream) {
    hipLaunchKernelGGL(k_simlocal, dim3(1), dim3(64), 0, stream, g, feats, nfeat, off, cx, cy, out);
    return hipGetLastError();
}

hipError_t copy_probe(const uint4* src, uint4* dst, long long n16, hipStream_t stream) {
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, stream, src, dst, n16);
    return hipGetLastError();
}

hipError_t bin_table(int lo, int hi, uint8_t* out, hipStream_t stream) {
    if (lo < -1020 || hi > 1020 || hi < lo) return hipErrorInvalidValue;  // the Sobel range of 8-bit images
    const long long n = (long long)(hi - lo + 1) * (hi - lo + 1);
    hipLaunchKernelGGL(k_bin_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, lo, hi, out);
    return hipGetLastError();
}

hipError_t division_check(int* mismatches, hipStream_t stream) {
    hipLaunchKernelGGL(k_div_check, dim3((1021 * 1021 + 255) / 256), dim3(256), 0, stream, mismatches);
    return hipGetLastError();
}

}  // namespace kern
}  // namespace sbm
