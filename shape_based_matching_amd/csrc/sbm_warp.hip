// shapeInfo_producer::transform on the GPU (line2Dup.h:245-253): cv::warpAffine with INTER_LINEAR and
// a constant zero border, output the size of the input, for a batch of affine maps of one source.
//
// The arithmetic is OpenCV's fixed-point warp, operation for operation, so results are bit-exact:
//   inverse map per destination row y and column x, in units of 1/1024 pixel:
//     X0 = rint((iM[1]*y + iM[2]) * 1024) + 16,   adelta[x] = rint(iM[0]*x*1024)
//     X  = (X0 + adelta[x]) >> 5                  (5 sub-pixel bits)
//     sx = X >> 5, fx = X & 31                    (Y likewise with iM[3..5])
//   weights (32-fy)(32-fx)*32, (32-fy)fx*32, fy(32-fx)*32, fy*fx*32 (sum 2^15), taps outside the
//   source read 0, result (sum + 2^14) >> 15.
// The doubles are IEEE operations in the same order as the host formula (no contraction), and rint
// rounds half to even like the host's lrint.
//
// One thread per destination pixel; grid.z runs over the maps.  The source is small and re-read
// through L2; destinations are written once, coalesced.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sbm_internal.h"

#pragma clang fp contract(off)

namespace sbm {
namespace {

constexpr int WX = 64, WY = 4;

__device__ __forceinline__ int round_sat_int(double v) {
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (-2147483647 - 1);
    return (int)__builtin_rint(v);
}
__device__ __forceinline__ int sat_short(int v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : v); }

template <int CH>
__global__ void __launch_bounds__(WX * WY) k_warp_affine(const uint8_t* __restrict__ src, int rows, int cols, int stride,
                                                         const double* __restrict__ inv, uint8_t* __restrict__ dst,
                                                         long long dst_image_bytes, int binary_mask) {
    const int x = blockIdx.x * WX + threadIdx.x;
    const int y = blockIdx.y * WY + threadIdx.y;
    if (x >= cols || y >= rows) return;
    const double* M = inv + 6 * (size_t)blockIdx.z;
    const double m0 = M[0], m1 = M[1], m2 = M[2], m3 = M[3], m4 = M[4], m5 = M[5];
    const int X0 = round_sat_int((m1 * (double)y + m2) * 1024.0) + 16;
    const int Y0 = round_sat_int((m4 * (double)y + m5) * 1024.0) + 16;
    const int ad = round_sat_int(m0 * (double)x * 1024.0);
    const int bd = round_sat_int(m3 * (double)x * 1024.0);
    const int X = (X0 + ad) >> 5, Y = (Y0 + bd) >> 5;
    const int sx = sat_short(X >> 5), sy = sat_short(Y >> 5);
    const int fx = X & 31, fy = Y & 31;
    const int w0 = (32 - fy) * (32 - fx) * 32, w1 = (32 - fy) * fx * 32, w2 = fy * (32 - fx) * 32, w3 = fy * fx * 32;
    const bool in_x0 = sx >= 0 && sx < cols, in_x1 = sx + 1 >= 0 && sx + 1 < cols;
    const bool in_y0 = sy >= 0 && sy < rows, in_y1 = sy + 1 >= 0 && sy + 1 < rows;
    const uint8_t* r0 = src + (size_t)(in_y0 ? sy : 0) * stride;
    const uint8_t* r1 = src + (size_t)(in_y1 ? sy + 1 : 0) * stride;
    uint8_t* out = dst + (size_t)blockIdx.z * dst_image_bytes + ((size_t)y * cols + x) * CH;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
        const int v0 = (in_y0 && in_x0) ? r0[sx * CH + k] : 0;
        const int v1 = (in_y0 && in_x1) ? r0[(sx + 1) * CH + k] : 0;
        const int v2 = (in_y1 && in_x0) ? r1[sx * CH + k] : 0;
        const int v3 = (in_y1 && in_x1) ? r1[(sx + 1) * CH + k] : 0;
        int o = (v0 * w0 + v1 * w1 + v2 * w2 + v3 * w3 + (1 << 14)) >> 15;
        o = o > 255 ? 255 : o;
        out[k] = binary_mask ? (o > 0 ? 255 : 0) : (uint8_t)o;
    }
}

}  // namespace

namespace kern {
hipError_t warp_affine(const uint8_t* src, int rows, int cols, int stride, int channels, const double* inv, int n,
                       uint8_t* dst, bool binary_mask, hipStream_t stream) {
    if (n <= 0 || rows <= 0 || cols <= 0) return hipSuccess;
    if (n > 65535) return hipErrorInvalidValue;
    const dim3 grid((cols + WX - 1) / WX, (rows + WY - 1) / WY, n);
    const long long img_bytes = (long long)rows * cols * channels;
    if (channels == 1)
        hipLaunchKernelGGL((k_warp_affine<1>), grid, dim3(WX, WY), 0, stream, src, rows, cols, stride, inv, dst, img_bytes,
                           binary_mask ? 1 : 0);
    else if (channels == 3)
        hipLaunchKernelGGL((k_warp_affine<3>), grid, dim3(WX, WY), 0, stream, src, rows, cols, stride, inv, dst, img_bytes,
                           binary_mask ? 1 : 0);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}
}  // namespace kern

}  // namespace sbm
