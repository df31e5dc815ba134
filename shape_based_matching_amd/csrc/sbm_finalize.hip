// The end of Detector::match and of the demos' NMS step on the GPU, in one block:
//   1. scores of the refined candidates, (raw * 100.f) / (4 * n)               line2Dup.cpp:1177, 1260
//   2. canonical order (similarity desc, template_id, class, y, x) and std::unique with
//      Match::operator== (x, y, similarity, class)                            line2Dup.cpp:1113-1116,
//                                                                              line2Dup.h:117-129
//   3. cv_dnn::NMSBoxes(boxes, scores, score_threshold, nms_threshold, eta = 1, top_k) with
//      box = (x, y, templ[0].width, templ[0].height)                           test.cpp:31-121, 435-451
// The match list is already sorted by score, so NMS's stable sort by score is the identity; the
// suppression matrix (bit j of row i: overlap(i, j) > threshold, j > i) is built by all threads and
// scanned greedily by one wavefront.  Lists longer than kFinalMax are left to the host (flag).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sbm_internal.h"

#pragma clang fp contract(off)

namespace sbm {
namespace {

constexpr int kThreads = 1024;

__device__ __forceinline__ float score_dev(int raw, int n) {  // same operations as the host's score_of
    const float num = (float)raw * 100.f;
    return num / (float)(4 * n);
}

// rectOverlap over jaccardDistance__ (test.cpp:88-105): 1 - (float)(1 - |a&b| / (|a| + |b| - |a&b|))
__device__ __forceinline__ float rect_overlap_dev(int ax, int ay, int aw, int ah, int bx, int by, int bw, int bh) {
    const int aa = aw * ah, ab = bw * bh;
    double dist;
    if (aa + ab <= 0) {
        dist = 0.0;
    } else {
        const int x1 = max(ax, bx), y1 = max(ay, by);
        const int x2 = min(ax + aw, bx + bw), y2 = min(ay + ah, by + bh);
        const int w = x2 - x1, h = y2 - y1;
        const double aab = (double)((w <= 0 || h <= 0) ? 0 : w * h);
        dist = 1.0 - aab / ((double)(aa + ab) - aab);
    }
    return 1.f - (float)dist;
}

struct Lists {
    uint32_t sim[kFinalMax];  // float bits; scores are >= 0, so bit order is value order
    int tid[kFinalMax], cls[kFinalMax], y[kFinalMax], x[kFinalMax], w[kFinalMax], h[kFinalMax];
    int idx[kFinalMax];       // sort permutation
    int flag[kFinalMax];      // kept by std::unique
    int list[kFinalMax];      // entry of each match-list position
    int wave_sum[kThreads / 64];
    int total;
};

__device__ __forceinline__ bool canon_less(const Lists& s, int a, int b) {  // invalid entries (tid < 0) last
    const bool va = s.tid[a] >= 0, vb = s.tid[b] >= 0;
    if (va != vb) return va;
    if (s.sim[a] != s.sim[b]) return s.sim[a] > s.sim[b];
    if (s.tid[a] != s.tid[b]) return s.tid[a] < s.tid[b];
    if (s.cls[a] != s.cls[b]) return s.cls[a] < s.cls[b];
    if (s.y[a] != s.y[b]) return s.y[a] < s.y[b];
    return s.x[a] < s.x[b];
}

// Exclusive prefix sum of s.flag[0..n) into s.idx-compatible positions (returned per element in pos).
__device__ void block_scan(Lists& s, int n, int* pos) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    constexpr int per = kFinalMax / kThreads;  // elements per thread
    int v[per], sum = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const int i = t * per + k;
        v[k] = i < n ? s.flag[i] : 0;
        sum += v[k];
    }
    int incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s.wave_sum[wv] = incl;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int k = 0; k < kThreads / 64; ++k) {
            const int c = s.wave_sum[k];
            s.wave_sum[k] = run;
            run += c;
        }
        s.total = run;
    }
    __syncthreads();
    int base = s.wave_sum[wv] + incl - sum;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const int i = t * per + k;
        if (i < n) pos[i] = base;
        base += v[k];
    }
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_finalize(kern::FinalArgs a) {
    __shared__ Lists s;
    __shared__ int s_pos[kFinalMax];
    const int t = threadIdx.x;
    const int K = min(*a.count_in, a.cap_in);
    if (K > kFinalMax) {
        if (t == 0) a.counters[kern::kFinalFlagSlot] = 1;  // the host finishes this one
        return;
    }
    int P2 = 1;
    while (P2 < K) P2 <<= 1;
    for (int i = t; i < P2; i += kThreads) {
        if (i < K) {
            const Cand cd = a.cand[i];
            const TemplLevelDev t0 = a.tl[(long long)cd.g * a.levels];  // level 0: all features, box size
            const float sim = (a.levels == 1 || cd.raw > 0) ? score_dev(cd.raw, t0.nfeat) : 0.f;
            s.sim[i] = __float_as_uint(sim);
            s.tid[i] = a.gtid[cd.g];
            s.cls[i] = a.gclass[cd.g];
            s.x[i] = cd.x;
            s.y[i] = cd.y;
            s.w[i] = t0.width;
            s.h[i] = t0.height;
        } else {
            s.sim[i] = 0u;
            s.tid[i] = -1;
            s.cls[i] = s.x[i] = s.y[i] = s.w[i] = s.h[i] = 0;
        }
        s.idx[i] = i;
    }
    __syncthreads();
    // bitonic sort of the permutation
    for (int k = 2; k <= P2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < P2; i += kThreads) {
                const int l = i ^ j;
                if (l > i) {
                    const int a0 = s.idx[i], a1 = s.idx[l];
                    const bool up = (i & k) == 0;
                    const bool swap = up ? canon_less(s, a1, a0) : canon_less(s, a0, a1);
                    if (swap) {
                        s.idx[i] = a1;
                        s.idx[l] = a0;
                    }
                }
            }
            __syncthreads();
        }
    }
    // std::unique on (x, y, similarity, class) of neighbours in that order
    for (int i = t; i < K; i += kThreads) {
        const int c = s.idx[i];
        bool keep = true;
        if (i > 0) {
            const int p = s.idx[i - 1];
            keep = !(s.x[c] == s.x[p] && s.y[c] == s.y[p] && s.sim[c] == s.sim[p] && s.cls[c] == s.cls[p]);
        }
        s.flag[i] = keep ? 1 : 0;
    }
    __syncthreads();
    block_scan(s, K, s_pos);
    const int M = s.total;
    // the match list in canonical order: x, y, similarity bits, class, template id
    for (int i = t; i < K; i += kThreads) {
        if (!s.flag[i]) continue;
        const int c = s.idx[i], o = s_pos[i];
        int* r = a.out + 5 * o;
        r[0] = s.x[c];
        r[1] = s.y[c];
        r[2] = (int)s.sim[c];
        r[3] = s.cls[c];
        r[4] = s.tid[c];
    }
    __syncthreads();
    // the entries of the match list, in list order
    for (int i = t; i < K; i += kThreads)
        if (s.flag[i]) s.list[s_pos[i]] = s.idx[i];
    __syncthreads();
    // NMS candidates: scores above the threshold form a prefix of the list; top_k truncates it
    int Kn = 0;
    {
        const float thr = a.score_threshold;
        if (t == 0) {
            int n = 0;
            while (n < M && __uint_as_float(s.sim[s.list[n]]) > thr) ++n;
            if (a.top_k > 0 && a.top_k < n) n = a.top_k;
            s.wave_sum[0] = n;
        }
        __syncthreads();
        Kn = s.wave_sum[0];
    }
    const int W64 = (Kn + 63) >> 6;
    for (int e = t; e < Kn * W64; e += kThreads) {
        const int i = e / W64, wd = e - i * W64;
        const int ci = s.list[i];
        uint64_t bits = 0;
        for (int b = 0; b < 64; ++b) {
            const int j = wd * 64 + b;
            if (j <= i || j >= Kn) continue;
            const int cj = s.list[j];
            const float ov = rect_overlap_dev(s.x[ci], s.y[ci], s.w[ci], s.h[ci], s.x[cj], s.y[cj], s.w[cj], s.h[cj]);
            if (ov > a.nms_threshold) bits |= 1ull << b;
        }
        a.mask[(long long)i * W64 + wd] = bits;
    }
    __syncthreads();
    if (t < 64) {  // greedy scan: a box is kept unless a kept box before it suppresses it
        uint64_t removed = 0;  // lane holds word `lane` of the removed set
        int nk = 0;
        for (int i = 0; i < Kn; ++i) {
            const uint64_t word = __shfl(removed, i >> 6, 64);
            if ((word >> (i & 63)) & 1ull) continue;
            if (t == 0) a.kept[nk] = i;
            ++nk;
            if (t < W64) removed |= a.mask[(long long)i * W64 + t];
        }
        if (t == 0) {
            a.counters[kern::kFinalCountSlot] = M;
            a.counters[kern::kKeptCountSlot] = nk;
            a.counters[kern::kFinalFlagSlot] = 0;
        }
    }
}

}  // namespace

namespace kern {
hipError_t finalize(const FinalArgs& a, hipStream_t stream) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kThreads), 0, stream, a);
    return hipGetLastError();
}
}  // namespace kern

}  // namespace sbm
