/*
 * Run-time choice of the build of sbm_cpu_fast.c for this host's CPU -- TEST INFRASTRUCTURE ONLY (the timed
 * CPU baseline of bench.py).  sbm_cpu_fast.c is compiled three times (AVX-512, AVX2, generic x86-64); the
 * widest instruction set the CPU reports is used.
 */
#include <stdint.h>

#include "sbm_oracle.h"

#define FAST_DECL(sfx)                                                                                             \
    int orc_fast_match##sfx(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, int L, \
                            const int* T, float threshold, int num_templates, const int32_t* tinfo,                \
                            const int32_t* feats, const int32_t* tclass, const int32_t* tid, int max_out,          \
                            int32_t* out_i32, float* out_sim, int nthreads, double* stage_seconds);               \
    int orc_fast_build_lm_pyramid##sfx(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask,        \
                                       float weak, int L, const int* T, uint8_t** lm_out);                         \
    void orc_fast_bin_table##sfx(int lo, int hi, uint8_t* out);
FAST_DECL(_avx512)
FAST_DECL(_avx2)
FAST_DECL(_generic)

static int isa_level(void) {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") && __builtin_cpu_supports("avx512vl"))
        return 2;
    if (__builtin_cpu_supports("avx2")) return 1;
    return 0;
}

const char* orc_fast_isa(void) {
    static const char* names[3] = {"x86-64", "AVX2", "AVX-512"};
    return names[isa_level()];
}

int orc_fast_match(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, int L,
                   const int* T, float threshold, int num_templates, const int32_t* tinfo, const int32_t* feats,
                   const int32_t* tclass, const int32_t* tid, int max_out, int32_t* out_i32, float* out_sim,
                   int nthreads, double* stage_seconds) {
    switch (isa_level()) {
        case 2:
            return orc_fast_match_avx512(img, rows, cols, ch, mask, weak, L, T, threshold, num_templates, tinfo, feats,
                                         tclass, tid, max_out, out_i32, out_sim, nthreads, stage_seconds);
        case 1:
            return orc_fast_match_avx2(img, rows, cols, ch, mask, weak, L, T, threshold, num_templates, tinfo, feats,
                                       tclass, tid, max_out, out_i32, out_sim, nthreads, stage_seconds);
        default:
            return orc_fast_match_generic(img, rows, cols, ch, mask, weak, L, T, threshold, num_templates, tinfo,
                                          feats, tclass, tid, max_out, out_i32, out_sim, nthreads, stage_seconds);
    }
}

int orc_fast_build_lm_pyramid(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, int L,
                              const int* T, uint8_t** lm_out) {
    switch (isa_level()) {
        case 2: return orc_fast_build_lm_pyramid_avx512(img, rows, cols, ch, mask, weak, L, T, lm_out);
        case 1: return orc_fast_build_lm_pyramid_avx2(img, rows, cols, ch, mask, weak, L, T, lm_out);
        default: return orc_fast_build_lm_pyramid_generic(img, rows, cols, ch, mask, weak, L, T, lm_out);
    }
}

void orc_fast_bin_table(int lo, int hi, uint8_t* out) {
    switch (isa_level()) {
        case 2: orc_fast_bin_table_avx512(lo, hi, out); break;
        case 1: orc_fast_bin_table_avx2(lo, hi, out); break;
        default: orc_fast_bin_table_generic(lo, hi, out); break;
    }
}
