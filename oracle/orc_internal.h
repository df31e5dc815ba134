/*
 * Helpers shared by the oracle's translation units (sbm_oracle.c, sbm_oracle16.c).
 * TEST INFRASTRUCTURE ONLY, like the rest of oracle/.
 */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H
#include <stdint.h>

#define ORC_EINVAL (-1000000000)

typedef struct {
    int x, y, cls, tid;
    float sim;
} omatch;

typedef struct {
    omatch* v;
    int n, cap;
} omvec;

typedef struct { int x, y, label; float theta, score; } ocand;

void orc_i_omvec_push(omvec* a, omatch m);
/* canonical sort + std::unique; returns the kept count */
int orc_i_sort_unique(omatch* v, int n);
/* The candidate stage of extractTemplate (5x5 magnitude NMS over the eroded mask, strong threshold,
 * stable sort by score).  The quantized angles come as 8-bit (angle8) or 16-bit (angle16) single-bit
 * values; exactly one of the two is non-null.  *out is malloc'ed; returns the candidate count. */
int orc_i_extract_candidates(const float* mag, const uint8_t* angle8, const uint16_t* angle16,
                             const float* angle_ori, const uint8_t* mask, int rows, int cols, float strong,
                             ocand** out);

#endif
