/* Static per-context frequencies of the distance-code coder (Section 3 code stream 0x0C).
 * Wire-format constants: the encoder and every decoder must hold the same table, so the
 * values are those of core/distance/distance.go:69-93 (przm_dist_static_freq4 in the
 * reference C99 decoder).  Context c is the class of the previous code (0: < 6, 1: < 14,
 * 2: < 26, 3: the rest); each row sums to 4096. */
#ifndef PRZM_DIST_FREQ_H
#define PRZM_DIST_FREQ_H
#include <stdint.h>
#ifdef __cplusplus
#define PRZM_DIST_QUAL static constexpr
#else
#define PRZM_DIST_QUAL static const
#endif
PRZM_DIST_QUAL uint16_t przm_dist_freq[4][64] = {
  { /* context 0 */
      1,   1,   1,   1,   1,   1, 128,  43, 128, 256, 341, 687, 341, 256, 427,  43,
    256,  43, 128,   1,  43,  43,  85,  85,  85,  85,  85,  85, 171,   1,  85,  43,
     85,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,
      1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,
  },
  { /* context 1 */
      2,   1,   1,   1,   1,   1,  83,  61, 183, 145, 236, 214, 291, 248, 250, 186,
    218, 176, 192, 157, 189, 154, 179, 115, 164, 111, 136,  70,  98,  60,  54,  36,
     31,  18,   5,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,
      1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,
  },
  { /* context 2 */
      3,   1,   1,   1,   1,   1,  58,  40, 113, 100, 164, 141, 194, 174, 245, 214,
    232, 188, 226, 208, 245, 189, 245, 166, 217, 145, 144,  87, 102,  68,  52,  34,
     44,  18,   6,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,
      1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,
  },
  { /* context 3 */
      2,   1,   1,   1,   1,   1,  39,  38,  86,  79, 163, 124, 167, 143, 187, 151,
    179, 140, 151, 136, 168, 123, 213, 123, 179, 147, 253, 169, 254, 177, 161, 112,
    112,  60,  25,   2,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,
      1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,   1,
  },
};
#endif
