/*
 * ref_tables.c -- exposes the static dictionary tables of the reference C99
 * decoder (src/brotli_tables.h, compiled from the reference tree where it lies)
 * to the CPU restatement in przm_refcpu.c, which indexes them as
 * core/dict/static_index.go does.  Built by oracle/Makefile into
 * oracle/_ref/libprzm_ref_tables.so only when the reference tree is present;
 * test and baseline infrastructure, never part of the product.
 */
#include "brotli_tables.h"

const void *przm_ref_table(int which) {
    switch (which) {
    case 0: return przm_word_base;      /* uint16_t[22] */
    case 1: return przm_dict_offsets;   /* uint32_t[21] */
    case 2: return przm_transforms;     /* uint8_t[121][3]: prefix, op, suffix */
    case 3: return przm_compact_pairs;  /* uint16_t[2048][2]: word id, transform */
    case 4: return przm_ps_strings;     /* {const char *, uint8_t}[50] */
    case 5: return przm_top_words;      /* {const char *, uint8_t}[512] */
    case 6: return przm_brotli_data;    /* uint8_t[122784] */
    default: return 0;
    }
}
