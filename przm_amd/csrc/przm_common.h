/* Shared constants and layouts of the MI355X encode path (device + host). */
#ifndef PRZM_COMMON_H
#define PRZM_COMMON_H

#include <stdint.h>

#define PRZM_BLOCK 65536u       /* frame block == match window (64 KiB) */
#define PRZM_HALF 32768u
#define PRZM_MAX_MATCH 1024u    /* compress.go:28 */
#define PRZM_EXT_CAP 65530u     /* przm_tokenizer.cuh:296 */
#define PRZM_RLE_CAP 65535u     /* przm_tokenizer.cuh:241 */

/* per-block stream slot capacities (bytes) */
#define PRZM_CMD_CAP 36864u     /* >= n/2 + 64: every token consumes >= 4 input bytes */
#define PRZM_OFF_CAP 49152u     /* <= 3 LEB bytes per explicit match of >= 4 bytes */
#define PRZM_LIT_CAP 65536u
#define PRZM_OP_CAP 36864u
#define PRZM_META_CAP 18432u
#define PRZM_VAR_CAP 36864u

/* stream indices */
enum { PRZM_S_CMD = 0, PRZM_S_OFF, PRZM_S_LIT, PRZM_S_OP, PRZM_S_META, PRZM_S_VAR, PRZM_S_TOK, PRZM_S_N };

/* entropy jobs per block: literal V-ANS, and Meta8 / V-ANS for each command substream */
enum { PRZM_J_LIT = 0, PRZM_J_OP_M8, PRZM_J_OP_V, PRZM_J_META_M8, PRZM_J_META_V, PRZM_J_VAR_M8, PRZM_J_VAR_V, PRZM_J_N };

/* one 64 KiB (or shorter, last) block of some input buffer */
typedef struct {
    uint64_t src_off;   /* byte offset of the block in the device input */
    uint32_t len;       /* 0..65536 */
    uint32_t flags;     /* PRZM_BF_* */
} przm_block_t;

#define PRZM_BF_RAW 1u      /* emit an uncompressed frame (high-entropy sample or empty) */
/* PRZS multi-frame stream of a buffer of more than one block (decompress.go:842-866 header,
 * cmd/przm-viz/analyzer.go:242-256 chunk walk): every block's frame is a chunk behind its u32
 * length, and the buffer's first block also carries the 17-byte stream header before it,
 * with the buffer's chunk count in the flag word's high bits */
#define PRZM_BF_CHUNK 2u    /* the frame is a length-prefixed chunk of a PRZS stream */
#define PRZM_BF_HEAD 4u     /* the stream header precedes this chunk */
#define PRZM_BF_COUNT_SHIFT 8u
#define PRZM_PRZS_HDR 17u   /* 'P','R','Z','S', flags u8, chunk count u32, total size u64 */
#define PRZM_PRZS_PREFIX 4u /* u32 chunk length */

/* Byte offsets of the per-block stream slots inside one block workspace. */
#define PRZM_WS_CMD 0u
#define PRZM_WS_OFF (PRZM_WS_CMD + PRZM_CMD_CAP)
#define PRZM_WS_LIT (PRZM_WS_OFF + PRZM_OFF_CAP)
#define PRZM_WS_OP (PRZM_WS_LIT + PRZM_LIT_CAP + 64u)
#define PRZM_WS_META (PRZM_WS_OP + PRZM_OP_CAP)
#define PRZM_WS_VAR (PRZM_WS_META + PRZM_META_CAP)
#define PRZM_WS_STREAMS (PRZM_WS_VAR + PRZM_VAR_CAP)

/* Entropy outputs: each job's coded stream is written backward ending at
 * the end of its slot.  A coded stream is only ever chosen when it is
 * shorter than the raw stream, so a job stops (and reports overflow) once
 * its output would exceed the raw length: slot = raw cap + 64. */
#define PRZM_ENC_LIT_CAP (PRZM_LIT_CAP + 64u)
#define PRZM_ENC_OP_CAP (PRZM_OP_CAP + 64u)
#define PRZM_ENC_META_CAP (PRZM_META_CAP + 64u)
#define PRZM_ENC_VAR_CAP (PRZM_VAR_CAP + 64u)
#define PRZM_WS_ENC_LIT PRZM_WS_STREAMS
#define PRZM_WS_ENC_OP_M8 (PRZM_WS_ENC_LIT + PRZM_ENC_LIT_CAP)
#define PRZM_WS_ENC_OP_V (PRZM_WS_ENC_OP_M8 + PRZM_ENC_OP_CAP)
#define PRZM_WS_ENC_META_M8 (PRZM_WS_ENC_OP_V + PRZM_ENC_OP_CAP)
#define PRZM_WS_ENC_META_V (PRZM_WS_ENC_META_M8 + PRZM_ENC_META_CAP)
#define PRZM_WS_ENC_VAR_M8 (PRZM_WS_ENC_META_V + PRZM_ENC_META_CAP)
#define PRZM_WS_ENC_VAR_V (PRZM_WS_ENC_VAR_M8 + PRZM_ENC_VAR_CAP)
#define PRZM_WS_BYTES (PRZM_WS_ENC_VAR_V + PRZM_ENC_VAR_CAP)
#define PRZM_ENC_OVERFLOW 0xFFFFFFFFu

/* Per-block metadata written by the kernels (u32 words). */
typedef struct {
    uint32_t len[PRZM_S_N];        /* stream lengths + token count */
    uint32_t enc_len[PRZM_J_N];    /* coded stream bytes (no header) per job; 0xFFFFFFFF = not run */
    uint32_t enc_state[PRZM_J_N][4];
    uint32_t mat_len[4];           /* joint-nibble matrix bytes for lit, op, meta, var */
    uint32_t frame_len;            /* final frame bytes */
    uint32_t frame_kind;           /* 0 compressed, 1 raw (entropy/empty), 2 raw (size guard), 3 raw (S4 hazard) */
    uint32_t choice[3];            /* substream coders chosen: 0 flat, 1 Meta8, 2 V-ANS; -1 flat section */
    uint32_t s_len[3];             /* S2, S3, S4 bytes */
    uint32_t s4_vans;
    uint32_t s3_mode;              /* Section 3's first byte: 0x09 flat, 0x06 / 0x0A / 0x0B prefix, 0 none */
    uint32_t s3_code;              /* prefix code stream: 0 Mode 12, 1 adaptive V-ANS */
    uint32_t s3_code_len;          /* prefix code stream bytes (head included) */
} przm_meta_t;

/* Normalized frequency tables per V-ANS job (lit, op, meta, var). */
typedef struct {
    uint32_t fc[256];       /* freq | cdf << 16 */
    uint8_t mat[552];       /* serialized joint-nibble matrix */
} przm_ftab_t;

/*
 * Section 3 as a distance-prefix container (an encoder option): per-block
 * offsets workspace.  A block has at most 16384 explicit distances (each
 * match covers >= 4 bytes), each below 65536, so at most 14 extra bits and
 * 12 code bits (frequency >= 1 of 4096) each.  Code-stream slots are filled
 * backward from their 16-byte aligned ends.
 */
typedef struct {
    uint32_t n;             /* explicit distances (the match count) */
    uint32_t orig;          /* their LEB128 bytes (the offset stream's length) */
    uint32_t extra;         /* extra-bit bytes */
    uint32_t m12_len;       /* Mode-12 stream bytes after its 5-byte head, or PRZM_ENC_OVERFLOW */
    uint32_t m12_state;
    uint32_t v_len;         /* V-ANS stream bytes of the codes (n >= 64), or PRZM_ENC_OVERFLOW */
    uint32_t v_state[4];
    uint32_t mat_len;       /* joint-nibble matrix bytes of the codes' table */
    uint32_t pad[5];
} przm_offs_t;

#define PRZM_OFFS_MAX 16384u
#define PRZM_OWS_HDR 0u                                       /* przm_offs_t */
#define PRZM_OWS_FTAB 64u                                     /* przm_ftab_t of the codes */
#define PRZM_OWS_CODES 1664u                                  /* one code per distance */
#define PRZM_OWS_CODES_CAP (PRZM_OFFS_MAX + 64u)
#define PRZM_OWS_EXTRA (PRZM_OWS_CODES + PRZM_OWS_CODES_CAP)  /* extra bits, MSB first */
#define PRZM_OWS_EXTRA_CAP (PRZM_OFFS_MAX * 14u / 8u + 64u)
#define PRZM_OWS_M12 (PRZM_OWS_EXTRA + PRZM_OWS_EXTRA_CAP)    /* Mode-12 stream (slot end aligned) */
#define PRZM_OWS_M12_CAP (PRZM_OFFS_MAX * 12u / 8u + 64u)
#define PRZM_OWS_VANS (PRZM_OWS_M12 + PRZM_OWS_M12_CAP)       /* V-ANS stream of the codes */
#define PRZM_OWS_VANS_CAP (2u * PRZM_OFFS_MAX + 64u)
#define PRZM_OWS_BYTES (PRZM_OWS_VANS + PRZM_OWS_VANS_CAP)

/*
 * Section 4 with the sheaf trials (an encoder option): per-block sheaf
 * workspace.  For each of the four modes the literals are split into four
 * fibers, laid out back to back with each fiber starting 16-byte aligned; the
 * winning mode's coded fibers go to slots of align16(l + 64) bytes each,
 * filled backward from their ends.
 */
typedef struct {
    uint32_t n;             /* literals */
    uint32_t l[4][4];       /* fiber lengths per mode */
    uint32_t k[4][4];       /* coded stream bytes per fiber from the counting pass, or PRZM_ENC_OVERFLOW
                               (the fiber stays raw) */
    uint32_t mat[4][4];     /* matrix bytes of each fiber's table */
    uint32_t choice;        /* Section 4: 0 raw, 1 adaptive V-ANS, 2 + mode a sheaf container */
    uint32_t len;           /* Section 4 bytes */
    uint32_t ek[4];         /* the chosen mode's coded stream bytes per fiber */
    uint32_t st[4][4];      /* and their final states */
} przm_sheaf_t;

#define PRZM_SWS_HDR 0u                                 /* przm_sheaf_t */
#define PRZM_SWS_FTAB 512u                              /* przm_ftab_t [mode][fiber] */
#define PRZM_SWS_FIB (PRZM_SWS_FTAB + 16u * 1576u)      /* fibers of mode m at FIB + m * FIB_CAP */
#define PRZM_SWS_FIB_CAP (65536u + 64u)
#define PRZM_SWS_OUT (PRZM_SWS_FIB + 4u * PRZM_SWS_FIB_CAP)
#define PRZM_SWS_OUT_CAP (65536u + 4u * 80u + 64u)
#define PRZM_SWS_BYTES (PRZM_SWS_OUT + PRZM_SWS_OUT_CAP)
/* u32 class counts and reservations of the sheaf counting-job order, after the job orders in the
 * coder-order buffer (8-byte aligned: one word of room) */
#define PRZM_SHEAF_BINS 512u

#endif
