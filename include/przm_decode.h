/*
 * przm_decode.h -- a CPU decoder of PRZM frames (libprzm_decode.so).
 *
 * It reads every frame the encode path of include/przm_gpu.h writes, and the
 * same subset of the format from any encoder:
 *   - uncompressed frames (flag 0x04) and empty frames;
 *   - the 16- and 24-bit section directories (flag 0x08), minimum match 3 or 4
 *     (flag 0x40), Section 1 as the frame lays it out (8 + 2 * byte 5 bytes);
 *   - Section 2 as Mode 8 (op / meta / var substreams, each flat 0x09, length
 *     0x00 + u32, Meta8 or adaptive V-ANS), flat 0x09, length 0x00 + u32 or raw;
 *   - Section 3 as 0x09 + LEB128 offsets, 0x00 + u32 length, or a distance-prefix
 *     container 0x06 / 0x0A / 0x0B whose code stream is raw, Mode 12 (0x0C), adaptive
 *     V-ANS or flat (core/distance/distance.go:563-690);
 *   - Section 4 as 4-lane V-ANS (docs/PRZM_SPEC.md:582-647), a sheaf container of
 *     four context fibers, each raw or V-ANS (core/literals/literals.go:601-664,
 *     sheaf.go:20-202), or raw literals;
 *   - the opcodes 0x00-0x7F (fused repeat, registers 0..2), 0x80-0xBF (explicit
 *     offset), 0xC0-0xC2 (long repeat), 0xF0 (literal run), 0xF1 (byte run) and
 *     0xF3-0xFA (short literal runs), with the ring registers {1, 4, 8}
 *     (docs/PRZM_SPEC.md:468-526, decompress.go).
 * Anything else (stride, dictionaries, normalisation, delta masks, nibble-plane
 * literal sections, the other opcodes) is refused with
 * PRZM_DEC_ERR_UNSUPPORTED rather than guessed.
 * Decoded bytes equal the reference decoders' on every frame it accepts.
 */
#ifndef PRZM_DECODE_H
#define PRZM_DECODE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRZM_DEC_ERR_CORRUPT -1     /* malformed frame or section */
#define PRZM_DEC_ERR_MAGIC -2       /* not a PRZM frame */
#define PRZM_DEC_ERR_CAPACITY -3    /* dst too small */
#define PRZM_DEC_ERR_UNSUPPORTED -4 /* a feature of the format this decoder does not read */
#define PRZM_DEC_ERR_RANGE -5       /* a match distance of 0 or beyond the bytes written */
#define PRZM_DEC_ERR_SIZE -6        /* the commands do not produce exactly the frame's size */

/* The frame's uncompressed size (its 24-bit header field), or an error. */
int64_t przm_decode_size(const uint8_t *src, size_t n);
/* One frame (exactly n bytes) into dst; returns the bytes written or an error. */
int64_t przm_decode_frame(const uint8_t *src, size_t n, uint8_t *dst, size_t cap);
/* nframes frames back to back (sizes[i] bytes each) into dst, one after another;
 * returns the total bytes written or the first error. */
int64_t przm_decode_frames(const uint8_t *src, size_t n, const uint32_t *sizes, size_t nframes, uint8_t *dst,
                           size_t cap);

/* A PRZS multi-frame stream (decompress.go:842-866: 'P','R','Z','S', flags u8,
 * u32 chunk count, u64 total uncompressed size; then per chunk its u32 length
 * and one frame, cmd/przm-viz/analyzer.go:242-256), or a single frame: the
 * total uncompressed size from the stream header (the frame's size for a frame),
 * or an error. */
int64_t przm_decode_stream_size(const uint8_t *src, size_t n);
/* The whole stream (exactly n bytes: header, then every chunk) or a single
 * frame into dst; the bytes written, or an error: PRZM_DEC_ERR_SIZE when the
 * chunks do not fill n or do not sum to the header's total, and
 * PRZM_DEC_ERR_UNSUPPORTED for a stream flag other than 0 (none is defined). */
int64_t przm_decode_stream(const uint8_t *src, size_t n, uint8_t *dst, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
