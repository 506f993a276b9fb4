/*
 * bc250_calib.h - measurement calibrators (no reference counterpart; they
 * serve the benchmark, not the encoder).
 */
#ifndef BC250_CALIB_H
#define BC250_CALIB_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
/* blocks x 256 threads, each running 8 chains of iters*16 v_sad_u8:
 * blocks*256*iters*128 instructions, 4 byte-absdiffs each. out holds
 * blocks*256 uint32. */
int bc250_calib_sad_peak(uint32_t *out, int blocks, int iters, void *stream);
/* Copies `bytes` (multiple of 16) from src to dst. */
int bc250_calib_copy(const void *src, void *dst, size_t bytes, void *stream);
#ifdef __cplusplus
}
#endif
#endif
