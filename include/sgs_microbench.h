/*
 * Measurement kernels of libsgs_microbench.so: the ceilings the renderer's roofline figures
 * are read against on the GPU they run on. No reference counterpart (the reference has no
 * benchmark kernels); not part of the renderer's ABI (include/sgs.h).
 *
 * Device pointers of the current device; work is enqueued on the caller's stream (a
 * hipStream_t as void*, NULL = null stream); returns 0 or a negative code (as sgs.h), with
 * sgs_mb_last_error() describing the failure.
 */
#ifndef SGS_MICROBENCH_H
#define SGS_MICROBENCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* sgs_mb_last_error(void);

/* Stream copy: dst[i] = src[i] for bytes / 16 16-byte words (bytes a multiple of 16, both
 * pointers 16-byte aligned, no overlap), a grid-stride loop of 16-byte loads and stores.
 * Moves 2 * bytes of HBM traffic: the copy ceiling of the device. */
int sgs_mb_copy(void* dst, const void* src, size_t bytes, void* stream);

/* Random 64-bit atomicMin: n_ops operations atomicMin(&words[a], v), with a uniform in
 * [0, n_words) and v a uniform 64-bit value, both from a counter-based hash of (seed, op)
 * (SplitMix64 finaliser), spread over a saturating grid. The rate n_ops / time is the
 * device's random u64 atomicMin throughput on a framebuffer of n_words words. */
int sgs_mb_atomic_min_u64(uint64_t* words, uint64_t n_words, uint64_t n_ops, uint64_t seed, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SGS_MICROBENCH_H */
