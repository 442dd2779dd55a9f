/*
 * sampo_compact.h -- the compaction index of the reference's createNewAgents
 * step as one kernel: for every agent, its rank among the agents of the
 * same new state.  The reference computes it with preScan
 * (kernel/preScan.cl:21-64) and a segmented scan (src/abms.c:359-378,
 * src/boltScan.c, kernel/boltScan.cl), three times per iteration for three
 * groups of states (src/abms.c:404-415); oldToNewAgents reads the result as
 * the offset of an agent inside its new range (kernel/oldToNewAgents.cl:51-93).
 * Here all eight states are ranked in one single-pass scan over a 1-byte key
 * per agent, writing a 4-byte destination index: 5 bytes per agent.
 *
 * The fused engine (sampo_engine.h) does not call this: it ranks survivors
 * inside its pass.  The entry points stand alone for callers that keep the
 * reference's separate compaction step, and as its measurement.
 *
 * All functions return 0 or a negative SAMPO_E* code (sampo_engine.h).
 */
#ifndef SAMPO_COMPACT_H
#define SAMPO_COMPACT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAMPO_COMPACT_KEYS 8u          /* keys 0..7 are ranked (the states egg..gravid) */
#define SAMPO_COMPACT_NONE 0xffffffffu /* dest of a key >= 8 (dead or absent agent) */

/* Device bytes of workspace sampo_compact_index() needs for n keys (32 per
 * 16384 keys and per 2^24 keys; it needs no initialisation).  Calls that overlap on one
 * workspace race: order them on one stream. */
uint64_t sampo_compact_workspace_bytes(uint64_t n);

/* keys: n bytes on the device; dest: n uint32 on the device; counts: 8
 * uint32 on the device.  For every i, dest[i] = #{ j < i : keys[j] ==
 * keys[i] } if keys[i] < 8, else SAMPO_COMPACT_NONE; counts[k] = #{ keys ==
 * k }.  n < 2^32; the workspace 16-byte aligned.  Four launches,
 * asynchronous on `stream` (a hipStream_t; NULL: the default stream). */
int sampo_compact_index(const uint8_t *keys, uint64_t n, uint32_t *dest, uint32_t *counts, void *workspace,
                        uint64_t workspace_bytes, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  sampo_compact_index without the workspace argument:
 * the workspace is a fixed array the library keeps on each GPU it runs on
 * (525 KB, held until the process ends), so the call allocates nothing.
 * Runs on the current device (the stream must belong to it); n at most
 * SAMPO_COMPACT_NOALLOC_MAX_KEYS (SAMPO_E_ARG beyond).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race, so order them
 * on one stream.  The workspace needs no clearing (each call writes all of
 * it that it reads). */
#define SAMPO_COMPACT_NOALLOC_MAX_KEYS (1ull << 28)
int sampo_compact_index_noalloc(const uint8_t *keys, uint64_t n, uint32_t *dest, uint32_t *counts, void *stream);

/* The same on host buffers (allocates device memory, synchronises). */
int sampo_compact_index_host(int device, const uint8_t *keys, uint64_t n, uint32_t *dest, uint32_t counts[8]);

/* Time `iters` calls (after one untimed call) on n keys generated on the
 * device from `seed` (states weighted like a population: 5 % dead); HIP
 * events on one stream; average milliseconds per call. */
int sampo_compact_time(int device, uint64_t n, uint32_t iters, uint64_t seed, float *ms_per_call);

#ifdef __cplusplus
}
#endif
#endif
