/*
 * sampo_wave.h -- wave-wide integer scans of gfx9 (64-lane waves) with DPP
 * moves: no LDS traffic, one VALU add per step.
 */
#ifndef SAMPO_WAVE_H
#define SAMPO_WAVE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* Inclusive scan inside each 16-lane row: Kogge-Stone by row shifts of 1,
 * 2, 4 and 8.  Lanes whose DPP source lies outside the row add 0. */
__device__ static inline uint32_t row_incl_scan(uint32_t x)
{
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false); /* row_shr:1 */
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false); /* row_shr:2 */
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false); /* row_shr:4 */
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false); /* row_shr:8 */
	return x;
}

/* Inclusive scan over the 64 lanes of a wave: the row scans, then the last
 * lanes of row 0 and of rows 0-1 carried into the rows after them by the
 * row broadcasts of lanes 15 and 31 (rows masked off add 0). */
__device__ static inline uint32_t wave_incl_scan(uint32_t x)
{
	x = row_incl_scan(x);
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); /* row_bcast:15 */
	x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); /* row_bcast:31 */
	return x;
}

/* The same for 64-bit words of 16-bit fields that never carry into the
 * next field: the two 32-bit halves scan independently. */
__device__ static inline uint64_t wave_incl_scan_fields(uint64_t x)
{
	return ((uint64_t)wave_incl_scan((uint32_t)(x >> 32)) << 32) | wave_incl_scan((uint32_t)x);
}

/* lane 63's value (a wave total after an inclusive scan), in every lane */
__device__ static inline uint32_t wave_last(uint32_t x) { return (uint32_t)__builtin_amdgcn_readlane((int)x, 63); }

#endif
