/*
 * sampo_compact.hip -- per-state ranks of the compaction step
 * (include/sampo_compact.h) on gfx950.
 *
 * The reference ranks the agents of each new state with preScan
 * (kernel/preScan.cl:21-64) and a segmented scan run three times per
 * iteration (src/abms.c:359-415): ~12 bytes per agent and pass for the flag,
 * data and key arrays, three passes.  Here a 1-byte key per agent is read
 * and its 4-byte rank written: 5 bytes per agent.
 *
 * Reduce, then scan, in four launches and with no waits between work-groups:
 *   k_count      per tile of 16384 keys, the count of each key;
 *   k_scan_local per block of 1024 tiles: per key, the exclusive prefix of
 *                each tile inside its block (in place) and the block's sum;
 *   k_scan_top   one work-group: the exclusive prefix over blocks, and the
 *                totals;
 *   k_rank       per tile: the keys again (the second read usually comes
 *                from the 256 MB Infinity Cache after k_count), their ranks
 *                inside the tile, plus the tile's and its block's prefix.
 * A single-pass look-back scan was built first and measured at 1.1 TB/s:
 * every tile took a ticket from one device-wide counter, and those atomics,
 * not HBM, set the pace (~75 ns each).
 *
 * Tiles are four rounds of 4096 keys; in a round a thread takes 16
 * consecutive keys with one 16-byte load (a wave reads 1 KiB contiguous).
 * Per-thread counts of the eight keys are packed as 16-bit fields in two
 * 64-bit words (a tile holds at most 16384 of a key), so a work-group scan
 * is a scan of two integers.  The ranks of a round go through LDS, padded
 * against bank conflicts, so that every store instruction writes 1 KiB
 * contiguous per wave.
 */
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>

#include <mutex>

#include "../../include/sampo_engine.h"
#include "../../include/sampo_compact.h"
#include "sampo_wave.h"

#define CT 256u                 /* threads per tile */
#define CV 16u                  /* keys per thread and round */
#define CR 4u                   /* rounds per tile */
#define CSUB (CT * CV)          /* keys per round: 4096 */
#define CTILE (CSUB * CR)       /* keys per tile: 16384 */
#define CK SAMPO_COMPACT_KEYS   /* ranked keys */
#define CBLK (CT * 4u)          /* tiles per block of the tile scan: 1024 */
#define CCNT 1u                 /* tiles per work-group of k_count */

namespace {

__device__ inline uint32_t field16(uint64_t w, uint32_t i) { return (uint32_t)(w >> (16u * i)) & 0xffffu; }

__device__ inline uint32_t lds_slot(uint32_t e) { return e + (e >> 5); } /* one pad word per 32 */

/* 16 keys of round r of the tile at t0 for thread tid, 4 per word; keys past
 * n are 0xff (not ranked, not written) */
__device__ inline void load_keys(const uint8_t *__restrict__ keys, uint64_t n, uint64_t t0, uint32_t r, uint32_t tid,
                                 uint32_t vec_in, uint32_t (&kw)[4])
{
	const uint64_t base = t0 + r * CSUB + (uint64_t)tid * CV;
	if (vec_in && base + CV <= n) {
		const uint4 v = *(const uint4 *)(keys + base);
		kw[0] = v.x;
		kw[1] = v.y;
		kw[2] = v.z;
		kw[3] = v.w;
	} else { /* the last tile, or keys not 16-byte aligned */
#pragma unroll
		for (uint32_t q = 0; q < 4u; ++q) {
			uint32_t x = 0;
#pragma unroll
			for (uint32_t b = 0; b < 4u; ++b) {
				const uint64_t i = base + 4u * q + b;
				x |= (uint32_t)(i < n ? keys[i] : 0xffu) << (8u * b);
			}
			kw[q] = x;
		}
	}
}

/* counts of keys 0..7 among the 16 keys of kw as 8-bit fields of one
 * 64-bit word (key 8 and up add nothing).  A select, not a branch: branches
 * on which counter to add to are turned by the compiler into a pointer
 * select that puts the counters in scratch memory.  The keys outnumber the
 * bytes moved, so this loop, not HBM, bounds the counting kernel; 8-bit
 * fields in one word take fewer instructions than 16-bit fields in two. */
__device__ inline uint64_t count8(const uint32_t (&kw)[4])
{
	uint64_t a = 0;
#pragma unroll
	for (uint32_t j = 0; j < CV; ++j) {
		const uint32_t k = (kw[j >> 2] >> (8u * (j & 3u))) & 0xffu;
		a += k < 8u ? 1ull << (8u * k) : 0ull;
	}
	return a;
}

/* Counting with byte permutes.  For the key pair (2m, 2m + 1) the 8-byte
 * table {perm_hi(m):perm_lo(m)} holds 0x01 at byte 2m and 0x10 at byte
 * 2m + 1, zeros elsewhere: v_perm_b32 with a word of 4 keys as the selector
 * gives, per byte, that key's increment of the pair's two nibble counters --
 * 4 keys counted by one instruction per pair.  Selector bytes 8-11 read the
 * sign bits of table bytes (0 here) and 12 reads 0, so keys 8-12 count
 * nowhere; 13 and more would read 0xff, so the words of a thread that holds
 * such a key are counted by count8 instead (big_keys). */
__device__ inline uint32_t perm_lo(uint32_t m) { return m == 0u ? 0x00001001u : (m == 1u ? 0x10010000u : 0u); }
__device__ inline uint32_t perm_hi(uint32_t m) { return m == 2u ? 0x00001001u : (m == 3u ? 0x10010000u : 0u); }

/* nonzero if a byte of w is 13 or more: (b & 0x7f) + 0x73 carries into the
 * top bit exactly for 13 <= b < 128, and stays inside its byte */
__device__ inline uint32_t big_keys(uint32_t w) { return (((w & 0x7f7f7f7fu) + 0x73737373u) | w) & 0x80808080u; }

/* the 4 keys of w into the nibble counters acc (each counter takes one key
 * per word: 15 words at most between flushes) */
__device__ inline void nib_add(uint32_t w, uint32_t (&acc)[4])
{
#pragma unroll
	for (uint32_t m = 0; m < 4u; ++m) acc[m] += __builtin_amdgcn_perm(perm_hi(m), perm_lo(m), w);
}

/* the nibble counters summed over their 4 bytes into c[8], then cleared */
__device__ inline void nib_flush(uint32_t (&acc)[4], uint32_t (&c)[CK])
{
#pragma unroll
	for (uint32_t m = 0; m < 4u; ++m) {
		c[2u * m] = __builtin_amdgcn_sad_u8(acc[m] & 0x0f0f0f0fu, 0u, c[2u * m]);
		c[2u * m + 1u] = __builtin_amdgcn_sad_u8((acc[m] >> 4) & 0x0f0f0f0fu, 0u, c[2u * m + 1u]);
		acc[m] = 0u;
	}
}

/* counts c[0..3] / c[4..7] as the 16-bit fields of one word each */
__device__ inline void pack16(const uint32_t (&c)[CK], uint64_t &ca, uint64_t &cb)
{
	ca = ((uint64_t)(c[2] | (c[3] << 16)) << 32) | (c[0] | (c[1] << 16));
	cb = ((uint64_t)(c[6] | (c[7] << 16)) << 32) | (c[4] | (c[5] << 16));
}

/* counts of keys 0-3 and 4-7 among the R x 16 keys of kw, as 16-bit fields
 * (the counters are flushed every 2 rounds: 8 keys each at most) */
template <uint32_t R>
__device__ inline void count_keys(const uint32_t (&kw)[R][4], uint64_t &ca, uint64_t &cb)
{
	uint32_t c[CK] = {0, 0, 0, 0, 0, 0, 0, 0}, acc[4] = {0, 0, 0, 0}, big = 0;
#pragma unroll
	for (uint32_t r = 0; r < R; ++r) {
#pragma unroll
		for (uint32_t q = 0; q < 4u; ++q) {
			nib_add(kw[r][q], acc);
			big |= big_keys(kw[r][q]);
		}
		if (r % 2u == 1u || r + 1u == R) nib_flush(acc, c);
	}
	if (big) { /* a key of 13 or more: the exact count */
		uint64_t c8 = 0; /* at most 64 of a key: 8-bit fields hold it */
#pragma unroll 1
		for (uint32_t r = 0; r < R; ++r) c8 += count8(kw[r]); /* rolled: this rare path sets no register budget */
#pragma unroll
		for (uint32_t k = 0; k < CK; ++k) c[k] = (uint32_t)(c8 >> (8u * k)) & 0xffu;
	}
	pack16(c, ca, cb);
}

/* phase 1: tile_count[tile * 8 + k] = keys equal to k in the tile; a
 * work-group counts CCNT consecutive tiles (one: with all its 16 key words
 * loaded at once, a work-group per tile reads fastest) */
__global__ void __launch_bounds__(CT) k_count(const uint8_t *__restrict__ keys, uint32_t n, uint32_t ntiles,
                                              uint32_t *__restrict__ tile_count, uint32_t vec_in)
{
	__shared__ uint64_t s_w[CCNT][CT / 64u][2];
	const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
#pragma unroll 1
	for (uint32_t g = 0; g < CCNT; ++g) { /* one tile at a time: its 16 key words, counts, wave totals */
		const uint64_t t0 = ((uint64_t)blockIdx.x * CCNT + g) * CTILE;
		uint32_t kw[CR][4];
#pragma unroll
		for (uint32_t r = 0; r < CR; ++r) load_keys(keys, n, t0, r, tid, vec_in, kw[r]);
		uint64_t ta, tb;
		count_keys<CR>(kw, ta, tb);
		ta = wave_incl_scan_fields(ta); /* fields stay below 2^16 */
		tb = wave_incl_scan_fields(tb);
		if (lane == 63u) {
			s_w[g][w][0] = ta;
			s_w[g][w][1] = tb;
		}
	}
	__syncthreads();
	if (tid < CCNT * CK) {
		const uint32_t g = tid / CK, k = tid % CK, tile = blockIdx.x * CCNT + g;
		uint32_t c = 0;
		for (uint32_t q = 0; q < CT / 64u; ++q) c += field16(s_w[g][q][k >> 2], k & 3u);
		if (tile < ntiles) tile_count[(size_t)tile * CK + k] = c;
	}
}

/* block exclusive scan of eight u32 per thread over a work-group of CT
 * threads; returns the totals through `total` */
__device__ inline void block_scan8(uint32_t (&v)[CK], uint32_t (&total)[CK], uint32_t (*s_w)[CK])
{
	const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
	uint32_t x[CK];
#pragma unroll
	for (uint32_t k = 0; k < CK; ++k) {
		x[k] = wave_incl_scan(v[k]);
		if (lane == 63u) s_w[w][k] = x[k];
	}
	__syncthreads();
#pragma unroll
	for (uint32_t k = 0; k < CK; ++k) {
		uint32_t before = 0, tot = 0;
#pragma unroll
		for (uint32_t q = 0; q < CT / 64u; ++q) {
			if (q == w) before = tot;
			tot += s_w[q][k];
		}
		v[k] = x[k] - v[k] + before;
		total[k] = tot;
	}
}

/* phase 2a: per block of CBLK tiles (4 per thread), tile_count becomes the
 * exclusive prefix inside the block; block_sum receives the block's counts */
__global__ void __launch_bounds__(CT) k_scan_local(uint32_t *__restrict__ tile_count, uint32_t ntiles,
                                                   uint32_t *__restrict__ block_sum)
{
	__shared__ uint32_t s_w[CT / 64u][CK];
	const uint32_t tid = threadIdx.x;
	const uint32_t t0 = blockIdx.x * CBLK + tid * 4u;
	uint4 a[4], b[4];
#pragma unroll
	for (uint32_t q = 0; q < 4u; ++q) {
		a[q] = t0 + q < ntiles ? *(const uint4 *)(tile_count + (size_t)(t0 + q) * CK) : make_uint4(0, 0, 0, 0);
		b[q] = t0 + q < ntiles ? *(const uint4 *)(tile_count + (size_t)(t0 + q) * CK + 4u) : make_uint4(0, 0, 0, 0);
	}
	uint32_t v[CK] = {0, 0, 0, 0, 0, 0, 0, 0}, tot[CK];
#pragma unroll
	for (uint32_t q = 0; q < 4u; ++q) {
		v[0] += a[q].x; v[1] += a[q].y; v[2] += a[q].z; v[3] += a[q].w;
		v[4] += b[q].x; v[5] += b[q].y; v[6] += b[q].z; v[7] += b[q].w;
	}
	block_scan8(v, tot, s_w);
#pragma unroll
	for (uint32_t q = 0; q < 4u; ++q) {
		if (t0 + q < ntiles) {
			uint32_t *c = tile_count + (size_t)(t0 + q) * CK;
			*(uint4 *)c = make_uint4(v[0], v[1], v[2], v[3]);
			*(uint4 *)(c + 4u) = make_uint4(v[4], v[5], v[6], v[7]);
		}
		v[0] += a[q].x; v[1] += a[q].y; v[2] += a[q].z; v[3] += a[q].w;
		v[4] += b[q].x; v[5] += b[q].y; v[6] += b[q].z; v[7] += b[q].w;
	}
	if (tid < CK) block_sum[(size_t)blockIdx.x * CK + tid] = tot[tid];
}

/* phase 2b (one work-group, a block per thread): block_sum becomes the
 * exclusive prefix over blocks; counts receives the totals */
__global__ void __launch_bounds__(CT) k_scan_top(uint32_t *__restrict__ block_sum, uint32_t nblocks,
                                                 uint32_t *__restrict__ counts)
{
	__shared__ uint32_t s_w[CT / 64u][CK];
	const uint32_t tid = threadIdx.x;
	uint32_t v[CK], tot[CK];
	const uint4 a = tid < nblocks ? *(const uint4 *)(block_sum + (size_t)tid * CK) : make_uint4(0, 0, 0, 0);
	const uint4 b = tid < nblocks ? *(const uint4 *)(block_sum + (size_t)tid * CK + 4u) : make_uint4(0, 0, 0, 0);
	v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
	v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
	block_scan8(v, tot, s_w);
	if (tid < nblocks) {
		uint32_t *c = block_sum + (size_t)tid * CK;
		*(uint4 *)c = make_uint4(v[0], v[1], v[2], v[3]);
		*(uint4 *)(c + 4u) = make_uint4(v[4], v[5], v[6], v[7]);
	}
	if (tid < CK) counts[tid] = tot[tid];
}

/* phase 3: ranks of the tile's keys plus the tile's prefix */
__global__ void __launch_bounds__(CT) k_rank(const uint8_t *__restrict__ keys, uint32_t n,
                                             uint32_t *__restrict__ dest, const uint32_t *__restrict__ tile_prefix,
                                             const uint32_t *__restrict__ block_prefix, uint32_t vec_in,
                                             uint32_t vec_out)
{
	__shared__ uint64_t s_w[CT / 64u][CR][2];
	__shared__ uint64_t s_ea[CR][CT], s_eb[CR][CT]; /* per round: each thread's exclusive counts in the tile */
	__shared__ uint32_t s_pref[CK];
	__shared__ uint16_t s_base[CK][CT]; /* per key: this thread's exclusive count in the tile, this round */
	__shared__ uint32_t s_out[CSUB + CSUB / 32u];
	const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
	const uint64_t t0 = (uint64_t)blockIdx.x * CTILE;
	if (tid < CK)
		s_pref[tid] = tile_prefix[(size_t)blockIdx.x * CK + tid] + block_prefix[(size_t)(blockIdx.x / CBLK) * CK + tid];
	uint32_t kw[CR][4];
#pragma unroll
	for (uint32_t r = 0; r < CR; ++r) load_keys(keys, n, t0, r, tid, vec_in, kw[r]);
	/* per round, the work-group exclusive scan of the packed counts; the
	 * exclusive counts wait in LDS (in registers the four rounds' scans
	 * would be interleaved by the compiler into 256 VGPRs) */
#pragma unroll 1
	for (uint32_t r = 0; r < CR; ++r) {
		uint64_t ca, cb;
		uint32_t one[1][4] = {{kw[r][0], kw[r][1], kw[r][2], kw[r][3]}};
		count_keys<1>(one, ca, cb);
		const uint64_t xa = wave_incl_scan_fields(ca), xb = wave_incl_scan_fields(cb);
		if (lane == 63u) {
			s_w[w][r][0] = xa;
			s_w[w][r][1] = xb;
		}
		s_ea[r][tid] = xa - ca;
		s_eb[r][tid] = xb - cb;
	}
	__syncthreads();
	uint64_t ra = 0, rb = 0; /* keys of the rounds before */
#pragma unroll
	for (uint32_t r = 0; r < CR; ++r) {
		uint64_t wa = 0, wb = 0, ta = 0, tb = 0;
#pragma unroll
		for (uint32_t q = 0; q < CT / 64u; ++q) {
			if (q == w) {
				wa = ta;
				wb = tb;
			}
			ta += s_w[q][r][0];
			tb += s_w[q][r][1];
		}
		s_ea[r][tid] += ra + wa;
		s_eb[r][tid] += rb + wb;
		ra += ta;
		rb += tb;
	}
	/* per round: ranks into LDS in key order, then out in contiguous
	 * 16-byte pieces (a wave stores 1 KiB per instruction).  A key's rank is
	 * its tile's prefix, plus this thread's exclusive count of it in the
	 * tile (read from LDS by the key), plus its earlier copies among the
	 * thread's 16 keys of the round (nibble counters in one register: at
	 * most 15 before the 16th key). */
#pragma unroll
	for (uint32_t r = 0; r < CR; ++r) {
		const uint64_t xa = s_ea[r][tid], xb = s_eb[r][tid];
#pragma unroll
		for (uint32_t k = 0; k < CK; ++k) s_base[k][tid] = (uint16_t)field16(k < 4u ? xa : xb, k & 3u);
		uint32_t run = 0;
#pragma unroll
		for (uint32_t j = 0; j < CV; ++j) {
			const uint32_t k = __builtin_amdgcn_ubfe(kw[r][j >> 2], 8u * (j & 3u), 8u);
			const uint32_t sh = (k & 7u) << 2;
			const bool valid = k < 8u;
			const uint32_t d = s_pref[k & 7u] + s_base[k & 7u][tid] + __builtin_amdgcn_ubfe(run, sh, 4u);
			run += valid ? 1u << sh : 0u;
			s_out[lds_slot(tid * CV + j)] = valid ? d : SAMPO_COMPACT_NONE;
		}
		__syncthreads();
#pragma unroll
		for (uint32_t q = 0; q < CV / 4u; ++q) {
			const uint32_t e = q * (CT * 4u) + tid * 4u;
			const uint32_t a = lds_slot(e); /* e..e+3 share one 32-word row */
			const uint64_t i = t0 + r * CSUB + e;
			if (vec_out && i + 4u <= n) {
				*(uint4 *)(dest + i) = make_uint4(s_out[a], s_out[a + 1u], s_out[a + 2u], s_out[a + 3u]);
			} else {
#pragma unroll
				for (uint32_t c = 0; c < 4u; ++c)
					if (i + c < n) dest[i + c] = s_out[a + c];
			}
		}
		__syncthreads();
	}
}

__global__ void k_zero8(uint32_t *p)
{
	if (threadIdx.x < CK) p[threadIdx.x] = 0u;
}

/* cumulative percentages of keys 0..8 (8: dead) */
__constant__ uint8_t c_cut[9] = {20, 50, 58, 65, 75, 85, 90, 95, 100};

/* keys of a population: 5 % dead (key 8), the states weighted like a
 * mid-season mosquito population */
__global__ void k_gen_keys(uint8_t *keys, uint64_t n, uint64_t seed)
{
	const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1u); /* splitmix64 */
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	z ^= z >> 31;
	const uint32_t u = (uint32_t)(z >> 32) % 100u;
	uint8_t k = 0;
	while (u >= c_cut[k]) ++k;
	keys[i] = k; /* 0..7 states, 8 dead */
}

__thread char t_err[256];

int cfail(int code, const char *what, hipError_t e = hipSuccess)
{
	snprintf(t_err, sizeof t_err, "%s%s%s", what, e != hipSuccess ? ": " : "", e != hipSuccess ? hipGetErrorString(e) : "");
	fprintf(stderr, "sampo_compact: %s\n", t_err);
	return code;
}

#define CHK(x)                                              \
	do {                                                    \
		hipError_t e_ = (x);                                \
		if (e_ != hipSuccess) return cfail(SAMPO_E_HIP, #x, e_); \
	} while (0)

} // namespace

extern "C" uint64_t sampo_compact_workspace_bytes(uint64_t n)
{
	const uint64_t tiles = (n + CTILE - 1u) / CTILE, blocks = (tiles + CBLK - 1u) / CBLK;
	return (tiles + blocks) * CK * sizeof(uint32_t);
}

extern "C" int sampo_compact_index(const uint8_t *keys, uint64_t n, uint32_t *dest, uint32_t *counts, void *workspace,
                                   uint64_t workspace_bytes, void *stream)
{
	if (!counts || (n && (!keys || !dest || !workspace))) return cfail(SAMPO_E_ARG, "null argument");
	if (n >= (1ull << 32)) return cfail(SAMPO_E_ARG, "n must be below 2^32");
	hipStream_t s = (hipStream_t)stream;
	if (n == 0) {
		hipLaunchKernelGGL(k_zero8, dim3(1), dim3(64), 0, s, counts);
		CHK(hipGetLastError());
		return 0;
	}
	if (workspace_bytes < sampo_compact_workspace_bytes(n) || ((uintptr_t)workspace & 15u))
		return cfail(SAMPO_E_ARG, "workspace too small or not 16-byte aligned");
	const uint32_t tiles = (uint32_t)((n + CTILE - 1u) / CTILE);
	uint32_t *tc = (uint32_t *)workspace;
	const uint32_t vec_in = ((uintptr_t)keys & 15u) == 0u, vec_out = ((uintptr_t)dest & 15u) == 0u;
	const uint32_t blocks = (tiles + CBLK - 1u) / CBLK; /* at most 256 for n < 2^32 */
	uint32_t *bs = tc + (size_t)tiles * CK;
	hipLaunchKernelGGL(k_count, dim3((tiles + CCNT - 1u) / CCNT), dim3(CT), 0, s, keys, (uint32_t)n, tiles, tc, vec_in);
	hipLaunchKernelGGL(k_scan_local, dim3(blocks), dim3(CT), 0, s, tc, tiles, bs);
	hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(CT), 0, s, bs, blocks, counts);
	hipLaunchKernelGGL(k_rank, dim3(tiles), dim3(CT), 0, s, keys, (uint32_t)n, dest, (const uint32_t *)tc,
	                   (const uint32_t *)bs, vec_in, vec_out);
	CHK(hipGetLastError());
	return 0;
}

/* The workspace of sampo_compact_index_noalloc: per-tile counts and per-block
 * sums for SAMPO_COMPACT_NOALLOC_MAX_KEYS keys, one copy on each GPU. */
#define NOALLOC_TILES (SAMPO_COMPACT_NOALLOC_MAX_KEYS / CTILE)
#define NOALLOC_WORDS ((NOALLOC_TILES + (NOALLOC_TILES + CBLK - 1u) / CBLK) * CK)
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];
static_assert(NOALLOC_WORDS * 4ull == (NOALLOC_TILES + 16u) * 32u, "525 KB");

namespace {
constexpr int kMaxDevices = 64;
/* the device address of `scratch` on each device, looked up once per device
 * with that device current (the host-side name is not a device address) */
void *g_scratch[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];
std::once_flag g_scratch_once[kMaxDevices];

int scratch_on_current_device(void **out)
{
	int dev = -1;
	CHK(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices) return cfail(SAMPO_E_ARG, "device index beyond the scratch table");
	std::call_once(g_scratch_once[dev], [dev] {
		void *p = nullptr;
		g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		g_scratch[dev] = g_scratch_err[dev] == hipSuccess ? p : nullptr;
	});
	if (g_scratch_err[dev] != hipSuccess || !g_scratch[dev])
		return cfail(SAMPO_E_HIP, "hipGetSymbolAddress(scratch)", g_scratch_err[dev]);
	*out = g_scratch[dev];
	return 0;
}
} // namespace

extern "C" int sampo_compact_index_noalloc(const uint8_t *keys, uint64_t n, uint32_t *dest, uint32_t *counts,
                                           void *stream)
{
	if (n > SAMPO_COMPACT_NOALLOC_MAX_KEYS)
		return cfail(SAMPO_E_ARG, "n exceeds SAMPO_COMPACT_NOALLOC_MAX_KEYS (2^28) for the fixed workspace");
	void *ws = nullptr;
	const int rc = scratch_on_current_device(&ws);
	if (rc) return rc;
	return sampo_compact_index(keys, n, dest, counts, ws, (uint64_t)NOALLOC_WORDS * sizeof(uint32_t), stream);
}

extern "C" int sampo_compact_index_host(int device, const uint8_t *keys, uint64_t n, uint32_t *dest, uint32_t counts[8])
{
	if (!counts || (n && (!keys || !dest))) return cfail(SAMPO_E_ARG, "null argument");
	if (n >= (1ull << 32)) return cfail(SAMPO_E_ARG, "n must be below 2^32");
	CHK(hipSetDevice(device));
	uint8_t *dk = nullptr;
	uint32_t *dd = nullptr, *dc = nullptr;
	void *ws = nullptr;
	const uint64_t wsb = sampo_compact_workspace_bytes(n);
	hipError_t e = hipMalloc(&dc, CK * sizeof(uint32_t));
	if (e == hipSuccess && wsb) e = hipMalloc(&ws, wsb);
	if (e == hipSuccess && n) e = hipMalloc(&dk, n);
	if (e == hipSuccess && n) e = hipMalloc(&dd, n * sizeof(uint32_t));
	if (e == hipSuccess && n) e = hipMemcpy(dk, keys, n, hipMemcpyHostToDevice);
	int rc = e == hipSuccess ? 0 : cfail(SAMPO_E_NOMEM, "device allocation or upload failed", e);
	if (!rc) rc = sampo_compact_index(dk, n, dd, dc, ws, wsb, nullptr);
	if (!rc) {
		e = hipDeviceSynchronize();
		if (e == hipSuccess && n) e = hipMemcpy(dest, dd, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
		if (e == hipSuccess) e = hipMemcpy(counts, dc, CK * sizeof(uint32_t), hipMemcpyDeviceToHost);
		if (e != hipSuccess) rc = cfail(SAMPO_E_HIP, "compaction failed", e);
	}
	(void)hipFree(dk);
	(void)hipFree(dd);
	(void)hipFree(dc);
	(void)hipFree(ws);
	return rc;
}

extern "C" int sampo_compact_time(int device, uint64_t n, uint32_t iters, uint64_t seed, float *ms_per_call)
{
	if (!ms_per_call || n == 0 || n >= (1ull << 32) || iters == 0) return cfail(SAMPO_E_ARG, "bad argument");
	CHK(hipSetDevice(device));
	uint8_t *dk = nullptr;
	uint32_t *dd = nullptr, *dc = nullptr;
	void *ws = nullptr;
	hipStream_t s = nullptr;
	hipEvent_t a = nullptr, b = nullptr;
	const uint64_t wsb = sampo_compact_workspace_bytes(n);
	hipError_t e = hipMalloc(&dk, n);
	if (e == hipSuccess) e = hipMalloc(&dd, n * sizeof(uint32_t));
	if (e == hipSuccess) e = hipMalloc(&dc, CK * sizeof(uint32_t));
	if (e == hipSuccess) e = hipMalloc(&ws, wsb);
	if (e == hipSuccess) e = hipStreamCreate(&s);
	if (e == hipSuccess) e = hipEventCreate(&a);
	if (e == hipSuccess) e = hipEventCreate(&b);
	int rc = e == hipSuccess ? 0 : cfail(SAMPO_E_NOMEM, "device allocation failed", e);
	if (!rc) {
		hipLaunchKernelGGL(k_gen_keys, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, s, dk, n, seed);
		rc = sampo_compact_index(dk, n, dd, dc, ws, wsb, s);
	}
	if (!rc) {
		e = hipEventRecord(a, s);
		for (uint32_t i = 0; i < iters && !rc; ++i) rc = sampo_compact_index(dk, n, dd, dc, ws, wsb, s);
		if (e == hipSuccess) e = hipEventRecord(b, s);
		if (e == hipSuccess) e = hipEventSynchronize(b);
		float ms = 0.0f;
		if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
		if (e != hipSuccess && !rc) rc = cfail(SAMPO_E_HIP, "timing failed", e);
		*ms_per_call = ms / (float)iters;
	}
	if (a) (void)hipEventDestroy(a);
	if (b) (void)hipEventDestroy(b);
	if (s) (void)hipStreamDestroy(s);
	(void)hipFree(dk);
	(void)hipFree(dd);
	(void)hipFree(dc);
	(void)hipFree(ws);
	return rc;
}
