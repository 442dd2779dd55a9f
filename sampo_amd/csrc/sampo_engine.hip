/*
 * sampo_engine.hip -- fused, batched per-iteration agent step for gfx950.
 *
 * One iteration of the reference loop (src/abms.c:502-529) is two launches
 * on one stream:
 *
 *   k_pass  (persistent)     work-groups take chunks by ticket.  A chunk is a
 *                            contiguous run of the instance's agents in
 *                            reference index order; each agent is read once,
 *                            killed and updated (kernel/killAgents.cl:24-148,
 *                            kernel/updateAgents.cl:39-270) in registers,
 *                            ranked by (input state, stayed / moved on) with
 *                            64-lane ballots and an LDS scan, and written
 *                            once into its input block's slots in the other
 *                            buffer: stayers from the head, movers from the
 *                            tail.  No work-group waits on another;
 *   k_prep  (1 group per slot in use) folds the chunks' counts and sums
 *                            into the new segment list -- the compaction of
 *                            preScan + scan + oldToNewAgents
 *                            (kernel/preScan.cl, kernel/boltScan.cl,
 *                            kernel/oldToNewAgents.cl) -- computes the next
 *                            ranges, printStats row and calcStats sums
 *                            (kernel/calcL1de.cl, kernel/calcGender.cl), the
 *                            larval mortality table, and cuts the next
 *                            chunks.  A single slot plans the next pass from
 *                            its LDS record; with several, k_plan (one
 *                            work-group, launched after the prep) lays every
 *                            running instance's chunks end to end (plan_all).
 *
 * Survivors keep reference order: chunks are in old-index order, each run
 * (chunk c, input state, stayed / moved on) holds its agents in old-index
 * order, and the new state-s range is, chunk by chunk, the run moved in from
 * the previous state, the run that stayed (and for BMS the run moved in from
 * gravid) -- the stable rank the reference's scans compute.  New eggs are not
 * written (see sampo_model.h).
 */
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <type_traits>
#include <vector>

#include "sampo_model.h"
#include "sampo_wave.h"

/* ------------------------------------------------------------ device state */

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) void glob_void;
/* cache policy of a direct-to-LDS load: coherent at agent scope (sc1, the
 * lowering of an agent-scope atomic load) */
#define CPOL_SC1 16
/* Global-memory views of pointers read from LDS: accesses through them are
 * global (not flat) instructions, which the compiler knows cannot touch LDS
 * -- so they never make it wait for global memory before an LDS access, and
 * loads of one round trip are issued together. */
typedef __attribute__((address_space(1))) uint32_t gu32;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) u32x4 gu32x4;
template <class V> __device__ static inline __attribute__((address_space(1))) V *gmem(V *p)
{
	return (__attribute__((address_space(1))) V *)p;
}
/* Barrier for LDS traffic only: it waits for the wave's LDS instructions
 * (lgkmcnt) alone, not for its global loads and stores (a __syncthreads
 * waits for every store to complete).  For work-groups whose waves never
 * exchange data through global memory: the pass, whose direct-to-LDS loads
 * land in slots only the issuing lane reads after its own wait, so the next
 * tile's loads stay in flight across the current tile's barriers; and the
 * prep, whose global reads never follow a global write of the same launch
 * (a single instance is planned from its LDS record).  The empty asm
 * statements keep the compiler from moving memory accesses across it. */
__device__ static inline void lds_barrier()
{
	__asm__ __volatile__("" ::: "memory");
	__builtin_amdgcn_s_waitcnt(0xC07F); /* lgkmcnt(0) */
	__builtin_amdgcn_s_barrier();
	__asm__ __volatile__("" ::: "memory");
}

/* A pointer every lane holds the same value of, moved to scalar registers,
 * so that addressing is scalar base + 32-bit lane offset. */
__device__ static inline gu32 *uniform_ptr(const uint32_t *p)
{
	const uint64_t v = (uint64_t)p;
	const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
	const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
	return (gu32 *)(((uint64_t)hi << 32) | lo);
}

/* whole records through the global view, word by word */
template <class V> __device__ static inline void gput(V *dst, const V &v)
{
	static_assert(sizeof(V) % 4u == 0u, "word-sized records");
	gu32 *d = (gu32 *)dst;
	const uint32_t *s = (const uint32_t *)&v;
#pragma unroll
	for (unsigned i = 0; i < sizeof(V) / 4u; ++i) d[i] = s[i];
}
template <class V> __device__ static inline V gget(const V *src)
{
	static_assert(sizeof(V) % 4u == 0u, "word-sized records");
	V v;
	const gu32 *s = (const gu32 *)src;
	uint32_t *d = (uint32_t *)&v;
#pragma unroll
	for (unsigned i = 0; i < sizeof(V) / 4u; ++i) d[i] = s[i];
	return v;
}

#define PB_NEWBORN 0xffffffffu
#define SEG_REV 0x100u /* in Seg.state: the run is stored backwards from pbase */

struct Seg { /* a run of one state, stored contiguously (or the newborns) */
	uint32_t k0;     /* dense index of its first agent */
	uint32_t state;  /* state, | SEG_REV if agent k0 + r is at pbase - r */
	uint32_t pbase;  /* physical slot of its first agent, PB_NEWBORN for newborns */
	uint32_t vbase;  /* virtual index of its first agent */
};

__host__ __device__ static inline uint32_t seg_slot(uint32_t pbase, uint32_t state, uint32_t r)
{
	return (state & SEG_REV) ? pbase - r : pbase + r;
}

struct Chunk {
	uint32_t k0, k1; /* dense range: worked out by the pass from the iteration's cut, not stored */
	uint32_t seg0;   /* segment holding k0: the one field stored per chunk */
	uint32_t pad;
};

/* How an iteration's dense range is cut into chunks: nbig chunks of `big`
 * agents, nmid of `mid`, then chunks of `small` agents (equal chunks: nbig =
 * nmid = 0).  With few chunks per work-group, smaller last chunks let the
 * persistent work-groups of the pass finish together: one that is done
 * early takes another small one instead of the last big one holding the
 * whole launch. */
struct Cut {
	uint32_t big, mid, small, nbig, nmid;
};

__host__ __device__ static inline uint32_t cut_start(const Cut &k, uint32_t c)
{
	if (c <= k.nbig) return c * k.big;
	const uint32_t b1 = k.nbig * k.big;
	if (c <= k.nbig + k.nmid) return b1 + (c - k.nbig) * k.mid;
	return b1 + k.nmid * k.mid + (c - k.nbig - k.nmid) * k.small;
}

/* the first chunk starting at or after dense index x; cut_at_or_after(k, N)
 * is the number of chunks covering N agents */
__host__ __device__ static inline uint32_t cut_at_or_after(const Cut &k, uint32_t x)
{
	const uint32_t b1 = k.nbig * k.big, b2 = b1 + k.nmid * k.mid;
	if (x <= b1) return (x + k.big - 1u) / k.big;
	if (x <= b2) return k.nbig + (x - b1 + k.mid - 1u) / k.mid;
	return k.nbig + k.nmid + (x - b2 + k.small - 1u) / k.small;
}

/* Output layout.  The agents of chunk c that had state s at the start of the
 * iteration occupy the dense range [head, end) of that state within the
 * chunk; in the other buffer the same slot range receives their survivors:
 * those that keep state s from `head` upwards, those that move on to the
 * next state (s + 1, or BMS for gravid) from `end - 1` downwards, both in
 * old-index order.  Deaths leave the middle empty.  The next iteration's
 * state-s' range is, chunk by chunk, the run moved in from the previous
 * state, the run that stayed, and for BMS the run moved in from gravid --
 * the reference's stable order -- and an iteration needs exactly as many
 * output slots as it has agents.
 *
 * Per-chunk results, structure of arrays, cout[q * cmax + c]: q < 8 agents of
 * input state q that stayed, 8 + q agents of input state q that moved on,
 * then newEggs, bites, infectious bites, cycles, cycle sum, L1DE, females,
 * potentially infective. */
#define NOUT 24
#define CO_SUMS 16


struct IterState {
	uint32_t step, active, cur, n;
	uint32_t cnt[8], vstart[8], kstart[9];
	uint32_t nseg, nchunk, enew, nb_set;
	Cut cut;               /* chunks of this iteration */
	uint32_t seg0_done;    /* chunk seg0 already filled by the fold */
	uint32_t tb, h, seed_set;
	float temp;
	sampo_env env;
	float hmr_egg, hmr_pupa;
	/* the species functions (include/properties.h) at this iteration's
	 * temperature, environment and hour, from the instance's SpRow table */
	uint32_t open;         /* bit s: state s may leave now; bit 7: gravids may lay */
	float dev_rate;        /* larvaDevelopment(T) */
	float pupa_dly, imm_dly, bmd_dly;
	float nb_inc;          /* eggsIncubation at the newborns' laying temperature */
	float itn_kill, irs_kill; /* ITN x BITE_HUMAN x BITE_INDOOR, IRS x REST_INDOOR (updateAgents.cl:166,200) */
	/* species constants, written into both parities at admission; larva_* and
	 * batch_* are read as arrays through LDS pointers and must stay adjacent */
	float ms_dly, bms_dly, larva_mean, larva_sd, batch_mean, batch_sd, batch_decay, lay_p, bite_human;
	float hatch_thr[4], hatch_m[5], hatch_b[5];
	/* the iteration's 16 seeds (set seed_set) and seeds[2], seeds[3] of the
	 * set that laid the newborns (nb_set), four prepared sets of four
	 * (sm_taus_prepare) and two, copied here by the prep so that the pass
	 * reads them with the rest of this record */
	uint32_t seed[16], nbseed[8];
	double csd_inc;        /* 1/((111/(T-16))*14) of updateAgents.cl:69-70, in binary64 */
	float hmr_larva[100];
};

/* Species values of one iteration, evaluated on the host at admission from
 * its temperature, environment and hour of day: the k_prep work per
 * iteration stays one 32-byte load. */
struct SpRow {
	float dev_rate, pupa_dly, imm_dly, bmd_dly;
	float inc;             /* eggsIncubation(T), for the eggs laid in this iteration */
	float itn_kill, irs_kill;
	uint32_t open;
};

struct Inst {
	uint32_t *buf[2];      /* SoA: field f of buffer b at buf[b] + f * P */
	Seg *seg[2];           /* iteration with parity q reads seg[q] ... */
	Chunk *chunk[2];       /* ... and chunk[q] */
	uint32_t *cout;
	const uint32_t *seeds; /* (horizon + 1) x 4 sets of 4 seeds, each prepared by sm_taus_prepare */
	const float *temp;
	const sampo_env *env;
	sampo_row *rows;       /* horizon + 1 */
	const SpRow *sprow;    /* horizon */
	uint64_t max_agents;
	uint32_t P, horizon, mode, env_count;
	uint32_t P_max;        /* physical slots the buffers may grow to */
	uint32_t chunk_size, target_chunks, cmax, smax;
	uint32_t status, step, pending, needs_begin;
	uint32_t compact;      /* immature states store only their non-zero words (sm_fields) */
	uint32_t tile_class;   /* the pass launch (tile shape) that advances it */
	uint32_t even_cut;     /* cut into at most target_chunks equal chunks (day launches, chunk_cut) */
	uint32_t l1de, females, potinf, ncyc, scyc, bites, ibites; /* stats of the start of `step` */
	float hmr_stage[4];
	float hmr_adult[100];
	IterState st[2];       /* iteration `step` uses st[step & 1] */
};

/* Instances are admitted under one of up to MAX_CLASSES tile shapes of the
 * pass (work-group size, agents per thread, register cap); each iteration
 * launches one pass per shape, over the chunks of that shape's instances,
 * with a ticket counter and a chunk total of its own. */
#define MAX_CLASSES 4u

struct Ctx {
	uint32_t ticket[MAX_CLASSES];
	uint32_t total[MAX_CLASSES];
	unsigned long long agent_steps;
};

/* Day launches (k_day): one cooperative launch runs up to DAY_ITERS
 * iterations of a lone instance cut into at most DAY_CHUNKS equal chunks
 * (one launch per day stepped, up to four days per launch when several are:
 * successive cooperative launches are ~20 us apart).
 * Each work-group publishes its chunk's NOUT results as 64-bit words tagged
 * with the iteration's sequence number (high half) in the parity of the
 * iteration; every work-group waits for all of them and folds them itself.
 * A tag is the launch's first sequence number plus the iteration within the
 * launch, and the host never reuses one, so a word from another iteration or
 * launch never matches.  Every wait gives up after DAY_POLLS polls: it sets
 * `fault`, every other wait then gives up too, and the instance ends in
 * SAMPO_FAULT.
 *
 * Memory order inside a launch: the agents one iteration writes and the next
 * reads in another work-group, and the tagged words, all go through
 * agent-scope accesses (sc1: coherent across the XCDs' L2 caches), so no L2
 * write-back or invalidate is needed.  A work-group sends its tags only
 * after every wave's agent stores have completed (s_waitcnt and a barrier),
 * and reads agents only after it has seen every tag of the iteration. */
#define DAY_CHUNKS 256u
#define DAY_ITERS 96u /* four days: one launch per four days when several are stepped at once */
#define DAY_SEGMAX (16u * DAY_CHUNKS + 2u) /* segments of an iteration with <= DAY_CHUNKS chunks */
#define DAY_POLLS (1u << 20)

struct DaySync {
	unsigned long long w[2][NOUT][DAY_CHUNKS];
	uint32_t fault;
};

/* ------------------------------------------------------------ the agent */

struct Ag {
	uint32_t age, his;  /* ageInHours, hoursInState: whole hours below 2^16, as integers */
	float dly, csd;
	uint32_t aux, avail, cyc, misc; /* misc: humanBloodmealCount (0-15), numEggBatches (16-30), isFemale (31) */
};

/* the stored words of an agent (sampo_model.h) and back */
__host__ __device__ static inline void pack_agent(const Ag &a, uint32_t (&w)[NFIELD])
{
	w[F_AH] = a.age | (a.his << 16);
	w[F_DLY] = sc_bits_f(a.dly);
	w[F_AUX] = a.aux;
	w[F_CSD] = sc_bits_f(a.csd);
	w[F_AVAIL] = a.avail;
	w[F_CH] = (a.cyc & 0xffffu) | (a.misc << 16);
	w[F_NF] = ((a.misc >> 16) & 0x7fffu) | ((a.misc >> 31) << 15);
}

__host__ __device__ static inline void unpack_agent(const uint32_t (&w0)[NFIELD], Ag &a, uint32_t fields = SM_ALL_FIELDS)
{
	uint32_t w[NFIELD];
#pragma unroll
	for (int f = 0; f < NFIELD; ++f) w[f] = (fields >> f) & 1u ? w0[f] : 0u; /* words not stored are zero */
	a.age = w[F_AH] & 0xffffu;
	a.his = w[F_AH] >> 16;
	a.dly = sc_float(w[F_DLY]);
	a.aux = w[F_AUX];
	a.csd = sc_float(w[F_CSD]);
	a.avail = w[F_AVAIL];
	a.cyc = w[F_CH] & 0xffffu;
	a.misc = (w[F_CH] >> 16) | ((w[F_NF] & 0x7fffu) << 16) | ((w[F_NF] >> 15) << 31);
}

/* the event sums of agent_step, one LDS word each per wave: eggs laid,
 * bites, infective bites, completed cycles, their summed lengths */
enum { EV_EGGS, EV_BITES, EV_IBITES, EV_NCYC, EV_SCYC, NEV };
typedef __attribute__((address_space(3))) uint32_t lds_u32;

/* per-chunk copy of the instance's iteration parameters */
struct TileParams {
	IterState ip;
	float hmr_adult[100];
	uint32_t *bin, *bout;
	const Seg *seg;
	const Chunk *chk;     /* the instance's chunks of this iteration */
	uint32_t *cout;
	Chunk ch;
	uint32_t P, nseg, cmax, compact;
};

/* HybridTaus of virtual index g with a prepared seed set (sm_taus_prepare) */
__device__ static inline float rnd(const uint32_t *z, uint32_t g)
{
	return sm_taus_draw(z, g);
}

/* The transcendental paths run only on state transitions; keeping them out
 * of line keeps their binary64 temporaries out of the pass's register budget.
 * They read their species parameters from the LDS copy of the iteration
 * state through one 32-bit LDS address: passing the values themselves costs
 * the caller registers across the call (measured: 2% on the pass). */
typedef const __attribute__((address_space(3))) float *lds_fp;

/* larvaDelay: p = {mean, sd} */
__device__ __attribute__((noinline)) float bm_delay(float u0, float u1, lds_fp p)
{
	return sc_box_muller(u0, u1, p[0], p[1]);
}

/* generateEggs: p = {mean, sd, decay} */
__device__ __attribute__((noinline)) uint32_t gen_eggs(float u0, float u1, uint32_t nb, lds_fp p)
{
	return sm_egg_batch(u0, u1, p[0], p[1], p[2], nb);
}

__device__ static inline void ev_add(lds_u32 *ev, uint32_t i, uint32_t v)
{
	__hip_atomic_fetch_add(&ev[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

/* killAgents + updateAgents for one agent in state s at virtual index g.
 * Returns the new state, or S_DEAD.  The rare events add to the wave's LDS
 * sums ev (registers held across the tile loop for them were spilled, and
 * every spill reload waited for the next tile's loads in flight). */
__device__ static inline uint32_t agent_step(Ag &a, uint32_t s, uint32_t g, const TileParams &T, lds_u32 *ev)
{
	const IterState &ip = T.ip;
	/* killAgents.cl:55-147.  The age in days, (uint)min(age / 24.0f, 99.0f):
	 * ages are whole hours below 2^16 (admission checks it), for which the
	 * correctly rounded quotient truncates to the integer one (checked for
	 * every such age), so it is one integer division by a constant instead
	 * of a float division sequence per agent.  The hour counts are binary32
	 * in the reference; whole numbers below 2^16, they are exact as integers,
	 * and (float)his is the reference's value wherever it meets a delay. */
	const uint32_t dq = a.age / 24u;
	const uint32_t d = dq < 99u ? dq : 99u;
	float hmr = s == S_EGG ? ip.hmr_egg : s == S_LARVA ? ip.hmr_larva[d] : s == S_PUPA ? ip.hmr_pupa
	                                                                                 : T.hmr_adult[d];
	if (rnd(&T.ip.seed[0], g) <= hmr) return S_DEAD;

	/* updateAgents.cl:62-70 */
	const float t = ip.temp;
	a.age += 1u;
	a.his += 1u;
	const float his = (float)a.his;
	uint32_t hbc = a.misc & 0xffffu;
	/* updateAgents.cl:69-70: csd + cond * (1/((111/(T-16))*14)) in binary64; the
	 * increment depends on T only (evaluated once per iteration).  T is
	 * finite, so the increment is finite or a zero, with the sign of T - 16
	 * (T - 16 is exact and a zero only at 16, where the increment is +0).
	 * Without the condition the addend 0 * increment is a zero of that sign,
	 * and adding it in binary64 and rounding back is adding the same zero in
	 * binary32, bit for bit (-0 + -0 = -0, -0 + +0 = +0, NaN quieted), so
	 * the binary64 work is only done for the agents that have bitten. */
	if (t > 16.0f && hbc > 0u)
		a.csd = (float)((double)a.csd + ip.csd_inc);
	else
		a.csd = a.csd + __uint_as_float(__float_as_uint(t - 16.0f) & 0x80000000u);
	const uint32_t open = ip.open;
	switch (s) {
	case S_EGG: /* :72-87 */
		if (his >= a.dly && (open & (1u << S_EGG))) {
			a.age = 24u;
			a.his = 0u;
			a.dly = bm_delay(rnd(&T.ip.seed[4], g), rnd(&T.ip.seed[8], g), (lds_fp)&ip.larva_mean);
			return S_LARVA;
		}
		return S_EGG;
	case S_LARVA: { /* :89-113 */
		if (ip.env.larvacideValue > 0.0f && a.his == 1u && rnd(&T.ip.seed[8], g) <= ip.env.larvacideValue)
			return S_DEAD;
		float c = sc_float(a.aux) + ip.dev_rate;
		a.aux = sc_bits_f(c);
		if (c >= a.dly && (open & (1u << S_LARVA))) {
			a.his = 0u;
			a.dly = ip.pupa_dly;
			return S_PUPA;
		}
		return S_LARVA;
	}
	case S_PUPA: /* :115-129 */
		if (his >= a.dly && (open & (1u << S_PUPA))) {
			a.age = 0u;
			a.his = 0u;
			a.dly = ip.imm_dly;
			return S_IMM;
		}
		return S_PUPA;
	case S_IMM: /* :131-143 */
		if (his >= a.dly && (open & (1u << S_IMM))) {
			a.dly = ip.ms_dly;
			a.his = 0u;
			return S_MS;
		}
		return S_IMM;
	case S_MS: /* :145-157 */
		if (his >= a.dly && (a.misc >> 31) && (open & (1u << S_MS))) {
			a.his = 0u;
			a.dly = ip.bms_dly;
			return S_BMS;
		}
		return S_MS;
	case S_BMS: /* :159-197 */
		a.cyc += 1u;
		if (his >= a.dly && (open & (1u << S_BMS)) && rnd(&T.ip.seed[4], g) <= ip.env.bloodmealSuccess) {
			const float itn = ip.itn_kill;
			if (itn > 0.0f && rnd(&T.ip.seed[8], g) <= itn) return S_DEAD;
			const uint32_t human = ip.bite_human < 1.0f ? (rnd(&T.ip.seed[12], g) <= ip.bite_human) : 1u;
			a.misc = (a.misc & 0xffff0000u) | ((hbc + human) & 0xffffu);
			ev_add(ev, EV_BITES, 1u);
			a.his = 0u;
			a.dly = ip.bmd_dly;
			if (a.csd >= 1.0f && human) ev_add(ev, EV_IBITES, 1u);
			return S_BMD;
		}
		return S_BMS;
	case S_BMD: { /* :199-228 */
		const float irs = ip.irs_kill;
		if (irs > 0.0f && rnd(&T.ip.seed[8], g) <= irs) return S_DEAD;
		a.cyc += 1u;
		if (his >= a.dly && (open & (1u << S_BMD))) {
			if (a.avail == 0u && (a.misc >> 31)) { /* generateEggs, properties.h:257-266 */
				uint32_t nb = (a.misc >> 16) & 0x7fffu;
				a.avail = gen_eggs(rnd(&T.ip.seed[4], g), rnd(&T.ip.seed[8], g), nb, (lds_fp)&ip.batch_mean);
				a.misc = (a.misc & 0x8000ffffu) | (((nb + 1u) & 0x7fffu) << 16);
			}
			a.his = 0u;
			a.aux = 0u;
			return S_GRAVID;
		}
		return S_BMD;
	}
	default: { /* gravid :230-269 */
		a.cyc += 1u;
		if ((open & (1u << S_GRAVID)) && rnd(&T.ip.seed[4], g) < ip.lay_p) {
			int dead = 0;
			if (ip.env.oviTrapValue > 0.0f && rnd(&T.ip.seed[8], g) <= ip.env.oviTrapValue) {
				dead = 1;
				a.avail = 0u;
			}
			if (!dead) { /* layEggs, properties.h:296-307 */
				float nbm = (float)ip.tb / (((float)a.aux + 1.0f) * ip.env.carryingCapacity);
				float pbm = sc_maxf(0.0f, 1.0f - nbm);
				uint32_t pe = sc_f2u(sc_ceilf((float)a.avail * pbm));
				ev_add(ev, EV_EGGS, pe);
				a.avail -= pe;
			}
			a.aux += 1u;
			if (a.avail == 0u) {
				ev_add(ev, EV_NCYC, 1u);
				ev_add(ev, EV_SCYC, a.cyc);
				a.cyc = 0u;
				a.his = 0u;
				a.dly = ip.bms_dly;
				return dead ? S_DEAD : S_BMS;
			}
		}
		return S_GRAVID;
	}
	}
}

__device__ static inline void load_agent(Ag &a, const uint32_t *b, size_t P, uint32_t p, uint32_t fields)
{
	uint32_t w[NFIELD];
#pragma unroll
	for (int f = 0; f < NFIELD; ++f) w[f] = (fields >> f) & 1u ? b[(size_t)f * P + p] : 0u;
	unpack_agent(w, a, fields);
}

/* initAgents.cl:34-64 for newborn egg g, laid with seeds z (sets 2 and 3,
 * prepared by sm_taus_prepare);
 * inc is eggsIncubation at the laying temperature, thr/m/b the eggsHatching
 * bands */
__device__ static inline void newborn(Ag &a, const uint32_t *z, float inc, const float *thr, const float *m,
                                      const float *b, uint32_t g)
{
	a.age = 0u;
	a.his = 0u;
	a.csd = 0.0f;
	a.aux = 0u;
	a.avail = 0u;
	a.cyc = 0u;
	float u3 = sm_taus_draw(z + 4, g);
	float u2 = sm_taus_draw(z, g);
	a.misc = ((double)u3 <= 0.5) ? 0x80000000u : 0u;
	a.dly = inc + sm_hatching(thr, m, b, u2);
}

__device__ static inline uint32_t wave_sum(uint32_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

/* ------------------------------------------------------------ plan */

/* Lay every running instance's chunks end to end for the next pass, per
 * tile class c: pl_base[c * (ninst + 1) + i] = first ticket of instance i
 * among class c's chunks (instances of other classes count no chunks),
 * ctx->total[c] = class c's chunks; reset the tickets.  Run by one
 * work-group of 1024 threads: k_plan after the prep of several slots, after
 * a release, or k_hold.  The instances' fields are read with atomic loads at
 * agent scope. */
__device__ static void plan_all(Inst *inst, Ctx *ctx, uint32_t *pl_base, uint8_t *pl_par, uint32_t ninst,
                                uint32_t ncls)
{
	__shared__ uint32_t s_scan[1024];
	const uint32_t tid = threadIdx.x;
	auto chunks_of = [&](uint32_t i, uint32_t c) -> uint32_t {
		Inst *I = inst + i;
		const uint32_t status = __hip_atomic_load(&I->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		const uint32_t par = __hip_atomic_load(&I->step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1u;
		const uint32_t active = __hip_atomic_load(&I->st[par].active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		const uint32_t cl = __hip_atomic_load(&I->tile_class, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		pl_par[i] = (uint8_t)par;
		return status == SAMPO_RUNNING && active && cl == c
		           ? __hip_atomic_load(&I->st[par].nchunk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
		           : 0u;
	};
	for (uint32_t c = 0; c < ncls; ++c) {
		uint32_t *pb = pl_base + c * (ninst + 1u);
		if (ninst <= 64u) { /* one wave, no barriers */
			if (tid < 64u) {
				const uint32_t nc = tid < ninst ? chunks_of(tid, c) : 0u;
				uint32_t x = nc;
				for (int o = 1; o < 64; o <<= 1) {
					uint32_t y = __shfl_up(x, o, 64);
					if (tid >= (uint32_t)o) x += y;
				}
				if (tid < ninst) pb[tid] = x - nc;
				const uint32_t all = __shfl(x, 63, 64);
				if (tid == 0) {
					pb[ninst] = all;
					ctx->total[c] = all;
					ctx->ticket[c] = 0u;
				}
			}
			continue;
		}
		uint32_t carry = 0;
		for (uint32_t b = 0; b < ninst; b += 1024u) {
			uint32_t i = b + tid;
			uint32_t nc = 0;
			if (i < ninst) nc = chunks_of(i, c);
			s_scan[tid] = nc;
			__syncthreads();
			for (uint32_t off = 1; off < 1024u; off <<= 1) {
				uint32_t v = tid >= off ? s_scan[tid - off] : 0u;
				__syncthreads();
				s_scan[tid] += v;
				__syncthreads();
			}
			if (i < ninst) pb[i] = carry + s_scan[tid] - nc;
			carry += s_scan[1023];
			__syncthreads();
		}
		if (tid == 0) {
			pb[ninst] = carry;
			ctx->total[c] = carry;
			ctx->ticket[c] = 0u;
		}
	}
}

__global__ void __launch_bounds__(1024) k_plan(Inst *inst, Ctx *ctx, uint32_t *pl_base, uint8_t *pl_par,
                                               uint32_t ninst, uint32_t ncls)
{
	plan_all(inst, ctx, pl_base, pl_par, ninst, ncls);
}

/* Hold (hold[i] != 0) or resume instances 0..n-1, then plan the next pass.
 * Only running instances are held and only held ones resumed. */
__global__ void __launch_bounds__(1024) k_hold(Inst *inst, const uint8_t *hold, uint32_t n, Ctx *ctx,
                                               uint32_t *pl_base, uint8_t *pl_par, uint32_t ninst, uint32_t ncls)
{
	for (uint32_t i = threadIdx.x; i < n; i += 1024u) {
		const uint32_t st = __hip_atomic_load(&inst[i].status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		uint32_t to = st;
		if (hold[i] && st == SAMPO_RUNNING) to = SAMPO_HELD;
		if (!hold[i] && st == SAMPO_HELD) to = SAMPO_RUNNING;
		if (to != st) __hip_atomic_store(&inst[i].status, to, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
	__syncthreads();
	plan_all(inst, ctx, pl_base, pl_par, ninst, ncls);
}

/* ------------------------------------------------------------ k_prep */

#define PREP_WG 1024

/* The default least agents per chunk: small enough that the last chunks
 * handed out end the launch within a few tens of microseconds of each
 * other (a sweep of 4096 / 3072 / 2560 / 2048 / 1536 over 30 days of the
 * 32M-agent run: 3072 fastest, 1 % ahead of 4096) */
#define DEFAULT_CHUNK 3072u

/* Chunks of N agents: at least chunk_size agents each, about target_chunks
 * of them.  When that makes chunks of 3 x chunk_size or more, the first
 * target/2 chunks (one per work-group of the pass) take 80 % of the agents;
 * of the rest, target/8 chunks are of one size and up to 3 x target/4 of
 * half that size, the last ones handed out being the smallest: the launch
 * ends about when the last small chunk does. */
__device__ __attribute__((noinline)) static Cut chunk_cut(const Inst *I, uint32_t N)
{
	const uint32_t cs = I->chunk_size, T = I->target_chunks;
	if (I->even_cut) { /* equal chunks of a multiple of 64 agents, at most T of them */
		const uint32_t l = sm_roundup((uint32_t)(((uint64_t)N + T - 1u) / T), 64u);
		const uint32_t L = l > 64u ? l : 64u;
		return Cut{L, L, L, 0u, 0u};
	}
	Cut k = {cs, cs, cs, 0u, 0u};
	if (!T) return k;
	const uint32_t want = sm_roundup((uint32_t)(((uint64_t)N + T - 1u) / T), 256u);
	if (want <= cs) return k;
	k.big = k.mid = k.small = want;
	if (want >= 3u * cs && T >= 8u) { /* equal chunks at 3.75 x: the launch's last 20 % ran part-empty */
		const uint32_t nb = T / 2u, n1 = T / 8u, n2 = 2u * (T - nb - n1);
		const uint32_t big = sm_roundup((uint32_t)(((uint64_t)N * 4u + 5u * nb - 1u) / (5u * (uint64_t)nb)), 256u);
		if ((uint64_t)nb * big >= N) return k; /* rounding left no tail: equal chunks */
		const uint64_t rest = N - (uint64_t)nb * big;
		/* n1 chunks of m and n2 of m / 2 cover rest: m = rest / (n1 + n2 / 2) */
		uint32_t m = sm_roundup((uint32_t)((2u * rest + 2u * n1 + n2 - 1u) / (2u * (uint64_t)n1 + n2)), 512u);
		if (m < 2u * cs) m = 2u * cs;
		k.big = big;
		k.nbig = nb;
		k.mid = m;
		k.nmid = n1;
		k.small = m / 2u;
	}
	return k;
}

/* Exclusive scan of N values per thread over the work-group, all at once;
 * the totals land in tot (shared).  Waves scan with DPP and post their totals
 * to LDS; after one barrier, lane i < N of every wave adds up column i of
 * the waves before its own (its offset) and of all waves (the total), and
 * the offsets reach the wave's lanes by readlane.  Every column of every
 * wave is scanned, without branches, so the N DPP chains interleave (a
 * branch or a store between two would end each in a block of its own);
 * columns a wave holds no values in are zeros and scan to zeros. */
template <int N, int NTH = PREP_WG>
__device__ static inline void block_excl_n(const uint32_t (&v)[N], uint32_t (&ex)[N], uint32_t *tot)
{
	constexpr uint32_t NW = NTH / 64u;
	static_assert(N <= 64, "one lane per column");
	__shared__ uint32_t lds[NW][N];
	const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
	uint32_t x[N];
#pragma unroll
	for (int i = 0; i < N; ++i) x[i] = wave_incl_scan(v[i]);
	if (lane == 63u) { /* the wave totals */
#pragma unroll
		for (int i = 0; i < N; ++i) lds[w][i] = x[i];
	}
#pragma unroll
	for (int i = 0; i < N; ++i) ex[i] = x[i] - v[i];
	lds_barrier();
	uint32_t off = 0u, all = 0u;
	if (lane < (uint32_t)N) {
#pragma unroll
		for (uint32_t q = 0; q < NW; ++q) {
			const uint32_t y = lds[q][lane];
			all += y;
			off += q < w ? y : 0u;
		}
		if (w == 0) tot[lane] = all;
	}
#pragma unroll
	for (int i = 0; i < N; ++i) ex[i] += (uint32_t)__builtin_amdgcn_readlane((int)off, i);
	lds_barrier();
}

/* The runs of one chunk in the next iteration's dense order: for each new
 * state s', the run moved in from the previous state, the run that stayed,
 * and for BMS the run moved in from gravid.  Their counter rows in cout (q
 * stayed in input state q, 8 + q moved on from input state q): */
__device__ static inline uint32_t runs_of(uint32_t sn) { return sn == 0u ? 1u : (sn == (uint32_t)S_BMS ? 3u : 2u); }
__device__ static inline uint32_t run_row(uint32_t sn, uint32_t kind)
{
	return sn == 0u ? 0u : (kind == 0u ? 8u + sn - 1u : (kind == 1u ? sn : 8u + S_GRAVID));
}

/* The serial part of a fold: from the totals tot (16 per-state counts of
 * agents and of non-empty runs, then the 8 sums) of the finished iteration
 * ip, the next one's counts, ranges, segment count, parity, step and chunk
 * cut into np, the calcStats sums into I, and per new state its dense start
 * kst, virtual start vst and first segment index sb.  Returns its agents. */
__device__ static uint32_t fold_totals(Inst *I, const IterState &ip, IterState &np, const uint32_t *tot,
                                       uint32_t *kst, uint32_t *vst, uint32_t *sb)
{
	uint32_t t[24]; /* all reads first: the writes below may alias tot as far as the compiler knows */
#pragma unroll
	for (uint32_t i = 0; i < 24u; ++i) t[i] = tot[i];
	const uint32_t *sum = t + 16; /* newEggs, bites, ..., potentially infective */
	const uint32_t E = sum[0];
	uint32_t k = 0, vend = 0, b = E > 0u ? 1u : 0u;
#pragma unroll
	for (uint32_t s = 0; s < 8u; ++s) {
		const uint32_t c = t[s] + (s == 0u ? E : 0u);
		np.cnt[s] = c;
		kst[s] = k;
		k += c;
		const uint32_t vs = s == 0 ? 0u : ((vend + SAMPO_PAD) / SAMPO_PAD) * SAMPO_PAD;
		np.vstart[s] = vs;
		vst[s] = vs;
		vend = vs + c;
		sb[s] = b;
		b += t[8 + s];
	}
	kst[8] = k;
	np.nseg = b;
	np.enew = E;
	np.cur = ip.cur ^ 1u;
	np.step = ip.step + 1u;
	I->bites = sum[1];
	I->ibites = sum[2];
	I->ncyc = sum[3];
	I->scyc = sum[4];
	I->l1de = sum[5];
	I->females = sum[6];
	I->potinf = sum[7];
	np.cut = chunk_cut(I, k);
	np.seg0_done = 1u;
	return k;
}

/* Stage A: fold the finished iteration of I (parity par) into the segment
 * list, ranges and sums of the next one (parity par ^ 1).  Thread t owns
 * chunks [tK, tK + K): it loads their 16 run counts and 8 sums (consecutive
 * lanes, consecutive chunks), one scan over the work-group gives each
 * thread's offset in every new state's range, and each thread writes its own
 * chunks' segments.  Only non-empty runs get a segment, so a tile of the
 * pass never overlaps more segments than it has agents (plus one); the scan
 * also counts each thread's non-empty runs per new state.
 *
 * With at most 512 chunks the work-group has threads to spare: TPC = 2, 4 or
 * 8 parts of CW threads (whole waves) each take every chunk, part p only the
 * new states s' with s' % TPC == p (and part 0 the sums).  A state's runs of
 * a chunk are still all in one thread, so the scan's offsets are unchanged,
 * and each thread writes a TPC-th of the segments. */
__device__ static void fold_iteration(Inst *I, uint32_t par)
{
	constexpr uint32_t KR = 2; /* chunks per thread held in registers */
	__shared__ uint32_t s_kst[9], s_vst[8], s_sb[8], s_tot[24], s_C;
	__shared__ Cut s_cut;
	const uint32_t tid = threadIdx.x;
	const IterState &ip = I->st[par];
	IterState &np = I->st[par ^ 1u];
	const uint32_t C = ip.nchunk, cmax = I->cmax;
	const gu32 *co = uniform_ptr(I->cout); /* NOUT * cmax words: 32-bit indices */
	auto *nch = gmem(I->chunk[par ^ 1u]);
	Seg *seg = I->seg[par ^ 1u]; /* written with gput */
	const uint32_t K = (C + PREP_WG - 1u) / PREP_WG;
	const uint32_t CW = (C + 63u) & ~63u; /* threads of one part: whole waves */
	const uint32_t spare = CW ? PREP_WG / CW : 1u;
	const uint32_t TPC = spare >= 8u ? 8u : (spare >= 4u ? 4u : (spare >= 2u ? 2u : 1u));
	uint32_t c_lo, c_hi, own; /* this thread's chunks and new states (bit s') */
	if (TPC == 1u) {
		c_lo = tid * K < C ? tid * K : C;
		c_hi = c_lo + K < C ? c_lo + K : C;
		own = 0xffu;
	} else {
		const uint32_t p = tid / CW, c = tid - p * CW;
		const bool has = p < TPC && c < C;
		c_lo = has ? c : C;
		c_hi = has ? c + 1u : C;
		own = 0u;
		for (uint32_t sn = p; sn < 8u; sn += TPC) own |= 1u << sn;
	}
	const bool sums = TPC == 1u || tid < CW; /* part 0 (or every thread) adds the chunk sums */
	const bool in_regs = K <= KR;
	auto count = [&](uint32_t row, uint32_t c) -> uint32_t { return co[row * cmax + c]; };
	/* loads: the run counts of this thread's chunks (when few) and their sums.
	 * One 16-entry array per chunk: read by a uniform index, each stays in
	 * registers. */
	static_assert(KR == 2, "two register-held chunks per thread");
	uint32_t nr0[16], nr1[16];
	uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	if (in_regs) { /* all of this thread's loads in flight at once: one round trip.  They are
	                  unconditional (column 0 stands in past c_hi) and masked afterwards */
#pragma unroll
		for (uint32_t j = 0; j < KR; ++j) {
			const uint32_t cj = c_lo + j < c_hi ? c_lo + j : 0u;
#pragma unroll
			for (uint32_t r = 0; r < 16u; ++r) (j ? nr1 : nr0)[r] = count(r, cj); /* all rows: loads under a row mask measured slower */
			if (sums) {
#pragma unroll
				for (uint32_t q = 0; q < 8u; ++q) v[q] += c_lo + j < c_hi ? co[(CO_SUMS + q) * cmax + cj] : 0u;
			}
		}
#pragma unroll
		for (uint32_t r = 0; r < 16u; ++r) {
			nr0[r] = c_lo < c_hi ? nr0[r] : 0u;
			nr1[r] = c_lo + 1u < c_hi ? nr1[r] : 0u;
		}
	} else {
		for (uint32_t c = c_lo; c < c_hi; ++c) {
#pragma unroll
			for (uint32_t q = 0; q < 8u; ++q) v[q] += co[(CO_SUMS + q) * cmax + c];
		}
	}
	/* per new state: agents (wv[sn]) and non-empty runs (wv[8 + sn]) of this
	 * thread's chunks; then the 8 sums, scanned along for their totals */
	uint32_t wv[24];
#pragma unroll
	for (int i = 0; i < 24; ++i) wv[i] = i < 16 ? 0u : v[i - 16];
	if (in_regs) {
#pragma unroll
		for (uint32_t j = 0; j < KR; ++j)
#pragma unroll
			for (uint32_t sn = 0; sn < 8u; ++sn)
#pragma unroll
				for (uint32_t kind = 0; kind < 3u; ++kind) {
					if (kind >= runs_of(sn)) continue;
					const uint32_t n = (own >> sn) & 1u ? (j ? nr1 : nr0)[run_row(sn, kind)] : 0u; /* zero past c_hi */
					wv[sn] += n;
					wv[8u + sn] += n ? 1u : 0u;
				}
	} else {
		for (uint32_t c = c_lo; c < c_hi; ++c)
#pragma unroll
			for (uint32_t sn = 0; sn < 8u; ++sn)
#pragma unroll
				for (uint32_t kind = 0; kind < 3u; ++kind) {
					if (kind >= runs_of(sn)) continue;
					const uint32_t n = count(run_row(sn, kind), c);
					wv[sn] += n;
					wv[8u + sn] += n ? 1u : 0u;
				}
	}
	uint32_t ex[24];
	block_excl_n<24>(wv, ex, s_tot);
	const uint32_t E = s_tot[16]; /* newEggs */
	if (tid == 0) {
		const uint32_t k = fold_totals(I, ip, np, s_tot, s_kst, s_vst, s_sb);
		const uint32_t nc = cut_at_or_after(np.cut, k);
		s_cut = np.cut;
		s_C = nc < cmax ? nc : cmax;
	}
	lds_barrier();
	const Cut cut = s_cut, oc = ip.cut; /* chunks of the next and of the finished iteration */
	const uint32_t NC = s_C, on = ip.n;
	/* segments: the newborns, then per new state the non-empty runs of
	 * chunks 0, 1, ...  Each new chunk's first segment is recorded as its
	 * start is passed. */
	if (E > 0u) {
		if (tid == 0) {
			const Seg sg = {0u, 0u, PB_NEWBORN, 0u};
			gput(&seg[0], sg);
		}
		for (uint32_t c = tid; c < NC && cut_start(cut, c) < E; c += PREP_WG) nch[c].seg0 = 0u;
	}
	if (c_lo >= c_hi) return;
	/* one run: its segment at dense k0, index idx, and the chunk starts it covers */
	gu32x4 *segv = (gu32x4 *)seg;
	auto emit = [&](uint32_t sn, uint32_t row, uint32_t c, uint32_t n, uint32_t &k0, uint32_t &idx, uint32_t &nxc,
	                uint32_t &nxk) {
		const uint32_t src = row & 7u, mv = row >> 3;
		/* the source block: chunk c's part of the state-src range */
		const uint32_t ck0 = cut_start(oc, c), ce = cut_start(oc, c + 1u), ck1 = ce < on ? ce : on;
		const uint32_t bh = ip.kstart[src] > ck0 ? ip.kstart[src] : ck0;
		const uint32_t be = ip.kstart[src + 1u] < ck1 ? ip.kstart[src + 1u] : ck1;
		segv[idx] = u32x4{k0, sn | (mv ? SEG_REV : 0u), mv ? be - 1u : bh, s_vst[sn] - s_kst[sn] + k0}; /* Seg */
		for (; nxc < NC && nxk < k0 + n; ++nxc, nxk = cut_start(cut, nxc)) nch[nxc].seg0 = idx;
		++idx;
		k0 += n;
	};
	if (in_regs) {
#pragma unroll
		for (uint32_t sn = 0; sn < 8u; ++sn) {
			if (!((own >> sn) & 1u)) continue;
			uint32_t k0 = s_kst[sn] + (sn == 0u ? E : 0u) + ex[sn];
			uint32_t idx = s_sb[sn] + ex[8u + sn];
			uint32_t nxc = cut_at_or_after(cut, k0), nxk = cut_start(cut, nxc); /* next new chunk start at or after k0 */
#pragma unroll
			for (uint32_t j = 0; j < KR; ++j) {
				if (c_lo + j >= c_hi) continue;
#pragma unroll
				for (uint32_t kind = 0; kind < 3u; ++kind) {
					if (kind >= runs_of(sn)) continue;
					const uint32_t row = run_row(sn, kind);
					const uint32_t n = (j ? nr1 : nr0)[row];
					if (n) emit(sn, row, c_lo + j, n, k0, idx, nxc, nxk);
				}
			}
		}
	} else {
		for (uint32_t sn = 0; sn < 8u; ++sn) {
			uint32_t k0 = s_kst[sn] + (sn == 0u ? E : 0u) + ex[sn];
			uint32_t idx = s_sb[sn] + ex[8u + sn];
			uint32_t nxc = cut_at_or_after(cut, k0), nxk = cut_start(cut, nxc);
			for (uint32_t c = c_lo; c < c_hi; ++c)
				for (uint32_t kind = 0; kind < runs_of(sn); ++kind) {
					const uint32_t row = run_row(sn, kind);
					const uint32_t n = count(row, c);
					if (n) emit(sn, row, c, n, k0, idx, nxc, nxk);
				}
		}
	}
}

/* The per-iteration inputs begin_iteration reads from global memory, as
 * words: gathered one word per lane of a wave before the fold, so that their
 * round trip overlaps the fold's (BIN_* are word offsets). */
enum {
	BIN_TEMP = 0,                         /* temp[t] */
	BIN_SR = 1,                           /* sprow[t] */
	BIN_INC = BIN_SR + sizeof(SpRow) / 4, /* sprow[t - 1].inc: the newborns' laying temperature */
	BIN_ENV = BIN_INC + 1,                /* env[t] (env[0] when there is one) */
	BIN_SEED = BIN_ENV + sizeof(sampo_env) / 4, /* seed set step + 1 */
	BIN_NBSEED = BIN_SEED + 16,           /* seeds[2], seeds[3] of set step */
	BIN_WORDS = BIN_NBSEED + 8
};
static_assert(BIN_WORDS <= 64, "one lane per word");

/* word q of the inputs of iteration `step` (the seeds only while step <
 * horizon: both sets exist) */
__device__ static inline uint32_t begin_input(const Inst *I, uint32_t step, uint32_t q)
{
	const uint32_t t = step < I->horizon ? step : I->horizon - 1u;
	const uint32_t tp = step == 0u ? 0u : (step - 1u < I->horizon ? step - 1u : I->horizon - 1u);
	const uint32_t te = I->env_count == 1u ? 0u : t;
	const gu32 *p = nullptr;
	if (q == BIN_TEMP)
		p = (const gu32 *)I->temp + t;
	else if (q < BIN_INC)
		p = (const gu32 *)(I->sprow + t) + (q - BIN_SR);
	else if (q == BIN_INC)
		p = (const gu32 *)(I->sprow + tp) + offsetof(SpRow, inc) / 4u;
	else if (q < BIN_SEED)
		p = (const gu32 *)(I->env + te) + (q - BIN_ENV);
	else if (q < BIN_NBSEED)
		p = step < I->horizon ? (const gu32 *)I->seeds + (size_t)(step + 1u) * 16u + (q - BIN_SEED) : nullptr;
	else if (q < BIN_WORDS)
		p = step < I->horizon ? (const gu32 *)I->seeds + (size_t)step * 16u + 8u + (q - BIN_NBSEED) : nullptr;
	return p ? *p : 0u;
}

/* Stage B: calcStats of the iteration in st[par] (src/abms.c:230-310), its
 * printStats row (stored when write_row is set), checks, per-iteration
 * parameters and chunks.  bin holds the words begin_input gives for this
 * iteration. */
__device__ static void begin_iteration(Inst *I, uint32_t par, const uint32_t *bin, bool write_row = true)
{
	__shared__ uint32_t s_status;
	const uint32_t tid = threadIdx.x;
	IterState &ip = I->st[par];
	if (tid >= 128u && tid < 228u) {
		/* the larval mortality table of this iteration, beside tid 0's work:
		 * killAgents.cl's larval rate with L1DE and CC (reference mode), or
		 * the stage rate (unused when the iteration does not run) */
		const uint32_t d = tid - 128u;
		const float cc = sc_float(bin[BIN_ENV + offsetof(sampo_env, carryingCapacity) / 4u]);
		ip.hmr_larva[d] = I->mode == SAMPO_MODE_REFERENCE ? sm_hmr_larva(d, I->l1de, cc) : I->hmr_stage[1];
	}
	if (tid >= 64u && tid < 88u && ip.step < I->horizon) {
		/* the seed sets step + 1 (seed_set) and step (nb_set) of a running
		 * iteration (step < horizon: both exist), beside tid 0's work */
		const uint32_t q = tid - 64u;
		if (q < 16u)
			ip.seed[q] = bin[BIN_SEED + q];
		else
			ip.nbseed[q - 16u] = bin[BIN_NBSEED + (q - 16u)];
	}
	if (tid == 0) {
		/* every read first, then the arithmetic, then every write: the
		 * compiler cannot tell the record's fields apart through these
		 * pointers, so a read after a write would wait for the write */
		uint32_t cnt[8];
		for (int s = 0; s < 8; ++s) cnt[s] = ip.cnt[s];
		const uint32_t v0 = ip.vstart[0], v3 = ip.vstart[3], v7 = ip.vstart[7], step = ip.step;
		const uint32_t l1de = I->l1de, fem = I->females, pinf = I->potinf;
		const uint32_t ncyc = I->ncyc, scyc = I->scyc, bites = I->bites, ibites = I->ibites;
		const uint32_t horizon = I->horizon, P = I->P, P_max = I->P_max, mode = I->mode;
		const uint64_t max_agents = I->max_agents;
		const float hs0 = I->hmr_stage[0], hs2 = I->hmr_stage[2];
		sampo_row *const rows = I->rows;
		uint32_t bw[BIN_SEED];
		for (uint32_t q = 0; q < BIN_SEED; ++q) bw[q] = bin[q];
		uint32_t kst[9], n = 0;
		for (int s = 0; s < 8; ++s) {
			kst[s] = n;
			n += cnt[s];
		}
		kst[8] = n;
		sampo_row row;
		for (int s = 0; s < 8; ++s) row.count[s] = cnt[s];
		uint32_t adults_range = (v7 + cnt[7]) - v3;
		row.l1de = cnt[1] > 0 ? l1de : 0u;
		row.females = adults_range > 0 ? fem : 0u;
		row.potInfective = adults_range > 0 ? pinf : 0u;
		row.numCycles = ncyc;
		row.sumCycles = scyc;
		row.bites = bites;
		row.infectBites = ibites;
		row.valid = 1u;
		uint32_t st;
		if (step >= horizon) {
			st = SAMPO_DONE;
		} else if (!((double)n <= (double)max_agents * 0.95)) {
			st = SAMPO_CAPACITY;
			row.valid = 0u;
		} else if (n == 0u) {
			st = SAMPO_EXTINCT;
			row.valid = 0u;
		} else if (n > P) { /* the output would not fit: grow, or give up */
			st = P < P_max ? SAMPO_GROW : SAMPO_PHYS_FULL;
		} else {
			st = SAMPO_RUNNING;
		}
		const float temp = sc_float(bw[BIN_TEMP]);
		SpRow sr;
		for (uint32_t q = 0; q < sizeof(SpRow) / 4u; ++q) ((uint32_t *)&sr)[q] = bw[BIN_SR + q];
		double csd_inc;
		{
			float q = 111.0f / (temp - 16.0f);
			csd_inc = 1.0 / ((double)q * 14.0);
		}
		for (int s = 0; s < 9; ++s) ip.kstart[s] = kst[s];
		ip.n = n;
		if (write_row) gput(&rows[step], row);
		ip.temp = temp;
		ip.nb_inc = sc_float(bw[BIN_INC]);
		ip.csd_inc = csd_inc;
		for (uint32_t q = 0; q < sizeof(sampo_env) / 4u; ++q) ((uint32_t *)&ip.env)[q] = bw[BIN_ENV + q];
		ip.open = sr.open;
		ip.dev_rate = sr.dev_rate;
		ip.pupa_dly = sr.pupa_dly;
		ip.imm_dly = sr.imm_dly;
		ip.bmd_dly = sr.bmd_dly;
		ip.itn_kill = sr.itn_kill;
		ip.irs_kill = sr.irs_kill;
		ip.nb_set = step;
		ip.seed_set = step + 1u;
		ip.h = sc_f2u((float)step * 1.0f) % 24u;
		ip.tb = l1de + (v0 + cnt[0]) + cnt[2]; /* killAgents.cl:42 */
		if (mode == SAMPO_MODE_CONFIG) {
			ip.hmr_egg = hs0;
			ip.hmr_pupa = hs2;
		} else {
			ip.hmr_egg = sm_hmr(0.1f);
			ip.hmr_pupa = sm_hmr(0.1f);
		}
		s_status = st;
	}
	lds_barrier();
	uint32_t st = s_status;
	/* chunks (chunk_cut).  Their outputs go to the same dense slot range of
	 * the other buffer, so N <= P is all the room needed. */
	const uint32_t N = ip.n;
	/* after a fold, its cut of these N agents and each chunk's first segment
	 * are in place; otherwise both are worked out here */
	const bool have_seg0 = ip.seg0_done;
	const Cut cut = have_seg0 ? ip.cut : chunk_cut(I, N);
	const uint32_t C = st == SAMPO_RUNNING ? cut_at_or_after(cut, N) : 0u;
	auto *ch = gmem(I->chunk[par]);
	const auto *seg = gmem(I->seg[par]);
	const uint32_t nseg = ip.nseg;
	/* initial populations: segment holding each chunk's first agent, segment
	 * j holding the chunks whose start falls in [k0_j, k0_(j+1)) */
	if (!have_seg0)
		for (uint32_t j = tid; j < nseg && C; j += blockDim.x) {
			uint32_t a = seg[j].k0;
			uint32_t z = j + 1u < nseg ? seg[j + 1u].k0 : N;
			for (uint32_t c = cut_at_or_after(cut, a); c < C && cut_start(cut, c) < z; ++c) ch[c].seg0 = j;
		}
	lds_barrier();
	if (tid == 0) {
		ip.cut = cut;
		ip.active = st == SAMPO_RUNNING;
		ip.nchunk = ip.active ? C : 0u;
		I->status = st;
		I->step = ip.step;
	}
}

/* ------------------------------------------------------------ day launches */

/* The last index i < m with k0[i] <= k, for k0 ascending with k0[0] <= k:
 * the segment holding dense index k. */
__device__ static inline uint32_t seg_find(const uint32_t *k0, uint32_t m, uint32_t k)
{
	uint32_t lo = 0, hi = m;
	while (hi - lo > 1u) {
		const uint32_t mid = (lo + hi) >> 1;
		if (k0[mid] <= k) lo = mid; else hi = mid;
	}
	return lo;
}

/* The last index i in [lo, hi) with k0[i] <= k, for k0[lo] <= k, by one
 * wave, 64 probes a round (the result in every lane). */
__device__ static inline uint32_t seg_find_wave(const uint32_t *k0, uint32_t lo, uint32_t hi, uint32_t k)
{
	const uint32_t lane = threadIdx.x & 63u;
	while (hi - lo > 1u) {
		const uint32_t step = (hi - lo + 63u) / 64u;
		const uint32_t j = lo + lane * step;
		const uint64_t m64 = __ballot(j < hi && k0[j] <= k); /* lane 0 probes lo: never empty */
		const uint32_t J = 63u - (uint32_t)__clzll((unsigned long long)m64);
		const uint32_t nlo = lo + J * step;
		hi = nlo + step < hi ? nlo + step : hi;
		lo = nlo;
	}
	return lo;
}

/* The fold of a day launch, in every work-group: wait for every chunk's
 * results of iteration j (words tagged `tag`, parity j & 1), then fold the
 * finished iteration st[par] of the LDS record sI into st[par ^ 1] --
 * fold_iteration's results -- and the new segment list into the LDS arrays
 * k0/st/pb/vb (dense start, state | SEG_REV, physical base, virtual base).
 * Thread c < C polls chunk c's NOUT words into registers, all at once and
 * again only those not yet published, and places its runs after one scan
 * over the work-group.  *seg0 receives the index of the segment holding
 * the first agent of this work-group's chunk (blockIdx.x) of the next
 * iteration.  Returns false in every thread if a wait gave up. */
template <int NTH>
__device__ static bool day_fold(Inst &sI, uint32_t par, DaySync *sync, uint32_t j, uint32_t tag, uint32_t *sk0,
                                uint32_t *sst, uint32_t *spb, uint32_t *svb, uint32_t *seg0)
{
	static_assert(NTH >= (int)DAY_CHUNKS, "a thread per chunk");
	__shared__ uint32_t s_kst[9], s_vst[8], s_sb[8], s_tot[24];
	const uint32_t tid = threadIdx.x;
	const IterState &ip = sI.st[par];
	IterState &np = sI.st[par ^ 1u];
	const uint32_t C = ip.nchunk;
	const bool mine = tid < C;
	uint32_t v[NOUT];
	{
		unsigned long long *const base = &sync->w[j & 1u][0][tid];
		uint32_t need = mine ? (1u << NOUT) - 1u : 0u;
		bool gave_up = false;
		for (uint32_t poll = 0; need; ++poll) {
			unsigned long long x[NOUT];
#pragma unroll
			for (uint32_t q = 0; q < NOUT; ++q)
				x[q] = (need >> q) & 1u ? __hip_atomic_load(base + q * DAY_CHUNKS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
				                        : 0ull;
#pragma unroll
			for (uint32_t q = 0; q < NOUT; ++q)
				if (((need >> q) & 1u) && (uint32_t)(x[q] >> 32) == tag) {
					v[q] = (uint32_t)x[q];
					need &= ~(1u << q);
				}
			if (!need) break;
			if (poll >= DAY_POLLS || __hip_atomic_load(&sync->fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
				__hip_atomic_store(&sync->fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				gave_up = true;
				break;
			}
			__builtin_amdgcn_s_sleep(2);
		}
		/* every thread's words are in: the agents of the finished iteration
		 * are read from here on, with agent-scope loads */
		if (__syncthreads_or(gave_up)) return false;
	}
	uint32_t wv[24];
#pragma unroll
	for (uint32_t sn = 0; sn < 8u; ++sn) {
		wv[sn] = 0u;
		wv[8u + sn] = 0u;
		wv[16u + sn] = mine ? v[CO_SUMS + sn] : 0u;
#pragma unroll
		for (uint32_t kind = 0; kind < 3u; ++kind) {
			if (kind >= runs_of(sn)) continue;
			const uint32_t n = mine ? v[run_row(sn, kind)] : 0u;
			wv[sn] += n;
			wv[8u + sn] += n ? 1u : 0u;
		}
	}
	uint32_t ex[24];
	block_excl_n<24, NTH>(wv, ex, s_tot); /* zeros past C */
	const uint32_t E = s_tot[16];
	if (tid == 0) {
		(void)fold_totals(&sI, ip, np, s_tot, s_kst, s_vst, s_sb);
		*seg0 = 0u; /* the newborns', unless a run below holds the chunk start */
	}
	lds_barrier();
	if (tid == 0 && E > 0u) { /* the newborns */
		sk0[0] = 0u;
		sst[0] = 0u;
		spb[0] = PB_NEWBORN;
		svb[0] = 0u;
	}
	if (mine) {
		/* the bases, read before the first write (which may alias them as far
		 * as the compiler knows) */
		uint32_t kst[8], vst[8], sbs[8], okst[9];
#pragma unroll
		for (uint32_t q = 0; q < 8u; ++q) {
			kst[q] = s_kst[q];
			vst[q] = s_vst[q];
			sbs[q] = s_sb[q];
		}
#pragma unroll
		for (uint32_t q = 0; q < 9u; ++q) okst[q] = ip.kstart[q];
		const Cut oc = ip.cut;
		const uint32_t on = ip.n, km = cut_start(np.cut, blockIdx.x);
		const uint32_t ck0 = cut_start(oc, tid), ce = cut_start(oc, tid + 1u), ck1 = ce < on ? ce : on;
#pragma unroll
		for (uint32_t sn = 0; sn < 8u; ++sn) {
			uint32_t k0 = kst[sn] + (sn == 0u ? E : 0u) + ex[sn];
			uint32_t idx = sbs[sn] + ex[8u + sn];
#pragma unroll
			for (uint32_t kind = 0; kind < 3u; ++kind) {
				if (kind >= runs_of(sn)) continue;
				const uint32_t row = run_row(sn, kind), n = v[row];
				if (!n) continue;
				/* the source block: chunk tid's part of the state-src range */
				const uint32_t src = row & 7u, mv = row >> 3;
				const uint32_t bh = okst[src] > ck0 ? okst[src] : ck0;
				const uint32_t be = okst[src + 1u] < ck1 ? okst[src + 1u] : ck1;
				sk0[idx] = k0;
				sst[idx] = sn | (mv ? SEG_REV : 0u);
				spb[idx] = mv ? be - 1u : bh;
				svb[idx] = vst[sn] - kst[sn] + k0;
				if (km >= k0 && km - k0 < n) *seg0 = idx;
				++idx;
				k0 += n;
			}
		}
	}
	lds_barrier();
	return true;
}

/* The instance record is worked on in LDS: the bookkeeping is a chain of
 * dependent scalar updates that would otherwise each wait on global memory.
 * One work-group per slot.  A single slot plans the next pass from its LDS
 * record; with more, k_plan follows as its own launch (the launch boundary
 * publishes every work-group's record: no fences, no counter).  Work-groups
 * of slots with nothing to fold or begin (empty, held, finished) leave after
 * reading three words. */
__global__ void __launch_bounds__(PREP_WG) k_prep(Inst *insts, Ctx *ctx, uint32_t *pl_base, uint8_t *pl_par,
                                                  uint32_t ncls)
{
	__shared__ Inst sI;
	__shared__ uint32_t s_fold, s_begin, s_par, s_work;
	__shared__ uint32_t s_bin[BIN_WORDS];
	Inst *G = insts + blockIdx.x;
	if (gridDim.x > 1u) {
		if (threadIdx.x == 0)
			s_work = (G->status == SAMPO_RUNNING && G->pending) || G->needs_begin;
		lds_barrier();
		if (!s_work) return;
	}
	{
		const uint32_t *src = (const uint32_t *)G;
		uint32_t *dst = (uint32_t *)&sI;
		for (uint32_t q = threadIdx.x; q < sizeof(Inst) / 4u; q += PREP_WG) dst[q] = src[q];
	}
	lds_barrier();
	Inst *I = &sI;
	if (threadIdx.x == 0) {
		s_fold = I->status == SAMPO_RUNNING && I->pending;
		s_par = I->step & 1u;
		s_begin = I->needs_begin;
		/* the iteration being folded was run over this many agents */
		if (s_fold) atomicAdd(&ctx->agent_steps, (unsigned long long)I->st[I->step & 1u].n);
	}
	lds_barrier();
	if (s_fold || s_begin) {
		/* the inputs of the iteration about to begin, one word per lane of the
		 * last wave, loaded before the fold so that both round trips overlap */
		const uint32_t qb = threadIdx.x - (PREP_WG - 64u);
		uint32_t bw = 0u;
		if (threadIdx.x >= PREP_WG - 64u && qb < BIN_WORDS)
			bw = begin_input(I, I->st[s_par].step + (s_fold ? 1u : 0u), qb);
		if (s_fold) {
			fold_iteration(I, s_par);
			lds_barrier();
			if (threadIdx.x == 0) {
				I->pending = 0u;
				I->step = I->step + 1u;
				s_begin = 1u;
				s_par ^= 1u;
			}
		}
		if (threadIdx.x >= PREP_WG - 64u && qb < BIN_WORDS) s_bin[qb] = bw;
		lds_barrier();
		if (s_begin) {
			begin_iteration(I, s_par, s_bin);
			if (threadIdx.x == 0) I->needs_begin = 0u;
		}
		lds_barrier();
		const uint32_t *src = (const uint32_t *)&sI;
		uint32_t *dst = (uint32_t *)G;
		for (uint32_t q = threadIdx.x; q < sizeof(Inst) / 4u; q += PREP_WG) dst[q] = src[q];
	}
	if (gridDim.x == 1u && threadIdx.x == 0) { /* alone: the next pass is planned from the LDS record */
		const uint32_t par = sI.step & 1u;
		const uint32_t nc = sI.status == SAMPO_RUNNING && sI.st[par].active ? sI.st[par].nchunk : 0u;
		pl_par[0] = (uint8_t)par;
		for (uint32_t c = 0; c < ncls; ++c) { /* one slot: its class has all chunks */
			const uint32_t n = c == sI.tile_class ? nc : 0u;
			pl_base[2u * c] = 0u;
			pl_base[2u * c + 1u] = n;
			ctx->total[c] = n;
			ctx->ticket[c] = 0u;
		}
	}
}

/* ------------------------------------------------------------ k_pass */


/* &base[slot] for slot < 2^30 (P is capped there): a 32-bit byte offset */
__device__ static inline gu32 *at_slot(gu32 *base, uint32_t slot)
{
	return (gu32 *)((__attribute__((address_space(1))) char *)base + (uint32_t)(slot << 2));
}

/* One agent of a tile: where it comes from. */
struct Src {
	uint32_t g;     /* virtual index (keys the random numbers) */
	uint32_t s;     /* state, S_DEAD for no agent */
	uint32_t born;  /* newborn: regenerated, not loaded */
};

/* The pass over a chunk is software-pipelined by tiles of WG * APT agents:
 * while tile t is updated, ranked and written, the loads of tile t + 1 run
 * as direct-to-LDS loads (global_load_lds_dword) into the other of two LDS
 * buffers.  Each lane's slots in a buffer are its own; the updated agent
 * stays in the lane's registers until it is stored. */
/* What a day launch adds to the pass's arguments. */
struct DayArgs {
	DaySync *sync;
	uint32_t seq0, iters; /* the tag of its first iteration; iterations to run */
	uint32_t *pl_base;    /* the plan it leaves for a pass that follows (one slot) */
	uint8_t *pl_par;
	uint32_t ncls;
};

/* The fused pass (DAY false): one iteration over every running instance of
 * tile class cls, persistent work-groups taking chunks by ticket.  The day
 * launch (DAY true, k_day below): the same chunk and tile code, fed chunks
 * by work-group index in every iteration of the launch. */
template <int WG, int APT, int WAVES, bool DAY = false>
__global__ void __launch_bounds__(WG, WAVES) k_pass(Inst *insts, Ctx *ctx, const uint32_t *pl_base_all,
                                             const uint8_t *pl_par, uint32_t ninst, uint32_t cls, DayArgs da)
{
	constexpr int NW = WG / 64;
	constexpr uint32_t TILE = (uint32_t)WG * APT;
	/* the segment window: a tile overlaps <= TILE + 1 segments, staged from
	 * the list in global memory; a day launch holds an iteration's whole list */
	constexpr uint32_t SEGC = DAY ? DAY_SEGMAX : TILE + 64u;
	__shared__ TileParams T;
	__shared__ uint32_t s_ticket, s_inst, s_c;
	__shared__ uint32_t s_k0[SEGC], s_st[SEGC], s_pb[SEGC], s_vb[SEGC];
	__shared__ uint32_t s_wlo, s_wm, s_wend; /* window: global index, size, dense end */
	__shared__ uint32_t s_wc[APT][NW][16];    /* per (input state, moved) */
	__shared__ uint32_t s_run[16], s_tc[16];
	__shared__ uint32_t s_base[16];  /* per class: the first slot of its end of the chunk's block of its input state */
	__shared__ uint32_t s_acc[NW][3];   /* l1de, females, potentially infective: per wave */
	__shared__ uint32_t s_ev[NW][NEV]; /* agent_step's event sums of the chunk, per wave */
	__shared__ uint32_t bufA[NFIELD][TILE], bufB[NFIELD][TILE];

	const uint32_t tid = threadIdx.x, lane = tid & 63u;
	const uint32_t w = __builtin_amdgcn_readfirstlane(tid >> 6); /* scalar: addresses built on it need no VGPR */
	const uint32_t total = DAY ? 0u : ctx->total[cls];
	const uint32_t *const pl_base = pl_base_all + cls * (ninst + 1u); /* this class's tickets */
	const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
	uint32_t next = 0; /* global index of the segment holding the last issued tile's end */

	/* stage the segments from global index `from` on until the window covers
	 * dense index te - 1 (wave 0) */
	auto stage = [&](uint32_t from, uint32_t te) {
		uint32_t j = from, m = 0, kend = T.ip.n;
		for (;;) {
			uint32_t jj = j + lane;
			Seg sg = {0u, 0u, 0u, 0u};
			bool valid = jj < T.nseg;
			if (valid) {
				const gu32 *q = (const gu32 *)T.seg + (size_t)jj * 4u;
				sg.k0 = q[0];
				sg.state = q[1];
				sg.pbase = q[2];
				sg.vbase = q[3];
			}
			bool in = valid && sg.k0 < te;
			uint64_t bl = __ballot(in);
			uint32_t cnt = (uint32_t)__popcll(bl);
			if (in) {
				s_k0[m + lane] = sg.k0;
				s_st[m + lane] = sg.state;
				s_pb[m + lane] = sg.pbase;
				s_vb[m + lane] = sg.vbase;
			}
			m += cnt;
			if (cnt < 64u) { /* first lane past te holds the window end */
				uint64_t past = __ballot(valid && !in);
				if (past) kend = __shfl(sg.k0, (int)(__ffsll((unsigned long long)past) - 1), 64);
				break;
			}
			if (m + 64u > SEGC) { /* window full: it ends where the next segment starts */
				uint32_t nk = j + 64u < T.nseg ? ((const gu32 *)T.seg)[(size_t)(j + 64u) * 4u] : T.ip.n;
				kend = nk;
				break;
			}
			j += 64u;
		}
		if (lane == 0) {
			s_wlo = from;
			s_wm = m;
			s_wend = kend;
		}
		/* every load above is complete before any lane leaves: nothing
		 * loaded here stays pending past the next tile's loads */
		__builtin_amdgcn_s_waitcnt(0x0F70);
	};

	/* Sources of tile [ts, te) from the window and their loads into buf.
	 * `next` enters as the global index of a segment starting at or before
	 * ts (the one holding the previous tile's last agent, where the window
	 * also starts after a restage) and leaves as the one holding te - 1.
	 * The tile's segments are found by a uniform forward walk; each lane then
	 * picks its own among those few.  Returns the tile's input-state span. */
	uint32_t span_lo = 0, span_hi = 0;
	uint32_t nx_ss[APT], nx_sg[APT]; /* the next tile, this lane's agents: state | newborn << 8; virtual index */
	auto issue = [&](uint32_t ts, uint32_t te, uint32_t (&buf)[NFIELD][TILE], gu32 *const (&fb)[NFIELD]) {
		const uint32_t m = s_wm;
		uint32_t q = next - s_wlo;
		while (q + 1u < m && s_k0[q + 1u] <= ts) ++q;
		const uint32_t q_lo = q;
		/* the segment holding te - 1: a short walk, or in a day launch 64
		 * probes a round when the tile spans many segments (fragmented
		 * ranges: runs of a few agents) */
		if (DAY && q + 8u < m && s_k0[q + 8u] < te)
			q = seg_find_wave(s_k0, q, m, te - 1u);
		else
			while (q + 1u < m && s_k0[q + 1u] < te) ++q;
		const uint32_t q_hi = q;
		span_lo = s_st[q_lo] & 0xffu;
		span_hi = s_st[q_hi] & 0xffu;
		next = s_wlo + q_hi;
#pragma unroll
		for (int jx = 0; jx < APT; ++jx) {
			const uint32_t k = ts + (uint32_t)jx * WG + tid;
			const uint32_t o = (uint32_t)jx * WG + tid;
			nx_ss[jx] = S_DEAD;
			if (k < te) {
				uint32_t qq = q_lo;
				if (!DAY || q_hi - q_lo <= 8u) { /* (the extra test measured 2 % slower in k_pass) */
					for (uint32_t j = q_lo + 1u; j <= q_hi; ++j) qq += k >= s_k0[j] ? 1u : 0u;
				} else { /* binary search among the tile's segments */
					uint32_t hi = q_hi + 1u;
					while (hi - qq > 1u) {
						const uint32_t mid = (qq + hi) >> 1;
						if (s_k0[mid] <= k) qq = mid; else hi = mid;
					}
				}
				const uint32_t r = k - s_k0[qq];
				const uint32_t pb = s_pb[qq], st = s_st[qq];
				const bool born = pb == PB_NEWBORN;
				nx_ss[jx] = (st & 0xffu) | (born ? 0x100u : 0u);
				nx_sg[jx] = s_vb[qq] + r;
				if (!born) {
					const uint32_t p = seg_slot(pb, st, r);
					const uint32_t fm = sm_fields(st & 0xffu, T.compact);
#pragma unroll
					for (int f = 0; f < NFIELD; ++f)
						if ((fm >> f) & 1u)
							__builtin_amdgcn_global_load_lds((glob_void *)at_slot(fb[f], p),
							                                 (lds_void *)&buf[f][(uint32_t)jx * WG + w * 64u], 4, 0,
							                                 DAY ? CPOL_SC1 : 0);
				}
			}
		}
	};

	/* Chunks.  k_pass takes them by ticket: a work-group's first chunk is its
	 * own index, later ones come from the counter (from gridDim.x on), taken
	 * by tid 0 while the chunk's last tile is worked on; T keeps an
	 * instance's parameters across its chunks (within a launch they do not
	 * change).  A day launch takes them by work-group index, iteration after
	 * iteration; between two iterations every work-group folds the finished
	 * one. */
	uint32_t nt = blockIdx.x;
	uint32_t have = 0xffffffffu;
	/* a day launch: the lone instance in slot 0, its record in LDS in every
	 * work-group, and the inputs of the iterations the launch may begin */
	struct DayLds {
		Inst sI;
		uint32_t bin[DAY_ITERS][BIN_WORDS];
		uint32_t seg0; /* the segment holding this work-group's chunk start, from the fold (~0u: search) */
	};
	__shared__ std::conditional_t<DAY, DayLds, uint32_t> s_day;
	uint32_t par = 0, iters = 0, j = 0; /* day launch: parity, iterations to run, the current one */
	unsigned long long steps = 0;       /* day launch: agent-steps of the folded iterations */
	bool fault = false;
	/* a day launch: the record's iteration parameters into T */
	auto load_params = [&](auto &dl) {
		const IterState &ip = dl.sI.st[par];
		const uint32_t *src = (const uint32_t *)&ip;
		uint32_t *dst = (uint32_t *)&T.ip;
		for (uint32_t q = tid; q < sizeof(IterState) / 4u; q += WG) dst[q] = src[q];
		if (tid == 0) {
			T.bin = dl.sI.buf[ip.cur];
			T.bout = dl.sI.buf[ip.cur ^ 1u];
			T.seg = dl.sI.seg[par];
			T.chk = dl.sI.chunk[par];
			T.nseg = ip.nseg;
			T.cout = dl.sI.cout;
			T.cmax = dl.sI.cmax;
			T.P = dl.sI.P;
			T.compact = dl.sI.compact;
			s_wlo = 0u;
			s_wm = ip.nseg;
			s_wend = ip.n;
		}
	};
	if constexpr (DAY) {
		Inst &sI = s_day.sI;
		{
			const uint32_t *src = (const uint32_t *)insts;
			uint32_t *dst = (uint32_t *)&sI;
			for (uint32_t q = tid; q < sizeof(Inst) / 4u; q += WG) dst[q] = src[q];
		}
		__syncthreads();
		par = sI.step & 1u;
		iters = da.iters < DAY_ITERS ? da.iters : DAY_ITERS;
		{
			const IterState &ip = sI.st[par];
			if (!(sI.status == SAMPO_RUNNING && ip.active)) return; /* nothing to run */
			/* a day launch follows a prep (nothing pending or to begin) of an
			 * instance cut into at most DAY_CHUNKS chunks */
			if (sI.pending || sI.needs_begin || ip.nchunk > DAY_CHUNKS || ip.nseg > DAY_SEGMAX) {
				if (blockIdx.x == 0 && tid == 0)
					__hip_atomic_store(&insts->status, SAMPO_FAULT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				return;
			}
		}
		load_params(s_day);
		for (uint32_t q = tid; q < 100u; q += WG) T.hmr_adult[q] = sI.hmr_adult[q];
		{ /* the first iteration's segment list into the window (later ones come
		     from the folds), and the inputs of the iterations the launch may begin */
			const gu32 *sg = (const gu32 *)sI.seg[par];
			for (uint32_t i = tid; i < sI.st[par].nseg; i += WG) {
				s_k0[i] = sg[4u * i];
				s_st[i] = sg[4u * i + 1u];
				s_pb[i] = sg[4u * i + 2u];
				s_vb[i] = sg[4u * i + 3u];
			}
			if (tid == 0) s_day.seg0 = ~0u; /* no fold yet: the chunk's first segment is searched for */
			const uint32_t step0 = sI.st[par].step;
			for (uint32_t x = tid; x < iters * BIN_WORDS; x += WG) {
				const uint32_t jj = x / BIN_WORDS;
				s_day.bin[jj][x - jj * BIN_WORDS] = begin_input(&sI, step0 + 1u + jj, x - jj * BIN_WORDS);
			}
		}
		__syncthreads();
	}
	for (;;) {
		uint32_t ii, c;
		if constexpr (DAY) {
			if (nt >= T.ip.nchunk) { /* the iteration's chunks are done: fold it, begin the next */
				Inst &sI = s_day.sI;
				if (j + 1u == iters && blockIdx.x != 0u) return; /* work-group 0 alone folds the last */
				if (!day_fold<WG>(sI, par, da.sync, j, da.seq0 + j, s_k0, s_st, s_pb, s_vb, &s_day.seg0)) {
					fault = true;
					break;
				}
				steps += T.ip.n;
				begin_iteration(&sI, par ^ 1u, s_day.bin[j], blockIdx.x == 0u);
				lds_barrier();
				par ^= 1u;
				load_params(s_day);
				lds_barrier();
				if (sI.status != SAMPO_RUNNING || j + 1u == iters) {
					if (blockIdx.x != 0u) return;
					break;
				}
				if (T.ip.nchunk > DAY_CHUNKS) { /* the even cut keeps it within; never trust it blindly */
					fault = true;
					break;
				}
				nt = blockIdx.x;
				++j;
				continue;
			}
			ii = 0u;
			c = nt;
			nt += gridDim.x;
		} else {
			if (w == 0) { /* the instance holding ticket t: the last one whose base <= t, 64 probes a round */
				const uint32_t t = __builtin_amdgcn_readfirstlane(nt);
				uint32_t lo = 0, base = 0;
				if (t < total) {
					uint32_t hi = ninst;
					while (hi - lo > 1u) {
						const uint32_t step = (hi - lo + 63u) / 64u;
						const uint32_t j = lo + lane * step;
						const uint32_t b = j < hi ? pl_base[j] : 0xffffffffu; /* lane 0 probes lo: base <= t */
						const uint64_t m = __ballot(b <= t);
						const uint32_t J = 63u - (uint32_t)__clzll((unsigned long long)m);
						base = __shfl(b, (int)J, 64);
						const uint32_t nlo = lo + J * step;
						hi = nlo + step < hi ? nlo + step : hi;
						lo = nlo;
					}
				}
				if (lane == 0) {
					s_ticket = t;
					s_inst = lo;
					s_c = t - base;
				}
			}
			lds_barrier();
			const uint32_t ticket = s_ticket;
			if (ticket >= total) break;
			ii = s_inst;
			c = s_c;
			if (ii != have) { /* the instance's parameters into LDS; both parities are
			                     loaded at once, with the parity, and one is kept */
				have = ii;
				const Inst *I = insts + ii;
				const uint32_t par = pl_par[ii];
				const uint32_t *s0 = (const uint32_t *)&I->st[0], *s1 = (const uint32_t *)&I->st[1];
				uint32_t *dst = (uint32_t *)&T.ip;
				for (uint32_t q = tid; q < sizeof(IterState) / 4u; q += WG) {
					uint32_t x0 = s0[q], x1 = s1[q];
					/* opaque values: the select must not become a load at a parity-dependent address */
					__asm__("" : "+v"(x0), "+v"(x1));
					dst[q] = par ? x1 : x0;
				}
				for (uint32_t q = tid; q < 100u; q += WG) T.hmr_adult[q] = I->hmr_adult[q];
				if (tid == 0) {
					uint32_t *b0 = I->buf[0], *b1 = I->buf[1];
					const Seg *g0 = I->seg[0], *g1 = I->seg[1];
					const Chunk *c0 = I->chunk[0], *c1 = I->chunk[1];
					uint32_t cu0 = I->st[0].cur, cu1 = I->st[1].cur, ns0 = I->st[0].nseg, ns1 = I->st[1].nseg;
					__asm__("" : "+v"(b0), "+v"(b1), "+v"(g0), "+v"(g1), "+v"(c0), "+v"(c1));
					__asm__("" : "+v"(cu0), "+v"(cu1), "+v"(ns0), "+v"(ns1));
					const uint32_t cur = par ? cu1 : cu0;
					T.bin = cur ? b1 : b0;
					T.bout = cur ? b0 : b1;
					T.seg = par ? g1 : g0;
					T.chk = par ? c1 : c0;
					T.nseg = par ? ns1 : ns0;
					T.cout = I->cout;
					T.cmax = I->cmax;
					T.P = I->P;
					T.compact = I->compact;
				}
				lds_barrier();
			}
		}
		if constexpr (DAY) {
			if (w == 0) { /* the chunk's range from the iteration's cut; its first segment from the fold or the window */
				const uint32_t e = cut_start(T.ip.cut, c + 1u), ck0 = cut_start(T.ip.cut, c);
				const uint32_t pre = s_day.seg0;
				const uint32_t g = c == blockIdx.x && pre != ~0u ? pre : seg_find_wave(s_k0, 0u, s_wm, ck0);
				if (lane == 0) {
					T.ch.k0 = ck0;
					T.ch.k1 = e < T.ip.n ? e : T.ip.n;
					T.ch.seg0 = g;
				}
			}
		} else if (tid == 0) { /* the chunk's range from the iteration's cut; its first segment from the record */
			const uint32_t e = cut_start(T.ip.cut, c + 1u);
			T.ch.k0 = cut_start(T.ip.cut, c);
			T.ch.k1 = e < T.ip.n ? e : T.ip.n;
			T.ch.seg0 = *gmem(&T.chk[c].seg0);
		}
		if (tid < 16u) s_run[tid] = 0u;
		if (tid < NW * NEV) (&s_ev[0][0])[tid] = 0u;
		lds_barrier();
		if (tid < 8u) {
			uint32_t h = T.ip.kstart[tid] > T.ch.k0 ? T.ip.kstart[tid] : T.ch.k0;
			uint32_t z = T.ip.kstart[tid + 1u] < T.ch.k1 ? T.ip.kstart[tid + 1u] : T.ch.k1;
			/* stayers fill the block from its head up, movers from its end down */
			s_base[2u * tid] = h;
			s_base[2u * tid + 1u] = (z > h ? z : h) - 1u;
		}
		const uint32_t k0 = T.ch.k0, k1 = T.ch.k1;
		lds_u32 *const ev = (lds_u32 *)&s_ev[w][0]; /* this wave's event sums */
		uint32_t l1de = 0, fem = 0, pinf = 0;

		/* per-field bases of both buffers in scalar registers */
		gu32 *fin[NFIELD], *fout[NFIELD];
		{
			gu32 *bi = uniform_ptr(T.bin), *bo = uniform_ptr(T.bout);
			const uint32_t Pu = __builtin_amdgcn_readfirstlane(T.P);
#pragma unroll
			for (int f = 0; f < NFIELD; ++f) {
				fin[f] = bi + (size_t)f * Pu;
				fout[f] = bo + (size_t)f * Pu;
			}
		}
		/* prologue: tile 0's window and loads */
		next = T.ch.seg0;
		if constexpr (!DAY) {
			const uint32_t te0 = k0 + TILE < k1 ? k0 + TILE : k1;
			if (w == 0) stage(next, te0);
			lds_barrier();
		}
		issue(k0, k0 + TILE < k1 ? k0 + TILE : k1, bufA, fin);

		/* one tile: its agents are in cur (loaded) or regenerated; the next
		 * tile's loads go to nxt */
		auto tile = [&](uint32_t ts, uint32_t (&cur)[NFIELD][TILE], uint32_t (&nxt)[NFIELD][TILE]) {
			const uint32_t te = ts + TILE < k1 ? ts + TILE : k1;
			Ag a[APT];
			Src me[APT];
			/* this tile's loads into cur are the oldest memory operations in
			 * flight: wait for them (vmcnt(0)) */
			__builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
			for (int jx = 0; jx < APT; ++jx) {
				const uint32_t o = (uint32_t)jx * WG + tid;
				const uint32_t ss = nx_ss[jx];
				me[jx].s = ss & 0xffu;
				me[jx].born = ss >> 8;
				me[jx].g = nx_sg[jx];
				if (me[jx].s < S_DEAD) {
					if (me[jx].born) {
						newborn(a[jx], T.ip.nbseed, T.ip.nb_inc, T.ip.hatch_thr, T.ip.hatch_m, T.ip.hatch_b, me[jx].g);
					} else {
						uint32_t wv[NFIELD];
#pragma unroll
						for (int f = 0; f < NFIELD; ++f) wv[f] = cur[f][o];
						unpack_agent(wv, a[jx], sm_fields(me[jx].s, T.compact));
					}
				}
			}
			/* the input states the tile spans (dense order is state-sorted),
			 * found when its loads were issued */
			const uint32_t s_lo = __builtin_amdgcn_readfirstlane(span_lo);
			const uint32_t s_hi = __builtin_amdgcn_readfirstlane(span_hi);
			/* next tile: window (wave 0), barrier, loads into nxt */
			const uint32_t ts2 = te, te2 = ts2 + TILE < k1 ? ts2 + TILE : k1;
			const bool more = ts2 < k1;
			lds_barrier(); /* every lane has read the window for this tile */
			if (more) {
				if constexpr (!DAY) { /* a day launch's window holds the whole list */
					if (w == 0 && te2 > s_wend) stage(next, te2);
					lds_barrier();
				}
				issue(ts2, te2, nxt, fin);
			} else if (!DAY && tid == 0) {
				/* tickets past the first round: none when the grid covers every
				 * chunk (then no work-group waits on the one shared counter) */
				nt = total <= gridDim.x ? total : gridDim.x + atomicAdd(&ctx->ticket[cls], 1u);
			}
			uint32_t ns[APT], lr[APT];
			uint32_t keep[APT][NFIELD]; /* each survivor's words, in registers until they are stored */
#pragma unroll 1
			for (int jx = 0; jx < APT; ++jx) {
				uint32_t n2 = 0xffu; /* class 2 * input state + moved, 0xff: dead or no agent */
				if (me[jx].s < S_DEAD) {
					const uint32_t s = me[jx].s;
					const uint32_t o = (uint32_t)jx * WG + tid;
					Ag x = a[jx];
					const uint32_t ns2 = agent_step(x, s, me[jx].g, T, ev);
					if (ns2 < S_DEAD) {
						n2 = 2u * s + (ns2 != s ? 1u : 0u);
						pack_agent(x, keep[jx]);
						if (ns2 == S_LARVA) l1de += x.age / 24u; /* calcL1de.cl: (uint)(age / 24.0f), whole hours */
						if (ns2 >= S_IMM && (x.misc >> 31)) {
							fem += 1u;
							pinf += x.csd >= 1.0f ? 1u : 0u;
						}
					}
				}
				/* rank survivors by (input state, moved) over the input states
				 * the tile spans */
				uint32_t rank = 0;
				for (uint32_t cc = 2u * s_lo; cc <= 2u * s_hi + 1u; ++cc) {
					uint64_t mk = __ballot(n2 == cc);
					if (n2 == cc) rank = (uint32_t)__popcll(mk & lt);
					if (lane == 0) s_wc[jx][w][cc] = (uint32_t)__popcll(mk);
				}
				ns[jx] = n2;
				lr[jx] = rank;
			}
			lds_barrier();
			if (tid < 16u) { /* exclusive over (jx, wave) per class, after the chunk's earlier tiles */
				uint32_t run = s_run[tid];
				if ((tid >> 1) >= s_lo && (tid >> 1) <= s_hi)
					for (int jx = 0; jx < APT; ++jx)
						for (int q = 0; q < NW; ++q) {
							uint32_t cc = s_wc[jx][q][tid];
							s_wc[jx][q][tid] = run;
							run += cc;
						}
				s_tc[tid] = run;
			}
			lds_barrier();
#pragma unroll 1
			for (int jx = 0; jx < APT; ++jx) {
				const uint32_t cl = ns[jx];
				if (cl < 16u) {
					const uint32_t o = (uint32_t)jx * WG + tid;
					const uint32_t off = s_wc[jx][w][cl] + lr[jx];
					const uint32_t p = s_base[cl] + ((cl & 1u) ? 0u - off : off);
					uint32_t wv[NFIELD];
#pragma unroll
					for (int f = 0; f < NFIELD; ++f) wv[f] = keep[jx][f];
					auto put = [&](int f) {
						if constexpr (DAY) /* coherent at agent scope: other work-groups read it next */
							__hip_atomic_store((uint32_t *)at_slot(fout[f], p), wv[f], __ATOMIC_RELAXED,
							                   __HIP_MEMORY_SCOPE_AGENT);
						else
							*at_slot(fout[f], p) = wv[f];
					};
					/* the words the agent's new state keeps (sm_fields): a mover from
					 * state s is now in s + 1 (gravid: blood-meal seeking, an adult).
					 * Every state keeps AH, DLY and NF; AUX from the larva on; CSD,
					 * AVAIL and CH in the adult states, or every word when the
					 * instance stores the full record. */
					const uint32_t sn = (cl >> 1) + (cl & 1u);
					const bool full = !T.compact;
					put(F_AH);
					put(F_DLY);
					put(F_NF);
					if (full || sn != S_EGG) put(F_AUX);
					if (full || sn > S_PUPA) {
						put(F_CSD);
						put(F_AVAIL);
						put(F_CH);
					}
				}
			}
			/* the next tile rewrites s_wc only after its first barrier */
			if (tid < 16u) s_run[tid] = s_tc[tid];
		};

		for (uint32_t ts = k0; ts < k1; ts += 2u * TILE) {
			tile(ts, bufA, bufB);
			if (ts + TILE < k1) tile(ts + TILE, bufB, bufA);
		}
		{ /* chunk results */
			uint32_t v[3] = {l1de, fem, pinf};
#pragma unroll
			for (int q = 0; q < 3; ++q) {
				uint32_t r = wave_sum(v[q]);
				if (lane == 0) s_acc[w][q] = r;
			}
			if constexpr (DAY) {
				/* published once the chunk's agents are: every wave waits until its
				 * agent stores are complete (agent-scope stores, visible to every
				 * work-group once complete), then wave 0 sends the tagged words,
				 * rows in cout order */
				__asm__ __volatile__("" ::: "memory");
				__builtin_amdgcn_s_waitcnt(0);
				__syncthreads();
				if (w == 0) {
					if (lane < NOUT) {
						uint32_t r = 0;
						if (lane < 16u)
							r = s_run[2u * (lane & 7u) + (lane >> 3)];
						else
							for (int q = 0; q < NW; ++q)
								r += lane - 16u < NEV ? s_ev[q][lane - 16u] : s_acc[q][lane - 16u - NEV];
						__hip_atomic_store(&da.sync->w[j & 1u][lane][c], ((unsigned long long)(da.seq0 + j) << 32) | r,
						                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
					}
				}
			} else {
				lds_barrier();
				if (tid < 8u) { /* the sums in cout order: the NEV events, then l1de, fem, pinf */
					uint32_t r = 0;
					for (int q = 0; q < NW; ++q) r += tid < NEV ? s_ev[q][tid] : s_acc[q][tid - NEV];
					((gu32 *)T.cout)[(size_t)(CO_SUMS + tid) * T.cmax + c] = r;
				}
				if (tid < 16u) /* stayed: row q, moved on: row 8 + q */
					((gu32 *)T.cout)[(size_t)((tid & 1u) * 8u + (tid >> 1)) * T.cmax + c] = s_run[tid];
				if (tid == 0) insts[ii].pending = 1u;
			}
		}
		lds_barrier();
	}
	if constexpr (DAY) {
		Inst &sI = s_day.sI;
		if (blockIdx.x != 0u) return;
		if (fault) {
			if (tid == 0) __hip_atomic_store(&insts->status, SAMPO_FAULT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			return;
		}
		/* work-group 0 leaves what a prep would: the next iteration's segment
		 * list and chunk starts, the record, the plan of a following pass */
		const IterState &np = sI.st[par];
		{
			gu32x4 *sg = (gu32x4 *)sI.seg[par];
			for (uint32_t i = tid; i < np.nseg; i += WG) sg[i] = u32x4{s_k0[i], s_st[i], s_pb[i], s_vb[i]};
			auto *ch = gmem(sI.chunk[par]);
			for (uint32_t c = tid; c < np.nchunk; c += WG) ch[c].seg0 = seg_find(s_k0, np.nseg, cut_start(np.cut, c));
			const uint32_t *src = (const uint32_t *)&sI;
			gu32 *dst = (gu32 *)insts;
			for (uint32_t q = tid; q < sizeof(Inst) / 4u; q += WG) dst[q] = src[q];
		}
		if (tid == 0) {
			const uint32_t nc = sI.status == SAMPO_RUNNING && np.active ? np.nchunk : 0u;
			da.pl_par[0] = (uint8_t)par;
			for (uint32_t c = 0; c < da.ncls; ++c) {
				const uint32_t n = c == sI.tile_class ? nc : 0u;
				da.pl_base[2u * c] = 0u;
				da.pl_base[2u * c + 1u] = n;
				ctx->total[c] = n;
				ctx->ticket[c] = 0u;
			}
			atomicAdd(&ctx->agent_steps, steps);
		}
	}
}

/* A day launch: up to DAY_ITERS iterations of the lone instance in slot 0
 * in one cooperative launch (every work-group resident), a work-group per
 * chunk of its even cut.  Between iterations every work-group gathers all
 * chunk results and folds them itself, in place of two kernel boundaries and
 * a one-work-group prep; work-group 0 leaves the state a prep would. */
#define DAY_WG 512
static const void *k_day() { return (const void *)k_pass<DAY_WG, 1, 0, true>; }

/* ------------------------------------------------------------ export, histogram */

__device__ static inline uint32_t seg_find_global(const Seg *seg, uint32_t nseg, uint32_t k)
{
	uint32_t lo = 0, hi = nseg;
	while (hi - lo > 1u) {
		uint32_t mid = (lo + hi) >> 1;
		if (seg[mid].k0 <= k) lo = mid; else hi = mid;
	}
	return lo;
}

/* what newborn() needs of an iteration, by value */
struct Hatch {
	float inc, thr[4], m[5], b[5];
};

__global__ void k_export(const uint32_t *buf, uint32_t P, const Seg *seg, uint32_t nseg, uint32_t n,
                         const uint32_t *nbz, Hatch hb, uint32_t compact, sampo_agent *out)
{
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const Seg sg = seg[seg_find_global(seg, nseg, k)];
	Ag a;
	if (sg.pbase == PB_NEWBORN) newborn(a, nbz, hb.inc, hb.thr, hb.m, hb.b, sg.vbase + (k - sg.k0));
	else load_agent(a, buf, P, seg_slot(sg.pbase, sg.state, k - sg.k0), sm_fields(sg.state & 0xffu, compact));
	sampo_agent r;
	r.ageInHours = (float)a.age;
	r.hoursInState = (float)a.his;
	r.delay = a.dly;
	r.csd = a.csd;
	r.aux = a.aux;
	r.availableEggs = a.avail;
	r.cycleLength = a.cyc;
	r.humanBloodmealCount = a.misc & 0xffffu;
	r.numEggBatches = (a.misc >> 16) & 0x7fffu;
	r.isFemale = a.misc >> 31;
	out[k] = r;
}

/* ageHistogram.cl:13-35: LDS-private bins, one global atomic per bin per group */
__global__ void __launch_bounds__(256) k_histogram(const uint32_t *buf, const Seg *seg, uint32_t nseg, uint32_t k0,
                                                   uint32_t n, uint32_t *hist)
{
	__shared__ uint32_t bins[100];
	for (uint32_t q = threadIdx.x; q < 100u; q += blockDim.x) bins[q] = 0u;
	__syncthreads();
	for (uint32_t k = k0 + blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
		const Seg sg = seg[seg_find_global(seg, nseg, k)];
		float age = (float)(buf[seg_slot(sg.pbase, sg.state, k - sg.k0)] & 0xffffu); /* F_AH */
		uint32_t b = sc_f2u(age / 24.0f);
		atomicAdd(&bins[b < 99u ? b : 99u], 1u);
	}
	__syncthreads();
	for (uint32_t q = threadIdx.x; q < 100u; q += blockDim.x)
		if (bins[q]) atomicAdd(&hist[q], bins[q]);
}

/* Initial population: host records (dense, range order) into the SoA
 * buffer, and the calcStats sums of the start (kernel/calcL1de.cl,
 * kernel/calcGender.cl) into the instance. */
__global__ void __launch_bounds__(256) k_import(const sampo_agent *in, uint32_t n, uint32_t k_larva0,
                                                uint32_t k_larva1, uint32_t k_adult0, uint32_t *buf, uint32_t P,
                                                Inst *inst)
{
	__shared__ uint32_t s_sum[3][4];
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t l1de = 0, fem = 0, pinf = 0;
	if (k < n) {
		const sampo_agent a = in[k];
		Ag g;
		g.age = (uint32_t)a.ageInHours; /* whole hours in [0, 2^16): admission checks it */
		g.his = (uint32_t)a.hoursInState;
		g.dly = a.delay;
		g.aux = a.aux;
		g.csd = a.csd;
		g.avail = a.availableEggs;
		g.cyc = a.cycleLength;
		g.misc = (a.humanBloodmealCount & 0xffffu) | ((a.numEggBatches & 0x7fffu) << 16) |
		         (a.isFemale ? 0x80000000u : 0u);
		uint32_t w[NFIELD];
		pack_agent(g, w);
#pragma unroll
		for (int f = 0; f < NFIELD; ++f) buf[(size_t)f * P + k] = w[f];
		if (k >= k_larva0 && k < k_larva1) l1de = sc_f2u(a.ageInHours / 24.0f);
		if (k >= k_adult0 && a.isFemale) {
			fem = 1u;
			pinf = a.csd >= 1.0f ? 1u : 0u;
		}
	}
	l1de = wave_sum(l1de);
	fem = wave_sum(fem);
	pinf = wave_sum(pinf);
	const uint32_t w = threadIdx.x >> 6;
	if ((threadIdx.x & 63u) == 0) {
		s_sum[0][w] = l1de;
		s_sum[1][w] = fem;
		s_sum[2][w] = pinf;
	}
	__syncthreads();
	if (threadIdx.x < 3u) {
		uint32_t r = s_sum[threadIdx.x][0] + s_sum[threadIdx.x][1] + s_sum[threadIdx.x][2] + s_sum[threadIdx.x][3];
		uint32_t *dst = threadIdx.x == 0 ? &inst->l1de : threadIdx.x == 1 ? &inst->females : &inst->potinf;
		if (r) atomicAdd(dst, r);
	}
}

/* ============================================================ host side */

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, const char *a = "", long b = 0)
{
	snprintf(g_err, sizeof(g_err), fmt, a, b);
	return code;
}

#define HIPCHK(x)                                                                            \
	do {                                                                                     \
		hipError_t e_ = (x);                                                                 \
		if (e_ != hipSuccess) return fail(SAMPO_E_HIP, "HIP error %s at line %ld",          \
		                                  hipGetErrorString(e_), (long)__LINE__);            \
	} while (0)

struct Slot {
	bool used = false;
	uint32_t cls = 0;       /* its tile class */
	uint32_t day = 0;       /* work-groups of its day launches (its even cut); 0: none */
	void *mem = nullptr;    /* segments, chunks, counters, inputs, rows */
	void *agents = nullptr; /* the two SoA agent buffers, 2 * NFIELD * P dwords */
	uint32_t horizon = 0;
	uint32_t P = 0;
};

typedef void (*pass_fn)(Inst *, Ctx *, const uint32_t *, const uint8_t *, uint32_t, uint32_t, DayArgs);

template <int WG, int APT, int WAVES> pass_fn pass_ptr() { return k_pass<WG, APT, WAVES>; }

/* compiled tile shapes: work-group size, agents per thread, minimum waves
 * per SIMD (0: no register cap) */
pass_fn pick_pass(uint32_t wg, uint32_t apt, uint32_t waves)
{
#define SAMPO_TILE(W, A, V) if (wg == W && apt == A && waves == V) return pass_ptr<W, A, V>();
	SAMPO_TILE(64, 4, 0)
	SAMPO_TILE(128, 1, 0) SAMPO_TILE(128, 2, 0) SAMPO_TILE(128, 4, 0)
	SAMPO_TILE(256, 1, 0) SAMPO_TILE(256, 2, 0) SAMPO_TILE(256, 4, 0)
	SAMPO_TILE(256, 1, 6) SAMPO_TILE(256, 1, 8) SAMPO_TILE(256, 2, 6)
	SAMPO_TILE(512, 1, 0) SAMPO_TILE(512, 2, 0) SAMPO_TILE(512, 1, 6) SAMPO_TILE(512, 1, 8)
	SAMPO_TILE(768, 1, 6)
	SAMPO_TILE(1024, 1, 0) SAMPO_TILE(1024, 1, 6)
#undef SAMPO_TILE
	return nullptr;
}

size_t align256(size_t v) { return (v + 255u) & ~(size_t)255u; }

} // namespace

struct sampo_engine {
	int device = 0;
	hipStream_t stream = nullptr;
	uint32_t max_slots = 0;
	/* the launches (pass, prep, plan) cover slots 0..nslots-1: at least one
	 * past the last slot in use.  It grows at admission and shrinks at a
	 * release that halves it; either change drops the captured graphs,
	 * whose launches hold it. */
	uint32_t nslots = 1;
	Inst *d_inst = nullptr;
	Ctx *d_ctx = nullptr;
	uint32_t *d_pl_base = nullptr;
	uint8_t *d_pl_par = nullptr;
	uint8_t *d_hold = nullptr; /* sampo_engine_hold's mask */
	std::vector<Slot> slots;
	/* tile classes: one pass launch per class each iteration, over the
	 * instances admitted under it; class 0 starts as three work-groups of 8
	 * waves per CU.  cur: the class of the next admissions. */
	struct TileClass {
		uint32_t wg = 512, apt = 1, waves = 6;
		pass_fn pass = nullptr;
		int grid = 0;
		uint32_t n_inst = 0; /* instances admitted under it and not released */
	} cls[MAX_CLASSES];
	uint32_t ncls = 1, cur = 0;
	uint32_t chunk = DEFAULT_CHUNK;
	uint32_t target_chunks = 0; /* 0: 2 per resident work-group of the pass */
	uint32_t per_cu = 0; /* persistent work-groups per CU; 0: as many as fit */
	hipGraphExec_t day = nullptr;
	hipGraph_t day_graph = nullptr;
	/* day launches (k_day): work-groups for later admissions (0: off), the
	 * tagged words, the next tag */
	uint32_t day_wg = 0;
	DaySync *d_day = nullptr;
	uint32_t day_seq = 1;
	/* sampo_engine_capture: n iterations with an event pair around each pass */
	hipGraphExec_t cap = nullptr;
	hipGraph_t cap_graph = nullptr;
	std::vector<hipEvent_t> cap_ev;
};

extern "C" const char *sampo_last_error(void) { return g_err; }

extern "C" void sampo_seed_stream(uint32_t seed, uint32_t n, uint32_t *out)
{
	sc_glibc_rand g;
	sc_glibc_seed(&g, seed);
	for (uint32_t i = 0; i < n; ++i) out[i] = sc_glibc_next(&g);
}

static void drop_graph(sampo_engine *e)
{
	if (e->day) (void)hipGraphExecDestroy(e->day);
	if (e->day_graph) (void)hipGraphDestroy(e->day_graph);
	e->day = nullptr;
	e->day_graph = nullptr;
	if (e->cap) (void)hipGraphExecDestroy(e->cap);
	if (e->cap_graph) (void)hipGraphDestroy(e->cap_graph);
	for (auto &x : e->cap_ev) (void)hipEventDestroy(x);
	e->cap = nullptr;
	e->cap_graph = nullptr;
	e->cap_ev.clear();
}

static int configure_pass(sampo_engine *e, uint32_t c)
{
	int blocks = 0, cus = 0;
	auto &k = e->cls[c];
	HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void *)k.pass, (int)k.wg, 0));
	HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
	if (blocks < 1) blocks = 1;
	if (e->per_cu && (int)e->per_cu < blocks) blocks = (int)e->per_cu;
	k.grid = blocks * cus;
	return 0;
}

extern "C" int sampo_engine_create(int device, uint32_t max_slots, sampo_engine **out)
{
	if (!out || max_slots == 0 || max_slots > (1u << 20)) return fail(SAMPO_E_ARG, "bad arguments%s%ld");
	*out = nullptr;
	int ndev = 0;
	HIPCHK(hipGetDeviceCount(&ndev));
	if (device < 0 || device >= ndev) return fail(SAMPO_E_ARG, "no HIP device %s%ld", "", device);
	HIPCHK(hipSetDevice(device));
	sampo_engine *e = new sampo_engine();
	e->device = device;
	e->max_slots = max_slots;
	e->slots.resize(max_slots);
	HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
	HIPCHK(hipMalloc(&e->d_inst, sizeof(Inst) * max_slots));
	HIPCHK(hipMalloc(&e->d_ctx, sizeof(Ctx)));
	HIPCHK(hipMalloc(&e->d_pl_base, sizeof(uint32_t) * (max_slots + 1) * MAX_CLASSES));
	HIPCHK(hipMalloc(&e->d_pl_par, max_slots));
	HIPCHK(hipMalloc(&e->d_hold, max_slots));
	HIPCHK(hipMemset(e->d_ctx, 0, sizeof(Ctx)));
	std::vector<Inst> blank(max_slots);
	memset(blank.data(), 0, sizeof(Inst) * max_slots);
	for (auto &b : blank) b.status = SAMPO_EMPTY;
	HIPCHK(hipMemcpy(e->d_inst, blank.data(), sizeof(Inst) * max_slots, hipMemcpyHostToDevice));
	e->cls[0].pass = pick_pass(e->cls[0].wg, e->cls[0].apt, e->cls[0].waves);
	int rc = configure_pass(e, 0);
	if (rc) { sampo_engine_destroy(e); return rc; }
	*out = e;
	return 0;
}

/* k_prep over the slots in use; with more than one, k_plan after it */
static void enqueue_prep(sampo_engine *e)
{
	hipLaunchKernelGGL(k_prep, dim3(e->nslots), dim3(PREP_WG), 0, e->stream, e->d_inst, e->d_ctx, e->d_pl_base,
	                   e->d_pl_par, e->ncls);
	if (e->nslots > 1u)
		hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, e->stream, e->d_inst, e->d_ctx, e->d_pl_base, e->d_pl_par,
		                   e->nslots, e->ncls);
}

static void free_slot(Slot &s)
{
	if (s.mem) (void)hipFree(s.mem);
	if (s.agents) (void)hipFree(s.agents);
	s = Slot();
}

extern "C" void sampo_engine_destroy(sampo_engine *e)
{
	if (!e) return;
	(void)hipSetDevice(e->device);
	if (e->stream) (void)hipStreamSynchronize(e->stream);
	drop_graph(e);
	for (auto &s : e->slots) free_slot(s);
	(void)hipFree(e->d_inst);
	(void)hipFree(e->d_ctx);
	(void)hipFree(e->d_pl_base);
	(void)hipFree(e->d_pl_par);
	(void)hipFree(e->d_hold);
	if (e->d_day) (void)hipFree(e->d_day);
	if (e->stream) (void)hipStreamDestroy(e->stream);
	delete e;
}

extern "C" int sampo_engine_set_tile(sampo_engine *e, uint32_t wg, uint32_t apt, uint32_t waves)
{
	if (!e) return fail(SAMPO_E_ARG, "null engine%s%ld");
	pass_fn f = pick_pass(wg, apt, waves);
	if (!f) return fail(SAMPO_E_ARG, "tile shape not compiled in%s%ld");
	for (uint32_t c = 0; c < e->ncls; ++c) /* a class of this shape: later admissions join it */
		if (e->cls[c].wg == wg && e->cls[c].apt == apt && e->cls[c].waves == waves) {
			e->cur = c;
			return 0;
		}
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	uint32_t c = e->ncls;
	for (uint32_t q = 0; q < e->ncls && c == e->ncls; ++q) /* reuse a class no instance uses */
		if (e->cls[q].n_inst == 0u) c = q;
	if (c == MAX_CLASSES) return fail(SAMPO_E_FULL, "at most %s%ld tile shapes in use at once", "", (long)MAX_CLASSES);
	if (c == e->ncls) { /* one more pass launch per iteration */
		++e->ncls;
		drop_graph(e);
	} else if (e->cls[c].n_inst == 0u) {
		drop_graph(e); /* the captured launch of that class has the old shape */
	}
	auto &k = e->cls[c];
	k.wg = wg;
	k.apt = apt;
	k.waves = waves;
	k.pass = f;
	e->cur = c;
	return configure_pass(e, c);
}

extern "C" int sampo_engine_set_grid(sampo_engine *e, uint32_t per_cu)
{
	if (!e || per_cu > 64u) return fail(SAMPO_E_ARG, "per_cu must be 0..64%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	drop_graph(e);
	e->per_cu = per_cu;
	for (uint32_t c = 0; c < e->ncls; ++c) {
		const int rc = configure_pass(e, c);
		if (rc) return rc;
	}
	return 0;
}

extern "C" int sampo_engine_set_day(sampo_engine *e, uint32_t work_groups)
{
	if (!e || work_groups > DAY_CHUNKS) return fail(SAMPO_E_ARG, "work_groups must be 0..%s%ld", "", (long)DAY_CHUNKS);
	if (work_groups) {
		int coop = 0, blocks = 0, cus = 0;
		HIPCHK(hipSetDevice(e->device));
		HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, e->device));
		if (!coop) return fail(SAMPO_E_STATE, "device has no cooperative launch%s%ld");
		HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_day(), DAY_WG, 0));
		HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
		if ((long)blocks * cus < (long)work_groups)
			return fail(SAMPO_E_ARG, "at most %s%ld work-groups of a day launch fit at once", "", (long)blocks * cus);
		if (!e->d_day) {
			HIPCHK(hipStreamSynchronize(e->stream));
			HIPCHK(hipMalloc(&e->d_day, sizeof(DaySync)));
			HIPCHK(hipMemset(e->d_day, 0, sizeof(DaySync)));
		}
	}
	e->day_wg = work_groups;
	return 0;
}

extern "C" int sampo_engine_set_chunk(sampo_engine *e, uint32_t agents, uint32_t target_chunks)
{
	if (!e || agents < 64u || agents > (1u << 26)) return fail(SAMPO_E_ARG, "chunk size must be 64..2^26%s%ld");
	e->chunk = agents;
	e->target_chunks = target_chunks;
	return 0;
}

/* layout of one instance's bookkeeping allocation (all but the agent buffers) */
struct SlotLayout {
	size_t o_seg0, o_seg1, o_ch0, o_ch1, o_cout, o_seeds, o_temp, o_env, o_rows, o_sprow, total;
	uint32_t cmax, smax;
};

static SlotLayout slot_layout(uint64_t P_max, uint32_t chunk, uint32_t horizon, uint32_t env_count)
{
	SlotLayout L;
	L.cmax = (uint32_t)((P_max + chunk - 1u) / chunk) + 2u;
	L.smax = 16u * L.cmax + 2u;
	size_t off = 0;
	L.o_seg0 = off; off = align256(off + (size_t)L.smax * sizeof(Seg));
	L.o_seg1 = off; off = align256(off + (size_t)L.smax * sizeof(Seg));
	L.o_ch0 = off; off = align256(off + (size_t)L.cmax * sizeof(Chunk));
	L.o_ch1 = off; off = align256(off + (size_t)L.cmax * sizeof(Chunk));
	L.o_cout = off; off = align256(off + (size_t)NOUT * L.cmax * sizeof(uint32_t));
	L.o_seeds = off; off = align256(off + (size_t)(horizon + 1u) * 16u * sizeof(uint32_t));
	L.o_temp = off; off = align256(off + (size_t)horizon * sizeof(float));
	L.o_env = off; off = align256(off + (size_t)env_count * sizeof(sampo_env));
	L.o_rows = off; off = align256(off + (size_t)(horizon + 1u) * sizeof(sampo_row));
	L.o_sprow = off; off = align256(off + (size_t)horizon * sizeof(SpRow));
	L.total = off;
	return L;
}

static uint64_t round256(uint64_t v) { return ((v + 255ull) / 256ull) * 256ull; }

/* physical slots: what the caller asked for, else max_agents; never fewer
 * than the initial population, never more than max_agents */
static uint64_t initial_slots(const sampo_scenario *sc, uint64_t n0)
{
	uint64_t pmax = round256(sc->max_agents);
	uint64_t P = sc->phys_slots ? round256(sc->phys_slots) : pmax;
	if (P < round256(n0)) P = round256(n0);
	if (P < 256u) P = 256u;
	return P < pmax ? P : pmax;
}

extern "C" uint64_t sampo_engine_footprint(const sampo_scenario *sc, uint32_t chunk)
{
	if (!sc || sc->horizon == 0 || sc->max_agents == 0) return 0;
	uint64_t n0 = sc->mode == SAMPO_MODE_REFERENCE ? sc->n_eggs : 0u;
	if (sc->mode == SAMPO_MODE_CONFIG && sc->init_layout)
		for (int s = 0; s < 8; ++s) n0 += sc->init_layout->end[s] - sc->init_layout->start[s];
	const uint64_t P = initial_slots(sc, n0);
	const SlotLayout L = slot_layout(round256(sc->max_agents), chunk ? chunk : DEFAULT_CHUNK, sc->horizon,
	                                 sc->env_count ? sc->env_count : 1u);
	/* plus the staging of the initial population while it is imported */
	const uint64_t staging = sc->mode == SAMPO_MODE_CONFIG ? n0 * (uint64_t)sizeof(sampo_agent) : 0u;
	return (uint64_t)L.total + 2ull * NFIELD * 4ull * P + staging;
}

extern "C" int sampo_engine_admit(sampo_engine *e, const sampo_scenario *sc, uint32_t *slot_out)
{
	if (!e || !sc || !slot_out) return fail(SAMPO_E_ARG, "null argument%s%ld");
	if (sc->horizon == 0 || !sc->temp || !sc->env) return fail(SAMPO_E_ARG, "horizon, temp and env are required%s%ld");
	if (sc->env_count != 1 && sc->env_count != sc->horizon)
		return fail(SAMPO_E_ARG, "env_count must be 1 or horizon%s%ld");
	if (sc->mode != SAMPO_MODE_REFERENCE && sc->mode != SAMPO_MODE_CONFIG)
		return fail(SAMPO_E_ARG, "unknown mode%s%ld");
	if (sc->max_agents == 0) return fail(SAMPO_E_ARG, "max_agents is required%s%ld");
	for (uint32_t t = 0; t < sc->horizon; ++t) /* the species functions index tables by temperature */
		if (!__builtin_isfinite(sc->temp[t])) return fail(SAMPO_E_ARG, "temperature %s%ld is not finite", "", (long)t);
	/* slots are addressed with 32-bit byte offsets per field */
	if (round256(sc->max_agents) > (1ull << 30)) return fail(SAMPO_E_ARG, "max_agents exceeds 2^30%s%ld");
	uint32_t id = e->max_slots;
	for (uint32_t i = 0; i < e->max_slots; ++i)
		if (!e->slots[i].used) { id = i; break; }
	if (id == e->max_slots) return fail(SAMPO_E_FULL, "no free slot%s%ld");

	Inst h;
	memset(&h, 0, sizeof(h));
	h.horizon = sc->horizon;
	h.mode = sc->mode;
	h.env_count = sc->env_count;
	h.max_agents = sc->max_agents;
	h.chunk_size = e->chunk;
	h.target_chunks = e->target_chunks ? e->target_chunks : 2u * (uint32_t)e->cls[e->cur].grid;
	h.tile_class = e->cur;
	h.P_max = (uint32_t)round256(sc->max_agents);
	sampo_species sp;
	if (sc->species) sp = *sc->species;
	else sampo_species_default(&sp);
	if (sampo_species_check(&sp)) return fail(SAMPO_E_ARG, "invalid species table%s%ld");
	for (int q = 0; q < 2; ++q) { /* species constants of every iteration */
		IterState &x = h.st[q];
		x.ms_dly = sp.mate_seeking_delay;
		x.bms_dly = sp.bms_delay;
		x.larva_mean = sp.larva_delay[0];
		x.larva_sd = sp.larva_delay[1];
		x.batch_mean = sp.egg_batch[0];
		x.batch_sd = sp.egg_batch[1];
		x.batch_decay = sp.egg_batch[2];
		x.lay_p = sp.lay_probability;
		x.bite_human = sp.bite_human;
		memcpy(x.hatch_thr, sp.hatch_threshold, sizeof x.hatch_thr);
		memcpy(x.hatch_m, sp.hatch_slope, sizeof x.hatch_m);
		memcpy(x.hatch_b, sp.hatch_offset, sizeof x.hatch_b);
	}
	for (int c = 0; c < 4; ++c) h.hmr_stage[c] = sm_hmr(sc->dmr[c]);
	for (uint32_t d = 0; d < 100; ++d) h.hmr_adult[d] = sc->mode == SAMPO_MODE_CONFIG ? h.hmr_stage[3] : sm_hmr_adult(d);
	IterState &s0 = h.st[0];

	/* initial population (src/abms.c:185-228, or the given one) and its segments */
	std::vector<Seg> segs;
	uint64_t stored = 0;
	if (sc->mode == SAMPO_MODE_REFERENCE) {
		if ((uint64_t)sc->n_eggs > sc->max_agents) return fail(SAMPO_E_ARG, "n_eggs exceeds max_agents%s%ld");
		s0.cnt[0] = sc->n_eggs;
		s0.enew = sc->n_eggs;
		for (int s = 1; s < 8; ++s) s0.vstart[s] = sc->n_eggs + 1u;
		if (sc->n_eggs) segs.push_back(Seg{0u, 0u, PB_NEWBORN, 0u});
	} else {
		if (!sc->init_layout || !sc->init_agents) return fail(SAMPO_E_ARG, "config mode needs an initial population%s%ld");
		const sampo_layout *L = sc->init_layout;
		if (L->start[0] != 0) return fail(SAMPO_E_ARG, "egg range must start at 0%s%ld");
		uint64_t k = 0;
		for (int s = 0; s < 8; ++s) {
			if (L->end[s] < L->start[s]) return fail(SAMPO_E_ARG, "range %s%ld has end < start", "", s);
			if (s > 0 && L->start[s] <= L->end[s - 1]) return fail(SAMPO_E_ARG, "range %s%ld needs a gap before it", "", s);
			s0.cnt[s] = L->end[s] - L->start[s];
			s0.vstart[s] = L->start[s];
			if (s0.cnt[s]) segs.push_back(Seg{(uint32_t)k, (uint32_t)s, (uint32_t)k, L->start[s]});
			k += s0.cnt[s];
		}
		stored = k;
		if (stored > sc->max_agents) return fail(SAMPO_E_ARG, "initial population exceeds max_agents%s%ld");
	}
	{ /* humanBloodmealCount and numEggBatches are stored in 16 and 15 bits.
	     Each grows by at most one per gonotrophic cycle (blood meal seeking ->
	     digesting -> gravid -> seeking), which takes at least three
	     iterations since an agent changes state at most once per iteration,
	     so neither can pass start + ceil(horizon / 3).  Refuse a run in which
	     they could wrap. */
		uint64_t hbc0 = 0, nb0 = 0;
		float hours0 = 0.0f;
		uint64_t cyc0 = 0;
		for (uint64_t i = 0; i < stored; ++i) {
			const sampo_agent &a = sc->init_agents[i];
			if (a.humanBloodmealCount > hbc0) hbc0 = a.humanBloodmealCount;
			if (a.numEggBatches > nb0) nb0 = a.numEggBatches;
			if (a.cycleLength > cyc0) cyc0 = a.cycleLength;
			/* the hour counts are whole, non-negative numbers of hours */
			if (!(a.ageInHours >= 0.0f && a.hoursInState >= 0.0f && a.ageInHours < 65536.0f &&
			      a.hoursInState < 65536.0f && a.ageInHours == __builtin_floorf(a.ageInHours) &&
			      a.hoursInState == __builtin_floorf(a.hoursInState)))
				return fail(SAMPO_E_ARG, "agent %s%ld: ageInHours and hoursInState must be whole hours in [0, 2^16)",
				            "", (long)i);
			if (a.ageInHours > hours0) hours0 = a.ageInHours;
			if (a.hoursInState > hours0) hours0 = a.hoursInState;
		}
		const uint64_t cycles = ((uint64_t)sc->horizon + 2u) / 3u;
		if (hbc0 + cycles > 0xffffu || nb0 + cycles > 0x7fffu)
			return fail(SAMPO_E_ARG, "horizon too long: humanBloodmealCount or numEggBatches could exceed its "
			                         "16/15-bit field%s%ld");
		/* both hour counts and cycleLength grow by at most one an iteration */
		if ((uint64_t)hours0 + sc->horizon > 0xffffu || cyc0 + sc->horizon > 0xffffu)
			return fail(SAMPO_E_ARG, "horizon too long: ageInHours, hoursInState or cycleLength could exceed "
			                         "its 16-bit field%s%ld");
	}
	{ /* immatures keep only their non-zero words (sm_fields) when the initial
	     eggs, larvae and pupae have zero development sum, eggs, cycle length,
	     blood meals and batches (and the eggs no larval delay): the update
	     rules then keep those words zero (temperatures are finite, checked
	     below) */
		h.compact = 1u;
		uint64_t k = 0;
		for (uint32_t st = 0; st <= S_PUPA && stored && h.compact; ++st)
			for (uint32_t i = 0; i < s0.cnt[st]; ++i, ++k) {
				const sampo_agent &a = sc->init_agents[k];
				if (sc_bits_f(a.csd) || a.availableEggs || a.cycleLength || a.humanBloodmealCount ||
				    a.numEggBatches || (st == S_EGG && a.aux)) {
					h.compact = 0u;
					break;
				}
			}
	}
	const uint64_t P64 = initial_slots(sc, sc->mode == SAMPO_MODE_REFERENCE ? (uint64_t)sc->n_eggs : stored);
	const uint32_t P = (uint32_t)P64;
	h.P = P;
	s0.nseg = (uint32_t)segs.size();
	s0.step = 0;
	s0.cur = 0;
	h.step = 0;
	h.status = SAMPO_RUNNING;
	h.needs_begin = 1u;
	const SlotLayout L = slot_layout(h.P_max, e->chunk, sc->horizon, sc->env_count);
	h.cmax = L.cmax;
	h.smax = L.smax;
	if (e->day_wg) { /* day launches: a work-group per chunk, as many chunks as the counters hold */
		h.even_cut = 1u;
		h.target_chunks = e->day_wg < L.cmax - 2u ? e->day_wg : L.cmax - 2u;
	}

	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	/* Every allocation first: a failure here leaves nothing behind.  The
	 * instance record is uploaded as an empty slot and turned into a running
	 * one only after its inputs and initial population are in place, so on
	 * any later failure the device record is reset to empty before the
	 * memory it points into is freed (no kernel reads an empty slot's
	 * pointers). */
	Slot S;
	S.horizon = sc->horizon;
	S.P = P;
	sampo_agent *d_in = nullptr; /* staging of the initial population for k_import */
	hipError_t err = hipMalloc(&S.mem, L.total);
	if (err == hipSuccess) err = hipMalloc(&S.agents, (size_t)2 * NFIELD * P * sizeof(uint32_t));
	if (err == hipSuccess && stored) err = hipMalloc(&d_in, (size_t)stored * sizeof(sampo_agent));
	if (err != hipSuccess) {
		(void)hipGetLastError();
		if (d_in) (void)hipFree(d_in);
		free_slot(S);
		return fail(SAMPO_E_NOMEM, "device allocation failed: %s%ld", hipGetErrorString(err));
	}
	char *base = (char *)S.mem;
	const size_t seeds_n = (size_t)(sc->horizon + 1u) * 16u;
	std::vector<uint32_t> seeds(seeds_n);
	sampo_seed_stream(sc->seed, (uint32_t)seeds_n, seeds.data());
	for (size_t q = 0; q < seeds_n; q += 4) {
		uint32_t z[4] = {seeds[q], seeds[q + 1], seeds[q + 2], seeds[q + 3]};
		sm_taus_prepare(z, &seeds[q]);
	}
	h.buf[0] = (uint32_t *)S.agents;
	h.buf[1] = h.buf[0] + (size_t)NFIELD * P;
	h.seg[0] = (Seg *)(base + L.o_seg0);
	h.seg[1] = (Seg *)(base + L.o_seg1);
	h.chunk[0] = (Chunk *)(base + L.o_ch0);
	h.chunk[1] = (Chunk *)(base + L.o_ch1);
	h.cout = (uint32_t *)(base + L.o_cout);
	h.seeds = (const uint32_t *)(base + L.o_seeds);
	h.temp = (const float *)(base + L.o_temp);
	h.env = (const sampo_env *)(base + L.o_env);
	h.rows = (sampo_row *)(base + L.o_rows);
	h.sprow = (const SpRow *)(base + L.o_sprow);
	std::vector<SpRow> rows(sc->horizon);
	for (uint32_t t = 0; t < sc->horizon; ++t) {
		const float T = sc->temp[t];
		const sampo_env &en = sc->env[sc->env_count == 1u ? 0u : t];
		const uint32_t hr = sc_f2u((float)t * 1.0f) % 24u; /* worldTime, src/abms.c:514 */
		SpRow &r = rows[t];
		r.dev_rate = sm_dev_rate(&sp, T);
		r.pupa_dly = sm_pupa_delay(&sp, T);
		r.imm_dly = sm_linear(sp.immature_delay, T);
		r.bmd_dly = sm_linear(sp.bmd_delay, T);
		r.inc = sm_linear(sp.egg_incubation, T);
		r.itn_kill = en.ITNValue * sp.bite_human * sp.bite_indoor;
		r.irs_kill = en.IRSValue * sp.rest_indoor;
		r.open = 0u;
		for (int q = 0; q < 8; ++q) r.open |= (uint32_t)sm_open(&sp, q, hr) << q;
	}
	const uint32_t run_status = h.status, run_begin = h.needs_begin;
	h.status = SAMPO_EMPTY; /* inert until the population is in place */
	h.needs_begin = 0u;
	bool uploaded = false;
	err = hipMemcpy((void *)h.seeds, seeds.data(), seeds_n * sizeof(uint32_t), hipMemcpyHostToDevice);
	if (err == hipSuccess) err = hipMemcpy((void *)h.temp, sc->temp, (size_t)sc->horizon * sizeof(float), hipMemcpyHostToDevice);
	if (err == hipSuccess) err = hipMemcpy((void *)h.env, sc->env, (size_t)sc->env_count * sizeof(sampo_env), hipMemcpyHostToDevice);
	if (err == hipSuccess) err = hipMemset(h.rows, 0, (size_t)(sc->horizon + 1u) * sizeof(sampo_row));
	if (err == hipSuccess) err = hipMemcpy((void *)h.sprow, rows.data(), rows.size() * sizeof(SpRow), hipMemcpyHostToDevice);
	if (err == hipSuccess && !segs.empty()) err = hipMemcpy(h.seg[0], segs.data(), segs.size() * sizeof(Seg), hipMemcpyHostToDevice);
	if (err == hipSuccess) {
		err = hipMemcpy(e->d_inst + id, &h, sizeof(Inst), hipMemcpyHostToDevice);
		uploaded = err == hipSuccess;
	}
	if (err == hipSuccess && stored) { /* host records -> SoA, and the start's calcStats sums, on the device */
		err = hipMemcpy(d_in, sc->init_agents, (size_t)stored * sizeof(sampo_agent), hipMemcpyHostToDevice);
		if (err == hipSuccess) {
			const uint32_t k_l0 = s0.cnt[0], k_l1 = k_l0 + s0.cnt[1];
			const uint32_t k_a0 = k_l1 + s0.cnt[2];
			hipLaunchKernelGGL(k_import, dim3((uint32_t)((stored + 255u) / 256u)), dim3(256), 0, e->stream, d_in,
			                   (uint32_t)stored, k_l0, k_l1, k_a0, h.buf[0], P, e->d_inst + id);
			err = hipGetLastError();
			if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
		}
	}
	if (d_in) (void)hipFree(d_in);
	if (err == hipSuccess) { /* now it may run */
		const uint32_t flags[2] = {run_status, run_begin};
		static_assert(offsetof(Inst, needs_begin) == offsetof(Inst, status) + 3 * sizeof(uint32_t),
		              "status, step, pending, needs_begin are adjacent");
		err = hipMemcpy(&e->d_inst[id].status, &flags[0], sizeof(uint32_t), hipMemcpyHostToDevice);
		if (err == hipSuccess) err = hipMemcpy(&e->d_inst[id].needs_begin, &flags[1], sizeof(uint32_t), hipMemcpyHostToDevice);
	}
	if (err == hipSuccess && id + 1u > e->nslots) {
		e->nslots = id + 1u;
		drop_graph(e);
	}
	if (err == hipSuccess) { /* begin iteration 0 on the device (calcStats of src/abms.c:503) */
		enqueue_prep(e);
		err = hipGetLastError();
		if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
	}
	if (err != hipSuccess) {
		(void)hipGetLastError();
		if (uploaded) { /* the slot must not be run on freed memory */
			const uint32_t empty[4] = {SAMPO_EMPTY, 0u, 0u, 0u};
			(void)hipStreamSynchronize(e->stream);
			(void)hipMemcpy(&e->d_inst[id].status, empty, sizeof(empty), hipMemcpyHostToDevice);
		}
		free_slot(S);
		return fail(SAMPO_E_HIP, "admission failed: %s%ld", hipGetErrorString(err));
	}
	S.used = true;
	S.cls = e->cur;
	S.day = h.even_cut ? e->day_wg : 0u;
	e->cls[e->cur].n_inst += 1u;
	e->slots[id] = S;
	*slot_out = id;
	return 0;
}

extern "C" int sampo_engine_release(sampo_engine *e, uint32_t slot)
{
	if (!e || slot >= e->max_slots || !e->slots[slot].used) return fail(SAMPO_E_ARG, "bad slot%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	uint32_t st = SAMPO_EMPTY;
	HIPCHK(hipMemcpy(&e->d_inst[slot].status, &st, sizeof(st), hipMemcpyHostToDevice));
	uint32_t hi = 1; /* one past the last slot still in use */
	for (uint32_t i = 0; i < e->max_slots; ++i)
		if (e->slots[i].used && i != slot) hi = i + 1u;
	if (hi <= e->nslots / 2u) {
		e->nslots = hi;
		drop_graph(e);
	}
	/* plan the next pass without it before its memory goes */
	hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, e->stream, e->d_inst, e->d_ctx, e->d_pl_base, e->d_pl_par,
	                   e->nslots, e->ncls);
	HIPCHK(hipGetLastError());
	HIPCHK(hipStreamSynchronize(e->stream));
	e->cls[e->slots[slot].cls].n_inst -= 1u;
	free_slot(e->slots[slot]);
	return 0;
}

extern "C" int sampo_engine_hold(sampo_engine *e, uint32_t n, const uint8_t *hold)
{
	if (!e || (n && !hold) || n > e->max_slots) return fail(SAMPO_E_ARG, "bad arguments%s%ld");
	HIPCHK(hipSetDevice(e->device));
	if (n == 0) return 0;
	HIPCHK(hipMemcpyAsync(e->d_hold, hold, n, hipMemcpyHostToDevice, e->stream));
	hipLaunchKernelGGL(k_hold, dim3(1), dim3(1024), 0, e->stream, e->d_inst, e->d_hold, n, e->d_ctx, e->d_pl_base,
	                   e->d_pl_par, e->nslots, e->ncls);
	HIPCHK(hipGetLastError());
	HIPCHK(hipStreamSynchronize(e->stream));
	return 0;
}

extern "C" int sampo_engine_poll(sampo_engine *e, uint32_t n, uint32_t *status, uint32_t *step, uint64_t *n_agents)
{
	if (!e || n > e->max_slots) return fail(SAMPO_E_ARG, "bad arguments%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	std::vector<Inst> h(n);
	if (n) HIPCHK(hipMemcpy(h.data(), e->d_inst, sizeof(Inst) * n, hipMemcpyDeviceToHost));
	for (uint32_t i = 0; i < n; ++i) {
		const bool used = e->slots[i].used;
		if (status) status[i] = used ? h[i].status : SAMPO_EMPTY;
		if (step) step[i] = used ? h[i].step : 0u;
		if (n_agents) n_agents[i] = used ? h[i].st[h[i].step & 1u].n : 0u;
	}
	return 0;
}

/* Move one instance's agents into buffers of P2 slots per field. */
static int grow_slot(sampo_engine *e, uint32_t id, Inst &h, uint32_t P2)
{
	Slot &S = e->slots[id];
	void *mem = nullptr;
	hipError_t err = hipMalloc(&mem, (size_t)2 * NFIELD * P2 * sizeof(uint32_t));
	if (err != hipSuccess) {
		(void)hipGetLastError();
		return fail(SAMPO_E_NOMEM, "growing an instance failed: %s%ld", hipGetErrorString(err));
	}
	uint32_t *nb = (uint32_t *)mem;
	const uint32_t cur = h.st[h.step & 1u].cur;
	/* the current buffer holds the live agents in its first P slots per field */
	for (int f = 0; f < NFIELD; ++f)
		HIPCHK(hipMemcpyAsync(nb + ((size_t)cur * NFIELD + f) * P2, h.buf[cur] + (size_t)f * h.P,
		                      (size_t)h.P * sizeof(uint32_t), hipMemcpyDeviceToDevice, e->stream));
	HIPCHK(hipStreamSynchronize(e->stream));
	(void)hipFree(S.agents);
	S.agents = mem;
	S.P = P2;
	h.buf[0] = nb;
	h.buf[1] = nb + (size_t)NFIELD * P2;
	h.P = P2;
	return 0;
}

extern "C" int sampo_engine_service(sampo_engine *e, float headroom, uint32_t *grown)
{
	if (!e || !(headroom >= 1.0f)) return fail(SAMPO_E_ARG, "headroom must be >= 1%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	uint32_t hi = 0; /* slots past the last one in use are not read */
	for (uint32_t i = 0; i < e->max_slots; ++i)
		if (e->slots[i].used) hi = i + 1u;
	std::vector<Inst> h(hi);
	if (hi) HIPCHK(hipMemcpy(h.data(), e->d_inst, sizeof(Inst) * hi, hipMemcpyDeviceToHost));
	uint32_t count = 0, resumed = 0;
	int rc = 0;
	for (uint32_t i = 0; i < hi && !rc; ++i) {
		if (!e->slots[i].used) continue;
		Inst &I = h[i];
		const uint64_t n = I.st[I.step & 1u].n;
		const uint64_t want = (uint64_t)((double)n * headroom) + 1u;
		const bool paused = I.status == SAMPO_GROW;
		if (!paused && !(I.status == SAMPO_RUNNING && want > I.P && I.P < I.P_max)) continue;
		uint64_t P2 = round256(want + want / 4u);
		if (P2 < round256((uint64_t)I.P + I.P / 4u)) P2 = round256((uint64_t)I.P + I.P / 4u);
		if (P2 > I.P_max) P2 = I.P_max;
		rc = grow_slot(e, i, I, (uint32_t)P2);
		if (rc) break;
		if (paused) { /* run its calcStats again and continue */
			I.status = SAMPO_RUNNING;
			I.needs_begin = 1u;
			++resumed;
		}
		HIPCHK(hipMemcpy(e->d_inst + i, &I, sizeof(Inst), hipMemcpyHostToDevice));
		++count;
	}
	if (resumed) {
		enqueue_prep(e);
		HIPCHK(hipGetLastError());
		HIPCHK(hipStreamSynchronize(e->stream));
	}
	if (grown) *grown = count;
	return rc;
}

/* One iteration: the pass over every running instance, then k_prep, which
 * folds it, begins the next iteration and plans the next pass. */
/* One iteration: the fused pass and k_prep, optionally between two events.
 * Under stream capture the events are recorded as external event nodes of
 * the graph, so that they keep their timestamps on replay. */
static int enqueue_one(sampo_engine *e, hipEvent_t before_pass, hipEvent_t after_pass, bool capturing = false)
{
	const unsigned int ef = capturing ? hipEventRecordExternal : hipEventRecordDefault;
	if (before_pass) HIPCHK(hipEventRecordWithFlags(before_pass, e->stream, ef));
	for (uint32_t c = 0; c < e->ncls; ++c) { /* one pass per tile class, each over its own chunks */
		uint32_t cc = c;
		DayArgs none = {};
		void *args[] = {&e->d_inst, &e->d_ctx, &e->d_pl_base, &e->d_pl_par, &e->nslots, &cc, &none};
		HIPCHK(hipLaunchKernel((const void *)e->cls[c].pass, dim3(e->cls[c].grid), dim3(e->cls[c].wg), args, 0,
		                       e->stream));
	}
	if (after_pass) HIPCHK(hipEventRecordWithFlags(after_pass, e->stream, ef));
	enqueue_prep(e);
	return 0;
}

/* The lone instance advances by day launches: it is in slot 0 (the
 * launches cover one slot) and was admitted for them. */
static bool day_mode(const sampo_engine *e)
{
	return e->nslots == 1u && e->slots[0].used && e->slots[0].day && e->d_day;
}

/* n iterations as day launches of at most DAY_ITERS iterations each */
static int enqueue_days(sampo_engine *e, uint32_t n)
{
	const uint32_t G = e->slots[0].day;
	while (n) {
		const uint32_t it = n < DAY_ITERS ? n : DAY_ITERS;
		if (e->day_seq > 0xffffffffu - DAY_ITERS) { /* tags are never reused: clear them and start over */
			HIPCHK(hipMemsetAsync(e->d_day->w, 0, sizeof(e->d_day->w), e->stream));
			e->day_seq = 1u;
		}
		DayArgs da = {e->d_day, e->day_seq, it, e->d_pl_base, e->d_pl_par, e->ncls};
		uint32_t one = 1u, cls0 = 0u;
		void *args[] = {&e->d_inst, &e->d_ctx, &e->d_pl_base, &e->d_pl_par, &one, &cls0, &da};
		HIPCHK(hipLaunchCooperativeKernel(k_day(), dim3(G), dim3(DAY_WG), args, 0, e->stream));
		e->day_seq += it;
		n -= it;
	}
	return 0;
}

static int enqueue(sampo_engine *e, uint32_t n)
{
	if (day_mode(e)) return enqueue_days(e, n);
	for (uint32_t i = 0; i < n; ++i) {
		int rc = enqueue_one(e, nullptr, nullptr);
		if (rc) return rc;
	}
	HIPCHK(hipGetLastError());
	return 0;
}

extern "C" int sampo_engine_step(sampo_engine *e, uint32_t n)
{
	if (!e) return fail(SAMPO_E_ARG, "null engine%s%ld");
	HIPCHK(hipSetDevice(e->device));
	return enqueue(e, n);
}

extern "C" int sampo_engine_step_days(sampo_engine *e, uint32_t days)
{
	if (!e) return fail(SAMPO_E_ARG, "null engine%s%ld");
	HIPCHK(hipSetDevice(e->device));
	if (day_mode(e)) return enqueue_days(e, 24u * days);
	if (!e->day) {
		HIPCHK(hipStreamSynchronize(e->stream));
		HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
		int rc = enqueue(e, 24);
		hipGraph_t g = nullptr;
		hipError_t ce = hipStreamEndCapture(e->stream, &g);
		if (rc) return rc;
		HIPCHK(ce);
		e->day_graph = g;
		HIPCHK(hipGraphInstantiate(&e->day, g, nullptr, nullptr, 0));
	}
	for (uint32_t d = 0; d < days; ++d) HIPCHK(hipGraphLaunch(e->day, e->stream));
	return 0;
}

extern "C" int sampo_engine_sync(sampo_engine *e)
{
	if (!e) return fail(SAMPO_E_ARG, "null engine%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	return 0;
}

extern "C" int sampo_engine_time(sampo_engine *e, uint32_t n, int use_graph, float *ms)
{
	if (!e || !ms) return fail(SAMPO_E_ARG, "null argument%s%ld");
	if (use_graph && n % 24u) return fail(SAMPO_E_ARG, "graph timing needs whole days%s%ld");
	HIPCHK(hipSetDevice(e->device));
	if (use_graph) { int rc = sampo_engine_step_days(e, 0); if (rc) return rc; }
	hipEvent_t a, b;
	HIPCHK(hipEventCreate(&a));
	HIPCHK(hipEventCreate(&b));
	HIPCHK(hipEventRecord(a, e->stream));
	int rc = use_graph ? sampo_engine_step_days(e, n / 24u) : enqueue(e, n);
	if (rc) return rc;
	HIPCHK(hipEventRecord(b, e->stream));
	HIPCHK(hipEventSynchronize(b));
	HIPCHK(hipEventElapsedTime(ms, a, b));
	(void)hipEventDestroy(a);
	(void)hipEventDestroy(b);
	return 0;
}

static int read_inst(sampo_engine *e, uint32_t slot, Inst *h)
{
	if (!e || slot >= e->max_slots || !e->slots[slot].used) return fail(SAMPO_E_ARG, "bad slot%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	HIPCHK(hipMemcpy(h, e->d_inst + slot, sizeof(Inst), hipMemcpyDeviceToHost));
	return 0;
}

extern "C" int sampo_engine_status(sampo_engine *e, uint32_t slot, uint32_t *status, uint32_t *step,
                                   uint64_t *n_agents)
{
	Inst h;
	int rc = read_inst(e, slot, &h);
	if (rc) return rc;
	if (status) *status = h.status;
	if (step) *step = h.step;
	if (n_agents) *n_agents = h.st[h.step & 1u].n;
	return 0;
}

extern "C" int sampo_engine_rows(sampo_engine *e, uint32_t slot, uint32_t first, uint32_t n, sampo_row *out)
{
	Inst h;
	int rc = read_inst(e, slot, &h);
	if (rc) return rc;
	if (!out || (uint64_t)first + n > (uint64_t)h.horizon + 1u) return fail(SAMPO_E_ARG, "rows out of range%s%ld");
	HIPCHK(hipMemcpy(out, h.rows + first, (size_t)n * sizeof(sampo_row), hipMemcpyDeviceToHost));
	return 0;
}

extern "C" int sampo_engine_export(sampo_engine *e, uint32_t slot, sampo_layout *layout, sampo_agent *agents,
                                   uint64_t max_agents)
{
	Inst h;
	int rc = read_inst(e, slot, &h);
	if (rc) return rc;
	const uint32_t par = h.step & 1u;
	const IterState &ip = h.st[par];
	if (layout)
		for (int s = 0; s < 8; ++s) {
			layout->start[s] = ip.vstart[s];
			layout->end[s] = ip.vstart[s] + ip.cnt[s];
		}
	if (!agents) return 0;
	if (ip.n > max_agents) return fail(SAMPO_E_ARG, "output holds fewer than n agents%s%ld");
	if (ip.n == 0) return 0;
	sampo_agent *d = nullptr;
	HIPCHK(hipMalloc(&d, (size_t)ip.n * sizeof(sampo_agent)));
	const uint32_t *nbz = h.seeds + (size_t)ip.nb_set * 16u + 8u;
	Hatch hb;
	hb.inc = ip.nb_inc;
	memcpy(hb.thr, ip.hatch_thr, sizeof hb.thr);
	memcpy(hb.m, ip.hatch_m, sizeof hb.m);
	memcpy(hb.b, ip.hatch_b, sizeof hb.b);
	hipLaunchKernelGGL(k_export, dim3((ip.n + 255u) / 256u), dim3(256), 0, e->stream, h.buf[ip.cur], h.P, h.seg[par],
	                   ip.nseg, ip.n, nbz, hb, h.compact, d);
	hipError_t err = hipGetLastError();
	if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
	if (err == hipSuccess) err = hipMemcpy(agents, d, (size_t)ip.n * sizeof(sampo_agent), hipMemcpyDeviceToHost);
	(void)hipFree(d);
	HIPCHK(err);
	return 0;
}

extern "C" int sampo_engine_histogram(sampo_engine *e, uint32_t slot, uint32_t hist[100])
{
	Inst h;
	int rc = read_inst(e, slot, &h);
	if (rc) return rc;
	const uint32_t par = h.step & 1u;
	const IterState &ip = h.st[par];
	uint32_t *d = nullptr;
	HIPCHK(hipMalloc(&d, 100 * sizeof(uint32_t)));
	hipError_t err = hipMemsetAsync(d, 0, 100 * sizeof(uint32_t), e->stream);
	uint32_t adults = ip.n - ip.kstart[S_IMM];
	if (err == hipSuccess && adults > 0) {
		uint32_t blocks = (adults + 255u) / 256u;
		if (blocks > 2048u) blocks = 2048u;
		hipLaunchKernelGGL(k_histogram, dim3(blocks), dim3(256), 0, e->stream, h.buf[ip.cur], h.seg[par], ip.nseg,
		                   ip.kstart[S_IMM], ip.n, d);
		err = hipGetLastError();
	}
	if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
	if (err == hipSuccess) err = hipMemcpy(hist, d, 100 * sizeof(uint32_t), hipMemcpyDeviceToHost);
	(void)hipFree(d);
	HIPCHK(err);
	return 0;
}

extern "C" int sampo_engine_agent_steps(sampo_engine *e, uint64_t *total, int reset)
{
	if (!e || !total) return fail(SAMPO_E_ARG, "null argument%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	Ctx c;
	HIPCHK(hipMemcpy(&c, e->d_ctx, sizeof(Ctx), hipMemcpyDeviceToHost));
	*total = c.agent_steps;
	if (reset) {
		unsigned long long z = 0;
		HIPCHK(hipMemcpy(&e->d_ctx->agent_steps, &z, sizeof(z), hipMemcpyHostToDevice));
	}
	return 0;
}

/* ============================================================ workloads
 * Synthetic scenario inputs for the configuration mode: a host splitmix64
 * stream, independent of the reference seed stream, draws the initial
 * population and the daily temperature series. */

static uint64_t sm64_next(uint64_t *x)
{
	uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

static double sm64_unit(uint64_t *x) { return (double)(sm64_next(x) >> 11) * (1.0 / 9007199254740992.0); }

/* standard normal by Box-Muller in binary64 with the canonical log/sin */
static double sm64_normal(uint64_t *x)
{
	double u1 = 1.0 - sm64_unit(x); /* (0, 1] */
	double u2 = sm64_unit(x);
	double r = __builtin_sqrt(-2.0 * sc_log_d(u1));
	return r * sc_sin_d(6.283185307179586 * u2);
}

extern "C" int sampo_gen_temperature(uint64_t seed, uint32_t days, float mean, float amp, float noise,
                                     float *hourly)
{
	if (!hourly) return fail(SAMPO_E_ARG, "null output%s%ld");
	uint64_t x = seed ^ 0x5445u;
	for (uint32_t d = 0; d < days; ++d) {
		double t = (double)mean + (double)amp * sc_sin_d(6.283185307179586 * (double)d / 365.0) +
		           (double)noise * sm64_normal(&x);
		float tf = (float)t;
		for (uint32_t h = 0; h < 24u; ++h) hourly[(size_t)d * 24u + h] = tf;
	}
	return 0;
}

/* Initial population of n0 agents: stage egg / larva / pupa / immature adult
 * with probability .3 / .3 / .1 / .3, sex 50/50, each in its entry state with
 * a random part of that state's delay already elapsed (integer hours). */
extern "C" int sampo_gen_population(uint64_t seed, uint32_t n0, float t0, sampo_layout *layout,
                                    sampo_agent *agents)
{
	if (!layout || (!agents && n0)) return fail(SAMPO_E_ARG, "null output%s%ld");
	uint64_t x = seed ^ 0x504f50u;
	std::vector<sampo_agent> by[4];
	const int stage_of[4] = {S_EGG, S_LARVA, S_PUPA, S_IMM};
	sampo_species sp; /* the stages' delays of include/properties.h */
	sampo_species_default(&sp);
	const float dev = sm_dev_rate(&sp, t0);
	for (uint32_t i = 0; i < n0; ++i) {
		double us = sm64_unit(&x);
		int st = us < 0.3 ? 0 : us < 0.6 ? 1 : us < 0.7 ? 2 : 3;
		sampo_agent a;
		memset(&a, 0, sizeof(a));
		a.isFemale = sm64_unit(&x) < 0.5 ? 1u : 0u;
		double u = sm64_unit(&x);
		switch (st) {
		case 0: {
			a.delay = sm_egg_delay(&sp, t0, (float)sm64_unit(&x));
			float h = __builtin_floorf((float)u * a.delay);
			a.ageInHours = a.hoursInState = h > 0.0f ? h : 0.0f;
			break;
		}
		case 1: {
			a.delay = (float)(1.0 + 0.1 * sm64_normal(&x));
			float h = __builtin_floorf((float)(u * 200.0));
			a.hoursInState = h;
			a.ageInHours = 24.0f + h;
			a.aux = sc_bits_f(h * dev);
			break;
		}
		case 2: {
			a.delay = sm_pupa_delay(&sp, t0);
			float h = __builtin_floorf((float)u * a.delay);
			a.hoursInState = h > 0.0f ? h : 0.0f;
			a.ageInHours = 264.0f + a.hoursInState;
			break;
		}
		default: {
			a.delay = sm_linear(sp.immature_delay, t0);
			float h = __builtin_floorf((float)u * a.delay);
			a.hoursInState = a.ageInHours = h > 0.0f ? h : 0.0f;
			break;
		}
		}
		by[st].push_back(a);
	}
	uint32_t end = 0;
	size_t k = 0;
	int q = 0;
	for (int s = 0; s < 8; ++s) {
		uint32_t start = s == 0 ? 0u : ((end + SAMPO_PAD) / SAMPO_PAD) * SAMPO_PAD;
		uint32_t n = 0;
		if (q < 4 && stage_of[q] == s) {
			n = (uint32_t)by[q].size();
			for (auto &a : by[q]) agents[k++] = a;
			++q;
		}
		layout->start[s] = start;
		layout->end[s] = start + n;
		end = start + n;
	}
	return 0;
}

/* ============================================================ measurement */

extern "C" int sampo_engine_capture(sampo_engine *e, uint32_t n)
{
	if (!e || n == 0) return fail(SAMPO_E_ARG, "bad arguments%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	if (e->cap) (void)hipGraphExecDestroy(e->cap);
	if (e->cap_graph) (void)hipGraphDestroy(e->cap_graph);
	for (auto &x : e->cap_ev) (void)hipEventDestroy(x);
	e->cap = nullptr;
	e->cap_graph = nullptr;
	e->cap_ev.assign(2 * (size_t)n, nullptr);
	for (auto &x : e->cap_ev) HIPCHK(hipEventCreate(&x));
	HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
	int rc = 0;
	for (uint32_t i = 0; i < n && !rc; ++i) rc = enqueue_one(e, e->cap_ev[2 * i], e->cap_ev[2 * i + 1], true);
	hipGraph_t g = nullptr;
	hipError_t ce = hipStreamEndCapture(e->stream, &g);
	if (rc) return rc;
	HIPCHK(ce);
	e->cap_graph = g;
	HIPCHK(hipGraphInstantiate(&e->cap, g, nullptr, nullptr, 0));
	return 0;
}

extern "C" int sampo_engine_replay(sampo_engine *e)
{
	if (!e || !e->cap) return fail(SAMPO_E_STATE, "nothing captured%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipGraphLaunch(e->cap, e->stream));
	return 0;
}

extern "C" int sampo_engine_replay_times(sampo_engine *e, float *wall_ms, float *pass_ms)
{
	if (!e || !wall_ms || !pass_ms) return fail(SAMPO_E_ARG, "null argument%s%ld");
	if (!e->cap) return fail(SAMPO_E_STATE, "nothing captured%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	const size_t n = e->cap_ev.size() / 2u;
	float sum = 0.0f;
	for (size_t i = 0; i < n; ++i) {
		float ms = 0.0f;
		HIPCHK(hipEventElapsedTime(&ms, e->cap_ev[2 * i], e->cap_ev[2 * i + 1]));
		sum += ms;
	}
	HIPCHK(hipEventElapsedTime(wall_ms, e->cap_ev[0], e->cap_ev[2 * n - 1]));
	*pass_ms = sum;
	return 0;
}

extern "C" int sampo_engine_time_days(sampo_engine *e, uint32_t days, float *wall_ms, float *launch_ms)
{
	if (!e || !wall_ms || !launch_ms || days == 0) return fail(SAMPO_E_ARG, "bad arguments%s%ld");
	HIPCHK(hipSetDevice(e->device));
	{ /* a day graph is captured before the first event */
		const int rc = sampo_engine_step_days(e, 0);
		if (rc) return rc;
	}
	std::vector<hipEvent_t> ev(2 * (size_t)days);
	for (auto &x : ev) HIPCHK(hipEventCreate(&x));
	for (uint32_t d = 0; d < days; ++d) {
		HIPCHK(hipEventRecord(ev[2 * d], e->stream));
		const int rc = sampo_engine_step_days(e, 1);
		if (rc) return rc;
		HIPCHK(hipEventRecord(ev[2 * d + 1], e->stream));
	}
	HIPCHK(hipEventSynchronize(ev[2 * (size_t)days - 1]));
	float sum = 0.0f;
	for (uint32_t d = 0; d < days; ++d) {
		float ms = 0.0f;
		HIPCHK(hipEventElapsedTime(&ms, ev[2 * d], ev[2 * d + 1]));
		sum += ms;
	}
	HIPCHK(hipEventElapsedTime(wall_ms, ev[0], ev[2 * (size_t)days - 1]));
	for (auto &x : ev) (void)hipEventDestroy(x);
	*launch_ms = sum;
	return 0;
}

extern "C" int sampo_engine_time_pass(sampo_engine *e, uint32_t n, float *wall_ms, float *pass_ms)
{
	if (!e || !wall_ms || !pass_ms || n == 0) return fail(SAMPO_E_ARG, "bad arguments%s%ld");
	HIPCHK(hipSetDevice(e->device));
	std::vector<hipEvent_t> ev(2 * (size_t)n);
	for (auto &x : ev) HIPCHK(hipEventCreate(&x));
	for (uint32_t i = 0; i < n; ++i) {
		int rc = enqueue_one(e, ev[2 * i], ev[2 * i + 1]);
		if (rc) return rc;
	}
	HIPCHK(hipGetLastError());
	HIPCHK(hipEventSynchronize(ev[2 * (size_t)n - 1]));
	float sum = 0.0f;
	for (uint32_t i = 0; i < n; ++i) {
		float ms = 0.0f;
		HIPCHK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
		sum += ms;
	}
	HIPCHK(hipEventElapsedTime(wall_ms, ev[0], ev[2 * (size_t)n - 1]));
	HIPCHK(hipStreamSynchronize(e->stream));
	for (auto &x : ev) (void)hipEventDestroy(x);
	*pass_ms = sum;
	return 0;
}

/* plain device-to-device copy, 16 bytes per lane per step: the achievable
 * HBM ceiling and the reference point for traffic counters */
__global__ void __launch_bounds__(256) k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		dst[i] = src[i];
}

extern "C" int sampo_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes)
{
	HIPCHK(hipSetDevice(device));
	size_t f = 0, t = 0;
	HIPCHK(hipMemGetInfo(&f, &t));
	if (free_bytes) *free_bytes = f;
	if (total_bytes) *total_bytes = t;
	return 0;
}

extern "C" int sampo_stream_copy(int device, uint64_t bytes, uint32_t iters, float *ms_per_copy)
{
	if (!ms_per_copy || bytes < 16 || iters == 0) return fail(SAMPO_E_ARG, "bad arguments%s%ld");
	HIPCHK(hipSetDevice(device));
	size_t n = (size_t)(bytes / 16);
	uint4 *a = nullptr, *b = nullptr;
	HIPCHK(hipMalloc(&a, n * 16));
	HIPCHK(hipMalloc(&b, n * 16));
	HIPCHK(hipMemset(a, 1, n * 16));
	/* one 16-byte element per lane, a work-group per 4 KiB: the fastest plain
	 * copy measured (6.2 TB/s for 1 GiB; persistent grid-stride grids ran at
	 * 4.5-5.4 TB/s) */
	const size_t groups = (n + 255u) / 256u;
	dim3 grid((unsigned)(groups < 0x7fffffffu ? groups : 0x7fffffffu));
	hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, 0, a, b, n);
	hipEvent_t x, y;
	HIPCHK(hipEventCreate(&x));
	HIPCHK(hipEventCreate(&y));
	HIPCHK(hipEventRecord(x, 0));
	for (uint32_t i = 0; i < iters; ++i) hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, 0, a, b, n);
	HIPCHK(hipEventRecord(y, 0));
	HIPCHK(hipEventSynchronize(y));
	float ms = 0.0f;
	HIPCHK(hipEventElapsedTime(&ms, x, y));
	*ms_per_copy = ms / (float)iters;
	(void)hipEventDestroy(x);
	(void)hipEventDestroy(y);
	(void)hipFree(a);
	(void)hipFree(b);
	return 0;
}
