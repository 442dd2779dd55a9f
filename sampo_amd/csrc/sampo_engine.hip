/*
 * sampo_engine.hip -- fused, batched per-iteration agent step for gfx950.
 *
 * One iteration of the reference loop (src/abms.c:502-529) is two launches:
 *
 *   k_plan  (1 work-group)  lays every running instance's tiles end to end
 *                           and resets the tile ticket;
 *   k_pass  (persistent)    work-groups take tiles in ticket order; each tile
 *                           reads its agents once, applies killAgents and
 *                           updateAgents (kernel/killAgents.cl:24-148,
 *                           kernel/updateAgents.cl:39-270), ranks the
 *                           survivors by new state with 64-lane ballots and
 *                           an LDS scan, gets the counts of all earlier tiles
 *                           of its instance by decoupled look-back, and writes
 *                           each survivor once to its final slot -- the
 *                           compaction of preScan + scan + oldToNewAgents
 *                           (kernel/preScan.cl, kernel/boltScan.cl,
 *                           kernel/oldToNewAgents.cl) without materialising
 *                           any prefix array.
 *
 * The look-back also carries the sums that calcStats (kernel/calcL1de.cl,
 * kernel/calcGender.cl) and the bites/cycles/newEggs atomics need, so the
 * instance's last tile holds the totals after its own look-back and prepares
 * the next iteration in place (layout, printStats row, mortality table).
 * New eggs are not written (see sampo_model.h); the next pass regenerates
 * them from the seeds that laid them, which is what initAgents computes.
 */
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "sampo_model.h"

/* ------------------------------------------------------------ device state */

struct Inst {
	uint32_t *buf[2];      /* SoA: field f of buffer b at buf[b] + f * P */
	uint32_t P;
	uint32_t horizon;
	uint32_t mode;
	uint32_t env_count;
	uint64_t max_agents;
	const uint32_t *seeds; /* (horizon + 1) sets of 16 values */
	const float *temp;
	const sampo_env *env;
	sampo_row *rows;       /* horizon + 1 */
	float hmr_stage[4];
	float hmr_adult[100];
	uint32_t status;
	uint32_t step;
	uint32_t pending;      /* tot holds the totals of a finished iteration */
	uint32_t tot[16];      /* 8 new-state counts, newEggs, bites, infectious bites,
	                          cycles, cycle sum, L1DE, females, potentially infective */
	sampo_iter p[2];       /* iteration step uses p[step & 1] */
};

struct Ctx {
	uint32_t ticket;
	uint32_t total;
	uint32_t epoch;
	uint32_t pad;
	unsigned long long agent_steps;
};

#define LB_AGG 1ull
#define LB_INC 2ull
#define NV 16 /* look-back values per tile: 8 state counts + 8 sums */

__device__ static inline void lb_store(uint64_t *p, uint64_t kind, uint32_t epoch, uint32_t v)
{
	uint64_t w = (kind << 62) | ((uint64_t)(epoch & 0x3fffffffu) << 32) | v;
	__hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* ------------------------------------------------------------ the agent */

/* devRates, include/properties.h:94-98: hourly larval development, 12..43 degC */
__constant__ float c_dev_rates[32] = {
	0.000175044f, 0.000549094f, 0.00109031f, 0.001494158f, 0.001752487f, 0.001962866f,
	0.002171238f, 0.002392989f, 0.002633689f, 0.00289619f, 0.003182685f, 0.003495287f,
	0.003836204f, 0.004207789f, 0.004612566f, 0.005053243f, 0.005532729f, 0.006054144f,
	0.006620836f, 0.007236393f, 0.007904657f, 0.008629745f, 0.009416051f, 0.010267388f,
	0.011113704f, 0.011960019f, 0.012806335f, 0.010401163f, 0.007995991f, 0.00033845f,
	0.00000533086f, 0.0000000840986f};

struct Ag {
	float age, his, dly, csd;
	uint32_t aux, avail, cyc, misc;
};

struct Acc {
	uint32_t eggs, bites, ibites, ncyc, scyc;
};

/* per-tile copy of the instance's iteration parameters */
struct TileParams {
	sampo_iter ip;
	float hmr_adult[100];
	uint32_t seed[16];
	uint32_t nbseed[8];   /* seeds[2], seeds[3] of the set that laid the newborns */
	uint32_t *bin, *bout;
	uint32_t P;
	uint32_t mode;
};

__device__ static inline float rnd(const uint32_t *z, uint32_t g)
{
	return sc_hybrid_taus(z[0], z[1], z[2], z[3], g);
}

__device__ static inline int night(uint32_t h) { return h <= 6u || h >= 18u; }

/* The transcendental paths run only on state transitions; keeping them out
 * of line keeps their binary64 temporaries out of the pass's register budget. */
__device__ __attribute__((noinline)) float bm_delay(float u0, float u1, float mean, float sd)
{
	return sc_box_muller(u0, u1, mean, sd);
}

__device__ __attribute__((noinline)) uint32_t gen_eggs(float u0, float u1, uint32_t nb)
{
	uint32_t eggs = sc_f2u(sc_box_muller(u0, u1, 170.0f, 30.0f) + 0.5f);
	return sc_f2u((float)eggs * sc_powf(0.8f, (float)nb) + 0.5f);
}

/* killAgents + updateAgents for one agent in state s at virtual index g.
 * Returns the new state, or S_DEAD. */
__device__ static inline uint32_t agent_step(Ag &a, uint32_t s, uint32_t g, const TileParams &T, Acc &acc)
{
	const sampo_iter &ip = T.ip;
	/* killAgents.cl:55-147 */
	uint32_t d = sc_f2u(sc_minf(a.age / 24.0f, 99.0f));
	float hmr = s == S_EGG ? ip.hmr_egg : s == S_LARVA ? ip.hmr_larva[d] : s == S_PUPA ? ip.hmr_pupa
	                                                                                 : T.hmr_adult[d];
	if (rnd(&T.seed[0], g) <= hmr) return S_DEAD;

	/* updateAgents.cl:62-70 */
	const float t = ip.temp;
	a.age += 1.0f;
	a.his += 1.0f;
	uint32_t hbc = a.misc & 0xffffu;
	{
		int cond = (t > 16.0f && hbc > 0u);
		float q = 111.0f / (t - 16.0f);
		double x = 1.0 / ((double)q * 14.0);
		a.csd = (float)((double)a.csd + (double)cond * x);
	}
	const uint32_t h = ip.h;
	switch (s) {
	case S_EGG: /* :72-87 */
		if (a.his >= a.dly) {
			a.age = 24.0f;
			a.his = 0.0f;
			a.dly = bm_delay(rnd(&T.seed[4], g), rnd(&T.seed[8], g), 1.0f, 0.1f);
			return S_LARVA;
		}
		return S_EGG;
	case S_LARVA: { /* :89-113 */
		if (ip.env.larvacideValue > 0.0f && a.his == 1.0f && rnd(&T.seed[8], g) <= ip.env.larvacideValue)
			return S_DEAD;
		/* larvaDevelopment, properties.h:105-109; index clamped to the table */
		float tc = sc_minf(sc_maxf(t, 12.0f), 48.0f);
		int ix = (int)tc - 12;
		if (ix > 31) ix = 31;
		float c = sc_float(a.aux) + c_dev_rates[ix];
		a.aux = sc_bits_f(c);
		if (c >= a.dly && night(h)) {
			a.his = 0.0f;
			a.dly = (float)(-0.923 * (double)t + 60.923);
			return S_PUPA;
		}
		return S_LARVA;
	}
	case S_PUPA: /* :115-129 */
		if (a.his >= a.dly && night(h)) {
			a.age = 0.0f;
			a.his = 0.0f;
			a.dly = -2.667f * t + 120.0f;
			return S_IMM;
		}
		return S_PUPA;
	case S_IMM: /* :131-143 */
		if (a.his >= a.dly) {
			a.dly = 1.0f;
			a.his = 0.0f;
			return S_MS;
		}
		return S_IMM;
	case S_MS: /* :145-157 */
		if (a.his >= a.dly && (a.misc >> 31) && h == 18u) {
			a.his = 0.0f;
			a.dly = 0.0f;
			return S_BMS;
		}
		return S_MS;
	case S_BMS: /* :159-197 */
		a.cyc += 1u;
		if (a.his >= a.dly && night(h) && rnd(&T.seed[4], g) <= ip.env.bloodmealSuccess) {
			float itn = ip.env.ITNValue * 1.0f * 1.0f;
			if (itn > 0.0f && rnd(&T.seed[8], g) <= itn) return S_DEAD;
			a.misc = (a.misc & 0xffff0000u) | ((hbc + 1u) & 0xffffu);
			acc.bites += 1u;
			a.his = 0.0f;
			a.dly = -1.231f * t + 77.231f;
			if (a.csd >= 1.0f) acc.ibites += 1u;
			return S_BMD;
		}
		return S_BMS;
	case S_BMD: { /* :199-228 */
		float irs = ip.env.IRSValue * 1.0f;
		if (irs > 0.0f && rnd(&T.seed[8], g) <= irs) return S_DEAD;
		a.cyc += 1u;
		if (a.his >= a.dly) {
			if (a.avail == 0u && (a.misc >> 31)) { /* generateEggs, properties.h:257-266 */
				uint32_t nb = (a.misc >> 16) & 0x7fffu;
				a.avail = gen_eggs(rnd(&T.seed[4], g), rnd(&T.seed[8], g), nb);
				a.misc = (a.misc & 0x8000ffffu) | (((nb + 1u) & 0x7fffu) << 16);
			}
			a.his = 0.0f;
			a.aux = 0u;
			return S_GRAVID;
		}
		return S_BMD;
	}
	default: { /* gravid :230-269 */
		a.cyc += 1u;
		if (night(h) && rnd(&T.seed[4], g) < 0.25f) {
			int dead = 0;
			if (ip.env.oviTrapValue > 0.0f && rnd(&T.seed[8], g) <= ip.env.oviTrapValue) {
				dead = 1;
				a.avail = 0u;
			}
			if (!dead) { /* layEggs, properties.h:296-307 */
				float nbm = (float)ip.tb / (((float)a.aux + 1.0f) * ip.env.carryingCapacity);
				float pbm = sc_maxf(0.0f, 1.0f - nbm);
				uint32_t pe = sc_f2u(sc_ceilf((float)a.avail * pbm));
				acc.eggs += pe;
				a.avail -= pe;
			}
			a.aux += 1u;
			if (a.avail == 0u) {
				acc.ncyc += 1u;
				acc.scyc += a.cyc;
				a.cyc = 0u;
				a.his = 0.0f;
				a.dly = 0.0f;
				return dead ? S_DEAD : S_BMS;
			}
		}
		return S_GRAVID;
	}
	}
}

/* ------------------------------------------------------------ k_plan */

__global__ void __launch_bounds__(1024) k_plan(Inst *inst, Ctx *ctx, uint32_t *pl_base, uint8_t *pl_par,
                                               uint32_t ninst, uint32_t tile)
{
	__shared__ uint32_t s_scan[1024];
	__shared__ unsigned long long s_steps[1024];
	const uint32_t tid = threadIdx.x;
	uint32_t carry = 0;
	unsigned long long steps = 0;
	for (uint32_t b = 0; b < ninst; b += 1024u) {
		uint32_t i = b + tid;
		uint32_t nt = 0;
		if (i < ninst) {
			const Inst &I = inst[i];
			uint32_t par = I.step & 1u;
			const sampo_iter &ip = I.p[par];
			if (I.status == SAMPO_RUNNING && ip.active) {
				nt = (ip.n + tile - 1u) / tile;
				steps += ip.n;
			}
			pl_par[i] = (uint8_t)par;
		}
		s_scan[tid] = nt;
		__syncthreads();
		for (uint32_t off = 1; off < 1024u; off <<= 1) {
			uint32_t v = tid >= off ? s_scan[tid - off] : 0u;
			__syncthreads();
			s_scan[tid] += v;
			__syncthreads();
		}
		if (i < ninst) pl_base[i] = carry + s_scan[tid] - nt;
		carry += s_scan[1023];
		__syncthreads();
	}
	s_steps[tid] = steps;
	__syncthreads();
	for (uint32_t off = 512; off > 0; off >>= 1) {
		if (tid < off) s_steps[tid] += s_steps[tid + off];
		__syncthreads();
	}
	if (tid == 0) {
		pl_base[ninst] = carry;
		ctx->total = carry;
		ctx->ticket = 0u;
		ctx->epoch = ctx->epoch + 1u;
		ctx->agent_steps += s_steps[0];
	}
}

/* ------------------------------------------------------------ k_pass */

__device__ static inline uint32_t range_of(const uint32_t *kstart, uint32_t k)
{
	uint32_t s = 0;
#pragma unroll
	for (int r = 1; r < 8; ++r) s += (k >= kstart[r]) ? 1u : 0u;
	return s;
}

__device__ static inline void load_agent(Ag &a, const uint32_t *b, size_t P, uint32_t p)
{
	a.age = sc_float(b[p]);
	a.his = sc_float(b[P + p]);
	a.dly = sc_float(b[2 * P + p]);
	a.aux = b[3 * P + p];
	a.csd = sc_float(b[4 * P + p]);
	a.avail = b[5 * P + p];
	a.cyc = b[6 * P + p];
	a.misc = b[7 * P + p];
}

__device__ static inline void store_agent(const Ag &a, uint32_t *b, size_t P, uint32_t p)
{
	b[p] = sc_bits_f(a.age);
	b[P + p] = sc_bits_f(a.his);
	b[2 * P + p] = sc_bits_f(a.dly);
	b[3 * P + p] = a.aux;
	b[4 * P + p] = sc_bits_f(a.csd);
	b[5 * P + p] = a.avail;
	b[6 * P + p] = a.cyc;
	b[7 * P + p] = a.misc;
}

/* initAgents.cl:34-64 for newborn egg g, laid with seeds z (sets 2 and 3) at
 * temperature t */
__device__ __host__ static inline void newborn(Ag &a, const uint32_t *z, float t, uint32_t g)
{
	a.age = 0.0f;
	a.his = 0.0f;
	a.csd = 0.0f;
	a.aux = 0u;
	a.avail = 0u;
	a.cyc = 0u;
	float u3 = sc_hybrid_taus(z[4], z[5], z[6], z[7], g);
	float u2 = sc_hybrid_taus(z[0], z[1], z[2], z[3], g);
	a.misc = ((double)u3 <= 0.5) ? 0x80000000u : 0u;
	a.dly = sm_egg_delay(t, u2);
}

__device__ static inline uint32_t wave_sum(uint32_t v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

/* k_prep: one work-group per instance whose iteration just finished.  From
 * the totals the last tile left it builds the layout that oldToNewAgents.cl
 * :121-160 writes to popD[1] and runs calcStats (src/abms.c:230-310) of the
 * next iteration: printStats row, capacity and extinction checks, biomass,
 * and the larval mortality table of killAgents.cl:117-133. */
__global__ void __launch_bounds__(128) k_prep(Inst *insts)
{
	__shared__ uint32_t s_go, s_l1de;
	__shared__ float s_cc;
	Inst *I = insts + blockIdx.x;
	if (threadIdx.x == 0) {
		s_go = 0;
		if (I->status == SAMPO_RUNNING && I->pending) {
			const uint32_t par = I->step & 1u;
			const sampo_iter &ip = I->p[par];
			const uint32_t *tot = I->tot;
			sampo_prep_in in;
			in.step = ip.step + 1u;
			in.cur = ip.cur ^ 1u;
			uint32_t e = tot[8];
			for (int s = 0; s < 8; ++s) {
				in.cnt[s] = tot[s] + (s == 0 ? e : 0u);
				in.phys_in[s] = ip.pout[s];
			}
			uint32_t vend = in.cnt[0];
			in.vstart[0] = 0;
			for (int s = 1; s < 8; ++s) {
				in.vstart[s] = ((vend + SAMPO_PAD) / SAMPO_PAD) * SAMPO_PAD;
				vend = in.vstart[s] + in.cnt[s];
			}
			in.enew = e;
			in.bites = tot[9];
			in.ibites = tot[10];
			in.ncyc = tot[11];
			in.scyc = tot[12];
			in.l1de = tot[13];
			in.females = tot[14];
			in.potinf = tot[15];
			uint32_t tn = in.step < I->horizon ? in.step : I->horizon - 1u;
			in.temp_t = I->temp[tn];
			in.temp_prev = ip.temp;
			in.env_t = I->env[I->env_count == 1u ? 0u : tn];
			sampo_iter *o = &I->p[par ^ 1u];
			uint32_t st = sm_prep(I->mode, I->horizon, I->max_agents, I->P, I->hmr_stage, &in, o,
			                      &I->rows[in.step]);
			I->status = st;
			I->step = in.step;
			I->pending = 0u;
			s_go = st == SAMPO_RUNNING ? 1u + (I->mode == SAMPO_MODE_REFERENCE ? 1u : 0u) : 0u;
			s_l1de = in.l1de;
			s_cc = in.env_t.carryingCapacity;
		}
	}
	__syncthreads();
	if (s_go) {
		sampo_iter *o = &I->p[I->step & 1u];
		for (uint32_t d = threadIdx.x; d < 100u; d += blockDim.x)
			o->hmr_larva[d] = s_go == 2u ? sm_hmr_larva(d, s_l1de, s_cc) : I->hmr_stage[1];
	}
}

template <int WG, int APT>
__global__ void __launch_bounds__(WG) k_pass(Inst *insts, Ctx *ctx, const uint32_t *pl_base,
                                             const uint8_t *pl_par, uint32_t ninst, uint64_t *lb)
{
	constexpr int NW = WG / 64;
	constexpr uint32_t TILE = (uint32_t)WG * APT;
	__shared__ TileParams T;
	__shared__ uint32_t s_ticket, s_inst;
	__shared__ uint32_t s_wc[APT][NW][8];
	__shared__ uint32_t s_acc[NW][8];
	__shared__ uint32_t s_agg[NV];
	__shared__ uint32_t s_excl[NV];

	const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
	const uint32_t epoch = ctx->epoch & 0x3fffffffu;
	const uint32_t total = ctx->total;

	for (;;) {
		if (tid == 0) {
			uint32_t t = atomicAdd(&ctx->ticket, 1u);
			s_ticket = t;
			if (t < total) { /* last instance whose base <= t */
				uint32_t lo = 0, hi = ninst;
				while (hi - lo > 1u) {
					uint32_t mid = (lo + hi) >> 1;
					if (pl_base[mid] <= t) lo = mid; else hi = mid;
				}
				s_inst = lo;
			}
		}
		__syncthreads();
		const uint32_t ticket = s_ticket;
		if (ticket >= total) break;
		const uint32_t ii = s_inst;
		Inst *I = insts + ii;
		const uint32_t par = pl_par[ii];
		{ /* parameters into LDS */
			const uint32_t *src = (const uint32_t *)&I->p[par];
			uint32_t *dst = (uint32_t *)&T.ip;
			for (uint32_t q = tid; q < sizeof(sampo_iter) / 4u; q += WG) dst[q] = src[q];
			for (uint32_t q = tid; q < 100u; q += WG) T.hmr_adult[q] = I->hmr_adult[q];
			if (tid < 16u) T.seed[tid] = I->seeds[(size_t)(I->p[par].seed_set) * 16u + tid];
			if (tid >= 32u && tid < 40u)
				T.nbseed[tid - 32u] = I->seeds[(size_t)(I->p[par].nb_set) * 16u + 8u + (tid - 32u)];
			if (tid == 64u % WG) {
				uint32_t cur = I->p[par].cur;
				T.bin = I->buf[cur];
				T.bout = I->buf[cur ^ 1u];
				T.P = I->P;
				T.mode = I->mode;
			}
		}
		__syncthreads();
		const uint32_t first = pl_base[ii];
		const uint32_t ntile_i = pl_base[ii + 1] - first; /* pl_base[ninst] = total */
		const uint32_t tl = ticket - first;
		const uint32_t N = T.ip.n;
		const size_t P = T.P;

		Ag a[APT];
		uint32_t ns[APT], lr[APT];
		Acc acc = {0, 0, 0, 0, 0};
		uint32_t l1de = 0, fem = 0, pinf = 0;
#pragma unroll
		for (int j = 0; j < APT; ++j) {
			const uint32_t k = tl * TILE + (uint32_t)j * WG + tid;
			ns[j] = S_DEAD;
			if (k < N) {
				const uint32_t s = range_of(T.ip.kstart, k);
				const uint32_t g = k + T.ip.voff[s];
				if (s == S_EGG && k < T.ip.enew) newborn(a[j], T.nbseed, T.ip.nb_temp, g);
				else load_agent(a[j], T.bin, P, k + T.ip.off_in[s]);
				uint32_t n2 = agent_step(a[j], s, g, T, acc);
				ns[j] = n2;
				if (n2 == S_LARVA) l1de += sc_f2u(a[j].age / 24.0f);
				if (n2 >= S_IMM && n2 < S_DEAD && (a[j].misc >> 31)) {
					fem += 1u;
					pinf += a[j].csd >= 1.0f ? 1u : 0u;
				}
			}
		}
		/* rank survivors by new state: ballot per state, wave counts to LDS */
		const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
#pragma unroll
		for (int j = 0; j < APT; ++j) {
			lr[j] = 0;
#pragma unroll
			for (uint32_t s = 0; s < 8u; ++s) {
				uint64_t m = __ballot(ns[j] == s);
				if (ns[j] == s) lr[j] = (uint32_t)__popcll(m & lt);
				if (lane == 0) s_wc[j][w][s] = (uint32_t)__popcll(m);
			}
		}
		{
			uint32_t v[8] = {acc.eggs, acc.bites, acc.ibites, acc.ncyc, acc.scyc, l1de, fem, pinf};
#pragma unroll
			for (int q = 0; q < 8; ++q) {
				uint32_t r = wave_sum(v[q]);
				if (lane == 0) s_acc[w][q] = r;
			}
		}
		__syncthreads();
		if (tid < 8u) { /* exclusive scan over (j, wave) per state */
			uint32_t run = 0;
			for (int j = 0; j < APT; ++j)
				for (int q = 0; q < NW; ++q) {
					uint32_t c = s_wc[j][q][tid];
					s_wc[j][q][tid] = run;
					run += c;
				}
			s_agg[tid] = run;
		} else if (tid < 16u) {
			uint32_t r = 0;
			for (int q = 0; q < NW; ++q) r += s_acc[q][tid - 8u];
			s_agg[tid] = r;
		}
		__syncthreads();
		if (w == 0) { /* decoupled look-back over this instance's tiles */
			uint64_t *mine = lb + (size_t)ticket * NV;
			const uint32_t v = lane & 15u;
			uint32_t excl = 0;
			if (ticket == first) {
				if (lane < 16u) lb_store(mine + lane, LB_INC, epoch, s_agg[lane]);
			} else {
				if (lane < 16u) lb_store(mine + lane, LB_AGG, epoch, s_agg[lane]);
				uint32_t resolved = 0;
				int64_t look = (int64_t)ticket - 1;
				const uint32_t dpt = lane >> 4;
				for (;;) {
					uint32_t res_v = __shfl(resolved, (int)v, 64);
					int64_t tt = look - (int64_t)dpt;
					uint32_t kind = 2u, val = 0u;
					if (!res_v && tt >= (int64_t)first) {
						const uint64_t *pw = lb + (size_t)tt * NV + v;
						uint64_t word;
						for (;;) {
							word = __hip_atomic_load(pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
							if ((uint32_t)((word >> 32) & 0x3fffffffu) == epoch && (word >> 62) != 0ull) break;
							__builtin_amdgcn_s_sleep(1);
						}
						kind = (uint32_t)(word >> 62);
						val = (uint32_t)word;
					}
					uint32_t sum = 0, found = 0;
#pragma unroll
					for (int dd = 0; dd < 4; ++dd) {
						uint32_t kv = __shfl(kind, dd * 16 + (int)v, 64);
						uint32_t vv = __shfl(val, dd * 16 + (int)v, 64);
						if (!found) {
							sum += vv;
							if (kv == 2u) found = 1;
						}
					}
					if (!resolved) {
						excl += sum;
						resolved = found;
					}
					uint64_t open = __ballot(lane < 16u && !resolved);
					if (open == 0ull) break;
					look -= 4;
				}
				if (lane < 16u) lb_store(mine + lane, LB_INC, epoch, excl + s_agg[lane]);
			}
			if (lane < 16u) s_excl[lane] = excl;
		}
		__syncthreads();
		/* scatter survivors to their final slots */
#pragma unroll
		for (int j = 0; j < APT; ++j) {
			const uint32_t s = ns[j];
			if (s < S_DEAD) {
				uint32_t dst = T.ip.pout[s] + s_excl[s] + s_wc[j][w][s] + lr[j];
				store_agent(a[j], T.bout, P, dst);
			}
		}
		if (tl == ntile_i - 1u && tid < NV) { /* instance totals, for k_prep */
			I->tot[tid] = s_excl[tid] + s_agg[tid];
			if (tid == 0) I->pending = 1u;
		}
		__syncthreads();
	}
}

/* ------------------------------------------------------------ export, histogram */

__global__ void k_export(const uint32_t *buf, uint32_t P, sampo_iter ip, const uint32_t *nbz, sampo_agent *out)
{
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= ip.n) return;
	uint32_t s = range_of(ip.kstart, k);
	Ag a;
	if (s == S_EGG && k < ip.enew) newborn(a, nbz, ip.nb_temp, k + ip.voff[s]);
	else load_agent(a, buf, P, k + ip.off_in[s]);
	sampo_agent r;
	r.ageInHours = a.age;
	r.hoursInState = a.his;
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
__global__ void __launch_bounds__(256) k_histogram(const uint32_t *buf, uint32_t P, sampo_iter ip, uint32_t *hist)
{
	__shared__ uint32_t bins[100];
	for (uint32_t q = threadIdx.x; q < 100u; q += blockDim.x) bins[q] = 0u;
	__syncthreads();
	const uint32_t k0 = ip.kstart[S_IMM];
	for (uint32_t k = k0 + blockIdx.x * blockDim.x + threadIdx.x; k < ip.n; k += gridDim.x * blockDim.x) {
		uint32_t s = range_of(ip.kstart, k);
		float age = sc_float(buf[k + ip.off_in[s]]);
		uint32_t b = sc_f2u(age / 24.0f);
		atomicAdd(&bins[b < 99u ? b : 99u], 1u);
	}
	__syncthreads();
	for (uint32_t q = threadIdx.x; q < 100u; q += blockDim.x)
		if (bins[q]) atomicAdd(&hist[q], bins[q]);
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
	uint32_t *buf = nullptr;
	uint32_t *seeds = nullptr;
	float *temp = nullptr;
	sampo_env *env = nullptr;
	sampo_row *rows = nullptr;
	uint32_t horizon = 0;
	uint32_t P = 0;
};

typedef void (*pass_fn)(Inst *, Ctx *, const uint32_t *, const uint8_t *, uint32_t, uint64_t *);

template <int WG, int APT> pass_fn pass_ptr() { return k_pass<WG, APT>; }

pass_fn pick_pass(uint32_t wg, uint32_t apt)
{
#define SAMPO_TILE(W, A) if (wg == W && apt == A) return pass_ptr<W, A>();
	SAMPO_TILE(64, 4) SAMPO_TILE(64, 8)
	SAMPO_TILE(128, 1) SAMPO_TILE(128, 2) SAMPO_TILE(128, 4)
	SAMPO_TILE(256, 1) SAMPO_TILE(256, 2) SAMPO_TILE(256, 4)
	SAMPO_TILE(512, 1) SAMPO_TILE(512, 2)
#undef SAMPO_TILE
	return nullptr;
}

} // namespace

struct sampo_engine {
	int device = 0;
	hipStream_t stream = nullptr;
	uint32_t max_slots = 0;
	Inst *d_inst = nullptr;
	Ctx *d_ctx = nullptr;
	uint32_t *d_pl_base = nullptr;
	uint8_t *d_pl_par = nullptr;
	uint64_t *d_lb = nullptr;
	size_t lb_tiles = 0;
	std::vector<Slot> slots;
	uint32_t wg = 256, apt = 4;
	pass_fn pass = nullptr;
	int grid = 0;
	hipGraphExec_t day = nullptr;
	hipGraph_t day_graph = nullptr;
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
}

static int configure_pass(sampo_engine *e)
{
	int blocks = 0, cus = 0;
	HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void *)e->pass, (int)e->wg, 0));
	HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
	if (blocks < 1) blocks = 1;
	e->grid = blocks * cus;
	return 0;
}

/* look-back words for the largest tile count the admitted instances can reach */
static int ensure_lb(sampo_engine *e)
{
	size_t tiles = 1;
	uint32_t tile = e->wg * e->apt;
	for (auto &s : e->slots)
		if (s.used) tiles += (s.P + tile - 1u) / tile + 1u;
	if (tiles <= e->lb_tiles) return 0;
	HIPCHK(hipStreamSynchronize(e->stream));
	drop_graph(e);
	if (e->d_lb) HIPCHK(hipFree(e->d_lb));
	e->d_lb = nullptr;
	size_t words = tiles * NV;
	HIPCHK(hipMalloc(&e->d_lb, words * sizeof(uint64_t)));
	HIPCHK(hipMemset(e->d_lb, 0, words * sizeof(uint64_t)));
	e->lb_tiles = tiles;
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
	HIPCHK(hipMalloc(&e->d_pl_base, sizeof(uint32_t) * (max_slots + 1)));
	HIPCHK(hipMalloc(&e->d_pl_par, max_slots));
	HIPCHK(hipMemset(e->d_ctx, 0, sizeof(Ctx)));
	std::vector<Inst> blank(max_slots);
	memset(blank.data(), 0, sizeof(Inst) * max_slots);
	for (auto &b : blank) b.status = SAMPO_EMPTY;
	HIPCHK(hipMemcpy(e->d_inst, blank.data(), sizeof(Inst) * max_slots, hipMemcpyHostToDevice));
	e->pass = pick_pass(e->wg, e->apt);
	int rc = configure_pass(e);
	if (rc) { sampo_engine_destroy(e); return rc; }
	rc = ensure_lb(e);
	if (rc) { sampo_engine_destroy(e); return rc; }
	*out = e;
	return 0;
}

static void free_slot(Slot &s)
{
	if (s.buf) (void)hipFree(s.buf);
	if (s.seeds) (void)hipFree(s.seeds);
	if (s.temp) (void)hipFree(s.temp);
	if (s.env) (void)hipFree(s.env);
	if (s.rows) (void)hipFree(s.rows);
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
	(void)hipFree(e->d_lb);
	if (e->stream) (void)hipStreamDestroy(e->stream);
	delete e;
}

extern "C" int sampo_engine_set_tile(sampo_engine *e, uint32_t wg, uint32_t apt)
{
	if (!e) return fail(SAMPO_E_ARG, "null engine%s%ld");
	pass_fn f = pick_pass(wg, apt);
	if (!f) return fail(SAMPO_E_ARG, "tile shape not compiled in%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	drop_graph(e);
	e->wg = wg;
	e->apt = apt;
	e->pass = f;
	int rc = configure_pass(e);
	if (rc) return rc;
	return ensure_lb(e);
}

static uint32_t next_start(uint32_t end) { return ((end + SAMPO_PAD) / SAMPO_PAD) * SAMPO_PAD; }

extern "C" int sampo_engine_admit(sampo_engine *e, const sampo_scenario *sc, uint32_t *slot_out)
{
	if (!e || !sc || !slot_out) return fail(SAMPO_E_ARG, "null argument%s%ld");
	if (sc->horizon == 0 || !sc->temp || !sc->env) return fail(SAMPO_E_ARG, "horizon, temp and env are required%s%ld");
	if (sc->env_count != 1 && sc->env_count != sc->horizon)
		return fail(SAMPO_E_ARG, "env_count must be 1 or horizon%s%ld");
	if (sc->mode != SAMPO_MODE_REFERENCE && sc->mode != SAMPO_MODE_CONFIG)
		return fail(SAMPO_E_ARG, "unknown mode%s%ld");
	if (sc->max_agents == 0) return fail(SAMPO_E_ARG, "max_agents is required%s%ld");
	uint64_t P64 = sc->phys_slots ? sc->phys_slots : 3ull * sc->max_agents + 1024ull;
	P64 = ((P64 + 255ull) / 256ull) * 256ull;
	if (P64 > (1ull << 31)) return fail(SAMPO_E_ARG, "physical slots exceed 2^31%s%ld");
	const uint32_t P = (uint32_t)P64;
	uint32_t id = e->max_slots;
	for (uint32_t i = 0; i < e->max_slots; ++i)
		if (!e->slots[i].used) { id = i; break; }
	if (id == e->max_slots) return fail(SAMPO_E_FULL, "no free slot%s%ld");

	/* ---- initial layout (src/abms.c:185-228, or the given population) ---- */
	sampo_prep_in in;
	memset(&in, 0, sizeof(in));
	std::vector<uint32_t> soa;
	if (sc->mode == SAMPO_MODE_REFERENCE) {
		if ((uint64_t)sc->n_eggs > sc->max_agents) return fail(SAMPO_E_ARG, "n_eggs exceeds max_agents%s%ld");
		in.cnt[0] = sc->n_eggs;
		in.enew = sc->n_eggs;
		in.vstart[0] = 0;
		for (int s = 1; s < 8; ++s) in.vstart[s] = sc->n_eggs + 1u;
	} else {
		if (!sc->init_layout || (!sc->init_agents)) return fail(SAMPO_E_ARG, "config mode needs an initial population%s%ld");
		const sampo_layout *L = sc->init_layout;
		uint64_t pos = 0, k = 0;
		if (L->start[0] != 0) return fail(SAMPO_E_ARG, "egg range must start at 0%s%ld");
		for (int s = 0; s < 8; ++s) {
			if (L->end[s] < L->start[s]) return fail(SAMPO_E_ARG, "range %s%ld has end < start", "", s);
			if (s > 0 && L->start[s] <= L->end[s - 1]) return fail(SAMPO_E_ARG, "range %s%ld needs a gap before it", "", s);
			in.cnt[s] = L->end[s] - L->start[s];
			in.vstart[s] = L->start[s];
			in.phys_in[s] = (uint32_t)pos;
			pos += ((in.cnt[s] + SAMPO_PAD - 1u) / SAMPO_PAD) * SAMPO_PAD;
		}
		if (pos > P) return fail(SAMPO_E_ARG, "initial population exceeds physical slots%s%ld");
		soa.assign((size_t)NFIELD * pos, 0u);
		for (int s = 0; s < 8; ++s) {
			for (uint32_t r = 0; r < in.cnt[s]; ++r, ++k) {
				const sampo_agent &a = sc->init_agents[k];
				size_t p = in.phys_in[s] + r;
				soa[F_AGE * pos + p] = sc_bits_f(a.ageInHours);
				soa[F_HIS * pos + p] = sc_bits_f(a.hoursInState);
				soa[F_DLY * pos + p] = sc_bits_f(a.delay);
				soa[F_AUX * pos + p] = a.aux;
				soa[F_CSD * pos + p] = sc_bits_f(a.csd);
				soa[F_AVAIL * pos + p] = a.availableEggs;
				soa[F_CYC * pos + p] = a.cycleLength;
				soa[F_MISC * pos + p] = (a.humanBloodmealCount & 0xffffu) | ((a.numEggBatches & 0x7fffu) << 16) |
				                        (a.isFemale ? 0x80000000u : 0u);
				if (s == S_LARVA) in.l1de += sc_f2u(a.ageInHours / 24.0f);
				if (s >= S_IMM && a.isFemale) {
					in.females += 1u;
					in.potinf += a.csd >= 1.0f ? 1u : 0u;
				}
			}
		}
	}
	in.step = 0;
	in.cur = 0;
	in.temp_t = sc->temp[0];
	in.temp_prev = sc->temp[0];
	in.env_t = sc->env[0];

	Inst h;
	memset(&h, 0, sizeof(h));
	h.P = P;
	h.horizon = sc->horizon;
	h.mode = sc->mode;
	h.env_count = sc->env_count;
	h.max_agents = sc->max_agents;
	for (int c = 0; c < 4; ++c) h.hmr_stage[c] = sm_hmr(sc->dmr[c]);
	for (uint32_t d = 0; d < 100; ++d) h.hmr_adult[d] = sc->mode == SAMPO_MODE_CONFIG ? h.hmr_stage[3] : sm_hmr_adult(d);
	sampo_row row0;
	h.status = sm_prep(h.mode, h.horizon, h.max_agents, P, h.hmr_stage, &in, &h.p[0], &row0);
	for (uint32_t d = 0; d < 100; ++d)
		h.p[0].hmr_larva[d] = sc->mode == SAMPO_MODE_CONFIG ? h.hmr_stage[1]
		                                                    : sm_hmr_larva(d, in.l1de, in.env_t.carryingCapacity);
	h.step = 0;

	/* ---- device allocations and uploads ---- */
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	Slot S;
	S.horizon = sc->horizon;
	S.P = P;
	std::vector<uint32_t> seeds((size_t)(sc->horizon + 1u) * 16u);
	sampo_seed_stream(sc->seed, (uint32_t)seeds.size(), seeds.data());
	hipError_t err = hipMalloc(&S.buf, (size_t)2 * NFIELD * P * sizeof(uint32_t));
	if (err == hipSuccess) err = hipMalloc(&S.seeds, seeds.size() * sizeof(uint32_t));
	if (err == hipSuccess) err = hipMalloc(&S.temp, (size_t)sc->horizon * sizeof(float));
	if (err == hipSuccess) err = hipMalloc(&S.env, (size_t)sc->env_count * sizeof(sampo_env));
	if (err == hipSuccess) err = hipMalloc(&S.rows, (size_t)(sc->horizon + 1u) * sizeof(sampo_row));
	if (err != hipSuccess) {
		free_slot(S);
		return fail(SAMPO_E_NOMEM, "device allocation failed: %s%ld", hipGetErrorString(err));
	}
	HIPCHK(hipMemcpy(S.seeds, seeds.data(), seeds.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(S.temp, sc->temp, (size_t)sc->horizon * sizeof(float), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(S.env, sc->env, (size_t)sc->env_count * sizeof(sampo_env), hipMemcpyHostToDevice));
	HIPCHK(hipMemset(S.rows, 0, (size_t)(sc->horizon + 1u) * sizeof(sampo_row)));
	HIPCHK(hipMemcpy(S.rows, &row0, sizeof(row0), hipMemcpyHostToDevice));
	if (!soa.empty()) {
		size_t pos = soa.size() / NFIELD;
		for (int f = 0; f < NFIELD; ++f)
			HIPCHK(hipMemcpy(S.buf + (size_t)f * P, soa.data() + (size_t)f * pos, pos * sizeof(uint32_t),
			                 hipMemcpyHostToDevice));
	}
	h.buf[0] = S.buf;
	h.buf[1] = S.buf + (size_t)NFIELD * P;
	h.seeds = S.seeds;
	h.temp = S.temp;
	h.env = S.env;
	h.rows = S.rows;
	HIPCHK(hipMemcpy(e->d_inst + id, &h, sizeof(Inst), hipMemcpyHostToDevice));
	S.used = true;
	e->slots[id] = S;
	int rc = ensure_lb(e);
	if (rc) return rc;
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
	free_slot(e->slots[slot]);
	return 0;
}

static int enqueue(sampo_engine *e, uint32_t n)
{
	const uint32_t tile = e->wg * e->apt;
	for (uint32_t i = 0; i < n; ++i) {
		hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, e->stream, e->d_inst, e->d_ctx, e->d_pl_base,
		                   e->d_pl_par, e->max_slots, tile);
		void *args[] = {&e->d_inst, &e->d_ctx, &e->d_pl_base, &e->d_pl_par, &e->max_slots, &e->d_lb};
		HIPCHK(hipLaunchKernel((const void *)e->pass, dim3(e->grid), dim3(e->wg), args, 0, e->stream));
		hipLaunchKernelGGL(k_prep, dim3(e->max_slots), dim3(128), 0, e->stream, e->d_inst);
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
	if (n_agents) *n_agents = h.p[h.step & 1u].n;
	return 0;
}

extern "C" int sampo_engine_rows(sampo_engine *e, uint32_t slot, uint32_t first, uint32_t n, sampo_row *out)
{
	if (!e || slot >= e->max_slots || !e->slots[slot].used || !out) return fail(SAMPO_E_ARG, "bad slot%s%ld");
	const Slot &S = e->slots[slot];
	if ((uint64_t)first + n > (uint64_t)S.horizon + 1u) return fail(SAMPO_E_ARG, "rows out of range%s%ld");
	HIPCHK(hipSetDevice(e->device));
	HIPCHK(hipStreamSynchronize(e->stream));
	HIPCHK(hipMemcpy(out, S.rows + first, (size_t)n * sizeof(sampo_row), hipMemcpyDeviceToHost));
	return 0;
}

extern "C" int sampo_engine_export(sampo_engine *e, uint32_t slot, sampo_layout *layout, sampo_agent *agents,
                                   uint64_t max_agents)
{
	Inst h;
	int rc = read_inst(e, slot, &h);
	if (rc) return rc;
	const sampo_iter &ip = h.p[h.step & 1u];
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
	hipLaunchKernelGGL(k_export, dim3((ip.n + 255u) / 256u), dim3(256), 0, e->stream, h.buf[ip.cur], h.P, ip, nbz, d);
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
	const sampo_iter &ip = h.p[h.step & 1u];
	uint32_t *d = nullptr;
	HIPCHK(hipMalloc(&d, 100 * sizeof(uint32_t)));
	hipError_t err = hipMemsetAsync(d, 0, 100 * sizeof(uint32_t), e->stream);
	uint32_t adults = ip.n - ip.kstart[S_IMM];
	if (err == hipSuccess && adults > 0) {
		uint32_t blocks = (adults + 255u) / 256u;
		if (blocks > 2048u) blocks = 2048u;
		hipLaunchKernelGGL(k_histogram, dim3(blocks), dim3(256), 0, e->stream, h.buf[ip.cur], h.P, ip, d);
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
	float tc = sc_minf(sc_maxf(t0, 12.0f), 48.0f);
	int ix = (int)tc - 12;
	if (ix > 31) ix = 31;
	float dev;
	{ /* devRates entry for t0 (host copy of the device table) */
		static const float rates[32] = {
			0.000175044f, 0.000549094f, 0.00109031f, 0.001494158f, 0.001752487f, 0.001962866f,
			0.002171238f, 0.002392989f, 0.002633689f, 0.00289619f, 0.003182685f, 0.003495287f,
			0.003836204f, 0.004207789f, 0.004612566f, 0.005053243f, 0.005532729f, 0.006054144f,
			0.006620836f, 0.007236393f, 0.007904657f, 0.008629745f, 0.009416051f, 0.010267388f,
			0.011113704f, 0.011960019f, 0.012806335f, 0.010401163f, 0.007995991f, 0.00033845f,
			0.00000533086f, 0.0000000840986f};
		dev = rates[ix];
	}
	for (uint32_t i = 0; i < n0; ++i) {
		double us = sm64_unit(&x);
		int st = us < 0.3 ? 0 : us < 0.6 ? 1 : us < 0.7 ? 2 : 3;
		sampo_agent a;
		memset(&a, 0, sizeof(a));
		a.isFemale = sm64_unit(&x) < 0.5 ? 1u : 0u;
		double u = sm64_unit(&x);
		switch (st) {
		case 0: {
			a.delay = sm_egg_delay(t0, (float)sm64_unit(&x));
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
			a.delay = (float)(-0.923 * (double)t0 + 60.923);
			float h = __builtin_floorf((float)u * a.delay);
			a.hoursInState = h > 0.0f ? h : 0.0f;
			a.ageInHours = 264.0f + a.hoursInState;
			break;
		}
		default: {
			a.delay = -2.667f * t0 + 120.0f;
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

extern "C" int sampo_engine_time_pass(sampo_engine *e, uint32_t n, float *wall_ms, float *pass_ms)
{
	if (!e || !wall_ms || !pass_ms || n == 0) return fail(SAMPO_E_ARG, "bad arguments%s%ld");
	HIPCHK(hipSetDevice(e->device));
	const uint32_t tile = e->wg * e->apt;
	std::vector<hipEvent_t> ev(2 * (size_t)n);
	for (auto &x : ev) HIPCHK(hipEventCreate(&x));
	for (uint32_t i = 0; i < n; ++i) {
		hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, e->stream, e->d_inst, e->d_ctx, e->d_pl_base,
		                   e->d_pl_par, e->max_slots, tile);
		HIPCHK(hipEventRecord(ev[2 * i], e->stream));
		void *args[] = {&e->d_inst, &e->d_ctx, &e->d_pl_base, &e->d_pl_par, &e->max_slots, &e->d_lb};
		HIPCHK(hipLaunchKernel((const void *)e->pass, dim3(e->grid), dim3(e->wg), args, 0, e->stream));
		HIPCHK(hipEventRecord(ev[2 * i + 1], e->stream));
		hipLaunchKernelGGL(k_prep, dim3(e->max_slots), dim3(128), 0, e->stream, e->d_inst);
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

extern "C" int sampo_stream_copy(int device, uint64_t bytes, uint32_t iters, float *ms_per_copy)
{
	if (!ms_per_copy || bytes < 16 || iters == 0) return fail(SAMPO_E_ARG, "bad arguments%s%ld");
	HIPCHK(hipSetDevice(device));
	size_t n = (size_t)(bytes / 16);
	uint4 *a = nullptr, *b = nullptr;
	HIPCHK(hipMalloc(&a, n * 16));
	HIPCHK(hipMalloc(&b, n * 16));
	HIPCHK(hipMemset(a, 1, n * 16));
	int cus = 0;
	HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
	dim3 grid((unsigned)(cus * 16));
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
