/*
 * sampo_icl_kernels.hip -- the reference kernel contract on gfx950.
 *
 * One HIP kernel per reference OpenCL kernel, same name, same argument list,
 * same per-work-item semantics, on the reference's array-of-structs buffers
 * (sizes 24/16/8 bytes per agent as the driver allocates them).  These are
 * what icl_create_kernel() hands out; the fused engine (sampo_engine.hip) is
 * the fast path for whole runs.  Numerics follow sampo_canon.h.
 */
#include <hip/hip_runtime.h>

#include "sampo_canon.h"
#include "sampo_icl_types.h"
#include "sampo_model.h"

#define LOCAL 64u /* LOCAL_SIZE, include/agent.h:16 */

__device__ static inline uint32_t gid1() { return blockIdx.x * blockDim.x + threadIdx.x; }

__device__ static inline float taus(const IclSeeds *s, uint32_t g) { return sc_hybrid_taus(s->z1, s->z2, s->z3, s->z4, g); }

/* gap predicate of killAgents.cl:28-35 and its copies */
__device__ static inline bool in_range(const IclPopulation *p, uint32_t g)
{
	for (int s = 0; s < 7; ++s)
		if (g >= p->r[s].end && g < p->r[s + 1].start) return false;
	return g < p->r[7].end;
}

__device__ static inline float hmr(float dmr) { return 1.0f - sc_powf(1.0f - dmr, 1.0f / 24.0f); }

/* kernel/initAgents.cl:26-65; sp is the species table (the -DSPECIES header) */
__global__ void initAgents(IclAgent *agents, IclAgentAge *ages, IclAgentState *states, const IclEnb *enb,
                           float temperature, const IclSeeds *seeds, const sampo_species *sp)
{
	uint32_t id = gid1();
	if (id >= enb->newEggs) return;
	states[id].dead = 0u;
	states[id].state = ICL_EGG;
	IclAgentAge a;
	a.ageInHours = 0.0f;
	a.hoursInState = 0.0f;
	a.isFemale = (double)taus(&seeds[3], id) <= 0.5 ? 1u : 0u;
	a.csd = 0.0f;
	ages[id] = a;
	IclAgent ag;
	ag.humanBloodmealCount = 0u;
	ag.availableEggs = 0u;
	ag.cycleLength = 0u;
	ag.aux = 0u;
	ag.delay = sm_egg_delay(sp, temperature, taus(&seeds[2], id));
	ag.numEggBatches = 0u;
	agents[id] = ag;
}

/* kernel/killAgents.cl:24-148 */
__global__ void killAgents(IclAgentAge *agents, IclAgentState *states, const IclPopulation *pop, IclEnb *enb,
                           IclBnc *bnc, const IclSeeds *seeds, float carryingCapacity)
{
	uint32_t g = gid1();
	if (g == 0) {
		enb->newEggs = 0u;
		enb->totalBiomass = (int32_t)(pop->numLarvae1DayEquiv + pop->r[0].end + (pop->r[2].end - pop->r[2].start));
		bnc->numBites = 0u;
		bnc->numInfectBites = 0u;
		bnc->numCycles = 0u;
		bnc->sumCycles = 0u;
	}
	if (!in_range(pop, g)) return;
	uint32_t d = sc_f2u(sc_minf(agents[g].ageInHours / 24.0f, 99.0f));
	float u = taus(&seeds[0], g);
	float rate;
	if (g >= pop->r[3].start) {
		float a = 0.1f, B = 25.0f, s = 0.1f;
		float e = sc_expf((float)d / B);
		rate = hmr((a * e) / (1.0f + (a * B * s * (e - 1.0f))));
	} else if (g < pop->r[0].end) {
		rate = hmr(0.1f);
	} else if (g < pop->r[1].end) {
		float dmr = 0.1f * sc_expf((float)pop->numLarvae1DayEquiv / ((float)d * carryingCapacity * 1.0f));
		rate = hmr(sc_minf(dmr, 0.8f));
	} else {
		rate = hmr(0.1f);
	}
	if (u <= rate) states[g].dead = 1u;
}

/* Configuration-mode mortality: the range classes of killAgents.cl:37-147
 * (adults from immatures.start, eggs below eggs.end, larvae below
 * larvae.end, pupae otherwise) with one hourly rate per class in place of
 * the reference formulas.  An addition of this library, used by the
 * driver loop for scenarios with constant per-stage daily mortality. */
__global__ void killAgentsStaged(IclAgentAge *agents, IclAgentState *states, const IclPopulation *pop, IclEnb *enb,
                                 IclBnc *bnc, const IclSeeds *seeds, float hmrEgg, float hmrLarva, float hmrPupa,
                                 float hmrAdult)
{
	uint32_t g = gid1();
	if (g == 0) {
		enb->newEggs = 0u;
		enb->totalBiomass = (int32_t)(pop->numLarvae1DayEquiv + pop->r[0].end + (pop->r[2].end - pop->r[2].start));
		bnc->numBites = 0u;
		bnc->numInfectBites = 0u;
		bnc->numCycles = 0u;
		bnc->sumCycles = 0u;
	}
	if (!in_range(pop, g)) return;
	float u = taus(&seeds[0], g);
	float rate = g >= pop->r[3].start ? hmrAdult : g < pop->r[0].end ? hmrEgg : g < pop->r[1].end ? hmrLarva : hmrPupa;
	if (u <= rate) states[g].dead = 1u;
}

/* include/properties.h:257-266 */
__device__ __attribute__((noinline)) static uint32_t gen_eggs(float u0, float u1, uint32_t nb, const sampo_species *sp)
{
	return sm_egg_batch(u0, u1, sp->egg_batch[0], sp->egg_batch[1], sp->egg_batch[2], nb);
}

/* include/properties.h:87-91 */
__device__ __attribute__((noinline)) static float larva_delay(float u0, float u1, const sampo_species *sp)
{
	return sc_box_muller(u0, u1, sp->larva_delay[0], sp->larva_delay[1]);
}

/* kernel/updateAgents.cl:39-270 */
__global__ void updateAgents(IclAgent *agents, IclAgentAge *ages, IclAgentState *states, const IclPopulation *pop,
                             IclEnvironment env, float t, IclBnc *bnc, IclEnb *enb, const IclSeeds *seeds,
                             float dt, uint32_t h, const sampo_species *sp)
{
	uint32_t g = gid1();
	if (!in_range(pop, g)) return;
	IclAgentState st = states[g];
	if (st.dead) return;
	IclAgent a = agents[g];
	IclAgentAge ae = ages[g];
	ae.ageInHours += dt;
	ae.hoursInState += dt;
	{
		int cond = (t > 16.0f && a.humanBloodmealCount > 0u);
		float q = 111.0f / (t - 16.0f);
		double x = 1.0 / ((double)q * 14.0);
		ae.csd = (float)((double)ae.csd + (double)cond * x);
	}
	if (g < pop->r[0].end) {
		if (ae.hoursInState >= a.delay && sm_open(sp, 0, h)) {
			ae.ageInHours = 24.0f;
			ae.hoursInState = 0.0f;
			a.delay = larva_delay(taus(&seeds[1], g), taus(&seeds[2], g), sp);
			st.state = ICL_LARVA;
			agents[g] = a;
			states[g] = st;
		}
		ages[g] = ae;
		return;
	}
	if (g < pop->r[1].end) {
		if (env.larvacideValue > 0.0f && ae.hoursInState == dt && taus(&seeds[2], g) <= env.larvacideValue) {
			st.dead = 1u;
			states[g] = st;
			return;
		}
		float cl = sc_float(a.aux) + sm_dev_rate(sp, t);
		a.aux = sc_bits_f(cl);
		if (cl >= a.delay && sm_open(sp, 1, h)) {
			ae.hoursInState = 0.0f;
			a.delay = sm_pupa_delay(sp, t);
			st.state = ICL_PUPA;
			states[g] = st;
		}
		agents[g] = a;
		ages[g] = ae;
		return;
	}
	if (g < pop->r[2].end) {
		if (ae.hoursInState >= a.delay && sm_open(sp, 2, h)) {
			ae.ageInHours = 0.0f;
			ae.hoursInState = 0.0f;
			a.delay = sm_linear(sp->immature_delay, t);
			st.state = ICL_IMMATURE;
			agents[g] = a;
			states[g] = st;
		}
		ages[g] = ae;
		return;
	}
	if (g < pop->r[3].end) {
		if (ae.hoursInState >= a.delay && sm_open(sp, 3, h)) {
			a.delay = sp->mate_seeking_delay;
			ae.hoursInState = 0.0f;
			st.state = ICL_MATESEEKING;
			agents[g] = a;
			states[g] = st;
		}
		ages[g] = ae;
		return;
	}
	if (g < pop->r[4].end) {
		if (ae.hoursInState >= a.delay && ae.isFemale && sm_open(sp, 4, h)) {
			ae.hoursInState = 0.0f;
			a.delay = sp->bms_delay;
			st.state = ICL_BMS;
			states[g] = st;
			agents[g] = a;
		}
		ages[g] = ae;
		return;
	}
	if (g < pop->r[5].end) {
		++a.cycleLength;
		if (ae.hoursInState >= a.delay && sm_open(sp, 5, h) && taus(&seeds[1], g) <= env.bloodmealSuccess) {
			float itn = env.ITNValue * sp->bite_human * sp->bite_indoor;
			if (itn > 0.0f && taus(&seeds[2], g) <= itn) {
				st.dead = 1u;
				states[g] = st;
				return;
			}
			uint32_t human = sp->bite_human < 1.0f ? (taus(&seeds[3], g) <= sp->bite_human) : 1u;
			a.humanBloodmealCount += human;
			atomicAdd(&bnc->numBites, 1u);
			ae.hoursInState = 0.0f;
			a.delay = sm_linear(sp->bmd_delay, t);
			st.state = ICL_BMD;
			if (ae.csd >= 1.0f && human) atomicAdd(&bnc->numInfectBites, 1u);
			states[g] = st;
		}
		agents[g] = a;
		ages[g] = ae;
		return;
	}
	if (g < pop->r[6].end) {
		float irs = env.IRSValue * sp->rest_indoor;
		if (irs > 0.0f && taus(&seeds[2], g) <= irs) {
			st.dead = 1u;
			states[g] = st;
			return;
		}
		++a.cycleLength;
		if (ae.hoursInState >= a.delay && sm_open(sp, 6, h)) {
			if (a.availableEggs == 0u && ae.isFemale) {
				a.availableEggs = gen_eggs(taus(&seeds[1], g), taus(&seeds[2], g), a.numEggBatches, sp);
				a.numEggBatches += 1u;
			}
			ae.hoursInState = 0.0f;
			a.aux = 0u;
			st.state = ICL_GRAVID;
			states[g] = st;
		}
		agents[g] = a;
		ages[g] = ae;
		return;
	}
	++a.cycleLength;
	if (sm_open(sp, 7, h) && taus(&seeds[1], g) < sp->lay_probability) {
		if (env.oviTrapValue > 0.0f && taus(&seeds[2], g) <= env.oviTrapValue) {
			st.dead = 1u;
			a.availableEggs = 0u;
		}
		if (!st.dead) {
			uint32_t tb = (uint32_t)enb->totalBiomass;
			float nb = (float)tb / (((float)a.aux + 1.0f) * env.carryingCapacity);
			float pb = sc_maxf(0.0f, 1.0f - nb);
			uint32_t pe = sc_f2u(sc_ceilf((float)a.availableEggs * pb));
			atomicAdd(&enb->newEggs, pe);
			a.availableEggs -= pe;
		}
		a.aux += 1u;
		if (a.availableEggs == 0u) {
			atomicAdd(&bnc->numCycles, 1u);
			atomicAdd(&bnc->sumCycles, a.cycleLength);
			a.cycleLength = 0u;
			ae.hoursInState = 0.0f;
			a.delay = sp->bms_delay;
			st.state = ICL_BMS;
			states[g] = st;
		}
	}
	agents[g] = a;
	ages[g] = ae;
}

__device__ static inline uint32_t block_sum(uint32_t v, uint32_t *lds)
{
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	if ((threadIdx.x & 63u) == 0) lds[threadIdx.x >> 6] = v;
	__syncthreads();
	uint32_t r = 0;
	for (uint32_t w = 0; w < (blockDim.x + 63u) / 64u; ++w) r += lds[w];
	__syncthreads();
	return r;
}

/* kernel/calcL1de.cl:17-44: per-group sums of (uint)(ageInHours/24) over pop[1]'s larvae */
__global__ void __launch_bounds__(1024) calcL1dePerGroup(const IclAgentAge *agents, const IclPopulation *pop,
                                                         uint32_t *buff)
{
	__shared__ uint32_t lds[16];
	const IclPopulation *p1 = pop + 1;
	if (!(p1->r[1].end > p1->r[1].start)) return;
	uint32_t v = 0;
	for (uint32_t i = p1->r[1].start + gid1(); i < p1->r[1].end; i += gridDim.x * blockDim.x)
		v += sc_f2u(agents[i].ageInHours / 24.0f);
	v = block_sum(v, lds);
	if (threadIdx.x == 0) buff[blockIdx.x] = v;
}

/* kernel/calcL1de.cl:53-93: total, then pop[0] = pop[1] with the total */
__global__ void __launch_bounds__(1024) calcL1deTotal(IclPopulation *pop, const uint32_t *buff, uint32_t slots)
{
	__shared__ uint32_t lds[16];
	bool any = (pop[1].r[1].end - pop[1].r[1].start) != 0u;
	uint32_t v = 0;
	if (any && threadIdx.x < slots)
		for (uint32_t i = threadIdx.x; i < slots && i < blockDim.x; i += blockDim.x) v += buff[i];
	v = block_sum(v, lds);
	if (threadIdx.x != 0) return;
	IclPopulation p = pop[1];
	p.numLarvae1DayEquiv = any ? v : 0u;
	pop[0] = p;
}

/* kernel/calcGender.cl:17-58 */
__global__ void __launch_bounds__(1024) calcFemalesPerGroup(const IclAgentAge *agents, const IclPopulation *pop,
                                                            uint32_t *buff)
{
	__shared__ uint32_t lds[16];
	uint32_t f = 0, pi = 0;
	for (uint32_t i = pop->r[3].start + gid1(); i < pop->r[7].end; i += gridDim.x * blockDim.x) {
		bool gap = false;
		for (int s = 3; s < 7; ++s)
			if (i >= pop->r[s].end && i < pop->r[s + 1].start) gap = true;
		if (!gap && agents[i].isFemale) {
			++f;
			pi += agents[i].csd >= 1.0f ? 1u : 0u;
		}
	}
	f = block_sum(f, lds);
	pi = block_sum(pi, lds);
	if (threadIdx.x == 0) {
		buff[blockIdx.x] = f;
		buff[blockIdx.x + gridDim.x] = pi;
	}
}

/* kernel/calcGender.cl:66-89: slots = LOCAL_SIZE partial sums of each kind */
__global__ void __launch_bounds__(1024) calcFemalesTotal(IclPopulation *pop, const uint32_t *buff)
{
	__shared__ uint32_t lds[16];
	uint32_t f = 0, pi = 0;
	if (threadIdx.x < LOCAL) {
		f = buff[threadIdx.x];
		pi = buff[threadIdx.x + LOCAL];
	}
	f = block_sum(f, lds);
	pi = block_sum(pi, lds);
	if (threadIdx.x == 0) {
		pop->numFemale = f;
		pop->numPotentiallyInfective = pi;
	}
}

/* kernel/preScan.cl:21-64 */
__global__ void preScan(const IclAgentState *states, const IclPopulation *pop, int32_t *data, uint32_t *flags,
                        uint32_t start, uint32_t b1, uint32_t b2, uint32_t end, int32_t match)
{
	uint32_t g = gid1();
	++end;
	++b1;
	++b2;
	if (g > end || g < start) return;
	uint32_t idx = g - start;
	flags[idx] = (g >= b1 ? 1u : 0u) + (g >= b2 ? 1u : 0u) + 1u;
	if (!in_range(pop, g)) { data[idx] = 0; return; }
	IclAgentState s = states[g];
	data[idx] = (!s.dead && (s.state & (uint32_t)match) != 0u) ? 1 : 0;
}

__device__ static inline uint32_t next64(uint32_t e) { return ((e + LOCAL) / LOCAL) * LOCAL; }

/* kernel/oldToNewAgents.cl:113-292 */
__global__ void oldToNewAgents(const IclAgent *oa, const IclAgentAge *og, const IclAgentState *os, IclAgent *na,
                               IclAgentAge *ng, IclAgentState *ns, const int32_t *p1, const int32_t *p2,
                               const int32_t *p3, IclPopulation *population, const IclEnb *enb)
{
	uint32_t g = gid1();
	const IclPopulation pop = population[0];
	const uint32_t ls = pop.r[1].start;
	if (g == pop.r[7].end) {
		IclPopulation o = population[1];
		o.r[0].start = 0;
		o.r[0].end = (uint32_t)(p1[pop.r[0].end] + 1) + enb->newEggs;
		o.r[1].start = next64(o.r[0].end);
		o.r[1].end = o.r[1].start + (uint32_t)(p2[pop.r[1].end] + 1);
		o.r[2].start = next64(o.r[1].end);
		o.r[2].end = o.r[2].start + (uint32_t)(p3[pop.r[2].end - ls] + 1);
		o.r[3].start = next64(o.r[2].end);
		o.r[3].end = o.r[3].start + (uint32_t)(p1[pop.r[3].end] + 1);
		o.r[4].start = next64(o.r[3].end);
		o.r[4].end = o.r[4].start + (uint32_t)(p2[pop.r[4].end] + 1);
		o.r[5].start = next64(o.r[4].end);
		o.r[5].end = o.r[5].start + (uint32_t)(p3[pop.r[7].end - ls] + 1);
		o.r[6].start = next64(o.r[5].end);
		o.r[6].end = o.r[6].start + (uint32_t)(p1[pop.r[6].end] + 1);
		o.r[7].start = next64(o.r[6].end);
		o.r[7].end = o.r[7].start + (uint32_t)(p2[pop.r[7].end] + 1);
		population[1] = o;
	}
	if (!in_range(&pop, g)) return;
	IclAgentState st = os[g];
	if (st.dead) return;
	uint32_t e_end = enb->newEggs;
	uint32_t idx;
	if (st.state == ICL_EGG) {
		idx = e_end + (uint32_t)p1[g];
	} else {
		e_end += pop.r[0].end > 0u ? (uint32_t)(p1[pop.r[0].end] + 1) : 0u;
		uint32_t s1 = next64(e_end);
		if (st.state == ICL_LARVA) {
			idx = s1 + (uint32_t)p2[g];
		} else {
			uint32_t s2 = next64(s1 + (uint32_t)(p2[pop.r[1].end] + 1));
			if (st.state == ICL_PUPA) {
				idx = s2 + (uint32_t)p3[g - ls];
			} else {
				uint32_t s3 = next64(s2 + (uint32_t)(p3[pop.r[2].end - ls] + 1));
				if (st.state == ICL_IMMATURE) {
					idx = s3 + (uint32_t)p1[g];
				} else {
					uint32_t s4 = next64(s3 + (uint32_t)(p1[pop.r[3].end] + 1));
					if (st.state == ICL_MATESEEKING) {
						idx = s4 + (uint32_t)p2[g];
					} else {
						uint32_t s5 = next64(s4 + (uint32_t)(p2[pop.r[4].end] + 1));
						if (st.state == ICL_BMS) {
							idx = s5 + (uint32_t)p3[g - ls];
						} else {
							uint32_t s6 = next64(s5 + (uint32_t)(p3[pop.r[7].end - ls] + 1));
							if (st.state == ICL_BMD) {
								idx = s6 + (uint32_t)p1[g];
							} else if (st.state == ICL_GRAVID) {
								uint32_t s7 = next64(s6 + (uint32_t)(p1[pop.r[6].end] + 1));
								idx = s7 + (uint32_t)p2[g];
							} else {
								return;
							}
						}
					}
				}
			}
		}
	}
	na[idx] = oa[g];
	ng[idx] = og[g];
	ns[idx] = st;
}

/* kernel/ageHistogram.cl:13-35 over pop[1]: LDS-private bins, one global
 * atomic per non-empty bin per work-group */
__global__ void __launch_bounds__(1024) ageHistogram(const IclAgentAge *agents, const IclPopulation *pop,
                                                     uint32_t *hist)
{
	__shared__ uint32_t bins[100];
	for (uint32_t q = threadIdx.x; q < 100u; q += blockDim.x) bins[q] = 0u;
	__syncthreads();
	const IclPopulation *p = pop + 1;
	uint32_t g = gid1();
	if (g >= p->r[3].start && g < p->r[7].end) {
		bool gap = false;
		for (int s = 3; s < 7; ++s)
			if (g >= p->r[s].end && g < p->r[s + 1].start) gap = true;
		if (!gap) {
			uint32_t b = sc_f2u(agents[g].ageInHours / 24.0f);
			atomicAdd(&bins[b < 99u ? b : 99u], 1u);
		}
	}
	__syncthreads();
	for (uint32_t q = threadIdx.x; q < 100u; q += blockDim.x)
		if (bins[q]) atomicAdd(&hist[q], bins[q]);
}

/* kernel/serialScan.cl:19-42: one work-item */
__global__ void serialScan(const IclAgentState *states, int32_t *prefix, const IclPopulation *pop, uint32_t offset,
                           uint32_t start, uint32_t upper, uint32_t match)
{
	if (gid1() != 0) return;
	int32_t cur = -1;
	for (uint32_t i = start; i <= upper; ++i) {
		if (in_range(pop, i) && !states[i].dead && states[i].state == match) ++cur;
		prefix[i - offset] = cur;
	}
}

/* ------------------------------------------------------------------------
 * Segmented scan of the scan module (include/scan.h): in place,
 * data[i] = sum of data over the run of equal consecutive keys up to i, - 1.
 * Single pass: tiles in ticket order, segmented (head-flag, value) pairs
 * combined by decoupled look-back.  The operator
 *   (f1, v1) + (f2, v2) = (f1 | f2, f2 ? v2 : v1 + v2)
 * is associative; f marks a run start inside the span. */

#define SCAN_WG 256
#define SCAN_IPT 4
#define SCAN_TILE (SCAN_WG * SCAN_IPT)

__device__ static inline void sp_combine(uint32_t &f1, uint32_t &v1, uint32_t f2, uint32_t v2)
{
	v1 = f2 ? v2 : v1 + v2;
	f1 = f1 | f2;
}

__global__ void scan_reset(IclScanCtx *c)
{
	c->ticket = 0u;
	c->epoch = c->epoch + 1u;
}

__global__ void __launch_bounds__(SCAN_WG) scan_by_key(int32_t *data, const uint32_t *keys, uint32_t n,
                                                       IclScanCtx *c, uint64_t *status)
{
	__shared__ uint32_t s_tile;
	__shared__ uint32_t s_wf[SCAN_WG / 64], s_wv[SCAN_WG / 64];
	__shared__ uint32_t s_ef, s_ev;
	const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
	if (tid == 0) s_tile = atomicAdd(&c->ticket, 1u);
	__syncthreads();
	const uint32_t tile = s_tile;
	const uint32_t epoch = c->epoch & 0x1fffffffu;
	const uint32_t base = tile * SCAN_TILE + tid * SCAN_IPT;
	uint32_t f[SCAN_IPT], v[SCAN_IPT];
	uint32_t prev = base > 0u && base - 1u < n ? keys[base - 1u] : 0xffffffffu;
	for (int j = 0; j < SCAN_IPT; ++j) {
		uint32_t i = base + j;
		if (i < n) {
			uint32_t k = keys[i];
			f[j] = (i == 0u || k != prev) ? 1u : 0u;
			v[j] = (uint32_t)data[i];
			prev = k;
		} else {
			f[j] = 0u;
			v[j] = 0u;
		}
	}
	/* thread-local inclusive */
	uint32_t tf = f[0], tv = v[0];
	for (int j = 1; j < SCAN_IPT; ++j) sp_combine(tf, tv, f[j], v[j]);
	/* wave inclusive scan of thread aggregates */
	uint32_t xf = tf, xv = tv;
	for (int o = 1; o < 64; o <<= 1) {
		uint32_t pf = __shfl_up(xf, o, 64), pv = __shfl_up(xv, o, 64);
		if (lane >= (uint32_t)o) {
			uint32_t nf = pf, nv = pv;
			sp_combine(nf, nv, xf, xv);
			xf = nf;
			xv = nv;
		}
	}
	if (lane == 63u) { s_wf[w] = xf; s_wv[w] = xv; }
	__syncthreads();
	/* tile aggregate and this wave's prefix from earlier waves */
	uint32_t af = 0, av = 0, wf = 0, wv = 0;
	for (uint32_t q = 0; q < SCAN_WG / 64; ++q) {
		if (q == w) { wf = af; wv = av; }
		sp_combine(af, av, s_wf[q], s_wv[q]);
	}
	if (w == 0) {
		uint32_t ef = 0, ev = 0;
		uint64_t *mine = status + tile;
		if (tile == 0) {
			if (lane == 0)
				__hip_atomic_store(mine, (2ull << 62) | ((uint64_t)epoch << 33) | ((uint64_t)af << 32) | av,
				                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		} else {
			if (lane == 0)
				__hip_atomic_store(mine, (1ull << 62) | ((uint64_t)epoch << 33) | ((uint64_t)af << 32) | av,
				                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			/* lanes look back at tiles tile-1-lane, 64 at a time */
			int64_t look = (int64_t)tile - 1;
			uint32_t done = 0;
			while (!done) {
				int64_t t = look - (int64_t)lane;
				uint32_t kind = 2u, wfl = 1u, val = 0u;
				if (t >= 0) {
					uint64_t word;
					for (;;) {
						word = __hip_atomic_load(status + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
						if ((uint32_t)((word >> 33) & 0x1fffffffu) == epoch && (word >> 62) != 0ull) break;
						__builtin_amdgcn_s_sleep(1);
					}
					kind = (uint32_t)(word >> 62);
					wfl = (uint32_t)(word >> 32) & 1u;
					val = (uint32_t)word;
				}
				/* lanes are ordered nearest first: stop at the first inclusive
				 * prefix or the first aggregate containing a run start */
				uint64_t stop = __ballot(kind == 2u || wfl != 0u);
				uint32_t first = stop ? (uint32_t)__ffsll((unsigned long long)stop) - 1u : 64u;
				/* combine farthest to nearest over lanes [0, first] */
				uint32_t cf = 0, cv = 0;
				for (int q = (int)(first < 64u ? first : 63u); q >= 0; --q) {
					uint32_t qf = __shfl(wfl, q, 64), qv = __shfl(val, q, 64);
					sp_combine(cf, cv, qf, qv);
				}
				/* (cf, cv) covers the looked-at span; it precedes what was gathered so far */
				uint32_t nf = cf, nv = cv;
				sp_combine(nf, nv, ef, ev);
				ef = nf;
				ev = nv;
				if (first < 64u) done = 1u;
				look -= 64;
			}
			if (lane == 0) {
				uint32_t pf = ef, pv = ev;
				sp_combine(pf, pv, af, av);
				__hip_atomic_store(mine, (2ull << 62) | ((uint64_t)epoch << 33) | ((uint64_t)pf << 32) | pv,
				                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			}
		}
		if (lane == 0) { s_ef = ef; s_ev = ev; }
	}
	__syncthreads();
	/* exclusive prefix of this thread = tile prefix + wave prefix + lanes below */
	uint32_t pf = s_ef, pv = s_ev;
	sp_combine(pf, pv, wf, wv);
	uint32_t lf = __shfl_up(xf, 1, 64), lv = __shfl_up(xv, 1, 64);
	if (lane > 0) sp_combine(pf, pv, lf, lv);
	for (int j = 0; j < SCAN_IPT; ++j) {
		uint32_t i = base + j;
		sp_combine(pf, pv, f[j], v[j]);
		if (i < n) data[i] = (int32_t)pv - 1;
	}
}
