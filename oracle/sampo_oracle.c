/*
 * sampo_oracle.c -- CPU restatement of SAMPO's per-iteration agent step.
 *
 * TEST INFRASTRUCTURE ONLY (see sampo_oracle.h).  Every function below cites
 * the reference file and lines it restates.  Work-item kernels are written as
 * loops over global ids; the per-agent kernels may run those loops with
 * OpenMP (integer atomics keep every sum order-independent).
 */
#include "sampo_oracle.h"
#include "../sampo_amd/csrc/sampo_canon.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LOCAL_SIZE 64u   /* include/agent.h:16 */
#define SCAN_WX 128u     /* src/abms.c:497 SCAN_SIZE = 2 * LOCAL_SIZE */
#define MARGIN 1024u     /* slack past the storage for the +1 scan slots */
/* Storage beyond maxAgents: the reference writes new eggs and survivors with
 * no bound check (src/abms.c:385-386) and only asserts capacity at the next
 * calcStats; the restatement keeps STORAGE x capacity slots so that it stops
 * at that same check, not at its own allocation. */
#define STORAGE 4u

enum { R_EGG, R_LARVA, R_PUPA, R_IMM, R_MS, R_BMS, R_BMD, R_GRAVID };

/* ------------------------------------------------------------ numerics */

float or_expf(float x) { return sc_expf(x); }
float or_logf(float x) { return sc_logf(x); }
float or_sinf(float x) { return sc_sinf(x); }
float or_sqrtf(float x) { return sc_sqrtf(x); }
float or_powf(float x, float y) { return sc_powf(x, y); }
uint32_t or_f2u(float x) { return sc_f2u(x); }

/* include/gpuRand.h:33-50 */
float or_hybrid_taus(const or_seeds *s, uint32_t gid)
{
	return sc_hybrid_taus(s->z1, s->z2, s->z3, s->z4, gid);
}

/* include/gpuRand.h:56-63 */
float or_box_muller(float u0, float u1, float mean, float sd)
{
	return sc_box_muller(u0, u1, mean, sd);
}

void or_seed_stream(uint32_t seed, uint32_t n, uint32_t *out)
{
	sc_glibc_rand g;
	sc_glibc_seed(&g, seed);
	for (uint32_t i = 0; i < n; ++i) out[i] = sc_glibc_next(&g);
}

/* ------------------------------------------------ species plug-in restated
 * include/properties.h, with each constant of the header as a table entry */

#define NIGHT_HOURS 0xfc007fu /* worldTime <= 6u || worldTime >= 18u */
#define ALL_HOURS 0xffffffu   /* return true; */

static const or_species k_properties_h = {
	{-0.923f, 60.923f},                                            /* properties.h:29-31 */
	{0.5f, 0.85f, 0.90f, 0.94f},                                   /* properties.h:39-60 */
	{48.0f, 68.5714f, 480.0f, 600.00f, 2400.0f},
	{0.0f, -10.2857f, -360.0f, -468.0f, -2160.0f},
	{1.0f, 0.1f},                                                  /* properties.h:87-91 */
	{12.0f, 48.0f},                                                /* properties.h:105-109 */
	{0.000175044f, 0.000549094f, 0.00109031f, 0.001494158f, 0.001752487f, 0.001962866f,
	 0.002171238f, 0.002392989f, 0.002633689f, 0.00289619f, 0.003182685f, 0.003495287f,
	 0.003836204f, 0.004207789f, 0.004612566f, 0.005053243f, 0.005532729f, 0.006054144f,
	 0.006620836f, 0.007236393f, 0.007904657f, 0.008629745f, 0.009416051f, 0.010267388f,
	 0.011113704f, 0.011960019f, 0.012806335f, 0.010401163f, 0.007995991f, 0.00033845f,
	 0.00000533086f, 0.0000000840986f},                             /* properties.h:94-98 */
	{-0.923, 60.923},                                              /* properties.h:133-135, unsuffixed */
	{-2.667f, 120.0f},                                             /* properties.h:159-161 */
	1.0f,                                                          /* properties.h:185-187 */
	0.0f,                                                          /* properties.h:212-214 */
	{-1.231f, 77.231f},                                            /* properties.h:239-241 */
	{170.0f, 30.0f, 0.8f},                                         /* properties.h:257-266 */
	0.25f,                                                         /* updateAgents.cl:233 */
	1.0f, 1.0f, 1.0f,                                              /* properties.h:12-14 */
	/* eggs :68-70, larvae :117-119, pupae :143-145, immatures :169-171,
	 * mate seeking :195-197 (== 18), BMS :222-224, BMD :249-251, lay :280-282 */
	{ALL_HOURS, NIGHT_HOURS, NIGHT_HOURS, ALL_HOURS, 1u << 18, NIGHT_HOURS, ALL_HOURS, NIGHT_HOURS},
	0u,
};

void or_species_default(or_species *out) { *out = k_properties_h; }

/* properties.h:29-31 */
static float eggs_incubation(const or_species *sp, float t) { return sp->egg_incubation[0] * t + sp->egg_incubation[1]; }

/* properties.h:39-60: piecewise-linear hatching time, rounded to an integer */
float or_sp_eggs_hatching(const or_species *sp, float p)
{
	float m, b;
	if (p < sp->hatch_threshold[0]) { m = sp->hatch_slope[0]; b = sp->hatch_offset[0]; }
	else if (p < sp->hatch_threshold[1]) { m = sp->hatch_slope[1]; b = sp->hatch_offset[1]; }
	else if (p < sp->hatch_threshold[2]) { m = sp->hatch_slope[2]; b = sp->hatch_offset[2]; }
	else if (p < sp->hatch_threshold[3]) { m = sp->hatch_slope[3]; b = sp->hatch_offset[3]; }
	else { m = sp->hatch_slope[4]; b = sp->hatch_offset[4]; }
	return (float)sc_f2u(m * p + b + 0.5f);
}

/* initAgents.cl:57 */
float or_sp_egg_delay(const or_species *sp, float t, float p)
{
	return eggs_incubation(sp, t) + or_sp_eggs_hatching(sp, p);
}

/* properties.h:105-109; the index is clamped to the table (the reference
 * reads past its end for T >= 44 degC). */
float or_sp_larva_development(const or_species *sp, float t)
{
	float c = sc_minf(sc_maxf(t, sp->dev_temp_range[0]), sp->dev_temp_range[1]);
	int i = (int)c - (int)sp->dev_temp_range[0];
	if (i > 31) i = 31;
	return sp->dev_rates[i];
}

/* properties.h:133-135: evaluated in binary64 (unsuffixed literals) */
static float pupa_delay(const or_species *sp, float t) { return (float)(sp->pupa_delay[0] * (double)t + sp->pupa_delay[1]); }
/* properties.h:159-161 */
static float immature_delay(const or_species *sp, float t) { return sp->immature_delay[0] * t + sp->immature_delay[1]; }
/* properties.h:239-241 */
static float bmd_delay(const or_species *sp, float t) { return sp->bmd_delay[0] * t + sp->bmd_delay[1]; }
/* the transition-time predicates, properties.h:68-70 ... 280-282 */
static int transition_time(const or_species *sp, int s, uint32_t h) { return (int)((sp->window[s] >> h) & 1u); }

float or_eggs_hatching(float p) { return or_sp_eggs_hatching(&k_properties_h, p); }
float or_larva_development(float t) { return or_sp_larva_development(&k_properties_h, t); }
float or_pupa_delay(float t) { return pupa_delay(&k_properties_h, t); }

/* ------------------------------------------------------------ mortality
 * kernel/killAgents.cl:84-147 */
static float hmr_from_dmr(float dmr)
{
	float dsr = 1.0f - dmr;
	float hsr = sc_powf(dsr, 1.0f / 24.0f);
	return 1.0f - hsr;
}

float or_hmr_reference(int cls, uint32_t d, uint32_t l1de, float cc)
{
	if (cls == 3) { /* killAgents.cl:85-101 */
		float a = 0.1f, B = 25.0f, s = 0.1f;
		float dmr = (a * sc_expf((float)d / B)) /
		            (1.0f + (a * B * s * (sc_expf((float)d / B) - 1.0f)));
		return hmr_from_dmr(dmr);
	}
	if (cls == 1) { /* killAgents.cl:117-133 */
		float rain = 1.0f, a = 0.1f;
		float dmr = a * sc_expf((float)l1de / ((float)d * cc * rain));
		dmr = sc_minf(dmr, 0.8f);
		return hmr_from_dmr(dmr);
	}
	return hmr_from_dmr(0.1f); /* eggs :103-115, pupae :135-147 */
}

/* ------------------------------------------------------------- layout */

static int is_valid(const or_pop *p, uint32_t g)
{
	/* the gap predicate shared by killAgents.cl:28-35, updateAgents.cl:45-52,
	 * preScan.cl:40-47, oldToNewAgents.cl:166-173, serialScan.cl:24-31 */
	for (int s = 0; s < 7; ++s)
		if (g >= p->r[s].end && g < p->r[s + 1].start) return 0;
	return g < p->r[7].end;
}

static uint32_t roundup(uint32_t v, uint32_t m) { return ((v + m - 1u) / m) * m; }

/* --------------------------------------------------------------- kernels */

/* kernel/initAgents.cl:26-65 */
static void k_init_agents(or_sim *s, int buf, uint32_t n_new, float t)
{
	const or_seeds *sd = s->seeds;
	or_agent *ag = s->agents[buf];
	or_age *ge = s->ages[buf];
	or_state *st = s->states[buf];
	const or_species *sp = &s->sp;
	float inc = eggs_incubation(sp, t);
	#pragma omp parallel for schedule(static) num_threads(s->cfg.threads > 0 ? s->cfg.threads : 1)
	for (int64_t id = 0; id < (int64_t)n_new; ++id) {
		uint32_t g = (uint32_t)id;
		st[g].dead = 0;
		st[g].state = OR_EGG;
		ge[g].ageInHours = 0.0f;
		ge[g].hoursInState = 0.0f;
		ge[g].isFemale = (double)or_hybrid_taus(&sd[3], g) <= 0.5;
		ge[g].csd = 0.0f;
		ag[g].humanBloodmealCount = 0u;
		ag[g].availableEggs = 0u;
		ag[g].cycleLength = 0u;
		ag[g].aux = sc_bits_f(0.0f);
		float prob = or_hybrid_taus(&sd[2], g);
		ag[g].delay = inc + or_sp_eggs_hatching(sp, prob);
		ag[g].numEggBatches = 0u;
	}
}

/* kernel/killAgents.cl:24-148 */
static void k_kill_agents(or_sim *s, uint32_t global, float cc)
{
	const or_pop *p = &s->pop[0];
	or_age *ge = s->ages[s->cur];
	or_state *st = s->states[s->cur];
	/* work-item 0 (:37-50) */
	s->enb.newEggs = 0;
	s->enb.totalBiomass = (int32_t)(p->numLarvae1DayEquiv + p->r[R_EGG].end +
	                                (p->r[R_PUPA].end - p->r[R_PUPA].start));
	memset(&s->bnc, 0, sizeof(s->bnc));
	/* mortality depends only on (range class, ageInDays): tabulate it */
	float hmr[4][100];
	for (uint32_t d = 0; d < 100; ++d)
		for (int c = 0; c < 4; ++c)
			hmr[c][d] = s->cfg.mode == 1 ? s->hmr_stage[c]
			                             : or_hmr_reference(c, d, p->numLarvae1DayEquiv, cc);
	#pragma omp parallel for schedule(static) num_threads(s->cfg.threads > 0 ? s->cfg.threads : 1)
	for (int64_t i = 0; i < (int64_t)global; ++i) {
		uint32_t g = (uint32_t)i;
		if (!is_valid(p, g)) continue;
		uint32_t d = sc_f2u(sc_minf(ge[g].ageInHours / 24.0f, 99.0f));
		float prob = or_hybrid_taus(&s->seeds[0], g);
		int c = g >= p->r[R_IMM].start ? 3 : g < p->r[R_EGG].end ? 0 : g < p->r[R_LARVA].end ? 1 : 2;
		if (prob <= hmr[c][d]) st[g].dead = 1;
	}
}

/* include/properties.h:257-266 */
static uint32_t generate_eggs(const or_species *sp, const or_seeds *sd, uint32_t g, uint32_t *neb)
{
	float u0 = or_hybrid_taus(&sd[0], g), u1 = or_hybrid_taus(&sd[1], g);
	uint32_t eggs = sc_f2u(sc_box_muller(u0, u1, sp->egg_batch[0], sp->egg_batch[1]) + 0.5f);
	eggs = sc_f2u((float)eggs * sc_powf(sp->egg_batch[2], (float)*neb) + 0.5f);
	++(*neb);
	return eggs;
}

/* kernel/updateAgents.cl:39-270 */
static void k_update_agents(or_sim *s, uint32_t global, const or_env *envp, float t, uint32_t h)
{
	const or_pop *p = &s->pop[0];
	const or_seeds *sd = s->seeds;
	or_agent *ag = s->agents[s->cur];
	or_age *ge = s->ages[s->cur];
	or_state *st = s->states[s->cur];
	const or_env env = *envp;
	const or_species *sp = &s->sp;
	const float dt = 1.0f; /* src/abms.c:15 */
	uint32_t tb = (uint32_t)s->enb.totalBiomass;
	uint32_t new_eggs = 0, bites = 0, ibites = 0, ncyc = 0, scyc = 0;
	#pragma omp parallel for schedule(static) num_threads(s->cfg.threads > 0 ? s->cfg.threads : 1) \
		reduction(+ : new_eggs, bites, ibites, ncyc, scyc)
	for (int64_t i = 0; i < (int64_t)global; ++i) {
		uint32_t g = (uint32_t)i;
		if (!is_valid(p, g)) continue;
		or_state as = st[g];
		if (as.dead) continue;
		or_agent a = ag[g];
		or_age ae = ge[g];
		ae.ageInHours += dt;
		ae.hoursInState += dt;
		{ /* :69-70, binary64 by C promotion */
			int cond = (t > 16.0f && a.humanBloodmealCount > 0u);
			float q = 111.0f / (t - 16.0f);
			double x = (double)1.0f / ((double)q * 14.0);
			ae.csd = (float)((double)ae.csd + (double)cond * x);
		}
		if (g < p->r[R_EGG].end) { /* :72-87 */
			if (ae.hoursInState >= a.delay && transition_time(sp, 0, h)) {
				ae.ageInHours = 24.0f;
				ae.hoursInState = 0.0f;
				a.delay = sc_box_muller(or_hybrid_taus(&sd[1], g), or_hybrid_taus(&sd[2], g), sp->larva_delay[0],
				                        sp->larva_delay[1]);
				as.state = OR_LARVA;
				ag[g] = a;
				st[g] = as;
			}
			ge[g] = ae;
			continue;
		}
		if (g < p->r[R_LARVA].end) { /* :89-113 */
			if (env.larvacideValue > 0.0f && ae.hoursInState == dt) {
				if (or_hybrid_taus(&sd[2], g) <= env.larvacideValue) {
					as.dead = 1;
					st[g] = as;
					continue;
				}
			}
			float cl = sc_float(a.aux) + or_sp_larva_development(sp, t);
			a.aux = sc_bits_f(cl);
			if (cl >= a.delay && transition_time(sp, 1, h)) {
				ae.hoursInState = 0.0f;
				a.delay = pupa_delay(sp, t);
				as.state = OR_PUPA;
				st[g] = as;
			}
			ag[g] = a;
			ge[g] = ae;
			continue;
		}
		if (g < p->r[R_PUPA].end) { /* :115-129 */
			if (ae.hoursInState >= a.delay && transition_time(sp, 2, h)) {
				ae.ageInHours = 0.0f;
				ae.hoursInState = 0.0f;
				a.delay = immature_delay(sp, t);
				as.state = OR_IMMATURE;
				ag[g] = a;
				st[g] = as;
			}
			ge[g] = ae;
			continue;
		}
		if (g < p->r[R_IMM].end) { /* :131-143 */
			if (ae.hoursInState >= a.delay && transition_time(sp, 3, h)) {
				a.delay = sp->mate_seeking_delay;
				ae.hoursInState = 0.0f;
				as.state = OR_MATESEEKING;
				ag[g] = a;
				st[g] = as;
			}
			ge[g] = ae;
			continue;
		}
		if (g < p->r[R_MS].end) { /* :145-157 */
			if (ae.hoursInState >= a.delay && ae.isFemale && transition_time(sp, 4, h)) {
				ae.hoursInState = 0.0f;
				a.delay = sp->bms_delay;
				as.state = OR_BMS;
				st[g] = as;
				ag[g] = a;
			}
			ge[g] = ae;
			continue;
		}
		if (g < p->r[R_BMS].end) { /* :159-197 */
			++a.cycleLength;
			if (ae.hoursInState >= a.delay && transition_time(sp, 5, h)) {
				if (or_hybrid_taus(&sd[1], g) <= env.bloodmealSuccess) {
					float itn = env.ITNValue * sp->bite_human * sp->bite_indoor;
					if (itn > 0.0f && or_hybrid_taus(&sd[2], g) <= itn) {
						as.dead = 1;
						st[g] = as;
						continue;
					}
					int human = sp->bite_human < 1.0f ? (or_hybrid_taus(&sd[3], g) <= sp->bite_human) : 1;
					a.humanBloodmealCount += 1u * (uint32_t)human;
					bites += 1u;
					ae.hoursInState = 0.0f;
					a.delay = bmd_delay(sp, t);
					as.state = OR_BMD;
					if (ae.csd >= 1.0f && human) ibites += 1u;
					st[g] = as;
				}
			}
			ag[g] = a;
			ge[g] = ae;
			continue;
		}
		if (g < p->r[R_BMD].end) { /* :199-228 */
			float irs = env.IRSValue * sp->rest_indoor;
			if (irs > 0.0f && or_hybrid_taus(&sd[2], g) <= irs) {
				as.dead = 1;
				st[g] = as;
				continue;
			}
			++a.cycleLength;
			if (ae.hoursInState >= a.delay && transition_time(sp, 6, h)) {
				if (a.availableEggs == 0u && ae.isFemale)
					a.availableEggs = generate_eggs(sp, sd + 1, g, &a.numEggBatches);
				ae.hoursInState = 0.0f;
				a.aux = 0u; /* ovipositionAttemps = 0 overwrites the union */
				as.state = OR_GRAVID;
				st[g] = as;
			}
			ag[g] = a;
			ge[g] = ae;
			continue;
		}
		/* gravid :230-269 */
		++a.cycleLength;
		if (transition_time(sp, 7, h)) {
			if (or_hybrid_taus(&sd[1], g) < sp->lay_probability) {
				if (env.oviTrapValue > 0.0f && or_hybrid_taus(&sd[2], g) <= env.oviTrapValue) {
					as.dead = 1;
					a.availableEggs = 0;
				}
				if (!as.dead) { /* layEggs, properties.h:296-307 */
					float nb = (float)tb / (((float)a.aux + 1.0f) * env.carryingCapacity);
					float pb = sc_maxf(0.0f, 1.0f - nb);
					uint32_t pe = sc_f2u(sc_ceilf((float)a.availableEggs * pb));
					new_eggs += pe;
					a.availableEggs = a.availableEggs - pe;
				}
				a.aux += 1u;
				if (a.availableEggs == 0u) {
					ncyc += 1u;
					scyc += a.cycleLength;
					a.cycleLength = 0;
					ae.hoursInState = 0.0f;
					a.delay = sp->bms_delay;
					as.state = OR_BMS;
					st[g] = as;
				}
			}
		}
		ag[g] = a;
		ge[g] = ae;
	}
	s->enb.newEggs += new_eggs;
	s->bnc.numBites += bites;
	s->bnc.numInfectBites += ibites;
	s->bnc.numCycles += ncyc;
	s->bnc.sumCycles += scyc;
}

/* kernel/preScan.cl:21-64 */
static void k_pre_scan(or_sim *s, int32_t *data, uint32_t *flags, uint32_t start, uint32_t b1,
                       uint32_t b2, uint32_t end, uint32_t match)
{
	const or_pop *p = &s->pop[0];
	const or_state *st = s->states[s->cur];
	uint32_t global = ((end + LOCAL_SIZE) / LOCAL_SIZE) * LOCAL_SIZE; /* abms.c:361 */
	++end; ++b1; ++b2;
	for (uint32_t g = start; g < global && g <= end; ++g) {
		uint32_t idx = g - start;
		flags[idx] = (g >= b1) + (g >= b2) + 1u;
		if (!is_valid(p, g)) { data[idx] = 0; continue; }
		data[idx] = (!st[g].dead && (st[g].state & match) != 0u) ? 1 : 0;
	}
}

/* kernel/boltScan.cl:22-220 restated work-group by work-group; host sizing
 * from src/boltScan.c:36-126.  Lanes past the end of the input hold key 0,
 * value 0 (the reference leaves them unset); they feed only outputs that no
 * caller reads. */
void or_bolt_scan(const uint32_t *keys, int32_t *vals, uint32_t n, uint32_t wx)
{
	if (n == 0) return;
	uint32_t nwg = (n + wx - 1u) / wx;
	uint32_t sbuf = ((nwg + wx - 1u) / wx) * wx;
	uint32_t *key_sum = (uint32_t *)calloc(sbuf, sizeof(uint32_t));
	int32_t *pre_sum = (int32_t *)calloc(sbuf, sizeof(int32_t));
	int32_t *post_sum = (int32_t *)calloc(sbuf, sizeof(int32_t));
	uint32_t *lk = (uint32_t *)malloc(wx * sizeof(uint32_t));
	int32_t *lv = (int32_t *)malloc(wx * sizeof(int32_t));
	int32_t *sum = (int32_t *)malloc(wx * sizeof(int32_t));
	uint32_t *key = (uint32_t *)malloc(wx * sizeof(uint32_t));
	const int32_t init = -1;

	/* kernel 0: perBlockScanByKey */
	for (uint32_t grp = 0; grp < nwg; ++grp) {
		for (uint32_t l = 0; l < wx; ++l) {
			uint32_t gi = grp * wx + l;
			key[l] = gi < n ? keys[gi] : 0u;
			sum[l] = gi < n ? vals[gi] : 0;
			lk[l] = key[l];
			lv[l] = sum[l];
		}
		for (uint32_t off = 1; off < wx; off *= 2) {
			for (uint32_t l = wx; l-- > 0;) /* reads lower lanes' previous values */
				if (l >= off && key[l] == lk[l - off]) sum[l] += lv[l - off];
			for (uint32_t l = 0; l < wx; ++l) lv[l] = sum[l];
		}
		for (uint32_t l = 0; l < wx; ++l) {
			uint32_t gi = grp * wx + l;
			if (gi < n) vals[gi] = sum[l] + init;
		}
		key_sum[grp] = lk[wx - 1];
		pre_sum[grp] = lv[wx - 1];
	}

	/* kernel 1: intraBlockInclusiveScanByKey, one work-group of wx */
	uint32_t wpt = sbuf / wx;
	for (uint32_t l = 0; l < wx; ++l) {
		uint32_t map = l * wpt;
		uint32_t k = 0;
		int32_t ws = 0;
		if (map < nwg) {
			k = key_sum[map];
			ws = pre_sum[map];
			post_sum[map] = ws;
			for (uint32_t o = 1; o < wpt; ++o) {
				uint32_t pk = k;
				k = key_sum[map + o];
				if (map + o < nwg) {
					int32_t y = pre_sum[map + o];
					ws = (k == pk) ? ws + y : y;
					post_sum[map + o] = ws;
				}
			}
		}
		lk[l] = k;
		lv[l] = ws;
		sum[l] = ws;
	}
	for (uint32_t off = 1; off < wx; off *= 2) {
		for (uint32_t l = 0; l < wx; ++l) {
			if (l * wpt < nwg && l >= off) {
				if (lk[l] == lk[l - off]) sum[l] = sum[l] + lv[l - off];
				else sum[l] = lv[l];
			}
		}
		for (uint32_t l = 0; l < wx; ++l) lv[l] = sum[l];
	}
	for (uint32_t l = 1; l < wx; ++l) {
		uint32_t map = l * wpt;
		if (map >= nwg) continue;
		for (uint32_t o = 0; o < wpt; ++o) {
			int32_t y = post_sum[map + o];
			if (key_sum[map + o] == lk[l - 1]) y += lv[l - 1];
			post_sum[map + o] = y;
		}
	}

	/* kernel 2: perBlockAdditionByKey */
	for (uint32_t gi = 0; gi < n; ++gi) {
		uint32_t grp = gi / wx;
		if (grp > 0 && key_sum[grp - 1] == keys[gi]) vals[gi] += post_sum[grp - 1];
	}
	free(key_sum); free(pre_sum); free(post_sum);
	free(lk); free(lv); free(sum); free(key);
}

/* kernel/serialScan.cl:19-42 */
static void k_serial_scan(or_sim *s, int32_t *prefix, uint32_t offset, uint32_t start,
                          uint32_t upper, uint32_t match)
{
	const or_pop *p = &s->pop[0];
	const or_state *st = s->states[s->cur];
	int32_t curr = -1;
	for (uint32_t idx = start; idx <= upper; ++idx) {
		if (is_valid(p, idx) && !st[idx].dead && st[idx].state == match) ++curr;
		prefix[idx - offset] = curr;
	}
}

static uint32_t next_start(uint32_t end) { return ((end + LOCAL_SIZE) / LOCAL_SIZE) * LOCAL_SIZE; }

/* kernel/oldToNewAgents.cl:113-292.  Returns 0, or 1 when the new layout
 * would not fit the buffers. */
static int k_old_to_new(or_sim *s)
{
	const or_pop pop = s->pop[0];
	const int32_t *p1 = s->prefix[0], *p2 = s->prefix[1], *p3 = s->prefix[2];
	uint32_t ls = pop.r[R_LARVA].start;
	/* work-item g == gravids.end (:121-164) */
	or_pop *o = &s->pop[1];
	o->r[R_EGG].start = 0;
	o->r[R_EGG].end = (uint32_t)(p1[pop.r[R_EGG].end] + 1) + s->enb.newEggs;
	o->r[R_LARVA].start = next_start(o->r[R_EGG].end);
	o->r[R_LARVA].end = o->r[R_LARVA].start + (uint32_t)(p2[pop.r[R_LARVA].end] + 1);
	o->r[R_PUPA].start = next_start(o->r[R_LARVA].end);
	o->r[R_PUPA].end = o->r[R_PUPA].start + (uint32_t)(p3[pop.r[R_PUPA].end - ls] + 1);
	o->r[R_IMM].start = next_start(o->r[R_PUPA].end);
	o->r[R_IMM].end = o->r[R_IMM].start + (uint32_t)(p1[pop.r[R_IMM].end] + 1);
	o->r[R_MS].start = next_start(o->r[R_IMM].end);
	o->r[R_MS].end = o->r[R_MS].start + (uint32_t)(p2[pop.r[R_MS].end] + 1);
	o->r[R_BMS].start = next_start(o->r[R_MS].end);
	o->r[R_BMS].end = o->r[R_BMS].start + (uint32_t)(p3[pop.r[R_GRAVID].end - ls] + 1);
	o->r[R_BMD].start = next_start(o->r[R_BMS].end);
	o->r[R_BMD].end = o->r[R_BMD].start + (uint32_t)(p1[pop.r[R_BMD].end] + 1);
	o->r[R_GRAVID].start = next_start(o->r[R_BMD].end);
	o->r[R_GRAVID].end = o->r[R_GRAVID].start + (uint32_t)(p2[pop.r[R_GRAVID].end] + 1);
	if ((uint64_t)o->r[R_GRAVID].end + 2u > (uint64_t)STORAGE * s->cfg.capacity + MARGIN) return 1;

	const or_agent *oa = s->agents[s->cur];
	const or_age *og = s->ages[s->cur];
	const or_state *os = s->states[s->cur];
	int nb = s->cur ^ 1;
	or_agent *na = s->agents[nb];
	or_age *ng = s->ages[nb];
	or_state *ns = s->states[nb];
	uint32_t global = ((pop.r[R_GRAVID].end + LOCAL_SIZE) / LOCAL_SIZE) * LOCAL_SIZE;
	/* per-agent path (:166-290) */
	uint32_t e_end = s->enb.newEggs;
	uint32_t e_all = e_end + ((pop.r[R_EGG].end > 0) ? (uint32_t)(p1[pop.r[R_EGG].end] + 1) : 0u);
	uint32_t ls_n = next_start(e_all);
	uint32_t le_n = ls_n + (uint32_t)(p2[pop.r[R_LARVA].end] + 1);
	uint32_t ps_n = next_start(le_n);
	uint32_t pe_n = ps_n + (uint32_t)(p3[pop.r[R_PUPA].end - ls] + 1);
	uint32_t is_n = next_start(pe_n);
	uint32_t ie_n = is_n + (uint32_t)(p1[pop.r[R_IMM].end] + 1);
	uint32_t ms_n = next_start(ie_n);
	uint32_t me_n = ms_n + (uint32_t)(p2[pop.r[R_MS].end] + 1);
	uint32_t bs_n = next_start(me_n);
	uint32_t be_n = bs_n + (uint32_t)(p3[pop.r[R_GRAVID].end - ls] + 1);
	uint32_t ds_n = next_start(be_n);
	uint32_t de_n = ds_n + (uint32_t)(p1[pop.r[R_BMD].end] + 1);
	uint32_t gs_n = next_start(de_n);
	#pragma omp parallel for schedule(static) num_threads(s->cfg.threads > 0 ? s->cfg.threads : 1)
	for (int64_t i = 0; i < (int64_t)global; ++i) {
		uint32_t g = (uint32_t)i;
		if (!is_valid(&pop, g)) continue;
		or_state as = os[g];
		if (as.dead) continue;
		uint32_t idx;
		switch (as.state) {
		case OR_EGG: idx = e_end + (uint32_t)p1[g]; break;
		case OR_LARVA: idx = ls_n + (uint32_t)p2[g]; break;
		case OR_PUPA: idx = ps_n + (uint32_t)p3[g - ls]; break;
		case OR_IMMATURE: idx = is_n + (uint32_t)p1[g]; break;
		case OR_MATESEEKING: idx = ms_n + (uint32_t)p2[g]; break;
		case OR_BMS: idx = bs_n + (uint32_t)p3[g - ls]; break;
		case OR_BMD: idx = ds_n + (uint32_t)p1[g]; break;
		case OR_GRAVID: idx = gs_n + (uint32_t)p2[g]; break;
		default: continue;
		}
		na[idx] = oa[g];
		ng[idx] = og[g];
		ns[idx] = as;
	}
	(void)de_n;
	return 0;
}

/* -------------------------------------------------------- driver (abms.c) */

static int alloc_sim(or_sim *s, const or_config *cfg)
{
	memset(s, 0, sizeof(*s));
	s->cfg = *cfg;
	s->sp = cfg->species ? *cfg->species : k_properties_h;
	s->cfg.species = &s->sp;
	size_t n = (size_t)STORAGE * cfg->capacity + MARGIN;
	for (int b = 0; b < 2; ++b) {
		s->agents[b] = (or_agent *)calloc(n, sizeof(or_agent));
		s->ages[b] = (or_age *)calloc(n, sizeof(or_age));
		s->states[b] = (or_state *)calloc(n, sizeof(or_state));
		if (!s->agents[b] || !s->ages[b] || !s->states[b]) return -1;
	}
	for (int k = 0; k < 3; ++k) {
		s->prefix[k] = (int32_t *)calloc(n, sizeof(int32_t));
		if (!s->prefix[k]) return -1;
	}
	s->buff_len = (uint32_t)n;
	s->buff = (uint32_t *)calloc(n, sizeof(uint32_t));
	s->rng = malloc(sizeof(sc_glibc_rand));
	if (!s->buff || !s->rng) return -1;
	for (int c = 0; c < 4; ++c) s->hmr_stage[c] = hmr_from_dmr(cfg->dmr[c]);
	return 0;
}

/* src/abms.c:140-158 */
static void generate_seeds(or_sim *s)
{
	sc_glibc_rand *g = (sc_glibc_rand *)s->rng;
	for (int i = 0; i < 4; ++i) {
		s->seeds[i].z1 = sc_glibc_next(g);
		s->seeds[i].z2 = sc_glibc_next(g);
		s->seeds[i].z3 = sc_glibc_next(g);
		s->seeds[i].z4 = sc_glibc_next(g);
	}
}

/* src/abms.c:185-228 and main :606-672 */
int or_sim_init_reference(or_sim *s, const or_config *cfg, uint32_t seed, uint32_t n_eggs)
{
	if (alloc_sim(s, cfg)) return -1;
	if (n_eggs > cfg->capacity) return -1;
	sc_glibc_seed((sc_glibc_rand *)s->rng, seed);
	s->enb.newEggs = n_eggs;
	s->enb.totalBiomass = (int32_t)n_eggs;
	generate_seeds(s);
	k_init_agents(s, 0, n_eggs, cfg->temp[0]);
	s->cur = 0;
	or_pop *p = &s->pop[1];
	p->r[R_EGG].start = 0;
	p->r[R_EGG].end = n_eggs;
	for (int r = 1; r < 8; ++r) p->r[r].start = p->r[r].end = n_eggs + 1;
	return 0;
}

int or_sim_init_population(or_sim *s, const or_config *cfg, uint32_t seed, const or_pop *pop,
                           const or_agent *agents, const or_age *ages, const or_state *states)
{
	if (alloc_sim(s, cfg)) return -1;
	uint32_t span = pop->r[R_GRAVID].end;
	if (span > cfg->capacity) return -1;
	sc_glibc_seed((sc_glibc_rand *)s->rng, seed);
	generate_seeds(s); /* the set the reference spends on its initial eggs */
	memcpy(s->agents[0], agents, (size_t)span * sizeof(or_agent));
	memcpy(s->ages[0], ages, (size_t)span * sizeof(or_age));
	memcpy(s->states[0], states, (size_t)span * sizeof(or_state));
	s->pop[1] = *pop;
	s->cur = 0;
	return 0;
}

void or_sim_free(or_sim *s)
{
	for (int b = 0; b < 2; ++b) { free(s->agents[b]); free(s->ages[b]); free(s->states[b]); }
	for (int k = 0; k < 3; ++k) free(s->prefix[k]);
	free(s->buff);
	free(s->rng);
	memset(s, 0, sizeof(*s));
}

/* src/abms.c:230-310 with the reductions of kernel/calcL1de.cl:17-93 and
 * kernel/calcGender.cl:17-89.  Returns 0, 1 (capacity) or 2 (extinct). */
static int calc_stats(or_sim *s, or_stats *row)
{
	const or_age *ge = s->ages[s->cur];
	or_pop *p1 = &s->pop[1];
	uint32_t l1de = 0;
	if (p1->r[R_LARVA].end > p1->r[R_LARVA].start)
		for (uint32_t i = p1->r[R_LARVA].start; i < p1->r[R_LARVA].end; ++i)
			l1de += sc_f2u(ge[i].ageInHours / 24.0f);
	s->pop[0] = *p1;
	s->pop[0].numLarvae1DayEquiv = l1de;
	or_pop *p = &s->pop[0];
	uint32_t fem = 0, inf = 0;
	for (uint32_t i = p->r[R_IMM].start; i < p->r[R_GRAVID].end; ++i) {
		int gap = 0;
		for (int r = R_IMM; r < R_GRAVID; ++r)
			if (i >= p->r[r].end && i < p->r[r + 1].start) gap = 1;
		if (gap) continue;
		if (ge[i].isFemale) {
			++fem;
			inf += ge[i].csd >= 1.0f;
		}
	}
	p->numFemale = fem;
	p->numPotentiallyInfective = inf;
	uint32_t n = 0;
	for (int r = 0; r < 8; ++r) {
		row->count[r] = p->r[r].end - p->r[r].start;
		n += row->count[r];
	}
	uint32_t upper = p->r[R_GRAVID].end; /* max over adult range ends */
	uint32_t adults_range = upper - p->r[R_IMM].start;
	row->l1de = row->count[R_LARVA] > 0 ? l1de : 0;
	row->females = adults_range > 0 ? fem : 0;
	row->potInfective = adults_range > 0 ? inf : 0;
	row->numCycles = s->bnc.numCycles;
	row->sumCycles = s->bnc.sumCycles;
	row->bites = s->bnc.numBites;
	row->infectBites = s->bnc.numInfectBites;
	if (!((double)n <= (double)s->cfg.capacity * 0.95)) return 1;
	if (!(n > 0)) return 2;
	return 0;
}

int or_sim_run(or_sim *s, uint32_t n, or_stats *rows)
{
	for (uint32_t it = 0; it < n && s->status == 0; ++it) {
		uint32_t step = s->step;
		or_stats row;
		int st = calc_stats(s, &row);
		if (st) { s->status = st; break; }
		if (rows) rows[it] = row;
		generate_seeds(s);
		float t = s->cfg.temp[step];
		const or_env *env = &s->cfg.env[step];
		uint32_t h = sc_f2u((float)step * 1.0f) % 24u;
		/* update(): src/abms.c:312-341 */
		uint32_t global = roundup(s->pop[0].r[R_GRAVID].end, LOCAL_SIZE);
		k_kill_agents(s, global, env->carryingCapacity);
		k_update_agents(s, global, env, t, h);
		/* createNewAgents(): src/abms.c:381-462 */
		if ((uint64_t)s->enb.newEggs > (uint64_t)STORAGE * s->cfg.capacity) { s->status = 1; break; }
		k_init_agents(s, s->cur ^ 1, s->enb.newEggs, t);
		const or_pop *ph = &s->pop[0];
		if (s->cfg.scan_path == 0) {
			/* parallelPrefixScan x3, abms.c:407-414 */
			k_pre_scan(s, s->prefix[0], s->buff, 0, ph->r[R_EGG].end, ph->r[R_IMM].end,
			           ph->r[R_BMD].end, OR_EGG | OR_IMMATURE | OR_BMD);
			or_bolt_scan(s->buff, s->prefix[0], ph->r[R_BMD].end + 1u, SCAN_WX);
			k_pre_scan(s, s->prefix[1], s->buff, 0, ph->r[R_LARVA].end, ph->r[R_MS].end,
			           ph->r[R_GRAVID].end, OR_LARVA | OR_MATESEEKING | OR_GRAVID);
			or_bolt_scan(s->buff, s->prefix[1], ph->r[R_GRAVID].end + 1u, SCAN_WX);
			uint32_t ls = ph->r[R_LARVA].start;
			k_pre_scan(s, s->prefix[2], s->buff, ls, ph->r[R_PUPA].end, ph->r[R_GRAVID].end,
			           ph->r[R_GRAVID].end, OR_PUPA | OR_BMS);
			or_bolt_scan(s->buff, s->prefix[2], ph->r[R_GRAVID].end + 1u - ls, SCAN_WX);
		} else {
			/* sequentialPrefixScan x8, abms.c:417-439 */
			uint32_t ls = ph->r[R_LARVA].start;
			k_serial_scan(s, s->prefix[0], 0, ph->r[R_EGG].start, ph->r[R_EGG].end, OR_EGG);
			k_serial_scan(s, s->prefix[1], 0, ph->r[R_EGG].start, ph->r[R_LARVA].end, OR_LARVA);
			k_serial_scan(s, s->prefix[2], ls, ls, ph->r[R_PUPA].end, OR_PUPA);
			k_serial_scan(s, s->prefix[0], 0, ph->r[R_PUPA].start, ph->r[R_IMM].end, OR_IMMATURE);
			k_serial_scan(s, s->prefix[1], 0, ph->r[R_IMM].start, ph->r[R_MS].end, OR_MATESEEKING);
			k_serial_scan(s, s->prefix[2], ls, ph->r[R_MS].start, ph->r[R_GRAVID].end, OR_BMS);
			k_serial_scan(s, s->prefix[0], 0, ph->r[R_BMS].start, ph->r[R_BMD].end, OR_BMD);
			k_serial_scan(s, s->prefix[1], 0, ph->r[R_BMD].start, ph->r[R_GRAVID].end, OR_GRAVID);
		}
		if (k_old_to_new(s)) { s->status = 1; break; }
		s->cur ^= 1; /* swap, abms.c:519-521 */
		s->step = step + 1u;
	}
	return s->status;
}

void or_sim_population(const or_sim *s, or_pop *out) { *out = s->pop[1]; }

void or_sim_arrays(const or_sim *s, const or_agent **a, const or_age **g, const or_state **st)
{
	*a = s->agents[s->cur];
	*g = s->ages[s->cur];
	*st = s->states[s->cur];
}

/* kernel/ageHistogram.cl:13-35 over popD[1] */
void or_sim_histogram(const or_sim *s, uint32_t hist[100])
{
	const or_pop *p = &s->pop[1];
	const or_age *ge = s->ages[s->cur];
	memset(hist, 0, 100 * sizeof(uint32_t));
	for (uint32_t g = p->r[R_IMM].start; g < p->r[R_GRAVID].end; ++g) {
		int gap = 0;
		for (int r = R_IMM; r < R_GRAVID; ++r)
			if (g >= p->r[r].end && g < p->r[r + 1].start) gap = 1;
		if (gap) continue;
		uint32_t bin = sc_f2u(ge[g].ageInHours / 24.0f);
		if (bin > 99u) bin = 99u;
		hist[bin]++;
	}
}

/* opaque handles for the Python test wrapper */
or_sim *or_sim_create(void) { return (or_sim *)calloc(1, sizeof(or_sim)); }
void or_sim_destroy(or_sim *s) { if (s) { or_sim_free(s); free(s); } }
uint32_t or_sim_step(const or_sim *s) { return s->step; }
int or_sim_status(const or_sim *s) { return s->status; }
