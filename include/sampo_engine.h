/*
 * sampo_engine.h -- C ABI of the fused, batched agent-step engine (gfx950).
 *
 * One call of sampo_engine_step() advances every admitted scenario instance
 * by reference iterations (one simulated hour each, src/abms.c:15-16); each
 * iteration is the body of the reference loop src/abms.c:502-529:
 * calcStats -> generateSeeds -> killAgents -> updateAgents -> initAgents ->
 * 3x (preScan + segmented scan) -> oldToNewAgents.  The engine runs it as one
 * fused pass per iteration over structure-of-arrays storage, for any number
 * of instances at once, and keeps every observable result of the reference:
 * the per-iteration printStats row (src/abms.c:67-81), the population ranges
 * (struct Population, include/agent.h:109-123), every agent's fields in the
 * reference's padded, state-sorted index order, and the age histogram
 * (kernel/ageHistogram.cl).
 *
 * All functions return 0 on success or a negative SAMPO_E* code; the message
 * of the last failure is available from sampo_last_error().
 */
#ifndef SAMPO_ENGINE_H
#define SAMPO_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAMPO_E_ARG (-1)     /* invalid argument */
#define SAMPO_E_HIP (-2)     /* HIP runtime error */
#define SAMPO_E_NOMEM (-3)   /* device or host allocation failed */
#define SAMPO_E_FULL (-4)    /* no free instance slot */
#define SAMPO_E_STATE (-5)   /* call not valid in the current state */

/* instance status */
#define SAMPO_RUNNING 0u
#define SAMPO_DONE 1u          /* horizon reached */
#define SAMPO_CAPACITY 2u      /* nAgents > 0.95 * maxAgents (src/abms.c:305) */
#define SAMPO_EXTINCT 3u       /* nAgents == 0 (src/abms.c:307) */
#define SAMPO_PHYS_FULL 4u     /* physical slots of the instance exhausted */
#define SAMPO_GROW 5u          /* paused: its buffers need more slots; sampo_engine_service() grows them */
#define SAMPO_HELD 6u          /* paused by sampo_engine_hold() until it is resumed */
#define SAMPO_FAULT 7u         /* a day launch's wait for its work-groups gave up; the instance's state is lost */
#define SAMPO_EMPTY 0xffu      /* slot not in use */

#define SAMPO_MODE_REFERENCE 0u /* killAgents.cl formulas, initial eggs via initAgents */
#define SAMPO_MODE_CONFIG 1u    /* constant per-stage daily mortality, explicit initial population */

/* struct Environment (include/agent.h:22-35), same field order, 24 bytes */
typedef struct {
	float bloodmealSuccess;
	float ITNValue;
	float IRSValue;
	float carryingCapacity;
	float larvacideValue;
	float oviTrapValue;
} sampo_env;

/* One agent in host form: the union of struct Agent, AgentAge and the state
 * (include/agent.h:66-102).  `aux` holds the union ovipositionAttemps /
 * cumulativeLarvalDelay as raw bits.  State is implied by the range. */
typedef struct {
	float ageInHours;
	float hoursInState;
	float delay;
	float csd;           /* cumulativeSporogonicDevelopment */
	uint32_t aux;
	uint32_t availableEggs;
	uint32_t cycleLength;
	uint32_t humanBloodmealCount;
	uint32_t numEggBatches;
	uint32_t isFemale;
} sampo_agent;

/* Agent ranges of struct Population, in state order egg..gravid. */
typedef struct {
	uint32_t start[8];
	uint32_t end[8];
} sampo_layout;

/* printStats row (src/abms.c:67-81) taken at the start of an iteration. */
typedef struct {
	uint32_t count[8];     /* eggs, larvae, pupae, immature, mating, BMS, BMD, gravid */
	uint32_t l1de;         /* numLarvae1DazEquiv (0 without larvae) */
	uint32_t females;
	uint32_t potInfective;
	uint32_t numCycles;    /* BitesNcycles of the previous iteration */
	uint32_t sumCycles;
	uint32_t bites;
	uint32_t infectBites;
	uint32_t valid;        /* 1 if the reference would have printed this row */
} sampo_row;

/* Species behaviour: the runtime form of the reference's species plug-in,
 * the header chosen with -DSPECIES=<header> when the kernels are built
 * (src/abms.c:565-568, kernel/updateAgents.cl:6-23).  Every function of
 * include/properties.h is one member of the parametric family below;
 * sampo_species_default() gives include/properties.h itself.  Temperatures
 * are T in degC; binary32 unless noted, evaluated without contraction. */
typedef struct {
	float egg_incubation[2];   /* eggsIncubation (properties.h:29-31): [0]*T + [1] */
	float hatch_threshold[4];  /* eggsHatching (properties.h:39-60): first i with p < threshold[i], else 4; */
	float hatch_slope[5];      /*   (float)(uint)(slope[i]*p + offset[i] + 0.5f) */
	float hatch_offset[5];
	float larva_delay[2];      /* larvaDelay (properties.h:87-91): BoxMuller(mean [0], sd [1]) */
	float dev_temp_range[2];   /* larvaDevelopment (properties.h:105-109): T clamped to [lo, hi], */
	float dev_rates[32];       /*   dev_rates[min((int)T - (int)lo, 31)] per hour (properties.h:94-98) */
	double pupa_delay[2];      /* pupaDelay (properties.h:133-135): [0]*T + [1] in binary64, rounded once */
	float immature_delay[2];   /* immatureDelay (properties.h:159-161): [0]*T + [1] */
	float mate_seeking_delay;  /* mateSeekingDelay (properties.h:185-187) */
	float bms_delay;           /* bloodMealSeekingDelay (properties.h:212-214) */
	float bmd_delay[2];        /* bloodMealDigestingDelay (properties.h:239-241): [0]*T + [1] */
	float egg_batch[3];        /* generateEggs (properties.h:257-266): (uint)(BoxMuller(mean [0], sd [1]) + 0.5f),
	                              times pow([2], batches) */
	float lay_probability;     /* a gravid tries to lay with U < p (kernel/updateAgents.cl:233) */
	float bite_human;          /* BITE_HUMAN_PROBABILITY (properties.h:12, updateAgents.cl:15-17) */
	float bite_indoor;         /* BITE_INDOOR_PROBABILITY (properties.h:13) */
	float rest_indoor;         /* REST_INDOOR_PROBABILITY (properties.h:14) */
	uint32_t window[8];        /* transition-time predicates: bit h set if the transition out of state
	                              egg, larva, pupa, immature, mate seeking, BMS, BMD may happen at world
	                              hour h, and [7]: gravidsEggLayTime (properties.h:68-70 ... 280-282) */
	uint32_t reserved;         /* 0 */
} sampo_species;

/* include/properties.h as a table. */
void sampo_species_default(sampo_species *out);

/* Check a table: hatch thresholds ascending, dev_temp_range[0] integral and
 * <= [1], sds >= 0, probabilities in [0, 1], windows within 24 bits. */
int sampo_species_check(const sampo_species *sp);

/* Read a species file: properties.h values overridden by lines
 * "name v1 v2 ..." naming a member of sampo_species (windows as
 * "window_<egg|larva|pupa|immature|mate_seeking|bms|bmd|lay>" followed by
 * hours and ranges, e.g. "0-6 18-23", or "all" / "none"); '#' starts a
 * comment.  The result is checked with sampo_species_check(). */
int sampo_species_read(const char *path, sampo_species *out);

typedef struct {
	uint32_t mode;          /* SAMPO_MODE_* */
	uint32_t seed;          /* srand() seed of the host seed stream (src/abms.c:611) */
	uint32_t horizon;       /* iterations to run */
	uint32_t n_eggs;        /* reference mode: initial eggs (src/abms.c:18) */
	uint64_t max_agents;    /* maxAgents of the 95% capacity check (include/abms.h:16) */
	uint64_t phys_slots;    /* initial agent slots per buffer (grown on demand up to max_agents
	                           by sampo_engine_service); 0 = max_agents, never grown */
	float dmr[4];           /* config mode: daily mortality egg, larva, pupa, adult */
	const float *temp;      /* horizon temperatures, one per iteration */
	const sampo_env *env;   /* horizon environments, or one if env_count == 1 */
	uint32_t env_count;
	/* config mode initial population, reference layout (start multiple-of-64
	 * rule of kernel/oldToNewAgents.cl:140-160 or any layout with gaps >= 1);
	 * agents dense in range order (sum of range sizes). */
	const sampo_layout *init_layout;
	const sampo_agent *init_agents;
	const sampo_species *species; /* NULL: include/properties.h */
} sampo_scenario;

typedef struct sampo_engine sampo_engine;

const char *sampo_last_error(void);

/* Create an engine on HIP device `device` with room for max_slots instances. */
int sampo_engine_create(int device, uint32_t max_slots, sampo_engine **out);
void sampo_engine_destroy(sampo_engine *e);

/* Tile shape of the fused pass for the instances admitted from now on:
 * work-group size, agents per thread and the minimum waves per SIMD the
 * registers are capped for (0: no cap).  Instances of different shapes are
 * advanced by one pass launch per shape each iteration, up to 4 shapes in
 * use at once (SAMPO_E_FULL beyond; a shape no instance uses any more is
 * replaced).  Results do not depend on the shape.  Only shapes compiled into
 * the library are accepted (SAMPO_E_ARG otherwise). */
int sampo_engine_set_tile(sampo_engine *e, uint32_t wg, uint32_t apt, uint32_t waves);

/* Persistent work-groups of the pass per CU (0: as many as the tile shape
 * lets reside; otherwise at most that many). */
int sampo_engine_set_grid(sampo_engine *e, uint32_t per_cu);

/* Day launches for instances admitted from now on (no counterpart in the
 * reference, whose loop launches its kernels once per iteration,
 * src/abms.c:502-529).  work_groups > 0: such an instance is cut into at
 * most work_groups equal chunks, and while it is the only instance of the
 * engine, sampo_engine_step() and sampo_engine_step_days() advance it by
 * cooperative launches of work_groups work-groups, each running up to 96
 * iterations (four simulated days) with no kernel boundary between them.  For
 * instances small enough that a launch per iteration costs more than their
 * agents (about 1M agents and below).  Results do not depend on it.
 * work_groups must be at most 256 and at most the work-groups the device can
 * keep resident at once (SAMPO_E_ARG), and the device must support
 * cooperative launches (SAMPO_E_STATE).  0 (the default): off. */
int sampo_engine_set_day(sampo_engine *e, uint32_t work_groups);

/* Chunking of instances admitted from now on.  A chunk is the run of agents
 * one work-group takes and the granularity of output segments: at least
 * `agents` agents (default 3072), grown so that an instance has at most
 * `target_chunks` chunks (0: two per resident work-group of the pass). */
int sampo_engine_set_chunk(sampo_engine *e, uint32_t agents, uint32_t target_chunks);

/* Admit a scenario into a free slot: allocates its device state, uploads its
 * inputs and prepares iteration 0.  Returns the slot in *slot. */
int sampo_engine_admit(sampo_engine *e, const sampo_scenario *sc, uint32_t *slot);

/* Device bytes sampo_engine_admit() would allocate for this scenario with
 * the given chunk size (0: the default 3072), for admission by footprint. */
uint64_t sampo_engine_footprint(const sampo_scenario *sc, uint32_t chunk);

/* Free a slot (its instance must not be running in an unsynchronised step). */
int sampo_engine_release(sampo_engine *e, uint32_t slot);

/* Hold (hold[i] != 0) or resume (hold[i] == 0) the instances in slots
 * 0..n-1 from the next step on: a running instance becomes SAMPO_HELD and
 * takes no part in the steps until it is resumed; a held one runs again.
 * Instances in any other state are left as they are.  Holding changes when
 * an instance advances, never what it computes: its results stay those of a
 * run alone.  No counterpart in the reference (one run per process there):
 * it lets a server advance some admitted scenarios ahead of others, e.g.
 * the ones with the least work left. */
int sampo_engine_hold(sampo_engine *e, uint32_t n, const uint8_t *hold);

/* Enqueue n iterations for all running instances; asynchronous. */
int sampo_engine_step(sampo_engine *e, uint32_t n);

/* Capture one day (24 iterations) as a HIP graph on first use and launch it
 * `days` times; asynchronous. */
int sampo_engine_step_days(sampo_engine *e, uint32_t days);

int sampo_engine_sync(sampo_engine *e);

/* Time n iterations with HIP events on the engine's stream (synchronises). */
int sampo_engine_time(sampo_engine *e, uint32_t n, int use_graph, float *ms);

/* Enqueue n iterations with HIP events around every fused pass; returns the
 * wall time from the first to the last pass and the summed pass durations
 * (synchronises). */
int sampo_engine_time_pass(sampo_engine *e, uint32_t n, float *wall_ms, float *pass_ms);

/* Capture the next n iterations as one HIP graph, with a HIP event recorded
 * before and after every fused pass; sampo_engine_replay() launches it
 * (asynchronous, once per capture: the iterations are those n), and
 * sampo_engine_replay_times() then returns the time from the first pass's
 * start to the last pass's end and the summed pass durations (synchronises). */
int sampo_engine_capture(sampo_engine *e, uint32_t n);
int sampo_engine_replay(sampo_engine *e);
int sampo_engine_replay_times(sampo_engine *e, float *wall_ms, float *pass_ms);

/* Advance `days` days (sampo_engine_step_days) with a HIP event pair around
 * each day's launch -- the cooperative day launch, or the day graph -- and
 * return the time from the first day's start to the last day's end and the
 * summed per-day durations (synchronises). */
int sampo_engine_time_days(sampo_engine *e, uint32_t days, float *wall_ms, float *launch_ms);

/* Free and total device memory of HIP device `device`, for admission by footprint. */
int sampo_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes);

/* Device-to-device copy of `bytes` bytes with a plain 16-byte-per-lane
 * kernel, `iters` times; average milliseconds per copy. */
int sampo_stream_copy(int device, uint64_t bytes, uint32_t iters, float *ms_per_copy);

/* Status, current iteration and agents alive at its start for slots
 * 0..n-1 in one read (synchronises); empty slots report SAMPO_EMPTY.  Any
 * output pointer may be NULL. */
int sampo_engine_poll(sampo_engine *e, uint32_t n, uint32_t *status, uint32_t *step, uint64_t *n_agents);

/* Between steps (synchronises): give every instance that is paused in
 * SAMPO_GROW, or running with fewer than headroom x its agents in physical
 * slots, larger buffers (up to max_agents), and resume the paused ones.  A
 * paused instance stops advancing until this is called, so its step lags
 * the others'; keep stepping until every instance is done.  *grown receives
 * the number of instances given new buffers. */
int sampo_engine_service(sampo_engine *e, float headroom, uint32_t *grown);

/* Status, current iteration and agents alive at its start (synchronises). */
int sampo_engine_status(sampo_engine *e, uint32_t slot, uint32_t *status, uint32_t *step,
                        uint64_t *n_agents);

/* Copy rows [first, first+n) of the instance's stats series (synchronises). */
int sampo_engine_rows(sampo_engine *e, uint32_t slot, uint32_t first, uint32_t n, sampo_row *out);

/* Materialise the instance's current agents in reference index order:
 * layout receives the ranges; agents receives sum(end-start) records, dense
 * in range order, newborn eggs included (synchronises). */
int sampo_engine_export(sampo_engine *e, uint32_t slot, sampo_layout *layout, sampo_agent *agents,
                        uint64_t max_agents);

/* ageHistogram (kernel/ageHistogram.cl:13-35) of the current adults. */
int sampo_engine_histogram(sampo_engine *e, uint32_t slot, uint32_t hist[100]);

/* Sum over instances of agents alive at iteration starts since the last reset. */
int sampo_engine_agent_steps(sampo_engine *e, uint64_t *total, int reset);

/* Host-side reference seed stream: the first n glibc rand() values after
 * srand(seed) (src/abms.c:140-158,611). */
void sampo_seed_stream(uint32_t seed, uint32_t n, uint32_t *out);

/* Synthetic configuration-mode inputs (a splitmix64 stream independent of the
 * seed stream).  Daily temperature mean + amp*sin(2*pi*d/365) + noise*N(0,1),
 * held for the 24 iterations of each day. */
int sampo_gen_temperature(uint64_t seed, uint32_t days, float mean, float amp, float noise, float *hourly);

/* n0 agents: egg/larva/pupa/immature with probability .3/.3/.1/.3, sex 50/50,
 * part of each state's delay elapsed; laid out like the reference (ranges
 * start at the next multiple of 64 after the previous end).  `agents` must
 * hold n0 records; they come back dense in range order. */
int sampo_gen_population(uint64_t seed, uint32_t n0, float t0, sampo_layout *layout, sampo_agent *agents);

#ifdef __cplusplus
}
#endif
#endif
