/*
 * sampo_icl_types.h -- device layouts of the reference's kernel arguments.
 *
 * The driver allocates agent buffers as maxAgents * sizeof(struct Agent /
 * AgentAge / AgentState) with the host sizes 24 / 16 / 8 bytes
 * (include/agent.h:66-102, cl_bool being 4 bytes), and exchanges Population
 * (76 B), EggsNbiomass (8 B), BitesNcycles (16 B), Seeds (16 B) and
 * Environment (24 B) with the host.  These structs keep those sizes and field
 * orders; booleans are 32-bit.
 */
#ifndef SAMPO_ICL_TYPES_H
#define SAMPO_ICL_TYPES_H

#include <stdint.h>

enum { ICL_EGG = 1, ICL_LARVA = 2, ICL_PUPA = 4, ICL_IMMATURE = 8, ICL_MATESEEKING = 16,
       ICL_BMS = 32, ICL_BMD = 64, ICL_GRAVID = 128 };

typedef struct {
	uint32_t availableEggs, cycleLength, humanBloodmealCount;
	float delay;
	uint32_t numEggBatches;
	uint32_t aux; /* union ovipositionAttemps / cumulativeLarvalDelay */
} IclAgent;

typedef struct {
	float ageInHours, hoursInState;
	uint32_t isFemale;
	float csd;
} IclAgentAge;

typedef struct { uint32_t dead, state; } IclAgentState;
typedef struct { uint32_t start, end; } IclRange;
typedef struct {
	IclRange r[8];
	uint32_t numLarvae1DayEquiv, numFemale, numPotentiallyInfective;
} IclPopulation;
typedef struct { int32_t totalBiomass; uint32_t newEggs; } IclEnb;
typedef struct { uint32_t numCycles, sumCycles, numBites, numInfectBites; } IclBnc;
typedef struct { uint32_t z1, z2, z3, z4; } IclSeeds;
typedef struct {
	float bloodmealSuccess, ITNValue, IRSValue, carryingCapacity, larvacideValue, oviTrapValue;
} IclEnvironment;
typedef struct { uint32_t ticket, epoch; } IclScanCtx;

#ifdef __cplusplus
static_assert(sizeof(IclAgent) == 24 && sizeof(IclAgentAge) == 16 && sizeof(IclAgentState) == 8, "agent sizes");
static_assert(sizeof(IclPopulation) == 76 && sizeof(IclEnvironment) == 24, "struct sizes");
#endif

#endif
