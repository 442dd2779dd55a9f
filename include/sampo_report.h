/*
 * sampo_report.h -- the reference driver's text input and output around the
 * agent step: the per-iteration printStats line, the printPopulation range
 * table, histogram.txt, and the temperature and environment input files.
 * Host C in libsampo_icl.so; no device is touched.
 *
 * Formatting functions write a NUL-terminated string into buf (capacity
 * cap) and return its length, or a negative SAMPO_E* code if it does not
 * fit.  Readers return the number of values read, or a negative code.
 */
#ifndef SAMPO_REPORT_H
#define SAMPO_REPORT_H

#include <stddef.h>
#include <stdint.h>
#include "sampo_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* printStats (src/abms.c:67-81) for iteration `step` at temperature `temp`
 * from a stats row: time, temperature, adults, immature, mating, BMS, BMD,
 * gravid, females, potentially infective, males, eggs, larvae, pupae, L1DE,
 * biomass, mean cycle length (integer division), bites, infectious bites,
 * ", "-separated and ending in "\r\n".  hours_per_step is hoursInTimeStep
 * (src/abms.c:15). */
int sampo_format_stats(char *buf, size_t cap, uint32_t step, float hours_per_step, float temp,
                       const sampo_row *row);

/* printPopulation (src/abms.c:83-92): eight lines "<name> \t[start\tend]\tcount\r\n". */
int sampo_format_population(char *buf, size_t cap, const sampo_layout *layout);

/* histogram.txt of storeHistogram (src/abms.c:176-180): "%d, " for each of 100 bins. */
int sampo_format_histogram(char *buf, size_t cap, const uint32_t hist[100]);

/* readSingleValFile (src/abms.c:94-108): n whitespace-separated floats. */
int sampo_read_single_val_file(const char *path, uint32_t n, float *out);

/* readEnvironment (src/abms.c:110-137): per line the coverage and
 * effectiveness of IRS, ITN, larvicide and ovitrap, then the blood-meal
 * success, ", "-separated; each intervention value is coverage *
 * effectiveness in binary32, carrying capacity is the argument. */
int sampo_read_environment(const char *path, uint32_t n, float carrying_capacity, sampo_env *out);

#ifdef __cplusplus
}
#endif
#endif
