/*
 * sampo_icl.h -- drop-in replacement of the reference's device-runtime layer
 * (include/lib_icl.h, src/lib_icl.c) and its scan module (include/scan.h,
 * src/boltScan.c) over HIP on gfx950.
 *
 * The driver src/abms.c reaches the hot path only through these calls and
 * the kernel contract (the kernel/ *.cl names and argument lists).  Every entry
 * point below keeps the reference's name, argument meaning and fail-fast
 * error behaviour (a failed call prints a message and exits, as ICL_ASSERT
 * does, include/lib_icl.h:19-25); icl_finish() replaces the driver's direct
 * clFinish(dev->queue) (src/abms.c:676).  Kernels are not compiled at run
 * time: icl_create_kernel() looks the kernel name up in a registry of
 * gfx950 kernels built into libsampo_icl.so, one per reference kernel, with
 * the reference's argument list.  Scalar types keep the reference's host
 * names (include/host_types.h) so the driver compiles unchanged.
 */
#ifndef SAMPO_ICL_H
#define SAMPO_ICL_H

#include <stdarg.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

/* scalar types of the reference's host interface (include/host_types.h:9-12) */
typedef uint32_t cl_uint;
typedef int32_t cl_int;
typedef float cl_float;
typedef uint32_t cl_bool;
typedef uint64_t cl_ulong;
typedef uint64_t cl_mem_flags;
typedef uint64_t cl_device_type;
typedef uint64_t cl_map_flags;

/* the model's scalar names over them (include/host_types.h:9-12), so that
 * include/agent.h compiles with this header in place of host_types.h */
#ifndef REAL
#define REAL cl_float
#endif
#ifndef INT
#define INT cl_int
#endif
#ifndef UINT
#define UINT cl_uint
#endif
#ifndef BOOL
#define BOOL cl_bool
#endif

#define CL_FALSE 0u
#define CL_TRUE 1u
#define CL_MEM_READ_WRITE (1u << 0)
#define CL_MEM_WRITE_ONLY (1u << 1)
#define CL_MEM_READ_ONLY (1u << 2)

/* device classes (include/lib_icl.h:31-34) */
#define ICL_CPU (1u << 1)
#define ICL_GPU (1u << 2)
#define ICL_ACL (1u << 3)
#define ICL_ALL 0xFFFFFFFFu

/* include/lib_icl.h:19-25 */
#define ICL_ASSERT(__condition, __message, ...)                                   \
	if (!(__condition)) {                                                         \
		fprintf(stderr, "Device assertion failure in %s#%d:\n", __FILE__, __LINE__); \
		printf(__message, ##__VA_ARGS__);                                         \
		printf("\n");                                                             \
		exit(-1);                                                                 \
	}

#define ICL_INFO(__message, ...) { printf(__message, ##__VA_ARGS__); }

/* include/lib_icl.h:36-75: one HIP device with one in-order stream */
typedef struct _icl_device {
	int hip_device;
	void *stream;            /* hipStream_t: the in-order queue */
	cl_ulong mem_size;
	cl_ulong mem_available;  /* reduced by each buffer allocation */
	cl_ulong max_buffer_size;
	char *name;
	cl_device_type type;
	char *vendor;
	char *version;
	char *driver_version;
	char *profile;
	cl_uint max_compute_units;
	cl_uint max_clock_frequency;
	cl_uint max_work_item_dimensions;
	size_t *max_work_item_sizes;
	size_t max_work_group_size;
	cl_ulong local_mem_size;
} icl_device;

/* include/lib_icl.h:77-81 */
typedef struct _icl_buffer {
	void *mem;
	size_t size;
	icl_device *dev;
} icl_buffer;

typedef enum { ICL_SOURCE, ICL_BINARY, ICL_STRING, ICL_NO_CACHE } icl_create_kernel_flag;

/* include/lib_icl.h:85-88: a registry entry bound to a device, with the
 * device copy of the species table its build options selected */
typedef struct _icl_kernel {
	int id;
	icl_device *dev;
	void *species;
} icl_kernel;

/* include/lib_icl.h:90-93: start/stop HIP events of the last command */
typedef struct _icl_event {
	void *start;
	void *stop;
	cl_uint num_event;
} icl_event;

typedef enum { ICL_SEC, ICL_MILLI, ICL_NANO } icl_time_flag;
typedef long long time_int;

/* include/lib_icl.h:136-144 */
typedef struct _icl_timer {
	time_int start;
	time_int clocks;
	double current_time;
	icl_time_flag time_flag;
} icl_timer;

typedef enum { ICL_ENQUEUED, ICL_SUBMITTED, ICL_STARTED, ICL_FINISHED } icl_event_flag;

/* ---- devices: include/lib_icl.h:100-103 ---- */
void icl_init_devices(cl_device_type device_type);
cl_uint icl_get_num_devices(void);
icl_device *icl_get_device(cl_uint id);
void icl_release_devices(void);
void icl_print_device_short_info(icl_device *dev);
void icl_print_device_infos(icl_device *dev);
/* replaces clFinish(dev->queue) (src/abms.c:676) */
void icl_finish(icl_device *dev);

/* ---- buffers: include/lib_icl.h:105-119 ---- */
icl_buffer *icl_create_buffer(icl_device *dev, cl_mem_flags flags, size_t size);
void icl_read_buffer(const icl_buffer *buf, cl_bool blocking, size_t size, void *source_ptr,
                     icl_event *wait_event, icl_event *event);
void icl_read_buffer_offset(const icl_buffer *buf, cl_bool blocking, size_t offset, size_t size,
                            void *source_ptr, icl_event *wait_event, icl_event *event);
void icl_write_buffer(icl_buffer *buf, cl_bool blocking, size_t size, const void *source_ptr,
                      icl_event *event_wait, icl_event *event);
void icl_write_buffer_offset(icl_buffer *buf, cl_bool blocking, size_t offset, size_t size,
                             const void *source_ptr, icl_event *wait_event, icl_event *event);
void icl_copy_buffer(icl_buffer *src_buf, icl_buffer *dest_buf, size_t size, icl_event *event_wait,
                     icl_event *event);
void icl_release_buffer(icl_buffer *buf);
void icl_release_buffers(cl_uint num, ...);

/* ---- kernels: include/lib_icl.h:121-123 ----
 * kernel_name selects the registered gfx950 kernel and file_name, if given,
 * must name the reference file that holds it.  Of build_options only
 * -DSPECIES=<x> is read (src/abms.c:565-568): a name ending in properties.h
 * selects include/properties.h, anything else names a species file
 * (sampo_species_read, sampo_engine.h), looked up as given and then under
 * each -I directory of the options; without it, properties.h.  icl_run_kernel takes
 * num_args pairs (size_t size, void *value): size 0 passes an icl_buffer*,
 * (n, NULL) requests n bytes of local memory, anything else is a by-value
 * argument of that size (src/lib_icl.c:668-704). */
icl_kernel *icl_create_kernel(icl_device *dev, const char *file_name, const char *kernel_name,
                              const char *build_options, icl_create_kernel_flag flag);
void icl_run_kernel(const icl_kernel *kernel, cl_uint work_dim, const size_t *global_work_size,
                    const size_t *local_work_size, icl_event *wait_event, icl_event *event,
                    cl_uint num_args, ...);
void icl_release_kernel(icl_kernel *kernel);

/* ---- timers: include/lib_icl.h:146-150 ---- */
icl_timer *icl_init_timer(icl_time_flag time_flag);
void icl_start_timer(icl_timer *timer);
void icl_restart_timer(icl_timer *timer);
double icl_stop_timer(icl_timer *timer);
void icl_release_timer(icl_timer *timer);

/* ---- events: include/lib_icl.h:154-159 ---- */
icl_event *icl_create_event(void);
void icl_release_event(icl_event *event);
void icl_release_events(cl_uint num, ...);
void icl_wait_for_events(cl_uint num, ...);
double icl_profile_event(icl_event *event, icl_event_flag event_start, icl_event_flag event_end,
                         icl_time_flag time_flag);

/* ---- segmented scan: include/scan.h:5-9 ----
 * scan(data, flag, n): in place, data[i] becomes the inclusive sum of data
 * over the run of equal consecutive flag values containing i, minus one --
 * the result of src/boltScan.c:36-126 with init -1 -- computed by one
 * single-pass decoupled look-back kernel. */
void segmented_scan_init(size_t wx, cl_uint max_n, icl_device *dev, const char *build_options,
                         icl_create_kernel_flag flag);
void segmented_scan_release(void);
void scan(icl_buffer *data, icl_buffer *flag, cl_uint n);

#ifdef __cplusplus
}
#endif
#endif
