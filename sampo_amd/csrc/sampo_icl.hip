/*
 * sampo_icl.hip -- libsampo_icl.so: the reference's device-runtime API
 * (include/lib_icl.h) and scan API (include/scan.h) over HIP, with the
 * reference kernels registered under their own names.
 *
 * Semantics kept from src/lib_icl.c: one in-order queue per device;
 * enqueues are asynchronous and blocking transfers wait; every failure prints
 * a message and exits (ICL_ASSERT); buffers account for the device memory
 * they take.  What changes: kernels come precompiled for gfx950 from a
 * registry instead of clBuildProgram, varargs are type-checked against each
 * kernel's argument list before launch, and the scan module is one kernel.
 */
#include <hip/hip_runtime.h>

#include <time.h>

#include <string>
#include <vector>

#include "../../include/sampo_icl.h"
#include "sampo_icl_kernels.hip"

namespace {

std::vector<icl_device> g_devices;

void die(const char *what, const char *detail)
{
	fprintf(stderr, "sampo_icl: %s%s%s\n", what, detail && *detail ? ": " : "", detail ? detail : "");
	fflush(stderr);
	exit(-1);
}

#define HIP_OR_DIE(x)                                   \
	do {                                                \
		hipError_t e_ = (x);                            \
		if (e_ != hipSuccess) die(#x, hipGetErrorString(e_)); \
	} while (0)

hipStream_t stream_of(const icl_device *d) { return (hipStream_t)d->stream; }

/* one varargs pair of icl_run_kernel (src/lib_icl.c:679-693) */
struct Arg {
	size_t size;
	void *value;
};

void *buf_arg(const Arg &a, const char *kernel, int i)
{
	if (a.size != 0 || !a.value) {
		char m[128];
		snprintf(m, sizeof(m), "%s argument %d must be a buffer", kernel, i);
		die("kernel argument mismatch", m);
	}
	return ((icl_buffer *)a.value)->mem;
}

template <typename T> T val_arg(const Arg &a, const char *kernel, int i)
{
	if (a.size != sizeof(T) || !a.value) {
		char m[160];
		snprintf(m, sizeof(m), "%s argument %d must be a %zu-byte value, got %zu bytes", kernel, i, sizeof(T), a.size);
		die("kernel argument mismatch", m);
	}
	T v;
	memcpy(&v, a.value, sizeof(T));
	return v;
}

typedef void (*launcher)(dim3 grid, dim3 block, hipStream_t s, const Arg *a, const sampo_species *sp);

#define B(T, i) ((T *)buf_arg(a[i], name, i))
#define V(T, i) (val_arg<T>(a[i], name, i))

void l_initAgents(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "initAgents";
	hipLaunchKernelGGL(initAgents, g, b, 0, s, B(IclAgent, 0), B(IclAgentAge, 1), B(IclAgentState, 2),
	                   B(const IclEnb, 3), V(float, 4), B(const IclSeeds, 5), sp);
}
void l_killAgents(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "killAgents";
	hipLaunchKernelGGL(killAgents, g, b, 0, s, B(IclAgentAge, 0), B(IclAgentState, 1), B(const IclPopulation, 2),
	                   B(IclEnb, 3), B(IclBnc, 4), B(const IclSeeds, 5), V(float, 6));
}
void l_killAgentsStaged(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "killAgentsStaged";
	hipLaunchKernelGGL(killAgentsStaged, g, b, 0, s, B(IclAgentAge, 0), B(IclAgentState, 1),
	                   B(const IclPopulation, 2), B(IclEnb, 3), B(IclBnc, 4), B(const IclSeeds, 5), V(float, 6),
	                   V(float, 7), V(float, 8), V(float, 9));
}
void l_updateAgents(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "updateAgents";
	hipLaunchKernelGGL(updateAgents, g, b, 0, s, B(IclAgent, 0), B(IclAgentAge, 1), B(IclAgentState, 2),
	                   B(const IclPopulation, 3), V(IclEnvironment, 4), V(float, 5), B(IclBnc, 6), B(IclEnb, 7),
	                   B(const IclSeeds, 8), V(float, 9), V(uint32_t, 10), sp);
}
void l_calcL1dePerGroup(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "calcL1dePerGroup";
	hipLaunchKernelGGL(calcL1dePerGroup, g, b, 0, s, B(const IclAgentAge, 0), B(const IclPopulation, 1),
	                   B(uint32_t, 2));
}
void l_calcL1deTotal(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "calcL1deTotal";
	hipLaunchKernelGGL(calcL1deTotal, g, b, 0, s, B(IclPopulation, 0), B(const uint32_t, 1), V(uint32_t, 2));
}
void l_calcFemalesPerGroup(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "calcFemalesPerGroup";
	hipLaunchKernelGGL(calcFemalesPerGroup, g, b, 0, s, B(const IclAgentAge, 0), B(const IclPopulation, 1),
	                   B(uint32_t, 2));
}
void l_calcFemalesTotal(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "calcFemalesTotal";
	hipLaunchKernelGGL(calcFemalesTotal, g, b, 0, s, B(IclPopulation, 0), B(const uint32_t, 1));
}
void l_preScan(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "preScan";
	hipLaunchKernelGGL(preScan, g, b, 0, s, B(const IclAgentState, 0), B(const IclPopulation, 1), B(int32_t, 2),
	                   B(uint32_t, 3), V(uint32_t, 4), V(uint32_t, 5), V(uint32_t, 6), V(uint32_t, 7),
	                   V(int32_t, 8));
}
void l_oldToNewAgents(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "oldToNewAgents";
	hipLaunchKernelGGL(oldToNewAgents, g, b, 0, s, B(const IclAgent, 0), B(const IclAgentAge, 1),
	                   B(const IclAgentState, 2), B(IclAgent, 3), B(IclAgentAge, 4), B(IclAgentState, 5),
	                   B(const int32_t, 6), B(const int32_t, 7), B(const int32_t, 8), B(IclPopulation, 9),
	                   B(const IclEnb, 10));
}
void l_ageHistogram(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "ageHistogram";
	hipLaunchKernelGGL(ageHistogram, g, b, 0, s, B(const IclAgentAge, 0), B(const IclPopulation, 1),
	                   B(uint32_t, 2));
}
void l_serialScan(dim3 g, dim3 b, hipStream_t s, const Arg *a, const sampo_species *sp)
{
	const char *name = "serialScan";
	hipLaunchKernelGGL(serialScan, g, b, 0, s, B(const IclAgentState, 0), B(int32_t, 1), B(const IclPopulation, 2),
	                   V(uint32_t, 3), V(uint32_t, 4), V(uint32_t, 5), V(uint32_t, 6));
}
#undef B
#undef V

/* kernel registry: reference file, kernel name, argument count, launcher */
struct Entry {
	const char *file;
	const char *name;
	cl_uint nargs;
	launcher fn;
};

const Entry k_registry[] = {
	{"initAgents.cl", "initAgents", 6, l_initAgents},
	{"killAgents.cl", "killAgents", 7, l_killAgents},
	{"killAgents.cl", "killAgentsStaged", 10, l_killAgentsStaged}, /* configuration mode, not a reference kernel */
	{"updateAgents.cl", "updateAgents", 11, l_updateAgents},
	{"calcL1de.cl", "calcL1dePerGroup", 3, l_calcL1dePerGroup},
	{"calcL1de.cl", "calcL1deTotal", 3, l_calcL1deTotal},
	{"calcGender.cl", "calcFemalesPerGroup", 3, l_calcFemalesPerGroup},
	{"calcGender.cl", "calcFemalesTotal", 2, l_calcFemalesTotal},
	{"preScan.cl", "preScan", 9, l_preScan},
	{"oldToNewAgents.cl", "oldToNewAgents", 11, l_oldToNewAgents},
	{"ageHistogram.cl", "ageHistogram", 3, l_ageHistogram},
	{"serialScan.cl", "serialScan", 7, l_serialScan},
};
const int k_nreg = sizeof(k_registry) / sizeof(k_registry[0]);

const char *base_name(const char *p)
{
	const char *s = strrchr(p, '/');
	return s ? s + 1 : p;
}

char *dup(const char *s)
{
	size_t n = strlen(s) + 1;
	char *d = (char *)malloc(n);
	memcpy(d, s, n);
	return d;
}

time_int now_ns()
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (time_int)ts.tv_sec * 1000000000LL + ts.tv_nsec;
}

double scale(double ns, icl_time_flag f)
{
	return f == ICL_SEC ? ns * 1e-9 : f == ICL_MILLI ? ns * 1e-6 : ns;
}

/* scan module state (src/boltScan.c:14-27) */
struct ScanState {
	icl_device *dev = nullptr;
	IclScanCtx *ctx = nullptr;
	uint64_t *status = nullptr;
	cl_uint max_n = 0;
} g_scan;

void record_start(icl_event *ev, hipStream_t s)
{
	if (!ev) return;
	if (!ev->start) {
		HIP_OR_DIE(hipEventCreate((hipEvent_t *)&ev->start));
		HIP_OR_DIE(hipEventCreate((hipEvent_t *)&ev->stop));
	}
	HIP_OR_DIE(hipEventRecord((hipEvent_t)ev->start, s));
}

void record_stop(icl_event *ev, hipStream_t s)
{
	if (!ev) return;
	HIP_OR_DIE(hipEventRecord((hipEvent_t)ev->stop, s));
	ev->num_event = 1;
}

void wait_on(icl_event *ev, hipStream_t s)
{
	if (ev && ev->stop) HIP_OR_DIE(hipStreamWaitEvent(s, (hipEvent_t)ev->stop, 0));
}

} // namespace

/* ------------------------------------------------------------- devices */

extern "C" void icl_init_devices(cl_device_type device_type)
{
	g_devices.clear();
	if (!(device_type & ICL_GPU)) return; /* no CPU or accelerator devices behind this library */
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
	for (int i = 0; i < n; ++i) {
		hipDeviceProp_t p;
		HIP_OR_DIE(hipGetDeviceProperties(&p, i));
		icl_device d;
		memset(&d, 0, sizeof(d));
		d.hip_device = i;
		d.stream = nullptr; /* created by icl_get_device: only the devices a driver uses are touched */
		d.mem_size = p.totalGlobalMem;
		d.mem_available = p.totalGlobalMem;
		d.max_buffer_size = p.totalGlobalMem;
		d.name = dup(p.name);
		d.type = ICL_GPU;
		d.vendor = dup("AMD");
		d.version = dup(p.gcnArchName);
		int rt = 0;
		(void)hipRuntimeGetVersion(&rt);
		d.driver_version = dup(std::to_string(rt).c_str());
		d.profile = dup("HIP");
		d.max_compute_units = (cl_uint)p.multiProcessorCount;
		d.max_clock_frequency = (cl_uint)(p.clockRate / 1000);
		d.max_work_item_dimensions = 3;
		d.max_work_item_sizes = (size_t *)malloc(3 * sizeof(size_t));
		d.max_work_item_sizes[0] = (size_t)p.maxThreadsDim[0];
		d.max_work_item_sizes[1] = (size_t)p.maxThreadsDim[1];
		d.max_work_item_sizes[2] = (size_t)p.maxThreadsDim[2];
		d.max_work_group_size = (size_t)p.maxThreadsPerBlock;
		d.local_mem_size = p.sharedMemPerBlock;
		g_devices.push_back(d);
	}
}

extern "C" cl_uint icl_get_num_devices(void) { return (cl_uint)g_devices.size(); }

extern "C" icl_device *icl_get_device(cl_uint id)
{
	if (id >= g_devices.size()) die("icl_get_device", "no such device");
	icl_device &d = g_devices[id];
	HIP_OR_DIE(hipSetDevice(d.hip_device));
	if (!d.stream) {
		hipStream_t s;
		HIP_OR_DIE(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		d.stream = s;
	}
	return &d;
}

extern "C" void icl_release_devices(void)
{
	for (auto &d : g_devices) {
		if (d.stream) {
			(void)hipSetDevice(d.hip_device);
			(void)hipStreamSynchronize(stream_of(&d));
			(void)hipStreamDestroy(stream_of(&d));
		}
		free(d.name);
		free(d.vendor);
		free(d.version);
		free(d.driver_version);
		free(d.profile);
		free(d.max_work_item_sizes);
	}
	g_devices.clear();
}

extern "C" void icl_print_device_short_info(icl_device *dev)
{
	printf("%s | %s | %s | HIP %s | %u CUs\n", dev->name, dev->vendor, dev->version, dev->driver_version,
	       dev->max_compute_units);
}

extern "C" void icl_print_device_infos(icl_device *dev)
{
	icl_print_device_short_info(dev);
	printf("global memory %llu bytes, local memory %llu bytes, max work-group %zu\n",
	       (unsigned long long)dev->mem_size, (unsigned long long)dev->local_mem_size, dev->max_work_group_size);
}

extern "C" void icl_finish(icl_device *dev) { HIP_OR_DIE(hipStreamSynchronize(stream_of(dev))); }

/* ------------------------------------------------------------- buffers */

extern "C" icl_buffer *icl_create_buffer(icl_device *dev, cl_mem_flags flags, size_t size)
{
	(void)flags;
	if (size > dev->mem_available) die("icl_create_buffer", "not enough device memory left");
	icl_buffer *b = (icl_buffer *)malloc(sizeof(icl_buffer));
	HIP_OR_DIE(hipSetDevice(dev->hip_device));
	HIP_OR_DIE(hipMalloc(&b->mem, size ? size : 1));
	b->size = size;
	b->dev = dev;
	dev->mem_available -= size;
	return b;
}

extern "C" void icl_read_buffer_offset(const icl_buffer *buf, cl_bool blocking, size_t offset, size_t size,
                                       void *dst, icl_event *wait_event, icl_event *event)
{
	if (offset + size > buf->size) die("icl_read_buffer", "range exceeds the buffer");
	hipStream_t s = stream_of(buf->dev);
	wait_on(wait_event, s);
	record_start(event, s);
	HIP_OR_DIE(hipMemcpyAsync(dst, (const char *)buf->mem + offset, size, hipMemcpyDeviceToHost, s));
	record_stop(event, s);
	if (blocking) HIP_OR_DIE(hipStreamSynchronize(s));
}

extern "C" void icl_read_buffer(const icl_buffer *buf, cl_bool blocking, size_t size, void *dst,
                                icl_event *wait_event, icl_event *event)
{
	icl_read_buffer_offset(buf, blocking, 0, size, dst, wait_event, event);
}

extern "C" void icl_write_buffer_offset(icl_buffer *buf, cl_bool blocking, size_t offset, size_t size,
                                        const void *src, icl_event *wait_event, icl_event *event)
{
	if (offset + size > buf->size) die("icl_write_buffer", "range exceeds the buffer");
	hipStream_t s = stream_of(buf->dev);
	wait_on(wait_event, s);
	record_start(event, s);
	HIP_OR_DIE(hipMemcpyAsync((char *)buf->mem + offset, src, size, hipMemcpyHostToDevice, s));
	record_stop(event, s);
	if (blocking) HIP_OR_DIE(hipStreamSynchronize(s));
}

extern "C" void icl_write_buffer(icl_buffer *buf, cl_bool blocking, size_t size, const void *src,
                                 icl_event *wait_event, icl_event *event)
{
	icl_write_buffer_offset(buf, blocking, 0, size, src, wait_event, event);
}

extern "C" void icl_copy_buffer(icl_buffer *src, icl_buffer *dst, size_t size, icl_event *wait_event,
                                icl_event *event)
{
	if (size > src->size || size > dst->size) die("icl_copy_buffer", "range exceeds a buffer");
	hipStream_t s = stream_of(src->dev);
	wait_on(wait_event, s);
	record_start(event, s);
	HIP_OR_DIE(hipMemcpyAsync(dst->mem, src->mem, size, hipMemcpyDeviceToDevice, s));
	record_stop(event, s);
}

extern "C" void icl_release_buffer(icl_buffer *buf)
{
	if (!buf) return;
	(void)hipStreamSynchronize(stream_of(buf->dev));
	HIP_OR_DIE(hipFree(buf->mem));
	buf->dev->mem_available += buf->size;
	free(buf);
}

extern "C" void icl_release_buffers(cl_uint num, ...)
{
	va_list ap;
	va_start(ap, num);
	for (cl_uint i = 0; i < num; ++i) icl_release_buffer(va_arg(ap, icl_buffer *));
	va_end(ap);
}

/* ------------------------------------------------------------- kernels */

/* The species table of a build-options string: -DSPECIES=<x> as the
 * reference passes it (src/abms.c:565-568); see icl_create_kernel. */
static void species_of(const char *opts, sampo_species *sp)
{
	sampo_species_default(sp);
	if (!opts) return;
	std::vector<std::string> inc, toks;
	for (const char *p = opts; *p;) {
		while (*p == ' ' || *p == '\t') ++p;
		const char *q = p;
		while (*q && *q != ' ' && *q != '\t') ++q;
		if (q > p) toks.emplace_back(p, q);
		p = q;
	}
	std::string name;
	for (const std::string &t : toks) {
		if (t.rfind("-I", 0) == 0 && t.size() > 2) inc.push_back(t.substr(2));
		if (t.rfind("-DSPECIES=", 0) == 0) name = t.substr(10);
	}
	const std::string dflt = "properties.h";
	if (name.empty() ||
	    (name.size() >= dflt.size() && name.compare(name.size() - dflt.size(), dflt.size(), dflt) == 0))
		return;
	int rc = sampo_species_read(name.c_str(), sp);
	for (size_t i = 0; rc && i < inc.size(); ++i) rc = sampo_species_read((inc[i] + "/" + name).c_str(), sp);
	if (rc) die("icl_create_kernel: cannot read a valid species file", name.c_str());
}

extern "C" icl_kernel *icl_create_kernel(icl_device *dev, const char *file_name, const char *kernel_name,
                                         const char *build_options, icl_create_kernel_flag flag)
{
	(void)flag;
	const char *fb = base_name(file_name ? file_name : "");
	for (int i = 0; i < k_nreg; ++i) {
		if (strcmp(k_registry[i].name, kernel_name) != 0) continue;
		if (*fb && strcmp(fb, k_registry[i].file) != 0) die("icl_create_kernel: kernel not in that file", kernel_name);
		sampo_species sp;
		species_of(build_options, &sp);
		icl_kernel *k = (icl_kernel *)malloc(sizeof(icl_kernel));
		k->id = i;
		k->dev = dev;
		HIP_OR_DIE(hipSetDevice(dev->hip_device));
		HIP_OR_DIE(hipMalloc(&k->species, sizeof(sampo_species)));
		HIP_OR_DIE(hipMemcpy(k->species, &sp, sizeof(sp), hipMemcpyHostToDevice));
		return k;
	}
	die("icl_create_kernel: no such kernel", kernel_name);
	return nullptr;
}

extern "C" void icl_run_kernel(const icl_kernel *kernel, cl_uint work_dim, const size_t *global_work_size,
                               const size_t *local_work_size, icl_event *wait_event, icl_event *event,
                               cl_uint num_args, ...)
{
	const Entry &e = k_registry[kernel->id];
	if (work_dim != 1) die("icl_run_kernel: only 1-D ranges are registered", e.name);
	if (num_args != e.nargs) die("icl_run_kernel: wrong number of arguments for", e.name);
	Arg a[16];
	va_list ap;
	va_start(ap, num_args);
	for (cl_uint i = 0; i < num_args; ++i) {
		a[i].size = va_arg(ap, size_t);
		a[i].value = va_arg(ap, void *);
	}
	va_end(ap);
	size_t global = global_work_size[0];
	size_t local = local_work_size ? local_work_size[0] : 64;
	if (local == 0 || local > 1024 || global % local != 0)
		die("icl_run_kernel: global size must be a multiple of a local size <= 1024 for", e.name);
	if (global == 0) return;
	hipStream_t s = stream_of(kernel->dev);
	wait_on(wait_event, s);
	record_start(event, s);
	e.fn(dim3((unsigned)(global / local)), dim3((unsigned)local), s, a, (const sampo_species *)kernel->species);
	HIP_OR_DIE(hipGetLastError());
	record_stop(event, s);
}

extern "C" void icl_release_kernel(icl_kernel *kernel)
{
	if (!kernel) return;
	(void)hipStreamSynchronize(stream_of(kernel->dev));
	HIP_OR_DIE(hipFree(kernel->species));
	free(kernel);
}

/* ------------------------------------------------------------- events, timers */

extern "C" icl_event *icl_create_event(void) { return (icl_event *)calloc(1, sizeof(icl_event)); }

extern "C" void icl_release_event(icl_event *ev)
{
	if (!ev) return;
	if (ev->start) (void)hipEventDestroy((hipEvent_t)ev->start);
	if (ev->stop) (void)hipEventDestroy((hipEvent_t)ev->stop);
	free(ev);
}

extern "C" void icl_release_events(cl_uint num, ...)
{
	va_list ap;
	va_start(ap, num);
	for (cl_uint i = 0; i < num; ++i) icl_release_event(va_arg(ap, icl_event *));
	va_end(ap);
}

extern "C" void icl_wait_for_events(cl_uint num, ...)
{
	va_list ap;
	va_start(ap, num);
	for (cl_uint i = 0; i < num; ++i) {
		icl_event *ev = va_arg(ap, icl_event *);
		if (ev && ev->stop) HIP_OR_DIE(hipEventSynchronize((hipEvent_t)ev->stop));
	}
	va_end(ap);
}

/* elapsed time between the start and stop of the event's command; HIP
 * events mark execution, so ENQUEUED/SUBMITTED/STARTED all map to start */
extern "C" double icl_profile_event(icl_event *ev, icl_event_flag a, icl_event_flag b, icl_time_flag f)
{
	if (!ev || !ev->stop) return 0.0;
	HIP_OR_DIE(hipEventSynchronize((hipEvent_t)ev->stop));
	hipEvent_t x = a == ICL_FINISHED ? (hipEvent_t)ev->stop : (hipEvent_t)ev->start;
	hipEvent_t y = b == ICL_FINISHED ? (hipEvent_t)ev->stop : (hipEvent_t)ev->start;
	float ms = 0.0f;
	HIP_OR_DIE(hipEventElapsedTime(&ms, x, y));
	return scale((double)ms * 1e6, f);
}

extern "C" icl_timer *icl_init_timer(icl_time_flag f)
{
	icl_timer *t = (icl_timer *)calloc(1, sizeof(icl_timer));
	t->time_flag = f;
	return t;
}

extern "C" void icl_start_timer(icl_timer *t) { t->start = now_ns(); }

extern "C" void icl_restart_timer(icl_timer *t)
{
	t->current_time = 0.0;
	t->clocks = 0;
	t->start = now_ns();
}

extern "C" double icl_stop_timer(icl_timer *t)
{
	time_int d = now_ns() - t->start;
	t->clocks += d;
	t->current_time += scale((double)d, t->time_flag);
	return t->current_time;
}

extern "C" void icl_release_timer(icl_timer *t) { free(t); }

/* ------------------------------------------------------------- scan module */

extern "C" void segmented_scan_init(size_t wx, cl_uint max_n, icl_device *dev, const char *build_options,
                                    icl_create_kernel_flag flag)
{
	(void)wx;
	(void)build_options;
	(void)flag;
	segmented_scan_release();
	g_scan.dev = dev;
	g_scan.max_n = max_n;
	size_t tiles = (size_t)max_n / SCAN_TILE + 2;
	HIP_OR_DIE(hipSetDevice(dev->hip_device));
	HIP_OR_DIE(hipMalloc(&g_scan.ctx, sizeof(IclScanCtx)));
	HIP_OR_DIE(hipMalloc(&g_scan.status, tiles * sizeof(uint64_t)));
	HIP_OR_DIE(hipMemset(g_scan.ctx, 0, sizeof(IclScanCtx)));
	HIP_OR_DIE(hipMemset(g_scan.status, 0, tiles * sizeof(uint64_t)));
}

extern "C" void segmented_scan_release(void)
{
	if (g_scan.dev) {
		(void)hipStreamSynchronize(stream_of(g_scan.dev));
		(void)hipFree(g_scan.ctx);
		(void)hipFree(g_scan.status);
	}
	g_scan = ScanState();
}

extern "C" void scan(icl_buffer *data, icl_buffer *flag, cl_uint n)
{
	if (!g_scan.dev) die("scan", "segmented_scan_init was not called");
	if (n > g_scan.max_n + 1u) die("scan", "n exceeds the size given to segmented_scan_init");
	if ((size_t)n * 4u > data->size || (size_t)n * 4u > flag->size) die("scan", "n exceeds a buffer");
	if (n == 0) return;
	hipStream_t s = stream_of(g_scan.dev);
	hipLaunchKernelGGL(scan_reset, dim3(1), dim3(1), 0, s, g_scan.ctx);
	hipLaunchKernelGGL(scan_by_key, dim3((n + SCAN_TILE - 1) / SCAN_TILE), dim3(SCAN_WG), 0, s, (int32_t *)data->mem,
	                   (const uint32_t *)flag->mem, n, g_scan.ctx, g_scan.status);
	HIP_OR_DIE(hipGetLastError());
}
