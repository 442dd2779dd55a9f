/*
 * grid_wait_probe.hip -- what a grid-wide wait costs inside one launch on
 * this GPU, against a kernel boundary inside a HIP graph.
 *
 * A persistent launch per simulated day would replace the engine's two
 * kernels per iteration (the fused pass and k_prep) by one cooperative launch
 * with grid-wide waits between the phases.  This probe times such a wait,
 * four ways, for G work-groups of 512 threads (all resident: cooperative
 * launch), each having written `dirty` words of its own before the wait:
 *   counter       each work-group adds 1 to one device-scope counter
 *                 (release); work-group 0 waits for G arrivals, then
 *                 publishes the epoch; all wait for it (acquire);
 *   flags         each work-group stores the epoch to a flag of its own
 *                 (release); work-group 0's lanes wait for all flags;
 *   per-XCD       work-groups add to one of 8 counters by blockIdx % 8;
 *                 work-group 0 waits for the 8;
 *   last-arriver  the work-group whose add completes the count publishes
 *                 the epoch.
 * and, for comparison, a graph of 200 launches of a kernel of the same shape
 * doing the same writes: its time per launch is a kernel boundary.
 *
 * Every wait is bounded: after 2^22 polls a work-group sets a time-out flag
 * and every later wait returns at once, so the launch always ends; the
 * program then reports the variant as timed out.
 *
 * Build (by sampo_amd/build.py):  hipcc -O3 --offload-arch=gfx950 ... -o sampo_amd/grid_wait_probe
 * Run:                            ./sampo_amd/grid_wait_probe
 */
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                                       \
	do {                                                                                            \
		hipError_t e_ = (x);                                                                        \
		if (e_ != hipSuccess) {                                                                     \
			fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);          \
			exit(1);                                                                                \
		}                                                                                           \
	} while (0)

#define MAX_G 1024u
#define LINE 64u /* words: flags and counters on lines of their own */
#define POLLS (1u << 22)

struct Sync {
	uint32_t count[LINE];
	uint32_t go[LINE];
	uint32_t xcd[8][LINE];
	uint32_t timeout[LINE];
	uint32_t flags[MAX_G * LINE / 2u]; /* flag of work-group g at g * 32 */
};

__device__ static inline uint32_t ld_agent(uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

/* wait until *p >= v, at most POLLS polls (then flag the time-out) */
__device__ static void spin_ge(uint32_t *p, uint32_t v, Sync *b)
{
	for (uint32_t i = 0; i < POLLS; ++i) {
		if (ld_agent(p) >= v) return;
		if (ld_agent(b->timeout)) return;
		__builtin_amdgcn_s_sleep(1);
	}
	__hip_atomic_store(b->timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int MODE>
__device__ static void grid_wait(Sync *b, uint32_t e)
{
	const uint32_t G = gridDim.x, tid = threadIdx.x;
	__syncthreads();
	if (MODE == 1) { /* flags */
		if (tid == 0) __hip_atomic_store(&b->flags[blockIdx.x * 32u], e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
		if (blockIdx.x == 0) {
			for (uint32_t j = tid; j < G; j += blockDim.x) spin_ge(&b->flags[j * 32u], e, b);
			__syncthreads();
			if (tid == 0) __hip_atomic_store(b->go, e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
		}
		if (tid == 0) {
			spin_ge(b->go, e, b);
			__atomic_thread_fence(__ATOMIC_ACQUIRE);
		}
	} else if (tid == 0) {
		if (MODE == 0) { /* counter */
			__hip_atomic_fetch_add(b->count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
			if (blockIdx.x == 0) {
				spin_ge(b->count, G * e, b);
				__hip_atomic_store(b->go, e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
			}
		} else if (MODE == 2) { /* per-XCD counters */
			__hip_atomic_fetch_add(b->xcd[blockIdx.x & 7u], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
			if (blockIdx.x == 0) {
				for (uint32_t q = 0; q < 8u; ++q) spin_ge(b->xcd[q], (G / 8u + (q < G % 8u ? 1u : 0u)) * e, b);
				__hip_atomic_store(b->go, e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
			}
		} else { /* last arriver */
			const uint32_t old = __hip_atomic_fetch_add(b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
			if (old == G * e - 1u) __hip_atomic_store(b->go, e, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
		}
		spin_ge(b->go, e, b);
		__atomic_thread_fence(__ATOMIC_ACQUIRE);
	}
	__syncthreads();
}

template <int MODE>
__global__ void __launch_bounds__(512) k_waits(Sync *b, uint32_t iters, uint32_t dirty, uint32_t *buf,
                                               unsigned long long *t)
{
	const unsigned long long t0 = wall_clock64();
	for (uint32_t e = 1; e <= iters; ++e) {
		for (uint32_t q = threadIdx.x; q < dirty; q += blockDim.x) buf[(size_t)blockIdx.x * dirty + q] = e + q;
		grid_wait<MODE>(b, e);
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		t[0] = t0;
		t[1] = wall_clock64();
	}
}

__global__ void __launch_bounds__(512) k_step(uint32_t dirty, uint32_t *buf, uint32_t e)
{
	for (uint32_t q = threadIdx.x; q < dirty; q += blockDim.x) buf[(size_t)blockIdx.x * dirty + q] = e + q;
}

/* ns per wait (wall clock at 100 MHz), or -1 after a time-out */
template <int MODE>
static double time_waits(uint32_t G, uint32_t iters, uint32_t dirty, uint32_t *buf, Sync *b, unsigned long long *t)
{
	CK(hipMemset(b, 0, sizeof(Sync)));
	void *args[] = {&b, &iters, &dirty, &buf, &t};
	CK(hipLaunchCooperativeKernel((const void *)k_waits<MODE>, dim3(G), dim3(512), args, 0, 0));
	CK(hipDeviceSynchronize());
	unsigned long long h[2];
	uint32_t to;
	CK(hipMemcpy(h, t, sizeof(h), hipMemcpyDeviceToHost));
	CK(hipMemcpy(&to, b->timeout, sizeof(to), hipMemcpyDeviceToHost));
	return to ? -1.0 : (double)(h[1] - h[0]) * 10.0 / iters;
}

/* ns per launch of k_step in a graph of 200 launches, replayed 10 times */
static double time_launches(uint32_t G, uint32_t dirty, uint32_t *buf)
{
	hipStream_t s;
	CK(hipStreamCreate(&s));
	hipGraph_t g;
	hipGraphExec_t ge;
	CK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
	for (uint32_t e = 1; e <= 200u; ++e) k_step<<<G, 512, 0, s>>>(dirty, buf, e);
	CK(hipStreamEndCapture(s, &g));
	CK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
	CK(hipGraphLaunch(ge, s));
	CK(hipStreamSynchronize(s));
	hipEvent_t a, z;
	CK(hipEventCreate(&a));
	CK(hipEventCreate(&z));
	CK(hipEventRecord(a, s));
	for (int r = 0; r < 10; ++r) CK(hipGraphLaunch(ge, s));
	CK(hipEventRecord(z, s));
	CK(hipEventSynchronize(z));
	float ms;
	CK(hipEventElapsedTime(&ms, a, z));
	CK(hipEventDestroy(a));
	CK(hipEventDestroy(z));
	CK(hipGraphExecDestroy(ge));
	CK(hipGraphDestroy(g));
	CK(hipStreamDestroy(s));
	return ms * 1e6 / 2000.0;
}

int main()
{
	int coop = 0, ncu = 0, occ = 0;
	CK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, 0));
	CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
	CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)k_waits<0>, 512, 0));
	if (!coop) {
		fprintf(stderr, "no cooperative launch on this device\n");
		return 1;
	}
	printf("CUs %d, resident work-groups of 512 threads per CU %d\n", ncu, occ);
	const uint32_t maxdirty = 4096u, iters = 2000u;
	Sync *b;
	uint32_t *buf;
	unsigned long long *t;
	CK(hipMalloc(&b, sizeof(Sync)));
	CK(hipMalloc(&buf, (size_t)MAX_G * maxdirty * 4u));
	CK(hipMalloc(&t, 16));
	printf("%6s %8s | ns per wait: %8s %8s %8s %12s | ns per kernel in a graph\n", "G", "words/WG", "counter", "flags",
	       "per-XCD", "last-arriver");
	const uint32_t Gs[] = {64u, 256u, 512u, 768u}, Ds[] = {0u, 1024u, 4096u};
	for (uint32_t G : Gs) {
		if (G > (uint32_t)(occ * ncu) || G > MAX_G) continue;
		for (uint32_t D : Ds) {
			const double m0 = time_waits<0>(G, iters, D, buf, b, t), m1 = time_waits<1>(G, iters, D, buf, b, t);
			const double m2 = time_waits<2>(G, iters, D, buf, b, t), m3 = time_waits<3>(G, iters, D, buf, b, t);
			printf("%6u %8u | %21.0f %8.0f %8.0f %12.0f | %8.0f\n", G, D, m0, m1, m2, m3, time_launches(G, D, buf));
			fflush(stdout);
		}
	}
	CK(hipFree(b));
	CK(hipFree(buf));
	CK(hipFree(t));
	return 0;
}
