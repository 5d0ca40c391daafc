// hip_runtime.h -- TEST INFRASTRUCTURE: the HIP names the step library uses,
// implemented on the CPU, so that the library's own device sources run in this
// container, where there is no GPU (tests/emu/build_emu.py compiles them with
// g++ against this header; tests/emu/emu_runtime.cpp runs the kernels).
//
// Execution model.  A kernel launch runs every workgroup; each thread of a
// workgroup is a fiber on one host thread, switched only where the device
// code synchronises: __syncthreads (every live fiber of the workgroup), a
// cross-lane operation (__ballot, __shfl*: the live lanes of its 64-lane
// wave), or a bounded wait (__builtin_amdgcn_s_sleep).  __shared__ variables
// are thread_local statics, one copy per host thread and so per workgroup.
// Ordinary launches spread the workgroups over a few host threads; a
// cooperative launch gives every workgroup its own, so that grid-wide waits
// on atomics make progress.  Device memory is host memory; streams, events
// and copies are synchronous.
//
// Arithmetic is the device code's own, compiled for x86-64 with
// -ffp-contract=off: IEEE binary64 / binary32 operations in source order, as
// the gfx950 build performs them.
#pragma once

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <tuple>
#include <type_traits>
#include <utility>

// ---- qualifiers ----
#define __host__
#define __device__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
// (build_emu.py rewrites `__shared__` to `static thread_local` and `extern __shared__ T x[]`
// to a pointer to the launch's dynamic shared memory)

// ---- vector types ----
template <class T, unsigned N> struct HIP_vector_type;
template <class T> struct alignas(2 * sizeof(T)) HIP_vector_type<T, 2> { T x, y; };
template <class T> struct HIP_vector_type<T, 3> { T x, y, z; };
template <class T> struct alignas(4 * sizeof(T)) HIP_vector_type<T, 4> { T x, y, z, w; };
typedef HIP_vector_type<int, 2> int2;
typedef HIP_vector_type<int, 4> int4;
typedef HIP_vector_type<unsigned int, 2> uint2;
typedef HIP_vector_type<unsigned int, 4> uint4;
typedef HIP_vector_type<double, 2> double2;
typedef HIP_vector_type<double, 4> double4;
typedef HIP_vector_type<float, 2> float2;
typedef HIP_vector_type<float, 4> float4;
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }
inline uint2 make_uint2(unsigned x, unsigned y) { return uint2{x, y}; }
inline uint4 make_uint4(unsigned x, unsigned y, unsigned z, unsigned w) { return uint4{x, y, z, w}; }
inline double2 make_double2(double x, double y) { return double2{x, y}; }
inline float2 make_float2(float x, float y) { return float2{x, y}; }

struct dim3 {
	unsigned int x, y, z;
	dim3(unsigned int a = 1, unsigned int b = 1, unsigned int c = 1) : x(a), y(b), z(c) {}
};
struct uint3_ { unsigned int x, y, z; };

// ---- the emulator's state (emu_runtime.cpp) ----
namespace emu {
struct Block;
struct Lane {
	void *sp;                 // saved stack pointer while switched out
	uint3_ tid;
	int lane;                 // lane within its 64-lane wave
	int wave;
	int state;                // RUNNABLE, AT_BARRIER, AT_WAVE, EXITED
	Block *block;
	char *stack;
};
struct Block {
	uint3_ bid, bdim, gdim;
	void *dyn;                // dynamic shared memory of the launch
};
extern thread_local Lane *cur;

void *dyn_shared();
void sync_block();
void sleep_lane();
uint64_t wave_ballot(bool pred);
uint64_t wave_exchange(uint64_t bits, int src_kind, int arg, uint64_t mask, bool subset);
void wave_sync();
enum { X_IDX = 0, X_XOR = 1, X_UP = 2, X_DOWN = 3 };

template <class T> inline uint64_t to_bits(T v)
{
	static_assert(sizeof(T) <= 8 && std::is_trivially_copyable<T>::value, "lane value of at most 8 bytes");
	uint64_t b = 0;
	memcpy(&b, &v, sizeof(T));
	return b;
}
template <class T> inline T from_bits(uint64_t b)
{
	T v;
	memcpy(&v, &b, sizeof(T));
	return v;
}
} // namespace emu

#define threadIdx (emu::cur->tid)
#define blockIdx (emu::cur->block->bid)
#define blockDim (emu::cur->block->bdim)
#define gridDim (emu::cur->block->gdim)
static constexpr int warpSize = 64;

inline void __syncthreads() { emu::sync_block(); }
inline unsigned int __lane_id() { return (unsigned int)emu::cur->lane; }
inline unsigned long long __ballot(int pred) { return emu::wave_ballot(pred != 0); }
template <class T> inline T __shfl(T v, int src, int width = 64)
{
	(void)width;
	return emu::from_bits<T>(emu::wave_exchange(emu::to_bits(v), emu::X_IDX, src, 0, false));
}
template <class T> inline T __shfl_xor(T v, int m, int width = 64)
{
	(void)width;
	return emu::from_bits<T>(emu::wave_exchange(emu::to_bits(v), emu::X_XOR, m, 0, false));
}
template <class T> inline T __shfl_up(T v, unsigned int d, int width = 64)
{
	(void)width;
	return emu::from_bits<T>(emu::wave_exchange(emu::to_bits(v), emu::X_UP, (int)d, 0, false));
}
template <class T> inline T __shfl_down(T v, unsigned int d, int width = 64)
{
	(void)width;
	return emu::from_bits<T>(emu::wave_exchange(emu::to_bits(v), emu::X_DOWN, (int)d, 0, false));
}
// the lanes of a wave that call this together, in code whose other lanes are
// elsewhere (build_emu.py puts it where the device code takes the calling lanes
// with __ballot(1)): each calling lane on its own
inline unsigned long long emu_self_mask() { return 1ull << emu::cur->lane; }
// __shfl among the lanes of `mask` only
template <class T> inline T emu_shfl_in(T v, int src, unsigned long long mask)
{
	return emu::from_bits<T>(emu::wave_exchange(emu::to_bits(v), emu::X_IDX, src, mask, true));
}
// every live lane of the wave is here (the reconvergence the device code relies on)
inline void emu_wave_sync() { emu::wave_sync(); }

inline long long __double_as_longlong(double x)
{
	long long r;
	memcpy(&r, &x, 8);
	return r;
}
inline double __longlong_as_double(long long x)
{
	double r;
	memcpy(&r, &x, 8);
	return r;
}
inline int __float_as_int(float x)
{
	int r;
	memcpy(&r, &x, 4);
	return r;
}
inline float __int_as_float(int x)
{
	float r;
	memcpy(&r, &x, 4);
	return r;
}

inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __popc(unsigned int x) { return __builtin_popcount(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffs(int x) { return __builtin_ffs(x); }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((unsigned long long)x); }

// ---- memory model: every fiber of a host thread is sequential; atomics are the host's ----
#define __ATOMIC_WAVEFRONT_SCOPE 0
#define __HIP_MEMORY_SCOPE_SINGLETHREAD 1
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(__ATOMIC_SEQ_CST)
#define __builtin_amdgcn_s_sleep(n) emu::sleep_lane()
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add((p), (v), __ATOMIC_SEQ_CST)

template <class T> inline T atomicAdd(T *p, T v)
{
	if constexpr (std::is_floating_point<T>::value) {
		T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
		while (!__atomic_compare_exchange_n(p, &old, old + v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
		return old;
	} else {
		return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
	}
}
template <class T> inline T atomicSub(T *p, T v) { return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicOr(T *p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicAnd(T *p, T v) { return __atomic_fetch_and(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicExch(T *p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicCAS(T *p, T cmp, T v)
{
	__atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
	return cmp;
}
template <class T> inline T atomicMin(T *p, T v)
{
	T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
	while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
	return old;
}
template <class T> inline T atomicMax(T *p, T v)
{
	T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
	while (v > old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
	return old;
}

// device min / max / abs of HIP
template <class A, class B> inline typename std::common_type<A, B>::type min(A a, B b)
{
	typedef typename std::common_type<A, B>::type C;
	return (C)b < (C)a ? (C)b : (C)a;
}
template <class A, class B> inline typename std::common_type<A, B>::type max(A a, B b)
{
	typedef typename std::common_type<A, B>::type C;
	return (C)a < (C)b ? (C)b : (C)a;
}

// ---- runtime API ----
typedef int hipError_t;
enum {
	hipSuccess = 0,
	hipErrorInvalidValue = 1,
	hipErrorOutOfMemory = 2,
	hipErrorNotReady = 600,
	hipErrorNotSupported = 801,
	hipErrorCooperativeLaunchTooLarge = 720,
};
typedef struct emu_stream *hipStream_t;
typedef struct emu_event *hipEvent_t;
typedef struct emu_graph *hipGraph_t;
typedef struct emu_graph_exec *hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum { hipStreamDefault = 0, hipStreamNonBlocking = 1 };
enum { hipEventDefault = 0, hipEventDisableTiming = 2 };
enum { hipHostMallocDefault = 0 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipDeviceAttribute_t { hipDeviceAttributeMaxSharedMemoryPerBlock = 1, hipDeviceAttributeMultiprocessorCount = 2 };
struct hipDeviceProp_t {
	char name[256];
	int multiProcessorCount;
	size_t sharedMemPerBlock;
	int warpSize;
	int maxThreadsPerBlock;
	size_t totalGlobalMem;
	char gcnArchName[256];
};
struct hipFuncAttributes {
	size_t sharedSizeBytes;
	int maxThreadsPerBlock;
	int numRegs;
};

hipError_t hipMalloc(void **p, size_t bytes);
template <class T> inline hipError_t hipMalloc(T **p, size_t bytes) { return hipMalloc((void **)p, bytes); }
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags = 0);
template <class T> inline hipError_t hipHostMalloc(T **p, size_t bytes, unsigned flags = 0)
{
	return hipHostMalloc((void **)p, bytes, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s = nullptr);
hipError_t hipMemset(void *dst, int v, size_t bytes);
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t s = nullptr);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipGetDevice(int *d);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int d);
hipError_t hipFuncGetAttributes(hipFuncAttributes *a, const void *f);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *f, int threads, size_t dyn);

// ---- launches ----
namespace emu {
struct Launch {
	uint3_ grid, block;
	size_t shm;
	bool cooperative;
};
// runs body() as every thread of every workgroup of the launch
void run(const Launch &L, void (*body)(void *), void *ctx);
hipError_t set_launch_error(hipError_t e);

inline uint3_ to3(dim3 d) { return uint3_{d.x, d.y, d.z}; }
inline uint3_ to3(unsigned long long n) { return uint3_{(unsigned)n, 1u, 1u}; }
inline uint3_ to3(long long n) { return uint3_{(unsigned)n, 1u, 1u}; }
inline uint3_ to3(unsigned n) { return uint3_{n, 1u, 1u}; }
inline uint3_ to3(int n) { return uint3_{(unsigned)n, 1u, 1u}; }
inline uint3_ to3(long n) { return uint3_{(unsigned)n, 1u, 1u}; }
inline uint3_ to3(unsigned long n) { return uint3_{(unsigned)n, 1u, 1u}; }

bool refuse(const Launch &L);   // EMU_REFUSE_DYNAMIC_LDS: cooperative launches with dynamic LDS fail

template <class... P> struct Launcher {
	void (*k)(P...);
	Launch L;
	template <class... A> hipError_t operator()(A &&...a) const
	{
		if (refuse(L)) return set_launch_error(hipErrorCooperativeLaunchTooLarge);
		typedef std::tuple<typename std::decay<P>::type...> Args;
		Args args(std::forward<A>(a)...);
		struct Ctx { void (*k)(P...); Args *args; } ctx{k, &args};
		run(L, [](void *c) {
			Ctx *x = (Ctx *)c;
			std::apply(x->k, *x->args);
		}, &ctx);
		return hipSuccess;
	}
};
template <class... P, class G, class B>
inline Launcher<P...> launch(void (*k)(P...), G grid, B block, size_t shm = 0, hipStream_t = nullptr)
{
	return Launcher<P...>{k, Launch{to3(grid), to3(block), shm, false}};
}
template <class... P, class G, class B>
inline Launcher<P...> launch_cooperative(void (*k)(P...), G grid, B block, size_t shm = 0, hipStream_t = nullptr)
{
	return Launcher<P...>{k, Launch{to3(grid), to3(block), shm, true}};
}
} // namespace emu
