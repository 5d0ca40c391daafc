// emu_runtime.cpp -- TEST INFRASTRUCTURE: the CPU side of tests/emu/hip/hip_runtime.h.
// Fibers (one per device thread) switched by a small x86-64 context switch,
// workgroups on host threads, the HIP runtime calls the step library makes.
#include "hip/hip_runtime.h"

#include <sched.h>
#include <sys/mman.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

extern "C" void emu_switch(void **save_sp, void *new_sp);
asm(R"(
	.text
	.globl emu_switch
	.type emu_switch, @function
emu_switch:
	pushq %rbp
	pushq %rbx
	pushq %r12
	pushq %r13
	pushq %r14
	pushq %r15
	subq $8, %rsp
	stmxcsr (%rsp)
	fnstcw 4(%rsp)
	movq %rsp, (%rdi)
	movq %rsi, %rsp
	ldmxcsr (%rsp)
	fldcw 4(%rsp)
	addq $8, %rsp
	popq %r15
	popq %r14
	popq %r13
	popq %r12
	popq %rbx
	popq %rbp
	ret
	.size emu_switch, .-emu_switch
)");

namespace emu {

enum { RUNNABLE = 0, AT_BARRIER = 1, AT_WAVE = 2, EXITED = 3 };
enum { W_BALLOT = 1, W_EXCHANGE = 2, W_SYNC = 3 };

thread_local Lane *cur = nullptr;

namespace {

constexpr size_t STACK_BYTES = 256 * 1024;

struct WaveState {
	uint64_t arrived = 0;
	int kind[64];
	uint64_t val[64], res[64], mask[64];
	int xkind[64], arg[64];
	bool subset[64];
};

struct BlockRun {
	Block b;
	std::vector<Lane> lanes;
	std::vector<WaveState> waves;
	void *sched_sp = nullptr;
	void (*body)(void *) = nullptr;
	void *ctx = nullptr;
	int live = 0;
	int at_barrier = 0;
};

thread_local BlockRun *cur_run = nullptr;
thread_local std::vector<char *> stack_pool;
thread_local std::vector<char> dyn_buf;

char *get_stack(size_t i)
{
	while (stack_pool.size() <= i) {
		void *p = mmap(nullptr, STACK_BYTES, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
		if (p == MAP_FAILED) {
			fprintf(stderr, "emu: out of memory for fiber stacks\n");
			abort();
		}
		stack_pool.push_back((char *)p);
	}
	return stack_pool[i];
}

void yield_to_scheduler()
{
	Lane *l = cur;
	emu_switch(&l->sp, cur_run->sched_sp);
}

void lane_entry()
{
	BlockRun *r = cur_run;
	r->body(r->ctx);
	Lane *l = cur;
	l->state = EXITED;
	r->live--;
	emu_switch(&l->sp, r->sched_sp);
	abort();   // an exited fiber is never resumed
}

void init_lane(Lane &l)
{
	char *top = l.stack + STACK_BYTES;
	uintptr_t *sp = (uintptr_t *)top;
	*--sp = 0;                               // after ret: rsp = top - 8, as after a call
	*--sp = (uintptr_t)&lane_entry;          // return address of emu_switch
	for (int i = 0; i < 6; i++) *--sp = 0;   // rbp rbx r12 r13 r14 r15
	*--sp = 0;
	uint32_t mxcsr = 0x1F80;
	uint16_t fpucw = 0x037F;
	memcpy((char *)sp, &mxcsr, 4);
	memcpy((char *)sp + 4, &fpucw, 2);
	l.sp = sp;
	l.state = RUNNABLE;
}

// the wave ops whose lanes have all arrived: compute every lane's result
bool resolve_waves(BlockRun &r)
{
	bool any = false;
	const int nl = (int)r.lanes.size();
	for (size_t w = 0; w < r.waves.size(); w++) {
		WaveState &ws = r.waves[w];
		if (!ws.arrived) continue;
		uint64_t live = 0;
		for (int i = 0; i < 64; i++) {
			const int t = (int)w * 64 + i;
			if (t < nl && r.lanes[t].state != EXITED) live |= 1ull << i;
		}
		const int first = __builtin_ctzll(ws.arrived);
		const int kind = ws.kind[first];
		uint64_t need = live;
		if (kind == W_EXCHANGE && ws.subset[first]) need = ws.mask[first] & live;
		if ((ws.arrived & need) != need) continue;
		// every participant has arrived; all must be at the same kind of operation
		uint64_t part = ws.arrived & need;
		for (int i = 0; i < 64; i++)
			if ((part >> i & 1) && ws.kind[i] != kind) {
				fprintf(stderr, "emu: lanes of wave %zu of block %u meet different cross-lane operations\n", w,
				        r.b.bid.x);
				abort();
			}
		if (kind == W_BALLOT) {
			uint64_t m = 0;
			for (int i = 0; i < 64; i++)
				if ((part >> i & 1) && ws.val[i]) m |= 1ull << i;
			for (int i = 0; i < 64; i++)
				if (part >> i & 1) ws.res[i] = m;
		} else if (kind == W_EXCHANGE) {
			for (int i = 0; i < 64; i++) {
				if (!(part >> i & 1)) continue;
				int src = i;
				const int a = ws.arg[i];
				switch (ws.xkind[i]) {
				case X_IDX: src = a & 63; break;
				case X_XOR: src = i ^ a; break;
				case X_UP: src = i - a; break;
				case X_DOWN: src = i + a; break;
				}
				ws.res[i] = (src >= 0 && src < 64 && (part >> src & 1)) ? ws.val[src] : ws.val[i];
			}
		}
		for (int i = 0; i < 64; i++)
			if (part >> i & 1) r.lanes[w * 64 + i].state = RUNNABLE;
		ws.arrived &= ~part;
		any = true;
	}
	return any;
}

void run_block(BlockRun &r)
{
	BlockRun *saved_run = cur_run;
	Lane *saved_lane = cur;
	cur_run = &r;
	const int n = (int)r.lanes.size();
	for (int i = 0; i < n; i++) {
		Lane &l = r.lanes[i];
		l.stack = get_stack(i);
		l.block = &r.b;
		init_lane(l);
	}
	r.live = n;
	r.at_barrier = 0;
	while (r.live > 0) {
		bool ran = false;
		for (int i = 0; i < n; i++) {
			Lane &l = r.lanes[i];
			if (l.state != RUNNABLE) continue;
			cur = &l;
			emu_switch(&r.sched_sp, l.sp);
			ran = true;
		}
		if (r.live == 0) break;
		bool released = false;
		if (r.at_barrier > 0 && r.at_barrier == r.live) {
			for (int i = 0; i < n; i++)
				if (r.lanes[i].state == AT_BARRIER) r.lanes[i].state = RUNNABLE;
			r.at_barrier = 0;
			released = true;
		}
		if (resolve_waves(r)) released = true;
		if (!ran && !released) {
			int counts[4] = {0, 0, 0, 0};
			for (int i = 0; i < n; i++) counts[r.lanes[i].state]++;
			fprintf(stderr, "emu: block %u stuck: %d at a barrier, %d at a cross-lane operation, %d exited\n",
			        r.b.bid.x, counts[AT_BARRIER], counts[AT_WAVE], counts[EXITED]);
			abort();
		}
	}
	cur_run = saved_run;
	cur = saved_lane;
}

void setup_block(BlockRun &r, const Launch &L, uint3_ bid, void (*body)(void *), void *ctx)
{
	const int n = (int)(L.block.x * L.block.y * L.block.z);
	r.b.bid = bid;
	r.b.bdim = L.block;
	r.b.gdim = L.grid;
	r.body = body;
	r.ctx = ctx;
	r.lanes.resize(n);
	r.waves.assign((n + 63) / 64, WaveState());
	for (int i = 0; i < n; i++) {
		Lane &l = r.lanes[i];
		l.tid = uint3_{(unsigned)(i % L.block.x), (unsigned)((i / L.block.x) % L.block.y),
		               (unsigned)(i / (L.block.x * L.block.y))};
		l.lane = i % 64;
		l.wave = i / 64;
	}
}

void run_on_this_thread(const Launch &L, uint3_ bid, void (*body)(void *), void *ctx)
{
	if (dyn_buf.size() < L.shm + 64) dyn_buf.resize(L.shm + 64);
	BlockRun r;
	setup_block(r, L, bid, body, ctx);
	r.b.dyn = (void *)(((uintptr_t)dyn_buf.data() + 63) & ~(uintptr_t)63);
	run_block(r);
}

hipError_t last_error = hipSuccess;

} // namespace

void *dyn_shared() { return cur->block->dyn; }

void sync_block()
{
	cur->state = AT_BARRIER;
	cur_run->at_barrier++;
	yield_to_scheduler();
}

void sleep_lane()
{
	sched_yield();
	yield_to_scheduler();
}

static uint64_t wave_op(int kind, uint64_t bits, int xkind, int arg, uint64_t mask, bool subset)
{
	Lane *l = cur;
	WaveState &ws = cur_run->waves[l->wave];
	const int i = l->lane;
	ws.kind[i] = kind;
	ws.val[i] = bits;
	ws.xkind[i] = xkind;
	ws.arg[i] = arg;
	ws.mask[i] = mask;
	ws.subset[i] = subset;
	if (subset && mask == (1ull << i)) return bits;   // a wave of one lane
	ws.arrived |= 1ull << i;
	l->state = AT_WAVE;
	yield_to_scheduler();
	return ws.res[i];
}

uint64_t wave_ballot(bool pred) { return wave_op(W_BALLOT, pred ? 1 : 0, 0, 0, 0, false); }
uint64_t wave_exchange(uint64_t bits, int xkind, int arg, uint64_t mask, bool subset)
{
	return wave_op(W_EXCHANGE, bits, xkind, arg, mask, subset);
}
void wave_sync() { (void)wave_op(W_SYNC, 0, 0, 0, 0, false); }

// Host threads that live as long as the process, so that their fiber stacks
// and shared-memory copies are made once.  run(k, n, f): thread t < n calls f(t).
class Pool {
public:
	void run(int n, const std::function<void(int)> &f)
	{
		std::unique_lock<std::mutex> lk(m);
		while ((int)threads.size() < n) {
			const int id = (int)threads.size();
			threads.emplace_back([this, id]() { loop(id); });
			threads.back().detach();
		}
		job = &f;
		want = n;
		left = n;
		gen++;
		cv.notify_all();
		done.wait(lk, [&]() { return left == 0; });
		job = nullptr;
	}

private:
	void loop(int id)
	{
		unsigned long seen = 0;
		std::unique_lock<std::mutex> lk(m);
		for (;;) {
			cv.wait(lk, [&]() { return gen != seen; });
			seen = gen;
			if (id >= want) continue;
			const std::function<void(int)> *f = job;
			lk.unlock();
			(*f)(id);
			lk.lock();
			if (--left == 0) done.notify_all();
		}
	}
	std::mutex m;
	std::condition_variable cv, done;
	std::vector<std::thread> threads;
	const std::function<void(int)> *job = nullptr;
	int want = 0, left = 0;
	unsigned long gen = 0;
};

Pool &pool()
{
	static Pool *p = new Pool();   // never destroyed: its threads outlive static destruction
	return *p;
}

std::mutex launch_mutex;   // one launch at a time, as on one HIP stream

void run(const Launch &L, void (*body)(void *), void *ctx)
{
	const long nblocks = (long)L.grid.x * L.grid.y * L.grid.z;
	if (nblocks <= 0 || L.block.x * L.block.y * L.block.z == 0) return;
	std::lock_guard<std::mutex> g(launch_mutex);
	auto bid_of = [&](long k) {
		return uint3_{(unsigned)(k % L.grid.x), (unsigned)((k / L.grid.x) % L.grid.y),
		              (unsigned)(k / ((long)L.grid.x * L.grid.y))};
	};
	if (L.cooperative) {
		// every workgroup on a host thread of its own: grid-wide waits need them all running
		pool().run((int)nblocks, [&](int k) { run_on_this_thread(L, bid_of(k), body, ctx); });
		return;
	}
	const int workers = (int)std::min<long>(nblocks, std::max(1u, std::thread::hardware_concurrency()));
	std::atomic<long> next(0);
	pool().run(workers, [&](int) {
		for (long k = next++; k < nblocks; k = next++) run_on_this_thread(L, bid_of(k), body, ctx);
	});
}

hipError_t set_launch_error(hipError_t e) { return last_error = e; }

bool refuse(const Launch &L)
{
	static const bool on = getenv("EMU_REFUSE_DYNAMIC_LDS") != nullptr;
	return on && L.cooperative && L.shm > 0;
}

} // namespace emu

// ---- runtime API ----
using emu::last_error;

static hipError_t ok() { return hipSuccess; }

// allocation failures on demand: the n-th device allocation from now fails (0: none)
static std::atomic<long> malloc_countdown(0);
extern "C" void emu_fail_malloc_after(long n) { malloc_countdown = n; }
extern "C" long emu_malloc_countdown() { return malloc_countdown; }

hipError_t hipMalloc(void **p, size_t bytes)
{
	if (!p) return hipErrorInvalidValue;
	*p = nullptr;
	if (malloc_countdown > 0 && --malloc_countdown == 0) return hipErrorOutOfMemory;
	if (bytes == 0) return ok();
	void *q = nullptr;
#ifdef EMU_EXACT_ALLOC
	const size_t size = bytes;   // the AddressSanitizer build: a read past the allocation is reported
#else
	const size_t size = (bytes + 255) & ~(size_t)255;
#endif
	if (posix_memalign(&q, 256, size) != 0) return hipErrorOutOfMemory;
	memset(q, 0xA5, bytes);   // device memory holds no zeros it was not given
	*p = q;
	return ok();
}
hipError_t hipFree(void *p)
{
	free(p);
	return ok();
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned)
{
	*p = bytes ? malloc(bytes) : nullptr;
	return (bytes && !*p) ? hipErrorOutOfMemory : ok();
}
hipError_t hipHostFree(void *p)
{
	free(p);
	return ok();
}
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind)
{
	if (bytes) memmove(dst, src, bytes);
	return ok();
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t)
{
	return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemset(void *dst, int v, size_t bytes)
{
	if (bytes) memset(dst, v, bytes);
	return ok();
}
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t) { return hipMemset(dst, v, bytes); }

struct emu_stream { int unused; };
struct emu_event { double t; };
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
	*s = new emu_stream();
	return ok();
}
hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s)
{
	delete s;
	return ok();
}
hipError_t hipStreamSynchronize(hipStream_t) { return ok(); }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return ok(); }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return hipErrorNotSupported; }
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *, char *, size_t) { return hipErrorNotSupported; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphDestroy(hipGraph_t) { return ok(); }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return ok(); }
static double now_ms()
{
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
hipError_t hipEventCreate(hipEvent_t *e)
{
	*e = new emu_event{0.0};
	return ok();
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
	delete e;
	return ok();
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
	e->t = now_ms();
	return ok();
}
hipError_t hipEventQuery(hipEvent_t) { return ok(); }
hipError_t hipEventSynchronize(hipEvent_t) { return ok(); }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
	*ms = (float)(b->t - a->t);
	return ok();
}
hipError_t hipDeviceSynchronize() { return ok(); }
hipError_t hipGetLastError()
{
	const hipError_t e = last_error;
	last_error = hipSuccess;
	return e;
}
hipError_t hipPeekAtLastError() { return last_error; }
const char *hipGetErrorString(hipError_t e)
{
	switch (e) {
	case hipSuccess: return "no error";
	case hipErrorOutOfMemory: return "out of memory";
	case hipErrorNotSupported: return "operation not supported by the CPU emulation";
	default: return "emulated HIP error";
	}
}
hipError_t hipGetDevice(int *d)
{
	*d = 0;
	return ok();
}
hipError_t hipSetDevice(int) { return ok(); }
hipError_t hipGetDeviceCount(int *n)
{
	*n = 1;
	return ok();
}
// the figures of one MI355X (gfx950), so that the library sizes its launches as there
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int)
{
	memset(p, 0, sizeof(*p));
	snprintf(p->name, sizeof(p->name), "CPU emulation of an MI355X");
	snprintf(p->gcnArchName, sizeof(p->gcnArchName), "gfx950");
	p->multiProcessorCount = 256;
	p->sharedMemPerBlock = 160 * 1024;
	p->warpSize = 64;
	p->maxThreadsPerBlock = 1024;
	p->totalGlobalMem = (size_t)288 << 30;
	return ok();
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int)
{
	if (a == hipDeviceAttributeMaxSharedMemoryPerBlock) *v = 160 * 1024;
	else if (a == hipDeviceAttributeMultiprocessorCount) *v = 256;
	else return hipErrorInvalidValue;
	return ok();
}
// static LDS of a kernel: the figure the gfx950 build reports for the grid solve
// (EMU_GRID_STATIC_LDS, from hipcc's resource usage), 0 otherwise
#ifndef EMU_GRID_STATIC_LDS
#define EMU_GRID_STATIC_LDS 0
#endif
hipError_t hipFuncGetAttributes(hipFuncAttributes *a, const void *)
{
	a->sharedSizeBytes = EMU_GRID_STATIC_LDS;
	a->maxThreadsPerBlock = 1024;
	a->numRegs = 128;
	return ok();
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int, size_t)
{
	*n = 1;
	return ok();
}
