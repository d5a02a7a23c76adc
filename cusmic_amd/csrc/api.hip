/*
 * cusmic C ABI on HIP (include/cusmic.h, include/cusmic_amd.h).
 *
 * Host side of the cleaner: argument checks with the reference's messages and
 * status codes (src/api.cu:47-114), the host/device/batch entry points
 * (src/api.cu:155-192) and the ragged-batch/adjoint extensions. One call
 * enqueues the loop on the caller's stream in chunks of passes, reads the
 * device counters between chunks to stop once every frame has converged, and
 * returns once the device has posted the errors and pass counts after the
 * last kernel.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <new>
#include <stdexcept>
#include <vector>

#include "../../include/cusmic.h"
#include "../../include/cusmic_amd.h"
#include "kernels.hip"

#ifndef CUSMIC_VERSION
#define CUSMIC_VERSION "cusmic-amd 0.1.0"
#endif

using namespace cusmic;

/* Scratch of remove_cosmics_batch_device_noalloc: one block per device, held
 * by every device the library runs on until the process ends. Host code uses
 * it only through its device address (static_scratch), never by this name. */
constexpr size_t STATIC_SCRATCH_BYTES = size_t(CUSMIC_NOALLOC_SCRATCH_MIB) << 20;
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];


namespace {

void
hip_check(hipError_t status)
{
	if (status == hipErrorOutOfMemory || status == hipErrorMemoryAllocation)
		throw std::bad_alloc();
	if (status != hipSuccess)
		throw std::runtime_error(hipGetErrorString(status));
}

/* Exception firewall: no C++ exception crosses the ABI (src/api.cu:83-105). */
template <class Func>
int
guarded(Func run, char *msg, size_t len)
{
	if (msg && len)
		msg[0] = 0;
	try {
		run();
		return CUSMIC_OK;
	} catch (const std::invalid_argument &e) {
		if (msg && len)
			std::snprintf(msg, len, "%s", e.what());
		return CUSMIC_INVALID;
	} catch (const std::bad_alloc &e) {
		if (msg && len)
			std::snprintf(msg, len, "%s", e.what());
		return CUSMIC_NO_MEMORY;
	} catch (const std::exception &e) {
		if (msg && len)
			std::snprintf(msg, len, "%s", e.what());
		return CUSMIC_CUDA_ERROR;
	}
}

/* Settings and shapes (src/api.cu:47-81); same_shape selects the batch rule. */
size_t
check_arguments(const cusmic_image *ims, size_t nf, const cusmic_options *o, const double *clean,
	const uint8_t *mask, bool same_shape, const double *fg = nullptr,
	const double *frn = nullptr)
{
	if (!ims || !o || !clean || !mask || !nf || nf > 65535)
		throw std::invalid_argument("invalid images, batch size or output buffers");
	if (o->maxiter < 0 || o->border < 0 || o->border > CUSMIC_WRAP || !std::isfinite(o->contrast) ||
		o->contrast < 0 || !std::isfinite(o->cr_threshold) || o->cr_threshold < 0 ||
		!std::isfinite(o->neighbor_threshold) || o->neighbor_threshold < 0)
		throw std::invalid_argument("invalid detection settings");
	size_t total = 0;
	for (size_t f = 0; f < nf; ++f) {
		const cusmic_image &im = ims[f];
		size_t w = im.width, h = im.height;
		if (!w || !h || w > size_t(INT_MAX) / 4 / h)
			throw std::invalid_argument("invalid image dimensions");
		if (!im.data || (same_shape && (w != ims[0].width || h != ims[0].height)))
			throw std::invalid_argument("frames must have data and the same shape");
		double g = fg ? fg[f] : o->gain, rn = frn ? frn[f] : o->readnoise;
		if (!im.error) {
			bool needs_gain = o->maxiter || !std::isnan(g);
			if (!im.gain && needs_gain && (!std::isfinite(g) || g <= 0))
				throw std::invalid_argument("positive gain or error is required");
			if (!im.readnoise && (!std::isfinite(rn) || rn < 0))
				throw std::invalid_argument("read noise must be finite and nonnegative");
		}
		if (!im.background && !std::isnan(o->background) && !std::isfinite(o->background))
			throw std::invalid_argument("background must be finite");
		total += w * h;
	}
	if (total > size_t(UINT_MAX) || total > SIZE_MAX / sizeof(double))
		throw std::invalid_argument("batch is too large");
	return total;
}

/* Host map contents (src/api.cu:107-114), split over host threads for large
 * maps: the check reads every value once, at the host's memory bandwidth. */
void
check_values(const double *data, size_t n, const char *name, int minimum)
{
	if (!data)
		return;
	auto bad_in = [=](size_t lo, size_t hi) {
		bool bad = false;
		for (size_t i = lo; i < hi; ++i)
			bad |= !std::isfinite(data[i]) || (minimum == 1 && data[i] <= 0) ||
			       (minimum == 0 && data[i] < 0);
		return bad;
	};
	constexpr size_t PER_THREAD = size_t(1) << 20;
	const size_t threads = std::min<size_t>({16, std::max(1u, std::thread::hardware_concurrency()),
		(n + PER_THREAD - 1) / PER_THREAD});
	bool bad = false;
	if (threads <= 1) {
		bad = bad_in(0, n);
	} else {
		std::vector<char> part(threads, 0);
		std::vector<std::thread> pool;
		const size_t step = (n + threads - 1) / threads;
		for (size_t t = 1; t < threads; ++t)
			pool.emplace_back([&, t] { part[t] = bad_in(std::min(n, t * step), std::min(n, (t + 1) * step)); });
		part[0] = bad_in(0, std::min(n, step));
		for (auto &th : pool)
			th.join();
		for (char b : part)
			bad |= b != 0;
	}
	if (bad)
		throw std::invalid_argument(name);
}

/* Stream-ordered scratch comes from one memory pool per device that keeps
 * freed blocks, up to a quarter of the device's memory, for the next call.
 * (The default pool returns freed memory to the driver at every
 * synchronisation, and a call then waits for the driver to supply and clear
 * a fresh block: about half a second for the 17 GB of a large batch.) */
constexpr int POOL_DEVICES = 64;
std::mutex pool_lock;
hipMemPool_t pools[POOL_DEVICES];

hipMemPool_t
scratch_pool()
{
	int dev = 0;
	hip_check(hipGetDevice(&dev));
	if (dev < 0 || dev >= POOL_DEVICES)
		throw std::runtime_error("device index beyond the memory pool table");
	std::lock_guard<std::mutex> hold(pool_lock);
	if (!pools[dev]) {
		hipMemPoolProps props{};
		props.allocType = hipMemAllocationTypePinned;
		props.handleTypes = hipMemHandleTypeNone;
		props.location.type = hipMemLocationTypeDevice;
		props.location.id = dev;
		hipMemPool_t pool = nullptr;
		hip_check(hipMemPoolCreate(&pool, &props));
		size_t free_bytes = 0, total_bytes = 0;
		hip_check(hipMemGetInfo(&free_bytes, &total_bytes));
		uint64_t keep = total_bytes / 4;
		hip_check(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
		pools[dev] = pool;
	}
	return pools[dev];
}

/* A call's device scratch. Every call waits for its kernels before it
 * returns, so at scope exit a block is idle and goes back to a per-device
 * cache instead of being freed; the next call takes the smallest cached block
 * that fits. Only when none fits are the device's smaller cached blocks freed
 * and a new block allocated from the scratch pool. Returning a block is a
 * list insertion: a stream-ordered free took 0.36 ms per call inside torch's
 * autograd backward. The cache keeps at most a quarter of the device's memory
 * (beyond that a returned block is freed). A call that ends by an exception
 * first waits for its stream. */
struct cached_block {
	void *ptr;
	size_t bytes;
	int dev;
};
std::mutex block_lock;
std::vector<cached_block> block_cache;
size_t device_bytes[POOL_DEVICES]; /* total memory per device, read once */

struct device_block {
	void *ptr = nullptr;
	size_t bytes = 0;
	int dev = 0;
	hipStream_t stream;
	const int live_exceptions;
	device_block(size_t need, hipStream_t s) : stream(s), live_exceptions(std::uncaught_exceptions())
	{
		if (!need)
			return;
		hip_check(hipGetDevice(&dev));
		std::vector<void *> idle;
		{
			std::lock_guard<std::mutex> hold(block_lock);
			long best = -1;
			for (size_t k = 0; k < block_cache.size(); ++k)
				if (block_cache[k].dev == dev && block_cache[k].bytes >= need &&
					(best < 0 || block_cache[k].bytes < block_cache[size_t(best)].bytes))
					best = long(k);
			if (best >= 0) {
				ptr = block_cache[size_t(best)].ptr;
				bytes = block_cache[size_t(best)].bytes;
				block_cache.erase(block_cache.begin() + best);
				return;
			}
			for (size_t k = block_cache.size(); k-- > 0;)
				if (block_cache[k].dev == dev) {
					idle.push_back(block_cache[k].ptr);
					block_cache.erase(block_cache.begin() + long(k));
				}
		}
		for (void *q : idle)
			hip_check(hipFreeAsync(q, s));
		hip_check(hipMallocFromPoolAsync(&ptr, need, scratch_pool(), s));
		bytes = need;
	}
	~device_block()
	{
		if (!ptr)
			return;
		if (std::uncaught_exceptions() > live_exceptions)
			(void)hipStreamSynchronize(stream);
		std::unique_lock<std::mutex> hold(block_lock);
		size_t cached = bytes;
		for (const auto &c : block_cache)
			cached += c.dev == dev ? c.bytes : 0;
		if (dev >= 0 && dev < POOL_DEVICES && !device_bytes[dev]) {
			size_t free_bytes = 0, total_bytes = 0;
			if (hipMemGetInfo(&free_bytes, &total_bytes) == hipSuccess)
				device_bytes[dev] = total_bytes;
		}
		if (dev < 0 || dev >= POOL_DEVICES || cached > device_bytes[dev] / 4) {
			hold.unlock();
			(void)hipFreeAsync(ptr, stream);
			return;
		}
		block_cache.push_back({ptr, bytes, dev});
	}
	device_block(const device_block &) = delete;
	device_block &operator=(const device_block &) = delete;
};

/* Pinned host memory for a call's small transfers: the frame table goes up
 * from it, and the device posts the counters the host reads into it (word 0:
 * a sequence number, words 64 on: the counters). Copies between the device
 * and pageable memory are staged through the runtime's buffer, tens of
 * microseconds each, and a stream synchronisation costs about as much again
 * after the device has finished; polling a word that a kernel writes does
 * not. Blocks are taken from a pool, one per call in flight, and returned for
 * the next call; they are never freed (pinned memory goes back to the system
 * at process end). */
struct mailbox_block {
	void *ptr;
	size_t bytes;
	unsigned seq;
};
std::mutex mailbox_lock;
std::vector<mailbox_block> mailbox_pool;

struct mailbox {
	mailbox_block b{nullptr, 0, 0};
	hipStream_t stream;
	const int live_exceptions;
	mailbox(size_t need, hipStream_t s) : stream(s), live_exceptions(std::uncaught_exceptions())
	{
		{
			std::lock_guard<std::mutex> hold(mailbox_lock);
			for (size_t k = 0; k < mailbox_pool.size(); ++k)
				if (mailbox_pool[k].bytes >= need) {
					b = mailbox_pool[k];
					mailbox_pool.erase(mailbox_pool.begin() + long(k));
					return;
				}
		}
		b.bytes = std::max<size_t>(need, size_t(64) << 10);
		hip_check(hipHostMalloc(&b.ptr, b.bytes, hipHostMallocDefault));
		*static_cast<unsigned int *>(b.ptr) = 0;
	}
	~mailbox()
	{
		/* a call that ends by an exception may have left a copy from or to
		 * the block on its stream: let it finish before another call takes
		 * the block */
		if (std::uncaught_exceptions() > live_exceptions)
			(void)hipStreamSynchronize(stream);
		std::lock_guard<std::mutex> hold(mailbox_lock);
		mailbox_pool.push_back(b);
	}
	unsigned int *words() const { return static_cast<unsigned int *>(b.ptr) + 64; }
	/* Enqueue the post of n words at dev, then wait until they are here:
	 * poll the sequence word, asking the stream every few microseconds
	 * whether it still has work (a failed kernel ends the wait with its
	 * error); after POLL_MS the wait becomes a stream synchronisation, which
	 * gives the host's core back during long calls. Nothing is enqueued after
	 * the post before it arrives, so both waits end at the same point. */
	void
	post_and_wait(const unsigned int *dev, size_t n)
	{
		constexpr double POLL_MS = 20.0;
		const unsigned seq = ++b.seq;
		unsigned int *flag = static_cast<unsigned int *>(b.ptr);
		hipLaunchKernelGGL(post_counters, dim3(1), dim3(256), 0, stream, dev, flag, unsigned(n), seq);
		hip_check(hipGetLastError());
		const auto start = std::chrono::steady_clock::now();
		for (unsigned k = 1; __atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq; ++k) {
			if (k % 256 == 0) {
				const hipError_t e = hipStreamQuery(stream);
				const bool long_wait = std::chrono::duration<double, std::milli>(
					std::chrono::steady_clock::now() - start).count() > POLL_MS;
				if (e != hipErrorNotReady || long_wait) {
					if (e != hipErrorNotReady)
						hip_check(e);
					else
						hip_check(hipStreamSynchronize(stream));
					if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq)
						throw std::runtime_error("the device did not post its counters");
					break;
				}
			}
			__builtin_ia32_pause();
		}
	}
	mailbox(const mailbox &) = delete;
	mailbox &operator=(const mailbox &) = delete;
};

/* Device address of the static scratch array on the current device, looked
 * up once per device (each device holds its own copy of the array). */
constexpr int STATIC_DEVICES = 64;
std::mutex static_lock;
void *static_address[STATIC_DEVICES];

void *
static_scratch()
{
	int dev = 0;
	hip_check(hipGetDevice(&dev));
	if (dev < 0 || dev >= STATIC_DEVICES)
		throw std::runtime_error("device index beyond the static scratch table");
	std::lock_guard<std::mutex> hold(static_lock);
	if (!static_address[dev]) {
		void *ptr = nullptr;
		hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
		if (!ptr)
			throw std::runtime_error("no device address for the static scratch");
		static_address[dev] = ptr;
	}
	return static_address[dev];
}

size_t
align_up(size_t v, size_t a)
{
	return (v + a - 1) / a * a;
}

int
device_cus()
{
	int dev = 0, n = 0;
	hip_check(hipGetDevice(&dev));
	hip_check(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
	return n > 0 ? n : 1;
}

/* Host copies split over a few persistent threads: pageable memory moves to
 * and from pinned staging at several times one thread's memcpy rate. The
 * caller's thread takes the first slice. */
/* Copy bytes of doubles and tell whether one of them fails the map check of
 * check_values (minimum 1: finite and positive; 0: finite and nonnegative;
 * -1: finite). NO_CHECK copies any bytes unchecked. */
constexpr int NO_CHECK = -2;

bool
copy_checked(void *dst, const void *src, size_t bytes, int minimum)
{
	if (minimum == NO_CHECK) {
		std::memcpy(dst, src, bytes);
		return false;
	}
	const double *s = static_cast<const double *>(src);
	double *d = static_cast<double *>(dst);
	bool bad = false;
	for (size_t i = 0, n = bytes / sizeof(double); i < n; ++i) {
		const double v = s[i];
		d[i] = v;
		bad |= !std::isfinite(v) || (minimum == 1 && v <= 0) || (minimum == 0 && v < 0);
	}
	return bad;
}

class copy_threads {
public:
	static constexpr unsigned MAX_THREADS = 8;

	static copy_threads &get()
	{
		/* never destroyed: its workers stay parked until the process ends */
		static copy_threads *pool = new copy_threads;
		return *pool;
	}

	/* Copy on the workers and this thread; with a check (copy_checked),
	 * returns whether a value failed it. */
	bool copy(void *dst, const void *src, size_t bytes, int minimum = NO_CHECK)
	{
		const size_t slices = std::min<size_t>(workers_.size() + 1, std::max<size_t>(1, bytes >> 20));
		if (slices <= 1)
			return copy_checked(dst, src, bytes, minimum);
		std::lock_guard<std::mutex> one_job(job_lock_);
		/* slices of whole doubles */
		const size_t part = ((bytes + slices - 1) / slices + 7) & ~size_t(7);
		{
			std::lock_guard<std::mutex> l(m_);
			dst_ = static_cast<char *>(dst);
			src_ = static_cast<const char *>(src);
			bytes_ = bytes;
			part_ = part;
			slices_ = slices;
			minimum_ = minimum;
			bad_ = false;
			pending_ = slices - 1;
			++generation_;
		}
		go_.notify_all();
		const bool mine = copy_checked(dst, src, std::min(part, bytes), minimum);
		std::unique_lock<std::mutex> l(m_);
		done_.wait(l, [&] { return pending_ == 0; });
		return mine || bad_;
	}

private:
	copy_threads()
	{
		const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
		const unsigned n = std::min(MAX_THREADS, hw) - 1;
		for (unsigned k = 0; k < n; ++k)
			workers_.emplace_back([this, k] { work(k + 1); });
	}

	void work(size_t slice)
	{
		unsigned seen = 0;
		for (;;) {
			char *dst;
			const char *src;
			size_t lo, hi;
			int minimum;
			{
				std::unique_lock<std::mutex> l(m_);
				go_.wait(l, [&] { return generation_ != seen; });
				seen = generation_;
				if (slice >= slices_)
					continue;
				dst = dst_;
				src = src_;
				minimum = minimum_;
				lo = std::min(bytes_, slice * part_);
				hi = std::min(bytes_, lo + part_);
			}
			const bool bad = copy_checked(dst + lo, src + lo, hi - lo, minimum);
			std::lock_guard<std::mutex> l(m_);
			bad_ = bad_ || bad;
			if (--pending_ == 0)
				done_.notify_one();
		}
	}

	std::vector<std::thread> workers_;
	std::mutex job_lock_, m_;
	std::condition_variable go_, done_;
	char *dst_ = nullptr;
	const char *src_ = nullptr;
	size_t bytes_ = 0, part_ = 0, slices_ = 0, pending_ = 0;
	int minimum_ = NO_CHECK;
	bool bad_ = false;
	unsigned generation_ = 0;
};

/* Copies between pageable host memory and the device through two pinned
 * staging buffers: the host copy of one chunk overlaps the DMA of the other.
 * Copies below DIRECT bytes go straight from pageable memory, where the
 * hand-off would cost more than it saves. One staging set per process; calls
 * take turns on it. */
class staging {
public:
	static constexpr size_t CHUNK = size_t(32) << 20;
	static constexpr size_t DIRECT = size_t(4) << 20;

	static staging &get()
	{
		static staging *s = new staging;
		return *s;
	}

	std::mutex lock;

	/* With a check (copy_checked), the values are checked as they are
	 * copied, and the return value tells whether one failed; the whole plane
	 * is uploaded either way. */
	bool upload(void *dst, const void *src, size_t bytes, hipStream_t stream, int minimum = NO_CHECK)
	{
		bool bad = false;
		if (bytes < DIRECT) {
			if (minimum != NO_CHECK) {
				const double *v = static_cast<const double *>(src);
				for (size_t i = 0, n = bytes / sizeof(double); i < n; ++i)
					bad |= !std::isfinite(v[i]) || (minimum == 1 && v[i] <= 0) || (minimum == 0 && v[i] < 0);
			}
			hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
			return bad;
		}
		ensure();
		char *d = static_cast<char *>(dst);
		const char *s = static_cast<const char *>(src);
		for (size_t off = 0, k = 0; off < bytes; off += CHUNK, ++k) {
			const size_t len = std::min(CHUNK, bytes - off), b = k & 1;
			hip_check(hipEventSynchronize(done_[b]));
			bad |= copy_threads::get().copy(pin_[b], s + off, len, minimum);
			hip_check(hipMemcpyAsync(d + off, pin_[b], len, hipMemcpyHostToDevice, stream));
			hip_check(hipEventRecord(done_[b], stream));
		}
		return bad;
	}

	void download(void *dst, const void *src, size_t bytes, hipStream_t stream)
	{
		if (bytes < DIRECT) {
			hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
			return;
		}
		ensure();
		char *d = static_cast<char *>(dst);
		const char *s = static_cast<const char *>(src);
		const size_t chunks = (bytes + CHUNK - 1) / CHUNK;
		auto fetch = [&](size_t k) {
			const size_t off = k * CHUNK, len = std::min(CHUNK, bytes - off), b = k & 1;
			hip_check(hipMemcpyAsync(pin_[b], s + off, len, hipMemcpyDeviceToHost, stream));
			hip_check(hipEventRecord(done_[b], stream));
		};
		for (size_t k = 0; k < std::min<size_t>(2, chunks); ++k)
			fetch(k);
		for (size_t k = 0; k < chunks; ++k) {
			const size_t off = k * CHUNK, len = std::min(CHUNK, bytes - off), b = k & 1;
			hip_check(hipEventSynchronize(done_[b]));
			copy_threads::get().copy(d + off, pin_[b], len);
			if (k + 2 < chunks)
				fetch(k + 2);
		}
	}

private:
	void ensure()
	{
		if (pin_[0])
			return;
		for (int b = 0; b < 2; ++b) {
			hip_check(hipHostMalloc(&pin_[b], CHUNK, hipHostMallocDefault));
			hip_check(hipEventCreateWithFlags(&done_[b], hipEventDisableTiming));
		}
	}

	void *pin_[2] = {nullptr, nullptr};
	hipEvent_t done_[2] = {nullptr, nullptr};
};

/* Frame table of a batch: offsets and tile numbering. */
std::vector<Frame>
frame_table(const cusmic_image *ims, size_t nf, const cusmic_options &o, const double *fg,
	const double *frn, int *ntiles_out, int *nchunks_out = nullptr, int *nruns_out = nullptr,
	int *nblocks_out = nullptr, int run_tiles = RUN_TILES, int tile_w = TILE_W)
{
	std::vector<Frame> frames(nf);
	long long off = 0;
	int tiles = 0, chunks = 0, runs = 0, blocks = 0;
	for (size_t f = 0; f < nf; ++f) {
		const cusmic_image &im = ims[f];
		Frame &fr = frames[f];
		fr.data = im.data;
		fr.error = im.error;
		fr.gain = im.error ? nullptr : im.gain;
		fr.readnoise = im.error ? nullptr : im.readnoise;
		fr.background = im.background;
		fr.mask = im.mask;
		fr.g = fg ? fg[f] : o.gain;
		fr.rn = frn ? frn[f] : o.readnoise;
		fr.w = int(im.width);
		fr.h = int(im.height);
		fr.off = off;
		fr.tiles_x = (fr.w + tile_w - 1) / tile_w;
		fr.tiles_y = (fr.h + TILE_H - 1) / TILE_H;
		fr.tile0 = tiles;
		fr.chunk0 = chunks;
		fr.run0 = runs;
		fr.block0 = blocks;
		fr.blocks_x = (fr.w + BLK - 1) / BLK;
		long long px = (long long)fr.w * fr.h;
		long long nb = (long long)fr.blocks_x * ((fr.h + BLK - 1) / BLK);
		if ((long long)tiles + (long long)fr.tiles_x * fr.tiles_y > INT_MAX ||
			(long long)chunks + (px + CHUNK - 1) / CHUNK > INT_MAX || blocks + nb > INT_MAX)
			throw std::invalid_argument("batch is too large");
		blocks += int(nb);
		tiles += fr.tiles_x * fr.tiles_y;
		chunks += int((px + CHUNK - 1) / CHUNK);
		runs += fr.tiles_x * runs_y(fr.tiles_y, run_tiles);
		off += px;
	}
	*ntiles_out = tiles;
	if (nchunks_out)
		*nchunks_out = chunks;
	if (nruns_out)
		*nruns_out = runs;
	if (nblocks_out)
		*nblocks_out = blocks;
	return frames;
}

/* The whole call: prologue, hole fill, maxiter detection/replacement passes
 * and epilogue (src/cleaner.cuh:144-187), all device-driven. */
/* Optional per-kernel timing with events on the call's stream. */
struct timer {
	hipStream_t stream;
	bool on;
	std::vector<hipEvent_t> events;
	std::vector<int> kinds;
	explicit timer(hipStream_t s, bool enabled) : stream(s), on(enabled) {}
	~timer()
	{
		for (auto e : events)
			(void)hipEventDestroy(e);
	}
	void
	mark(int kind)
	{
		if (!on)
			return;
		hipEvent_t e;
		hip_check(hipEventCreate(&e));
		hip_check(hipEventRecord(e, stream));
		events.push_back(e);
		kinds.push_back(kind);
	}
	/* kind k of event i times the interval from event i-1 to i */
	void
	collect(double *out)
	{
		for (size_t i = 1; i < events.size(); ++i) {
			float ms = 0;
			hip_check(hipEventElapsedTime(&ms, events[i - 1], events[i]));
			if (kinds[i] >= 0)
				out[kinds[i]] += ms;
		}
	}
};

/* The scratch block is allocated per call, or is the caller's block
 * (static_base, static_bytes), in which case nothing is allocated and a batch
 * that needs more is rejected before any work is enqueued. */
void
clean_images(const cusmic_image *ims, size_t nf, const cusmic_options &o, const double *fg,
	const double *frn, int laplacian, double *clean, uint8_t *mask, int32_t *donors,
	int32_t *passes, hipStream_t stream, double *profile = nullptr, void *static_base = nullptr,
	size_t static_bytes = 0, int64_t *new_pixels = nullptr, size_t new_pixels_size = 0)
{
	for (size_t k = 0; new_pixels && k < new_pixels_size; ++k)
		new_pixels[k] = 0;
	/* the first pass's kernel, and how many of its workgroups fit at once */
	bool model = false; /* some frame uses the noise model instead of an error map */
	bool bg_map = false; /* some frame has a background map */
	for (size_t f = 0; f < nf; ++f) {
		model |= ims[f].error == nullptr;
		bg_map |= ims[f].background != nullptr;
	}
	const bool bg_any = bg_map || !std::isnan(o.background); /* some frame has a background */
	using kernel_t = void (*)(Params, int);
	kernel_t sig;
	if (model)
		sig = laplacian ? (bg_any ? significance_model<false, true> : significance_model<false, false>)
				: (bg_any ? significance_model<true, true> : significance_model<true, false>);
	else
		sig = laplacian ? (bg_map ? significance_maps<false, true, true>
				       : bg_any ? significance_maps<false, true, false> : significance_maps<false, false, false>)
				: (bg_map ? significance_maps<true, true, true>
				       : bg_any ? significance_maps<true, true, false> : significance_maps<true, false, false>);
	/* the error-map pass takes wide tiles */
	const int tile_w = model ? TILE_W : TILE_W2;
	const int cus = device_cus();
	int sig_per_cu = 0;
	hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&sig_per_cu, sig, BLOCK, 0));
	const int sig_resident = cus * std::max(sig_per_cu, 1);
	/* first-pass runs: the longest that still gives every resident
	 * workgroup four runs on average */
	int run_tiles = RUN_TILES;
	for (; run_tiles > 1; run_tiles /= 2) {
		long long runs = 0;
		for (size_t f = 0; f < nf; ++f) {
			const long long tx = ((long long)ims[f].width + tile_w - 1) / tile_w;
			const long long ty = ((long long)ims[f].height + TILE_H - 1) / TILE_H;
			runs += tx * ((ty + run_tiles - 1) / run_tiles);
		}
		if (runs >= 4LL * sig_resident)
			break;
	}
	int ntiles = 0, nchunks = 0, nruns = 0, nblocks = 0;
	std::vector<Frame> frames =
		frame_table(ims, nf, o, fg, frn, &ntiles, &nchunks, &nruns, &nblocks, run_tiles, tile_w);
	size_t total = 0;
	for (const auto &fr : frames)
		total += size_t(fr.w) * fr.h;
	int maxiter = o.maxiter;

	/* one scratch block: frame table | counters (zeroed) | flags | list */
	/* no part of the scratch grows with maxiter: the per-pass counters are a
	 * ring of PASS_RING slots */
	size_t n_counts = nf * (sizeof(Counts) / 4);
	size_t n_shard = nf * size_t(SHARDS * SHARD_FIELDS);
	size_t n_todo = size_t(PASS_RING) * nf;
	size_t n_queue = XCDS;
	size_t n_nlist = NL_GLOBALS + size_t(PASS_RING) * PASS_FIELDS;
	/* the words the host reads come first, in one run that one copy brings
	 * back: status (2) | processed pixels (64-bit) | counts | todo ring |
	 * pass counters */
	const size_t n_read = 4 + n_counts + n_todo + n_nlist;
	size_t words = n_read + n_shard + n_queue + 5 * size_t(nblocks) + size_t(ntiles);
	size_t off_frames = 0;
	size_t off_words = align_up(nf * sizeof(Frame), 256);
	size_t off_dlist = align_up(off_words + words * 4, 256);
	size_t off_clist = align_up(off_dlist + size_t(nblocks) * 16, 256); /* recompute + reseed */
	size_t off_cframe = align_up(off_clist + size_t(nblocks) * 8, 256);
	size_t off_rframe = align_up(off_cframe + size_t(nchunks) * 4, 256);
	size_t off_flags = align_up(off_rframe + size_t(nruns) * 4, 256);
	size_t planes = maxiter ? total : 0;
	size_t off_stamp = align_up(off_flags + planes, 256);
	size_t off_sprime = align_up(off_stamp + planes * 2, 256);
	size_t off_list = align_up(off_sprime + planes * 8, 256);
	size_t off_list_b = align_up(off_list + planes * 8, 256);
	size_t off_list_c = align_up(off_list_b + planes * 8, 256);
	size_t off_wide = align_up(off_list_c + planes * 8, 256);
	/* pixels within four of a frame edge: the seeds fine takes apart */
	size_t edge = 0;
	for (const auto &fr : frames)
		edge += size_t(fr.w) * fr.h - size_t(std::max(fr.w - 8, 0)) * std::max(fr.h - 8, 0);
	size_t off_edge = align_up(off_wide + planes * 8, 256);
	size_t bytes = off_edge + (maxiter ? edge : 0) * 8;
	if (static_base && bytes > static_bytes) {
		char text[160];
		std::snprintf(text, sizeof(text),
			"batch needs %zu bytes of scratch, more than the %zu of the static buffer", bytes,
			static_bytes);
		throw std::invalid_argument(text);
	}
	device_block mem(static_base ? 0 : bytes, stream);
	char *base = static_cast<char *>(static_base ? static_base : mem.ptr);
	/* the pinned block: sequence word and the counters the host reads | the
	 * frame table; a call that may not allocate copies through pageable
	 * memory and synchronises instead */
	const size_t box_table = align_up(256 + n_read * 4, 256);
	std::unique_ptr<mailbox> box(static_base ? nullptr : new mailbox(box_table + nf * sizeof(Frame), stream));
	std::vector<unsigned int> read_host(box ? 0 : n_read);
	unsigned int *read_words = box ? box->words() : read_host.data();
	const void *table = frames.data();
	if (box) {
		table = static_cast<char *>(box->b.ptr) + box_table;
		std::memcpy(const_cast<void *>(table), frames.data(), nf * sizeof(Frame));
	}
	hip_check(hipMemcpyAsync(base + off_frames, table, nf * sizeof(Frame), hipMemcpyHostToDevice,
		stream));

	Params p{};
	p.frames = reinterpret_cast<const Frame *>(base + off_frames);
	p.nf = int(nf);
	p.ntiles = ntiles;
	p.maxiter = maxiter;
	p.mode = o.border;
	p.laplacian = laplacian;
	p.run_tiles = run_tiles;
	p.tile_w = tile_w;
	p.contrast = o.contrast;
	p.threshold = o.cr_threshold;
	p.neighbor = o.neighbor_threshold;
	p.bg = o.background;
	p.clean = clean;
	p.mask = mask;
	p.donors = donors;
	unsigned int *w32 = reinterpret_cast<unsigned int *>(base + off_words);
	p.status = reinterpret_cast<int *>(w32);
	p.work = reinterpret_cast<unsigned long long *>(w32 + 2);
	p.counts = reinterpret_cast<Counts *>(w32 + 4);
	p.todo = reinterpret_cast<int *>(w32 + 4 + n_counts);
	p.nlist = w32 + 4 + n_counts + n_todo;
	p.shard = reinterpret_cast<int *>(w32 + n_read);
	p.queue = w32 + n_read + n_shard;
	p.block_pass = p.queue + n_queue;
	p.changed_tag = p.block_pass + nblocks;
	p.dirty_tag = p.changed_tag + nblocks;
	p.changed_sides = p.dirty_tag + nblocks;
	p.reseed_tag = p.changed_sides + nblocks;
	p.tile_skip = p.reseed_tag + nblocks;
	p.changed_list = reinterpret_cast<unsigned long long *>(base + off_clist);
	p.nblocks = nblocks;
	p.list_cap = unsigned(planes);
	p.dirty_list = reinterpret_cast<unsigned long long *>(base + off_dlist);
	p.chunk_frame = reinterpret_cast<int *>(base + off_cframe);
	p.run_frame = reinterpret_cast<int *>(base + off_rframe);
	p.nruns = nruns;
	p.nchunks = nchunks;
	p.flags = reinterpret_cast<uint8_t *>(base + off_flags);
	p.stamp = reinterpret_cast<uint16_t *>(base + off_stamp);
	p.sprime = reinterpret_cast<double *>(base + off_sprime);
	p.list = reinterpret_cast<unsigned long long *>(base + off_list);
	p.list_b = reinterpret_cast<unsigned long long *>(base + off_list_b);
	p.list_c = reinterpret_cast<unsigned long long *>(base + off_list_c);
	p.wide_all = reinterpret_cast<unsigned long long *>(base + off_wide);
	p.edge_list = reinterpret_cast<unsigned long long *>(base + off_edge);

	unsigned table_grid = unsigned(std::min<long long>(((long long)nchunks + nruns + 255) / 256, 4096));
	hipLaunchKernelGGL(set_tables, dim3(table_grid), dim3(256), 0, stream, p);
	if (!maxiter) {
		hipLaunchKernelGGL(copy_without_detection, dim3(nchunks), dim3(BLOCK), 0, stream, p);
		hip_check(hipGetLastError());
		hip_check(hipStreamSynchronize(stream));
		return;
	}

	p.reach = model ? REACH : REACH_MAPS;
	p.reach_s = model ? REACH_S : REACH_S_MAPS;
	timer tm(stream, profile != nullptr);
	/* Without a background the output is C except at holes, and the
	 * epilogue can touch only the flagged pixels and the holes, on a mask
	 * plane zeroed here with the flags. Not when the output mask overlaps an
	 * input mask, which the first pass has yet to read. */
	bool sparse_end = !bg_any;
	for (size_t f = 0; f < nf && sparse_end; ++f) {
		const uintptr_t in = reinterpret_cast<uintptr_t>(ims[f].mask), out = reinterpret_cast<uintptr_t>(mask);
		sparse_end = !in || in + ims[f].width * ims[f].height <= out || out + total <= in;
	}
	tm.mark(-1);
	hipLaunchKernelGGL(zero_words, dim3(unsigned(std::min<size_t>((words + 255) / 256, 4096))),
		dim3(256), 0, stream, w32, words);
	/* a few 16-byte stores per lane: the fill moves 6.5 TB/s, against 5.1 TB/s
	 * with 8 workgroups per CU striding over the planes */
	hipLaunchKernelGGL(zero_groups, dim3(unsigned(std::min<size_t>((total / 16 + 256) / 256, size_t(cus) * 128))),
		dim3(256), 0, stream, reinterpret_cast<uint4 *>(p.flags), (total + 15) / 16,
		sparse_end ? mask : nullptr, total);
	if (donors)
		hipLaunchKernelGGL(reset_donors, dim3(unsigned(std::min<size_t>((2 * total + 255) / 256, 8192))),
			dim3(256), 0, stream, donors, 2 * total);
	tm.mark(CUSMIC_PROFILE_PROLOGUE_MS);
	/* the first pass, with the prologue's work, in persistent workgroups: as
	 * many as fit on the device */
	hipLaunchKernelGGL(sig, dim3(unsigned(std::min(nruns, sig_resident))), dim3(BLOCK), 0, stream, p, 0);
	tm.mark(CUSMIC_PROFILE_FIRST_PASS_MS);
	const unsigned fold_grid = unsigned((nf * SHARDS + 255) / 256);
	hipLaunchKernelGGL(fold_counts, dim3(fold_grid), dim3(256), 0, stream, p, 0, 1);
	unsigned replace_grid = unsigned(cus) * 8;
	hipLaunchKernelGGL(replace_pass, dim3(replace_grid), dim3(BLOCK), 0, stream, p, 0, 1);
	hipLaunchKernelGGL(replace_widened, dim3(cus), dim3(BLOCK), 0, stream, p, 0, 1);
	tm.mark(CUSMIC_PROFILE_REPLACE_MS);
	unsigned list_grid = unsigned(cus) * 4;
	auto blocks = model ? (laplacian ? significance_blocks<true, false> : significance_blocks<true, true>)
			    : (laplacian ? significance_blocks<false, false> : significance_blocks<false, true>);
	int blocks_per_cu = 0;
	hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, blocks, BLOCK, 0));
	/* a multiple of XCDS: the block pass splits its list between the XCDs */
	const unsigned blocks_grid = unsigned(std::max(cus * std::max(blocks_per_cu, 1) / XCDS, 1) * XCDS);
	auto enqueue_pass = [&](int it) {
		/* pass 0: the blocks of tiles with a hole, after the fill; later
		 * passes: the blocks near a pixel the last replacement changed */
		if (it > 0)
			hipLaunchKernelGGL(expand_dirty, dim3(unsigned(cus) * 2), dim3(BLOCK), 0, stream, p, it);
		hipLaunchKernelGGL(blocks, dim3(blocks_grid), dim3(BLOCK), 0, stream, p, it);
		tm.mark(CUSMIC_PROFILE_SIGNIFICANCE_MS);
		hipLaunchKernelGGL(fine, dim3(list_grid), dim3(BLOCK), 0, stream, p, it);
		hipLaunchKernelGGL(fine_edge, dim3(list_grid), dim3(BLOCK), 0, stream, p, it);
		hipLaunchKernelGGL(grow_first, dim3(list_grid), dim3(BLOCK), 0, stream, p, it);
		hipLaunchKernelGGL(grow_second, dim3(list_grid), dim3(BLOCK), 0, stream, p, it);
		/* the pass's counts; after the first pass redo_scan folds them */
		if (it > 0)
			hipLaunchKernelGGL(redo_scan, dim3(list_grid), dim3(BLOCK), 0, stream, p, it);
		else
			hipLaunchKernelGGL(fold_counts, dim3(fold_grid), dim3(256), 0, stream, p, it, 0);
		tm.mark(CUSMIC_PROFILE_DETECT_MS);
		hipLaunchKernelGGL(replace_pass, dim3(replace_grid), dim3(BLOCK), 0, stream, p, it, 0);
		hipLaunchKernelGGL(replace_widened, dim3(cus), dim3(BLOCK), 0, stream, p, it, 0);
		tm.mark(CUSMIC_PROFILE_REPLACE_MS);
	};
	/* The loop in chunks of CHUNK_PASSES passes. Between chunks the host
	 * reads the pass counters (one synchronisation per chunk) and stops once
	 * no frame flagged a new pixel in the chunk's last pass, or a replacement
	 * found no donor. maxiter <= CHUNK_PASSES (the default 4) is one chunk
	 * and no synchronisation inside the loop. */
	/* the host's view of the counters, as of the last read_counters */
	const int *status_word = reinterpret_cast<const int *>(read_words);
	const Counts *counts = reinterpret_cast<const Counts *>(read_words + 4);
	const int *todo_ring = reinterpret_cast<const int *>(read_words + 4 + n_counts);
	const unsigned int *nl = read_words + 4 + n_counts + n_todo;
	auto read_counters = [&]() {
		hip_check(hipGetLastError());
		if (box) {
			box->post_and_wait(w32, n_read);
			/* the events of the stage timings must have completed */
			if (profile)
				hip_check(hipStreamSynchronize(stream));
		} else {
			hip_check(hipMemcpyAsync(read_words, w32, n_read * 4, hipMemcpyDeviceToHost, stream));
			hip_check(hipStreamSynchronize(stream));
		}
	};
	double list_sums[4] = {0, 0, 0, 0}; /* seeds, candidates, grown, widened */
	/* the lists and new pixels of passes [a, b), read from the ring */
	auto add_lists = [&](int a, int b) {
		for (int k = a; k < b; ++k) {
			list_sums[0] += nl[nl_pass(k, PF_SEED)];
			list_sums[1] += nl[nl_pass(k, PF_CAND)];
			list_sums[2] += nl[nl_pass(k, PF_GROW)];
			list_sums[3] += nl[nl_pass(k, PF_WIDE)];
			if (new_pixels && size_t(k) < new_pixels_size)
				for (size_t f = 0; f < nf; ++f)
					new_pixels[k] += todo_ring[size_t(pass_slot(k)) * nf + f];
		}
	};
	int passes_run = 0, listed = 0;
	while (passes_run < maxiter) {
		const int a = passes_run, b = std::min(maxiter, a + CHUNK_PASSES);
		if (a > 0) {
			/* the interval since the last event spans the host's read */
			tm.mark(-1);
			hipLaunchKernelGGL(clear_pass_slots, dim3(1), dim3(256), 0, stream, p, a + 1, b);
		}
		for (int it = a; it < b; ++it)
			enqueue_pass(it);
		passes_run = b;
		if (b == maxiter)
			break;
		read_counters();
		add_lists(listed, b);
		listed = b;
		if (!nl[nl_pass(b - 1, PF_ACTIVE)] || (*status_word & ERR_NO_DONORS))
			break;
	}
	if (sparse_end) {
		hipLaunchKernelGGL(epilogue_sparse, dim3(unsigned(cus) * 4), dim3(BLOCK), 0, stream, p);
	} else {
		int bg_maps = 0;
		for (size_t f = 0; f < nf; ++f)
			bg_maps |= ims[f].background != nullptr;
		long long groups = ((long long)total + 7) / 8;
		unsigned grid = unsigned(std::min<long long>((groups + BLOCK - 1) / BLOCK, (long long)cus * 16));
		hipLaunchKernelGGL(epilogue, dim3(grid), dim3(BLOCK), 0, stream, p, (long long)total, bg_maps);
	}
	tm.mark(CUSMIC_PROFILE_EPILOGUE_MS);
	read_counters();
	const unsigned long long work = read_words[2] | (unsigned long long)read_words[3] << 32;
	if (new_pixels && !profile)
		add_lists(listed, passes_run);
	/* frame f flagged new pixels in fresh_passes passes, each followed by a
	 * replacement; it ran one detection pass more, up to maxiter */
	auto detect_passes = [&](size_t f) { return std::min(counts[f].fresh_passes + 1, maxiter); };
	if (profile) {
		add_lists(listed, passes_run);
		for (int k = 0; k < CUSMIC_PROFILE_SIZE; ++k)
			profile[k] = 0;
		tm.collect(profile);
		profile[CUSMIC_PROFILE_DETECT_MS] +=
			profile[CUSMIC_PROFILE_SIGNIFICANCE_MS] + profile[CUSMIC_PROFILE_FIRST_PASS_MS];
		profile[CUSMIC_PROFILE_DETECT_LAUNCHES] = passes_run;
		for (size_t f = 0; f < nf; ++f) {
			double px = double(frames[f].w) * frames[f].h;
			profile[CUSMIC_PROFILE_DETECT_PASS_PIXELS] += px * detect_passes(f);
			profile[CUSMIC_PROFILE_REPLACE_PASS_PIXELS] += px * counts[f].fresh_passes;
		}
		profile[CUSMIC_PROFILE_DETECT_PIXELS] = double(work);
		profile[CUSMIC_PROFILE_FLAGGED] = nl[NL_LIST];
		profile[CUSMIC_PROFILE_SEEDS] = list_sums[0];
		profile[CUSMIC_PROFILE_CANDIDATES] = list_sums[1];
		profile[CUSMIC_PROFILE_GROWN] = list_sums[2];
		profile[CUSMIC_PROFILE_WIDENED] = list_sums[3];
	}
	if (*status_word & ERR_NO_DONORS)
		throw std::invalid_argument("no finite replacement donors");
	if (passes)
		for (size_t f = 0; f < nf; ++f)
			passes[f] = detect_passes(f);
}

/* Host path (src/api.cu:116-153): the maps' values are checked on the host,
 * in the reference's order and with its messages, before any kernel runs;
 * maxiter = 0 stays on the CPU. Otherwise the planes are uploaded through the
 * pinned staging buffers, each map checked as it is copied into them (one
 * read of it), and the results go straight into the caller's buffers once
 * the call has succeeded, so that a failed call leaves them untouched. */
void
clean_host(const cusmic_image &im, const cusmic_options &o, size_t n, double *clean_out,
	uint8_t *mask_out)
{
	/* the maps' checks, in the reference's order (src/api.cu:107-114) */
	struct map_check {
		const double *values;
		const char *message;
		int minimum;
	};
	const map_check checks[4] = {
		{im.error, "error must be finite and positive", 1},
		{im.error ? nullptr : im.gain, "gain must be finite and positive", 1},
		{im.error ? nullptr : im.readnoise, "read noise must be finite and nonnegative", 0},
		{im.background, "background must be finite", -1},
	};
	/* without a device to upload to (or with nothing to upload) the maps are
	 * checked first, as the reference checks them before any device work */
	int devices = 0;
	const bool device = hipGetDeviceCount(&devices) == hipSuccess && devices > 0;
	if (!o.maxiter || !device)
		for (const auto &c : checks)
			check_values(c.values, n, c.message, c.minimum);
	if (!o.maxiter) {
		if (clean_out != im.data)
			std::memmove(clean_out, im.data, n * sizeof(double));
		std::memset(mask_out, 0, n);
		if (im.background || !std::isnan(o.background))
			for (size_t i = 0; i < n; ++i) {
				if (!std::isfinite(clean_out[i]))
					continue;
				double bg = im.background ? im.background[i] : o.background;
				clean_out[i] += bg;
				clean_out[i] -= bg;
			}
		return;
	}
	hipStream_t stream = nullptr;
	staging &stage = staging::get();
	std::lock_guard<std::mutex> staged(stage.lock);
	const double *gmap = im.error ? nullptr : im.gain;
	const double *rnmap = im.error ? nullptr : im.readnoise;
	size_t planes = 2 + (im.error != nullptr) + (gmap != nullptr) + (rnmap != nullptr) +
			(im.background != nullptr);
	device_block mem(planes * n * sizeof(double) + 2 * align_up(n, 256) + 512, stream);
	char *base = static_cast<char *>(mem.ptr);
	double *d = reinterpret_cast<double *>(base);
	double *next = d;
	/* the maps are checked as they are copied to the staging buffers (one
	 * read of each); a failed check raises before any kernel runs, and the
	 * caller's buffers are untouched */
	auto plane = [&](const double *src, const map_check *check) -> const double * {
		if (!src)
			return nullptr;
		double *dst = next;
		next += n;
		if (stage.upload(dst, src, n * sizeof(double), stream, check ? check->minimum : NO_CHECK))
			throw std::invalid_argument(check->message);
		return dst;
	};
	cusmic_image gpu{};
	gpu.data = plane(im.data, nullptr);
	gpu.error = plane(im.error, &checks[0]);
	gpu.gain = plane(gmap, &checks[1]);
	gpu.readnoise = plane(rnmap, &checks[2]);
	gpu.background = plane(im.background, &checks[3]);
	double *clean = next;
	next += n;
	uint8_t *mask = reinterpret_cast<uint8_t *>(next);
	uint8_t *excl = mask + align_up(n, 256);
	if (im.mask) {
		stage.upload(excl, im.mask, n, stream);
		gpu.mask = excl;
	}
	gpu.width = im.width;
	gpu.height = im.height;
	clean_images(&gpu, 1, o, nullptr, nullptr, CUSMIC_LAPLACIAN_EXACT, clean, mask, nullptr,
		nullptr, stream);
	stage.download(clean_out, clean, n * sizeof(double), stream);
	stage.download(mask_out, mask, n, stream);
	hip_check(hipStreamSynchronize(stream));
}

} // namespace

extern "C" const char *
cusmic_version(void)
{
	return CUSMIC_VERSION;
}

extern "C" int
cusmic_cuda_version(void)
{
	return HIP_VERSION;
}

extern "C" void
cusmic_default_options(cusmic_options *options)
{
	if (options)
		*options = {3, 5, 3, NAN, 0, NAN, 4, CUSMIC_MIRROR};
}

extern "C" int
remove_cosmics(const cusmic_image *im, const cusmic_options *o, double *clean, uint8_t *mask,
	char *msg, size_t len)
{
	return guarded([&] {
		size_t n = check_arguments(im, 1, o, clean, mask, true);
		clean_host(*im, *o, n, clean, mask);
	}, msg, len);
}

extern "C" int
remove_cosmics_batch_device(const cusmic_image *ims, size_t nf, const cusmic_options *o,
	double *clean, uint8_t *mask, void *stream, char *msg, size_t len)
{
	return guarded([&] {
		check_arguments(ims, nf, o, clean, mask, true);
		clean_images(ims, nf, *o, nullptr, nullptr, CUSMIC_LAPLACIAN_EXACT, clean, mask, nullptr,
			nullptr, static_cast<hipStream_t>(stream));
	}, msg, len);
}

extern "C" int
remove_cosmics_batch_device_noalloc(const cusmic_image *ims, size_t nf, const cusmic_options *o,
	double *clean, uint8_t *mask, void *stream, char *msg, size_t len)
{
	return guarded([&] {
		check_arguments(ims, nf, o, clean, mask, true);
		clean_images(ims, nf, *o, nullptr, nullptr, CUSMIC_LAPLACIAN_EXACT, clean, mask, nullptr,
			nullptr, static_cast<hipStream_t>(stream), nullptr, static_scratch(),
			STATIC_SCRATCH_BYTES);
	}, msg, len);
}

extern "C" int
remove_cosmics_device(const cusmic_image *im, const cusmic_options *o, double *clean,
	uint8_t *mask, void *stream, char *msg, size_t len)
{
	return remove_cosmics_batch_device(im, 1, o, clean, mask, stream, msg, len);
}

extern "C" int
remove_cosmics_ragged_device(const cusmic_image *ims, size_t nf, const cusmic_options *o,
	const cusmic_ragged_options *ext, double *clean, uint8_t *mask, void *stream, char *msg,
	size_t len)
{
	return guarded([&] {
		const double *fg = ext ? ext->frame_gain : nullptr;
		const double *frn = ext ? ext->frame_readnoise : nullptr;
		int lap = ext ? ext->laplacian : CUSMIC_LAPLACIAN_EXACT;
		if (lap != CUSMIC_LAPLACIAN_EXACT && lap != CUSMIC_LAPLACIAN_CLOSE)
			throw std::invalid_argument("invalid Laplacian order");
		check_arguments(ims, nf, o, clean, mask, false, fg, frn);
		clean_images(ims, nf, *o, fg, frn, lap, clean, mask, ext ? ext->donors : nullptr,
			ext ? ext->passes : nullptr, static_cast<hipStream_t>(stream),
			ext ? ext->profile : nullptr, nullptr, 0, ext ? ext->new_pixels : nullptr,
			ext ? ext->new_pixels_size : 0);
	}, msg, len);
}

extern "C" int
cusmic_adjoint_device(const cusmic_image *ims, size_t nf, const uint8_t *mask,
	const int32_t *donors, const double *grad_clean, double *grad_data, void *stream, char *msg,
	size_t len)
{
	return guarded([&] {
		if (!ims || !nf || nf > 65535 || !mask || !donors || !grad_clean || !grad_data)
			throw std::invalid_argument("invalid images, batch size or output buffers");
		for (size_t f = 0; f < nf; ++f)
			if (!ims[f].width || !ims[f].height ||
				ims[f].width > size_t(INT_MAX) / 4 / ims[f].height)
				throw std::invalid_argument("invalid image dimensions");
		hipStream_t s = static_cast<hipStream_t>(stream);
		cusmic_options o;
		cusmic_default_options(&o);
		int ntiles = 0;
		std::vector<Frame> frames = frame_table(ims, nf, o, nullptr, nullptr, &ntiles);
		/* frame table | reach per frame, target count | the first
		 * ADJ_LIST targets of widened windows */
		constexpr unsigned ADJ_LIST = 8192;
		const size_t off_words = align_up(nf * sizeof(Frame), 256);
		const size_t n_words = nf + 1;
		const size_t off_list = align_up(off_words + n_words * 4, 256);
		device_block mem(off_list + ADJ_LIST * 8, s);
		char *base = static_cast<char *>(mem.ptr);
		/* the table goes up from a pinned block, which the target count
		 * comes back to */
		mailbox box(512 + nf * sizeof(Frame), s);
		void *table = static_cast<char *>(box.b.ptr) + 512; /* after the posted words */
		std::memcpy(table, frames.data(), nf * sizeof(Frame));
		hip_check(hipMemcpyAsync(base, table, nf * sizeof(Frame), hipMemcpyHostToDevice, s));
		int *reach = reinterpret_cast<int *>(base + off_words);
		unsigned int *count = reinterpret_cast<unsigned int *>(reach + nf);
		auto *list = reinterpret_cast<unsigned long long *>(base + off_list);
		hipLaunchKernelGGL(zero_words, dim3(unsigned((n_words + 255) / 256)), dim3(256), 0, s,
			reinterpret_cast<unsigned int *>(reach), n_words);
		const Frame *dframes = reinterpret_cast<const Frame *>(base);
		hipLaunchKernelGGL(adjoint_gather<true>, dim3(ntiles), dim3(BLOCK), 0, s, dframes, int(nf),
			mask, donors, grad_clean, grad_data, reach, count, list, ADJ_LIST);
		const unsigned wide_grid = unsigned(device_cus()) * 2;
		hipLaunchKernelGGL(adjoint_wide, dim3(wide_grid), dim3(BLOCK), 0, s, dframes, mask, donors,
			grad_clean, grad_data, reach, count, list, ADJ_LIST);
		hip_check(hipGetLastError());
		box.post_and_wait(count, 1);
		const unsigned n = box.words()[0];
		if (n > ADJ_LIST) {
			/* more targets than the list held: list them all and redo */
			device_block all(size_t(n) * 8 + 256, s);
			auto *big = static_cast<unsigned long long *>(all.ptr);
			unsigned int *count2 = reinterpret_cast<unsigned int *>(big + n);
			hipLaunchKernelGGL(zero_words, dim3(1), dim3(256), 0, s, count2, size_t(1));
			hipLaunchKernelGGL(adjoint_gather<false>, dim3(ntiles), dim3(BLOCK), 0, s, dframes,
				int(nf), mask, donors, grad_clean, grad_data, reach, count2, big, n);
			hipLaunchKernelGGL(adjoint_wide, dim3(wide_grid), dim3(BLOCK), 0, s, dframes, mask,
				donors, grad_clean, grad_data, reach, count2, big, n);
			hip_check(hipGetLastError());
			hip_check(hipStreamSynchronize(s));
		}
	}, msg, len);
}

extern "C" int
cusmic_selftest_arith(const double *a, const double *b, double *quotient, double *root, size_t n,
	void *stream)
{
	return guarded([&] {
		hipStream_t s = static_cast<hipStream_t>(stream);
		hipLaunchKernelGGL(selftest_arith, dim3(256), dim3(256), 0, s, a, b, quotient, root, n);
		hip_check(hipGetLastError());
		hip_check(hipStreamSynchronize(s));
	}, nullptr, 0);
}
