// common.hpp -- shared device/host helpers for the gfx950 NUFFT library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace bam {

// ---------------------------------------------------------------------------
// Errors.  BART's error() (misc/misc.c:95-128) prints and exits unless a
// catcher is armed; the library reproduces that: by default a failure prints
// and exits(1), and with error mode "return" it records the message, and the
// C-ABI entry point returns (NULL for constructors).
// ---------------------------------------------------------------------------

struct Error {
	std::string msg;
};

[[noreturn]] void fail(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

#define BAM_HIP(call)                                                                 \
	do {                                                                          \
		hipError_t e_ = (call);                                               \
		if (hipSuccess != e_)                                                 \
			::bam::fail("HIP error %s at %s:%d: %s", hipGetErrorName(e_), \
				    __FILE__, __LINE__, #call);                       \
	} while (0)

#define BAM_CHECK(cond, ...)                      \
	do {                                      \
		if (!(cond))                      \
			::bam::fail(__VA_ARGS__); \
	} while (0)

// ---------------------------------------------------------------------------
// Device buffer (owning, RAII).
// ---------------------------------------------------------------------------

template <class T> struct DevBuf {

	T* ptr = nullptr;
	size_t n = 0;

	DevBuf() = default;
	explicit DevBuf(size_t count) { alloc(count); }
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
	DevBuf& operator=(DevBuf&& o) noexcept
	{
		if (this != &o) {
			release();
			ptr = o.ptr;
			n = o.n;
			o.ptr = nullptr;
			o.n = 0;
		}
		return *this;
	}
	~DevBuf() { release(); }

	void alloc(size_t count)
	{
		release();
		if (count > 0)
			BAM_HIP(hipMalloc((void**)&ptr, count * sizeof(T)));
		n = count;
	}

	void release()
	{
		if (nullptr != ptr)
			(void)hipFree(ptr);
		ptr = nullptr;
		n = 0;
	}

	void upload(const T* h, size_t count, hipStream_t st)
	{
		BAM_CHECK(count <= n, "upload overflow (%zu > %zu)", count, n);
		BAM_HIP(hipMemcpyAsync(ptr, h, count * sizeof(T), hipMemcpyHostToDevice, st));
	}

	void upload(const std::vector<T>& h, hipStream_t st)
	{
		if (n < h.size())
			alloc(h.size());
		upload(h.data(), h.size(), st);
	}

	size_t bytes() const { return n * sizeof(T); }
};

// ---------------------------------------------------------------------------
// Complex helpers (float2 = complex float, interleaved re/im as in cfl files)
// ---------------------------------------------------------------------------

__host__ __device__ inline float2 cmul(float2 a, float2 b)
{
	return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// a * conj(b)
__host__ __device__ inline float2 cmulc(float2 a, float2 b)
{
	return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Work-group barrier that orders LDS only: a release fence and an acquire
// fence of work-group scope restricted to the "local" (LDS) address space
// around s_barrier.  The compiler orders every LDS access of the thread
// against them and emits s_waitcnt lgkmcnt(0) before the barrier, but no
// wait for vector memory operations, so that global stores (and loads whose
// results are not yet needed) stay in flight across it; __syncthreads()
// waits for those too.  For barriers that only separate LDS writes from LDS
// reads.
__device__ inline void lds_barrier()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
	__builtin_amdgcn_s_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Streams: the library launches on one stream per handle; default = 0.
hipStream_t default_stream();

}  // namespace bam
