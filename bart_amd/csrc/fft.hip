// fft.hip -- host side of the Stockham axis FFT: radix plans, shared twiddle
// tables, launch configuration, and a plain batched N-d FFT used by the
// exported fft()/ifft() (num/fft.c:387-417 counterparts).
#include "fft.hpp"
#include "volume.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

namespace bam {

static std::mutex tw_mutex;
static std::map<std::pair<int, int>, DevBuf<float2>*> tw_tables;  // (device, n)

static bool factorize(int n, AxisPlan& p)
{
	p.n = n;
	p.nr = 0;
	int m = n;

	auto push = [&](int r) {
		if (p.nr >= FFT_MAX_PASSES)
			return false;
		p.radix[p.nr++] = r;
		m /= r;
		return true;
	};

	while (0 == m % 8)
		if (!push(8)) return false;
	while (0 == m % 4)
		if (!push(4)) return false;
	while (0 == m % 2)
		if (!push(2)) return false;

	const int odd[] = { 3, 5, 7, 11, 13 };

	for (int f : odd)
		while (0 == m % f)
			if (!push(f)) return false;

	return 1 == m;
}

bool fft_size_supported(int n)
{
	AxisPlan p;
	return (n >= 1) && (n <= FFT_MAX_ELEMS) && factorize(n, p);
}

AxisPlan make_axis_plan(int n)
{
	AxisPlan p;

	BAM_CHECK(n >= 1 && n <= FFT_MAX_ELEMS && factorize(n, p),
		  "FFT length %d not supported (lengths up to %d with prime factors 2, 3, 5, 7, 11, 13)", n, FFT_MAX_ELEMS);

	int dev = 0;
	BAM_HIP(hipGetDevice(&dev));

	std::lock_guard<std::mutex> lock(tw_mutex);
	auto key = std::make_pair(dev, n);
	auto it = tw_tables.find(key);

	if (it == tw_tables.end()) {

		std::vector<float2> h(n);

		for (int k = 0; k < n; k++) {

			double a = -2. * M_PI * (double)k / (double)n;
			h[k] = make_float2((float)cos(a), (float)sin(a));
		}

		auto* buf = new DevBuf<float2>(n);
		BAM_HIP(hipMemcpy(buf->ptr, h.data(), n * sizeof(float2), hipMemcpyHostToDevice));
		it = tw_tables.emplace(key, buf).first;
	}

	p.tw = it->second->ptr;
	return p;
}

FftLaunchCfg fft_config(const AxisPlan& plan, long total, bool contig, int P_min, float sign1, float sign2)
{
	FftLaunchCfg c;
	const int n = plan.n;

	// BART_AMD_FFT_FIXED=off: runtime-plan kernels for every length
	static const bool fixed_on = [] {
		const char* e = getenv("BART_AMD_FFT_FIXED");
		return !(nullptr != e && 0 == strcmp(e, "off"));
	}();

	if (fixed_on && ct_supported(n) && ct_block_p(n) >= P_min) {

		const int P = ct_block_p(n);
		int logP = 0;
		while ((1 << logP) < P)
			logP++;

		c.L.plan = plan;
		c.L.P = P;
		c.L.logP = logP;
		c.L.contig = contig;
		c.L.total = total;
		c.L.sign1 = sign1;
		c.L.sign2 = sign2;
		c.L.pitch = n + 1;
		c.blocks = (unsigned)((total + P - 1) / P);
		c.lds_bytes = (size_t)(P * (n + 1) + n) * sizeof(float2);
		c.ept = 0;
		c.fixed = true;
		return c;
	}

	// transforms per block: aim at 2048-4096 elements; strided geometry
	// needs >= 16 adjacent columns for 128-byte coalescing
	int P = 1;

	while (P * n < 2048 && 2 * P * n <= FFT_MAX_ELEMS)
		P *= 2;

	if (!contig)
		while (P < 16 && 2 * P * n <= FFT_MAX_ELEMS)
			P *= 2;

	while (P < P_min)
		P *= 2;

	BAM_CHECK(P * n <= FFT_MAX_ELEMS, "FFT block of %d x %d exceeds LDS budget", P, n);

	int logP = 0;

	while ((1 << logP) < P)
		logP++;

	c.L.plan = plan;
	c.L.P = P;
	c.L.logP = logP;
	c.L.contig = contig;
	c.L.total = total;
	c.L.sign1 = sign1;
	c.L.sign2 = sign2;
	c.blocks = (unsigned)((total + P - 1) / P);
	c.L.pitch = n;
	c.lds_bytes = (size_t)P * n * sizeof(float2);
	c.ept = (P * n <= 2048) ? 8 : ((P * n <= 4096) ? 16 : 32);

	return c;
}

// ---------------------------------------------------------------------------
// Plain batched FFT over the set dims of a contiguous array whose first three
// dims are the transform dims (BART's fft(D, dims, flags, dst, src) with
// flags within 7 and the remaining dims as batch).
// ---------------------------------------------------------------------------

struct LdPlain {
	VolAxis g;
	const float2* src;
	__device__ float2 operator()(long t, int k) const { return src[g.offset(t, k)]; }
};

struct StPlain {
	VolAxis g;
	float2* dst;
	float scale;
	__device__ void store(const float2* lds, const FftLaunch& L, long t0) const
	{
		store_each(lds, L, t0, [&](long t, int k, float2 v) { dst[g.offset(t, k)] = cscale(v, scale); });
	}
};

// x-y slices of square size NX transformed in LDS (rows, then columns),
// work-group per slice; lanes keep their x (NTH a multiple of NX).  src may
// equal dst: a slice is read completely before it is written.
template <int NX>
__global__ __launch_bounds__((1024 / NX) * NX) void k_slice_plain(const float2* src, float2* dst, float sign,
								  float scale, const float2* __restrict__ tw_g)
{
	constexpr int NTH = (1024 / NX) * NX;
	constexpr int DY = NTH / NX;
	constexpr int LD = NX + 1;
	constexpr int NP = NX * NX;

	extern __shared__ float2 lds[];
	float2* tw = lds + NX * LD;

	const long off = (long)blockIdx.x * NP;
	const int x = threadIdx.x % NX;
	const int y0 = threadIdx.x / NX;

	for (int i = threadIdx.x; i < NX; i += NTH)
		tw[i] = tw_g[i];

	for (int y = y0; y < NX; y += DY)
		lds[y * LD + x] = src[off + y * NX + x];

	lds_barrier();

	slice_passes<NX, NX, 1, LD, NTH>(lds, tw, sign);
	slice_passes<NX, NX, LD, 1, NTH>(lds, tw, sign);

	for (int y = y0; y < NX; y += DY)
		dst[off + y * NX + x] = cscale(lds[y * LD + x], scale);
}

static bool slice_plain_size(int n)
{
	return 32 == n || 64 == n || 96 == n || 128 == n;
}

void fft_nd(const int dims[3], long batch, unsigned flags, float sign, float2* dst, const float2* src,
	    float scale, hipStream_t stream)
{
	// square x-y slices of a 3-d array: both in-plane axes in one pass
	if (7 == (flags & 7) && dims[0] == dims[1] && dims[2] > 1 && slice_plain_size(dims[0])) {

		const int n = dims[0];
		AxisPlan plan = make_axis_plan(n);
		const long planes = (long)dims[2] * batch;
		const size_t lds = sizeof(float2) * ((size_t)n * (n + 1) + n);

		{
			ProfScope prof("fft_slice", stream);
			const unsigned nth = (unsigned)((1024 / n) * n);

			switch (n) {
			case 32: hipLaunchKernelGGL(k_slice_plain<32>, dim3(planes), dim3(nth), lds, stream, src, dst, sign, 1.f, plan.tw); break;
			case 64: hipLaunchKernelGGL(k_slice_plain<64>, dim3(planes), dim3(nth), lds, stream, src, dst, sign, 1.f, plan.tw); break;
			case 96: hipLaunchKernelGGL(k_slice_plain<96>, dim3(planes), dim3(nth), lds, stream, src, dst, sign, 1.f, plan.tw); break;
			default: hipLaunchKernelGGL(k_slice_plain<128>, dim3(planes), dim3(nth), lds, stream, src, dst, sign, 1.f, plan.tw); break;
			}

			BAM_HIP(hipGetLastError());
		}

		VolAxis g = VolAxis::make(dims, 2, batch);
		AxisPlan pz = make_axis_plan(dims[2]);
		FftLaunchCfg cfg = fft_config(pz, g.transforms(), g.contig(), 1, sign, sign);
		LdPlain ld{ g, dst };
		StPlain st{ g, dst, scale };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
		return;
	}

	bool first = true;
	int last = -1;

	for (int a = 0; a < 3; a++)
		if ((flags & (1u << a)) && dims[a] > 1)
			last = a;

	if (-1 == last) {

		if (dst != src)
			BAM_HIP(hipMemcpyAsync(dst, src, sizeof(float2) * (size_t)dims[0] * dims[1] * dims[2] * batch,
					       hipMemcpyDeviceToDevice, stream));
		return;
	}

	for (int a = 0; a <= last; a++) {

		if (!((flags & (1u << a)) && dims[a] > 1))
			continue;

		VolAxis g = VolAxis::make(dims, a, batch);
		AxisPlan plan = make_axis_plan(dims[a]);
		FftLaunchCfg cfg = fft_config(plan, g.transforms(), g.contig(), 1, sign, sign);

		LdPlain ld{ g, first ? src : dst };
		StPlain st{ g, dst, (a == last) ? scale : 1.f };

		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
		first = false;
	}
}

}  // namespace bam
