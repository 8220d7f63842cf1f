// nufft.hip -- the decomposed 2x-oversampled NUFFT, its Toeplitz normal
// operator and PSF on gfx950.
//
// Algorithm of noncart/nufft.c (decomp = true, os = 2):
//   forward  (:1388-1458): roll -> x linphase_s -> FFT -> x fftmod -> sum_s gridH_s -> x weights
//   adjoint  (:1462-1540): x conj(weights) -> grid_s -> x conj(fftmod) -> IFFT -> sum_s conj(linphase_s) -> roll
//   normal   (:1546-1625): sum_s conj(linphase_s) IFFT(psf_s FFT(linphase_s x))
//   psf      (:727-776):   decompose(fftuc_2n(NUFFT^H_2n(|w|^2; 2 traj), periodic))
// Every element-wise factor is fused into the first or last FFT pass; the
// product with the coil maps (sense/modelnc.c, linops/fmac.c) and the sum
// over coils of S^H are fused there as well.
#include "cg.hpp"
#include "nufft_impl.hpp"
#include "prof.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>

namespace bam {

// ---------------------------------------------------------------------------
// Separable tables (host, FP64 maths)
// ---------------------------------------------------------------------------

// num/fft.c:64-90
double fftmod_phase(long length, long j)
{
	long center1 = length / 2;
	double shift = (double)center1 / (double)length;
	return ((double)j - (double)center1 / 2.) * shift;
}

std::complex<double> fftmod_value(double phase)
{
	double rem = phase - floor(phase);

	if (rem == 0.)
		return 1.;
	if (rem == 0.5)
		return -1.;
	if (rem == 0.25)
		return std::complex<double>(0., 1.);
	if (rem == 0.75)
		return std::complex<double>(0., -1.);

	return std::polar(1., 2. * M_PI * rem);
}

static double bessel_i0_d(double x)
{
	double q = 0.25 * x * x, term = 1., sum = 1.;

	for (int k = 1; k < 500; k++) {
		term *= q / ((double)k * (double)k);
		sum += term;
		if (term < 1.e-17 * sum)
			break;
	}

	return sum;
}

// noncart/grid.c:100-117, 522-531
double rolloff_1d(int d, int i, float os, double beta, double width)
{
	if (1 == d)
		return 1.;

	int od = os * d;
	int oi = i + (od / 2 - d / 2);
	float x = ((float)oi - (float)(od / 2)) / (float)od;

	double i0b = bessel_i0_d(beta);
	double a = beta * beta - pow(M_PI * x * width, 2.);
	double ft;

	if (0. == a)
		ft = 1. / i0b;
	else if (a > 0)
		ft = (sinh(sqrt(a)) / sqrt(a)) / i0b;
	else
		ft = (sin(sqrt(-a)) / sqrt(-a)) / i0b;

	return 1. / ft / width;
}

// x-y product tables, computed on the device with the same cmul as the
// per-element products they replace (bitwise equal)
__global__ void k_xy_tables(SepTables T, float2* out)
{
	const long nxy = (long)T.n[0] * T.n[1];

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < nxy; e += (long)gridDim.x * blockDim.x) {

		const int x = (int)(e % T.n[0]);
		const int y = (int)(e / T.n[0]);

		for (int b = 0; b < 4; b++)
			out[b * nxy + e] = cmul(T.lin[0][b & 1][x], T.lin[1][b >> 1][y]);

		out[4 * nxy + e] = cmul(T.fm[0][x], T.fm[1][y]);
	}
}

static void build_tables(Nufft& p)
{
	long tot = 0;
	long totr = 0;

	for (int d = 0; d < 3; d++) {
		tot += 3 * p.n[d];
		totr += p.n[d];
	}

	std::vector<float2> h2(tot);
	std::vector<float> hr(totr);
	long o2[3], orr[3];
	long a2 = 0, ar = 0;

	int bit = 0;

	for (int d = 0; d < 3; d++) {

		int n = p.n[d];
		o2[d] = a2;
		orr[d] = ar;
		a2 += 3 * n;
		ar += n;

		p.tabs.n[d] = n;
		p.tabs.bitpos[d] = (n > 1) ? bit++ : -1;

		for (int j = 0; j < n; j++) {

			hr[orr[d] + j] = (float)rolloff_1d(n, j, p.os, p.beta, p.width);

			std::complex<double> fm = (n > 1) ? fftmod_value(fftmod_phase(n, j)) : 1.;
			h2[o2[d] + 2 * n + j] = make_float2((float)fm.real(), (float)fm.imag());

			for (int b = 0; b < 2; b++) {

				std::complex<double> v = 1.;

				if (n > 1) {

					// compute_linphases (nufft.c:196-248) with linear_phase
					// (num/filter.c:241-265), one dim at a time
					double shift = -0.5 * b + (n / 2. - n / 2);
					double grad = 2. * M_PI * shift / (double)n;
					double ph = grad * (double)j - grad * (double)n / 2.;
					v = std::polar(1., ph) * fm / sqrt((double)n);
				}

				h2[o2[d] + b * n + j] = make_float2((float)v.real(), (float)v.imag());
			}
		}
	}

	p.tab2.alloc(tot);
	p.tabr.alloc(totr);
	BAM_HIP(hipMemcpy(p.tab2.ptr, h2.data(), sizeof(float2) * tot, hipMemcpyHostToDevice));
	BAM_HIP(hipMemcpy(p.tabr.ptr, hr.data(), sizeof(float) * totr, hipMemcpyHostToDevice));

	for (int d = 0; d < 3; d++) {
		p.tabs.lin[d][0] = p.tab2.ptr + o2[d];
		p.tabs.lin[d][1] = p.tab2.ptr + o2[d] + p.n[d];
		p.tabs.fm[d] = p.tab2.ptr + o2[d] + 2 * p.n[d];
		p.tabs.roll[d] = p.tabr.ptr + orr[d];
	}

	for (int b = 0; b < 4; b++)
		p.tabs.linxy[b] = nullptr;
	p.tabs.fmxy = nullptr;

	if (p.n[0] > 1 && p.n[1] > 1 && p.n[2] > 1) {

		const long nxy = (long)p.n[0] * p.n[1];
		p.tabxy.alloc(5 * nxy);
		hipLaunchKernelGGL(k_xy_tables, dim3((unsigned)std::min<long>((nxy + 255) / 256, 4096)), dim3(256), 0,
				   p.stream, p.tabs, p.tabxy.ptr);
		BAM_HIP(hipGetLastError());
		BAM_HIP(hipStreamSynchronize(p.stream));

		for (int b = 0; b < 4; b++)
			p.tabs.linxy[b] = p.tabxy.ptr + b * nxy;
		p.tabs.fmxy = p.tabxy.ptr + 4 * nxy;
	}
}

__device__ inline float2 lin_val(const SepTables& T, int s, const int c[3])
{
	float2 v = make_float2(1.f, 0.f);

#pragma unroll
	for (int d = 0; d < 3; d++)
		if (T.bitpos[d] >= 0)
			v = cmul(v, T.lin[d][(s >> T.bitpos[d]) & 1][c[d]]);

	return v;
}

__device__ inline float2 fm_val(const SepTables& T, const int c[3])
{
	float2 v = make_float2(1.f, 0.f);

#pragma unroll
	for (int d = 0; d < 3; d++)
		if (T.bitpos[d] >= 0)
			v = cmul(v, T.fm[d][c[d]]);

	return v;
}

__device__ inline float roll_val(const SepTables& T, const int c[3])
{
	return T.roll[0][c[0]] * T.roll[1][c[1]] * T.roll[2][c[2]];
}

__device__ inline long pix_index(const SepTables& T, const int c[3])
{
	return c[0] + (long)T.n[0] * (c[1] + (long)T.n[1] * c[2]);
}

// ---------------------------------------------------------------------------
// Geometry of the first axis with the sub-grid index fastest, then the coils
// of a coil group, so that the S x cc transforms of one image row and coil
// group share a work-group (for the sums over sub-grids and coils).
// t = s + S (cl + cc (g + G o)), coil c = cl + cc g;
// grid offset = (c + C s) N + o n + k.
// ---------------------------------------------------------------------------

struct GeomS {

	VolAxis va;  // first active axis, nvol = S C (grid layout [s][c])
	int S, C, cc, G;
	FastDiv fS, fcc, fG;

	static GeomS make(const int n[3], int axis, int S, int C, int cc)
	{
		GeomS g;
		g.va = VolAxis::make(n, axis, (long)S * C);
		g.S = S;
		g.C = C;
		g.cc = cc;
		g.G = C / cc;
		g.fS = FastDiv((uint32_t)S);
		g.fcc = FastDiv((uint32_t)cc);
		g.fG = FastDiv((uint32_t)g.G);
		return g;
	}

	__device__ inline void split(long t, uint32_t& s, uint32_t& o, uint32_t& c) const
	{
		uint32_t q, q2, cl, gi;
		fS.divmod((uint32_t)t, q, s);
		fcc.divmod(q, q2, cl);
		fG.divmod(q2, o, gi);
		c = cl + (uint32_t)cc * gi;
	}

	__device__ inline long grid_offset(uint32_t s, uint32_t o, uint32_t c, int k) const
	{
		return ((long)c + (long)C * s) * va.N + (long)o * va.n[va.axis] + k;
	}
};

// coils per reduction group: a power of two dividing C with S cc n <= 4096
static int coil_group(int S, int C, int n, bool reduce_coils)
{
	if (!reduce_coils)
		return 1;

	int cc = 1;

	while (0 == C % (2 * cc) && (long)S * 2 * cc * n <= 4096)
		cc *= 2;

	return cc;
}

// ---------------------------------------------------------------------------
// Pass functors
// ---------------------------------------------------------------------------

// First forward pass: image side -> grid, x roll x linphase_s
struct LdImgLin {
	GeomS g;
	SepTables T;
	const float2* src;   // coil images [C][N] or single image [N] (with sens)
	const float2* sens;  // [C][N] or null
	bool roll;

	__device__ float2 operator()(long t, int k) const
	{
		uint32_t s, o, c;
		g.split(t, s, o, c);
		int px[3];
		g.va.coords(0, o, k, px);
		long pi = pix_index(T, px);
		float2 v;

		if (nullptr != sens)
			v = cmul(src[pi], sens[(long)c * g.va.N + pi]);
		else
			v = src[(long)c * g.va.N + pi];

		if (roll)
			v = cscale(v, roll_val(T, px));

		return cmul(v, lin_val(T, (int)s, px));
	}
};

// Plain grid loads in S-grouped order
struct LdGridS {
	GeomS g;
	const float2* grid;

	__device__ float2 operator()(long t, int k) const
	{
		uint32_t s, o, c;
		g.split(t, s, o, c);
		return grid[g.grid_offset(s, o, c, k)];
	}
};

// Grid stores in S-grouped order (optionally x fftmod)
struct StGridS {
	GeomS g;
	SepTables T;
	float2* grid;
	bool fm;

	__device__ void put(long t, int k, float2 v) const
	{
		uint32_t s, o, c;
		g.split(t, s, o, c);
		if (fm) {
			int px[3];
			g.va.coords(0, o, k, px);
			v = cmul(v, fm_val(T, px));
		}
		grid[g.grid_offset(s, o, c, k)] = v;
	}

	__device__ void store(const float2* lds, const FftLaunch& L, long t0) const
	{
		store_each(lds, L, t0, [&](long t, int k, float2 v) { put(t, k, v); });
	}
};

// Standard-order grid loads / stores for the other axes
// Column-block access (used by the fixed-length kernel on strided axes):
// the P transforms of a block are P adjacent columns of one volume when P
// divides the transforms below the axis, so element (p, k) of the block is
// at base + p + inner k.
struct ColBase {
	long off;     // offset of element (0, 0)
	uint32_t i0;  // index below the axis of transform 0
	uint32_t o;   // index above the axis
	uint32_t v;   // volume
};

__device__ inline ColBase col_base(const VolAxis& g, long t0)
{
	ColBase b;
	g.split(t0, b.i0, b.o, b.v);
	b.off = g.offset_iov(b.i0, b.o, b.v, 0);
	return b;
}

struct LdGrid {
	static constexpr bool colblock = true;
	VolAxis g;
	const float2* grid;

	__device__ float2 operator()(long t, int k) const
	{
		uint32_t i, o, v;
		g.split(t, i, o, v);
		return grid[g.offset_iov(i, o, v, k)];
	}

	__device__ bool cb_ok(int P) const { return 0 == g.inner % (uint32_t)P; }
	__device__ ColBase base(long t0) const { return col_base(g, t0); }

	__device__ float2 at(const ColBase& b, int p, int k) const { return grid[b.off + p + (long)g.inner * k]; }
};

struct StGrid {
	static constexpr bool colblock = true;
	VolAxis g;
	SepTables T;
	float2* grid;
	bool fm;

	__device__ void store(const float2* lds, const FftLaunch& L, long t0) const
	{
		store_each(lds, L, t0, [&](long t, int k, float2 x) {
			uint32_t i, o, v;
			g.split(t, i, o, v);
			if (fm) {
				int px[3];
				g.coords(i, o, k, px);
				x = cmul(x, fm_val(T, px));
			}
			grid[g.offset_iov(i, o, v, k)] = x;
		});
	}

	__device__ bool cb_ok(int P) const { return 0 == g.inner % (uint32_t)P; }
	__device__ ColBase base(long t0) const { return col_base(g, t0); }

	__device__ void at(const ColBase& b, int p, int k, float2 x) const
	{
		if (fm) {
			if (2 == g.axis && nullptr != T.fmxy) {
				// fm_val = ((fm0 fm1) fm2), the x-y product from its table
				x = cmul(x, cmul(T.fmxy[b.i0 + p], T.fm[2][k]));
			} else {
				int px[3];
				g.coords(b.i0 + p, b.o, k, px);
				x = cmul(x, fm_val(T, px));
			}
		}

		grid[b.off + p + (long)g.inner * k] = x;
	}
};

// PSF multiply between forward and inverse transform (standard order,
// volumes v = c + C s)
struct MdPsf {
	static constexpr bool active = true;
	static constexpr bool colblock = true;
	VolAxis g;
	int C;
	const float2* psf;

	__device__ float2 operator()(long t, int k, float2 x) const
	{
		uint32_t i, o, v;
		g.split(t, i, o, v);
		uint32_t s = v / (uint32_t)C;
		long off = g.offset_iov(i, o, 0, k);
		return cmul(x, psf[(long)s * g.N + off]);
	}

	__device__ bool cb_ok(int P) const { return 0 == g.inner % (uint32_t)P; }

	// base: the PSF offset of element (0, 0) (sub-grid v / C)
	__device__ ColBase base(long t0) const
	{
		ColBase b = col_base(g, t0);
		b.off = (long)(b.v / (uint32_t)C) * g.N + g.offset_iov(b.i0, b.o, 0, 0);
		return b;
	}

	__device__ float2 at(const ColBase& b, int p, int k, float2 x) const
	{
		return cmul(x, psf[b.off + p + (long)g.inner * k]);
	}
};

// Toeplitz z-pass over pairs of sub-grids s, s + S/2 (the same x-y
// transform H, z factors lz_0 and lz_1): the stack holds the volumes
// v = c + C s, s < S/2, the PSF all S sub-grids
struct MdPsfPair {
	VolAxis g;
	int C;
	long half;  // PSF offset of sub-grid s + S/2 from sub-grid s: (S/2) N
	const float2* psf;
	const float2* lz0;
	const float2* lz1;

	__device__ bool cb_ok(int P) const { return 0 == g.inner % (uint32_t)P; }

	__device__ ColBase base(long t0) const
	{
		ColBase b = col_base(g, t0);
		b.off = (long)(b.v / (uint32_t)C) * g.N + g.offset_iov(b.i0, b.o, 0, 0);
		return b;
	}

	__device__ float2 at(const ColBase& b, int m, int p, int k, float2 x) const
	{
		return cmul(x, psf[b.off + m * half + p + (long)g.inner * k]);
	}

	__device__ float2 zf(int m, int k) const { return (0 == m) ? lz0[k] : lz1[k]; }
};

// PSF multiply in the S-grouped geometry (single active axis)
struct MdPsfS {
	static constexpr bool active = true;
	GeomS g;
	const float2* psf;

	__device__ float2 operator()(long t, int k, float2 x) const
	{
		uint32_t s, o, c;
		g.split(t, s, o, c);
		return cmul(x, psf[(long)s * g.va.N + (long)o * g.va.n[g.va.axis] + k]);
	}
};

// Last pass: sum_s conj(linphase_s) x (roll) -> image side.  With coil maps
// the block also sums conj(sens_c) x over the cc coils of its group, in a
// fixed order, and writes the group's partial image dst[g N + pixel] (the
// groups are summed afterwards, also in a fixed order); without maps it
// stores coil image c (cc = 1).  No atomics: the result is deterministic.
struct StReduceS {
	GeomS g;
	SepTables T;
	float2* dst;
	const float2* sens;
	bool roll;

	__device__ void store(const float2* lds, const FftLaunch& L, long t0) const
	{
		const int n = L.plan.n;
		const int S = g.S;
		const int SC = S * g.cc;
		const int groups = L.P / SC;

		for (int i = threadIdx.x; i < groups * n; i += FFT_THREADS) {

			int gi = i / n;
			int k = i - gi * n;
			long t = t0 + (long)gi * SC;

			if (t >= L.total)
				continue;

			uint32_t s, o, c;
			g.split(t, s, o, c);
			int px[3];
			g.va.coords(0, o, k, px);
			long pi = pix_index(T, px);
			const float r = roll ? roll_val(T, px) : 1.f;

			float2 out = make_float2(0.f, 0.f);

			for (int cl = 0; cl < g.cc; cl++) {

				float2 acc = make_float2(0.f, 0.f);

				for (int j = 0; j < S; j++)
					acc = cadd(acc, cmulc(lds[(gi * SC + cl * S + j) * L.pitch + k], lin_val(T, j, px)));

				if (roll)
					acc = cscale(acc, r);

				if (nullptr != sens)
					out = cadd(out, cmulc(acc, sens[(long)(c + cl) * g.va.N + pi]));
				else
					out = acc;
			}

			if (nullptr != sens)
				dst[(long)(c / g.cc) * g.va.N + pi] = out;
			else
				dst[(long)c * g.va.N + pi] = out;
		}
	}
};

// ---------------------------------------------------------------------------
// Slice kernels for 3-D grids whose x-y slice fits in LDS (NX x NY <= 128 x
// 128): one work-group transforms whole slices along x and y, so that the
// sub-grid stack makes one round trip through HBM for both in-plane axes.
// LDS holds the slice row-major with a row pitch of NX + 1 (bank spread of
// the column transforms) and the two twiddle tables.
//
// Sub-grids s and s + S/2 differ only in the z shift (the z decomposition
// bit is the top bit of s), and the z factor of linphase, lz_s(z), is one
// scalar per slice.  The x-y transform is linear, so
//   FFT_xy(v linxy_s lz_s(z)) = lz_s(z) FFT_xy(v linxy_s)
// and one transform serves both sub-grids of a pair; likewise the inverse
// transforms the combination conj(lz_s) g_s + conj(lz_s') g_s' once.
//
// k_slice_fwd: work-group (z, c).  The image-side values of slice z (coil
// image c, or image x map c; x rolloff when roll) are read once into
// registers; for every pair (s, s + S/2) they are multiplied by the x-y
// factor linxy_s, transformed along x and y, and stored to grid[s][c][z]
// times lz_s(z) and to grid[s + S/2][c][z] times lz_{s+S/2}(z).
//
// k_slice_inv: work-group (z, c).  For every pair, conj(lz_s) grid[s][c][z]
// + conj(lz_{s+S/2}) grid[s + S/2][c][z] is inverse-transformed along y and
// x, and conj(linxy_s) x value is summed over the pairs in order; the sum is
// scaled by the rolloff (when roll) and, with maps, multiplied by
// conj(map_c) into partial image c (the coils are summed in order by
// sum_groups), else stored as coil image c.  No atomics.
// ---------------------------------------------------------------------------

// threads of a slice work-group: a multiple of NX (every lane keeps its x)
template <int NX>
constexpr int slice_threads()
{
	return (1024 / NX) * NX;
}

template <int NX, int NY>
constexpr int slice_threads_inv()
{
	return slice_threads<NX>();
}

struct SliceGeom {
	int nz, S, C;
	long N;
};

// the first pass along y runs in registers when a lane's points y0 + i DY
// (i < EPT) are exactly the inputs of one radix-EPT butterfly
template <int NX, int NY>
constexpr bool slice_regpass()
{
	constexpr int DY = slice_threads<NX>() / NX;
	constexpr int EPT = NY / DY;
	return 0 == NY % DY && (2 == EPT || 4 == EPT || 8 == EPT || 16 == EPT);
}

template <int NX, int NY>
constexpr size_t slice_lds_bytes()
{
	return sizeof(float2) * ((size_t)NY * (NX + 1) + NX + NY);
}

// separable factors of one lane: x and z fixed, y = y0 + i dy
struct SliceLane {
	int x, y0, dy;
};

// x-y factor of sub-grid s's linphase, lin_x lin_y, from its table
__device__ inline float2 lin_xy(const SepTables& T, int s, int x, int y)
{
	const int bx = (s >> T.bitpos[0]) & 1, by = (s >> T.bitpos[1]) & 1;
	return T.linxy[bx + 2 * by][x + T.n[0] * y];
}

__device__ inline float roll_xyz(const SepTables& T, int x, int y, int z)
{
	return T.roll[0][x] * T.roll[1][y] * T.roll[2][z];
}

template <int NX, int NY>
__global__ __launch_bounds__(slice_threads<NX>()) void k_slice_fwd(SliceGeom sg, SepTables T,
								    const float2* __restrict__ src,
								    const float2* __restrict__ sens, bool roll,
								    float2* __restrict__ grid,
								    const float2* __restrict__ twx_g,
								    const float2* __restrict__ twy_g, const int* skip,
								    bool zpair)
{
	constexpr int NTH = slice_threads<NX>();
	constexpr int DY = NTH / NX;
	constexpr int EPT = (NY + DY - 1) / DY;
	constexpr int NP = NX * NY;
	constexpr int LD = NX + 1;

	extern __shared__ float2 lds[];
	float2* twx = lds + NY * LD;
	float2* twy = twx + NX;

	if (nullptr != skip && 0 != *skip)
		return;

	const int z = blockIdx.x;
	const int c = blockIdx.y;
	const int x = threadIdx.x % NX;
	const int y0 = threadIdx.x / NX;

	for (int i = threadIdx.x; i < NX; i += NTH)
		twx[i] = twx_g[i];

	for (int i = threadIdx.x; i < NY; i += NTH)
		twy[i] = twy_g[i];

	const long zoff = (long)z * NP;
	const float2* img = (nullptr != sens) ? src + zoff : src + (long)c * sg.N + zoff;
	const float2* map = (nullptr != sens) ? sens + (long)c * sg.N + zoff : nullptr;

	// image-side values of this lane's points, held across the sub-grids
	float2 vin[EPT];

#pragma unroll
	for (int i = 0; i < EPT; i++) {

		const int y = y0 + i * DY;
		vin[i] = make_float2(0.f, 0.f);

		if (0 == NY % DY || y < NY) {

			const int e = y * NX + x;
			float2 v = (nullptr != map) ? cmul(img[e], map[e]) : img[e];

			if (roll)
				v = cscale(v, roll_xyz(T, x, y, z));

			vin[i] = v;
		}
	}

	// z factors of the two sub-grids of a pair; sub-grid s + SH is s shifted in z
	const float2 lz0 = T.lin[2][0][z], lz1 = T.lin[2][1][z];
	const int SH = sg.S / 2;
	const long pair_off = (long)sg.C * SH * sg.N;

	for (int s = 0; s < SH; s++) {

		// per-iteration copies of the lane coordinates (keeps the element
		// addresses out of registers across the loop)
		int xs = x, ys = y0;
		asm volatile("" : "+v"(xs), "+v"(ys));

		if constexpr (slice_regpass<NX, NY>()) {

			// the lane's points y0 + i DY are the inputs of one butterfly of
			// the first radix-EPT pass along y: that pass runs in registers,
			// then the rest of y and the x passes in LDS
			float2 v[EPT];

#pragma unroll
			for (int i = 0; i < EPT; i++)
				v[i] = cmul(vin[i], lin_xy(T, s, xs, ys + i * DY));

			dft_any<EPT>(v, -1.f, twy, NY);

			lds_barrier();

#pragma unroll
			for (int q = 0; q < EPT; q++)
				lds[(ys * EPT + q) * LD + xs] = v[q];

			lds_barrier();

			slice_passes_rest<NY, NY / EPT, EPT, NX, LD, 1, NTH>(lds, twy, -1.f);

			// x passes; the last one stores its outputs straight to the grid
			// (16 lanes per row: 128-byte runs)
			constexpr int RLX = slice_last_radix(NX);
			slice_passes_range<NX, 1, NX / RLX, NY, 1, LD, NTH>(lds, twx, -1.f);
			float2* out = grid + ((long)c + (long)sg.C * s) * sg.N + zoff;
			slice_last_pass_rows<NX, RLX, NY, LD, NTH>(lds, twx, -1.f, [&](int, int, int p, int xo, float2 v) {
				if (zpair) {
					out[p * NX + xo] = v;
				} else {
					out[p * NX + xo] = cmul(v, lz0);
					out[pair_off + p * NX + xo] = cmul(v, lz1);
				}
			});
			continue;

		} else {

			lds_barrier();

#pragma unroll
			for (int i = 0; i < EPT; i++) {

				const int y = ys + i * DY;

				if (0 == NY % DY || y < NY)
					lds[y * LD + xs] = cmul(vin[i], lin_xy(T, s, xs, y));
			}

			lds_barrier();

			slice_passes<NX, NY, 1, LD, NTH>(lds, twx, -1.f);
			slice_passes<NY, NX, LD, 1, NTH>(lds, twy, -1.f);
		}

		float2* out = grid + ((long)c + (long)sg.C * s) * sg.N + zoff;

		xs = x;
		ys = y0;
		asm volatile("" : "+v"(xs), "+v"(ys));

#pragma unroll 4
		for (int i = 0; i < EPT; i++) {

			const int y = ys + i * DY;

			if (0 == NY % DY || y < NY) {
				const float2 l = lds[y * LD + xs];
				if (zpair) {
					out[y * NX + xs] = l;
				} else {
					out[y * NX + xs] = cmul(l, lz0);
					out[pair_off + y * NX + xs] = cmul(l, lz1);
				}
			}
		}
	}
}

template <int NX, int NY>
__global__ __launch_bounds__((slice_threads_inv<NX, NY>())) void k_slice_inv(SliceGeom sg, SepTables T,
								    const float2* __restrict__ grid,
								    const float2* __restrict__ sens, bool roll,
								    float2* __restrict__ dst,
								    const float2* __restrict__ twx_g,
								    const float2* __restrict__ twy_g, const int* skip,
								    bool zpair)
{
	constexpr int NTH = slice_threads_inv<NX, NY>();
	constexpr int DY = NTH / NX;
	constexpr int EPT = (NY + DY - 1) / DY;
	constexpr int NP = NX * NY;
	constexpr int LD = NX + 1;

	extern __shared__ float2 lds[];
	float2* twx = lds + NY * LD;
	float2* twy = twx + NX;

	if (nullptr != skip && 0 != *skip)
		return;

	const int z = blockIdx.x;
	const int c = blockIdx.y;
	const long zoff = (long)z * NP;
	const int x = threadIdx.x % NX;
	const int y0 = threadIdx.x / NX;

	for (int i = threadIdx.x; i < NX; i += NTH)
		twx[i] = twx_g[i];

	for (int i = threadIdx.x; i < NY; i += NTH)
		twy[i] = twy_g[i];

	float2 acc[EPT];

#pragma unroll
	for (int i = 0; i < EPT; i++)
		acc[i] = make_float2(0.f, 0.f);

	// z factors of the two sub-grids of a pair; sub-grid s + SH is s shifted in z
	const float2 lz0 = T.lin[2][0][z], lz1 = T.lin[2][1][z];
	const int SH = sg.S / 2;
	const long pair_off = (long)sg.C * SH * sg.N;

	for (int s = 0; s < SH; s++) {

		const float2* in = grid + ((long)c + (long)sg.C * s) * sg.N + zoff;

		// per-iteration copies of the lane coordinates (keeps the element
		// addresses out of registers across the loop)
		int xs = x, ys = y0;
		asm volatile("" : "+v"(xs), "+v"(ys));

		if constexpr (slice_regpass<NX, NY>()) {

			// first radix-EPT pass along y in registers (see k_slice_fwd)
			float2 v[EPT];

#pragma unroll
			for (int i = 0; i < EPT; i++) {
				const int e = (ys + i * DY) * NX + xs;
				v[i] = zpair ? in[e] : cadd(cmulc(in[e], lz0), cmulc(in[pair_off + e], lz1));
			}

			dft_any<EPT>(v, +1.f, twy, NY);

			lds_barrier();

#pragma unroll
			for (int q = 0; q < EPT; q++)
				lds[(ys * EPT + q) * LD + xs] = v[q];

			lds_barrier();

			slice_passes_rest<NY, NY / EPT, EPT, NX, LD, 1, NTH>(lds, twy, +1.f);

		} else {

			lds_barrier();

#pragma unroll 4
			for (int i = 0; i < EPT; i++) {

				const int y = ys + i * DY;

				if (0 == NY % DY || y < NY)
					lds[y * LD + xs] = zpair ? in[y * NX + xs]
								 : cadd(cmulc(in[y * NX + xs], lz0),
									cmulc(in[pair_off + y * NX + xs], lz1));
			}

			lds_barrier();

			slice_passes<NY, NX, LD, 1, NTH>(lds, twy, +1.f);
		}

		slice_passes<NX, NY, 1, LD, NTH>(lds, twx, +1.f);

		// sum over the pairs of conj(linxy_s) x value
		xs = x;
		ys = y0;
		asm volatile("" : "+v"(xs), "+v"(ys));

#pragma unroll
		for (int i = 0; i < EPT; i++) {

			const int y = ys + i * DY;

			if (0 == NY % DY || y < NY)
				acc[i] = cadd(acc[i], cmulc(lds[y * LD + xs], lin_xy(T, s, xs, y)));
		}
	}

	// rolloff, then conj(map_c) into partial image c, or coil image c
	const float2* map = (nullptr != sens) ? sens + (long)c * sg.N + zoff : nullptr;
	int xs = x, ys = y0;
	asm volatile("" : "+v"(xs), "+v"(ys));

#pragma unroll
	for (int i = 0; i < EPT; i++) {

		const int y = ys + i * DY;

		if (0 == NY % DY || y < NY) {

			float2 v = roll ? cscale(acc[i], roll_xyz(T, xs, y, z)) : acc[i];

			if (nullptr != map)
				v = cmulc(v, map[y * NX + xs]);

			dst[(long)c * sg.N + zoff + y * NX + xs] = v;
		}
	}
}

// slice sizes with a kernel (square slices)
static bool slice_supported(const int n[3])
{
	if (n[0] != n[1] || n[2] < 2)
		return false;

	switch (n[0]) {
	case 32: case 64: case 96: case 128:
		return true;
	default:
		return false;
	}
}

// the slice kernels pair sub-grid s with s + S/2: the z decomposition bit
// must be the top bit of s
static void check_slice_pairs(const Nufft& p)
{
	if (8 != p.S || 0 != p.tabs.bitpos[0] || 1 != p.tabs.bitpos[1] || 2 != p.tabs.bitpos[2])
		fail("slice FFT: needs the 3-D decomposition into 8 sub-grids (S = %d)", p.S);
}

// zpair: store only the x-y transform H of each pair (volumes s < S/2),
// without the z factors, for the pair z-pass of the Toeplitz operator
static void launch_slice_fwd(const Nufft& p, const float2* src, const float2* sens, bool roll, const int* skip,
			     const char* name, bool zpair = false)
{
	ProfScope prof(name, p.stream);
	SliceGeom sg{ p.n[2], p.S, p.C, p.N };
	dim3 blocks(p.n[2], p.C);
	check_slice_pairs(p);

#define BAM_SLICE(NN)                                                                                           	case NN:                                                                                                		hipLaunchKernelGGL((k_slice_fwd<NN, NN>), blocks, dim3(slice_threads<NN>()), (slice_lds_bytes<NN, NN>()), 				   p.stream, sg, p.tabs, src, sens, roll, p.grid.ptr, p.axis[0].tw, p.axis[1].tw,   				   skip, zpair);                                                                       		break

	switch (p.n[0]) {
		BAM_SLICE(32);
		BAM_SLICE(64);
		BAM_SLICE(96);
		BAM_SLICE(128);
	default:
		fail("slice FFT: unsupported size %d", p.n[0]);
	}
#undef BAM_SLICE

	BAM_HIP(hipGetLastError());
}

// zpair: the volumes s < S/2 already hold conj(lz_s) g_s + conj(lz_s') g_s'
// (pair z-pass output)
static void launch_slice_inv(const Nufft& p, float2* dst, const float2* sens, bool roll, const int* skip,
			     const char* name, bool zpair = false)
{
	ProfScope prof(name, p.stream);
	SliceGeom sg{ p.n[2], p.S, p.C, p.N };
	dim3 blocks(p.n[2], p.C);
	check_slice_pairs(p);

#define BAM_SLICE(NN)                                                                                           	case NN:                                                                                                		hipLaunchKernelGGL((k_slice_inv<NN, NN>), blocks, dim3(slice_threads_inv<NN, NN>()), (slice_lds_bytes<NN, NN>()), 				   p.stream, sg, p.tabs, p.grid.ptr, sens, roll, dst, p.axis[0].tw,                 				   p.axis[1].tw, skip, zpair);                                                         		break

	switch (p.n[0]) {
		BAM_SLICE(32);
		BAM_SLICE(64);
		BAM_SLICE(96);
		BAM_SLICE(128);
	default:
		fail("slice FFT: unsupported size %d", p.n[0]);
	}
#undef BAM_SLICE

	BAM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Small element-wise kernels
// ---------------------------------------------------------------------------

__global__ void k_fill(float2* p, long n, float2 v, const int* skip)
{
	if (nullptr != skip && 0 != *skip)
		return;

	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
		p[i] = v;
}

// dst[i] = sum_g part[g N + i], g = 0 .. G-1 in order
__global__ void k_sum_groups(float2* __restrict__ dst, const float2* __restrict__ part, long N, int G,
			     const int* skip)
{
	if (nullptr != skip && 0 != *skip)
		return;

	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {

		float2 a = part[i];

		for (int g = 1; g < G; g++)
			a = cadd(a, part[(long)g * N + i]);

		dst[i] = a;
	}
}

static void sum_groups(float2* dst, const float2* part, long N, int G, hipStream_t st, const int* skip)
{
	ProfScope prof("sum_coil_groups", st);
	unsigned blocks = (unsigned)std::min<long>((N + 255) / 256, 4096);
	hipLaunchKernelGGL(k_sum_groups, dim3(blocks), dim3(256), 0, st, dst, part, N, G, skip);
	BAM_HIP(hipGetLastError());
}

__global__ void k_axpy(float2* __restrict__ d, const float2* __restrict__ x, float a, long n, const int* skip)
{
	if (nullptr != skip && 0 != *skip)
		return;

	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
		const float2 v = x[i];
		d[i] = make_float2(d[i].x + a * v.x, d[i].y + a * v.y);
	}
}

void axpy(float2* d, const float2* x, float a, long n, hipStream_t st, const int* skip)
{
	const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 8192);
	hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(256), 0, st, d, x, a, n, skip);
	BAM_HIP(hipGetLastError());
}

void sum_images(float2* dst, const float2* part, long N, int G, hipStream_t st, const int* skip)
{
	sum_groups(dst, part, N, G, st, skip);
}

void fill(float2* p, long n, float2 v, hipStream_t st, const int* skip)
{
	if (n <= 0)
		return;
	ProfScope prof("fill", st);
	unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 8192);
	hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, p, n, v, skip);
	BAM_HIP(hipGetLastError());
}

static FftLaunchCfg fft_config_skip(const int* skip, const AxisPlan& plan, long total, bool contig, int P_min,
				   float sign1, float sign2)
{
	FftLaunchCfg c = fft_config(plan, total, contig, P_min, sign1, sign2);
	c.L.skip = skip;
	c.name = "fft_toeplitz";
	return c;
}

// ---------------------------------------------------------------------------
// Plan creation
// ---------------------------------------------------------------------------

void Nufft::create(const int dims[3], int C_, long M_, const float* traj, const float2* weights_host,
		   bool toeplitz_, bool periodic_, float width_, float os_, bool compute_psf)
{
	create_geometry(dims, C_, toeplitz_, periodic_, width_, os_);
	set_traj(M_, traj, weights_host, compute_psf);
}

void Nufft::create_geometry(const int dims[3], int C_, bool toeplitz_, bool periodic_, float width_, float os_)
{
	BAM_CHECK(2.f == os_, "only oversampling 2 is supported (got %f)", os_);
	BAM_CHECK(width_ > 0.f && width_ / 2.f <= (float)(MAX_TAPS - 1), "kernel width %f not supported", width_);

	N = 1;
	nact = 0;
	S = 1;

	for (int d = 0; d < 3; d++) {

		n[d] = dims[d];
		BAM_CHECK(n[d] >= 1, "bad image dims");
		N *= n[d];

		if (n[d] > 1) {
			BAM_CHECK(fft_size_supported(n[d]), "image dim %d not supported by the FFT", n[d]);
			act[nact++] = d;
			S *= 2;
		}
	}

	BAM_CHECK(nact >= 1, "NUFFT needs at least one non-trivial image dim");

	C = C_;
	M = 0;
	width = width_;
	os = os_;
	// BART's KB table is built by the first kb_init call, which
	// rolloff_correction makes with beta passed as float (noncart/nufft.c:898,
	// noncart/grid.c:537-540); the table, I0(beta) and the rolloff all use
	// the float-rounded beta
	beta = (double)(float)calc_beta(os, width);
	periodic = periodic_;
	toeplitz = toeplitz_;

	std::vector<float> kbh(KB_SIZE + 1);
	kb_table_host(beta, kbh.data());
	kb.alloc(KB_SIZE + 1);
	BAM_HIP(hipMemcpy(kb.ptr, kbh.data(), sizeof(float) * (KB_SIZE + 1), hipMemcpyHostToDevice));

	build_tables(*this);

	for (int j = 0; j < nact; j++)
		axis[j] = make_axis_plan(n[act[j]]);

	// BART_AMD_SLICE=off selects the per-axis passes also where the x-y
	// slices fit in LDS
	const char* sl = getenv("BART_AMD_SLICE");
	slice = (3 == nact) && slice_supported(n) && !(nullptr != sl && 0 == strcmp(sl, "off"));

	// BART_AMD_PAIRZ=off: Toeplitz z-pass over all S sub-grids instead of
	// over the pairs that share an x-y transform
	const char* pz = getenv("BART_AMD_PAIRZ");
	pair_zpass = slice && fft_pair_len_ok(n[2]) && !(nullptr != pz && 0 == strcmp(pz, "off"));

	grid.alloc((size_t)S * C * N);

	// partial images of the coil groups of S^H (allocated here so that no
	// allocation happens inside a captured CG iteration)
	// partial images of the S^H sum: one per coil with the slice kernels
	const int G = slice ? C : C / coil_group(S, C, n[act[0]], true);

	if (G > 1)
		part.alloc((size_t)G * N);
}

// Trajectory, weights, gridding plan and (Toeplitz) PSF -- nufft_set_traj
// (noncart/nufft.c:1002-1159).  Called again by nufft_update_traj: the
// sample count stays that of the k-space dims.
void Nufft::set_traj(long M_, const float* traj, const float2* weights_host, bool compute_psf)
{
	M = M_;
	has_traj = false;
	has_psf = false;
	has_weights = (nullptr != weights_host);

	if (has_weights) {
		weights.alloc(std::max<long>(M, 1));
		if (M > 0)
			BAM_HIP(hipMemcpy(weights.ptr, weights_host, sizeof(float2) * M, hipMemcpyHostToDevice));
	}

	// gridding geometry of the decomposition (nufft.c:147-188)
	GridGeom g = {};

	for (int d = 0; d < 3; d++)
		g.n[d] = n[d];

	g.N = N;
	g.C = C;
	g.M = M;
	g.S = S;
	g.os = 1.f;
	g.width = width / 2.f;
	g.periodic = periodic;

	for (int d = 0; d < 3; d++)
		g.fm[d] = tabs.fm[d];

	for (int s = 0; s < S; s++) {

		int idx = s;

		for (int d = 0; d < 3; d++) {

			int f = (n[d] > 1) ? 2 : 1;
			float sh = -(float)(idx % f) / f;
			idx /= f;

			if (f > 1)
				sh += (n[d] / 2.0 - n[d] / 2) / g.os;

			g.shift[s][d] = sh;
		}
	}

	gp.build(g, traj);
	has_traj = true;
	if (tune)
		tune_grid_adjoint(gp, kb.ptr, grid.ptr, stream);

	if (toeplitz && compute_psf)
		compute_psf_from_traj(traj, weights_host, true);
}

// ---------------------------------------------------------------------------
// Forward / adjoint / normal
// ---------------------------------------------------------------------------

// z axis of the sub-grid stack (slice path).  pair: the stack is read
// (forward) or written (adjoint) as the S/2 pair volumes of the slice
// kernels' zpair mode -- forward: volume s holds H, the x-y transform shared
// by sub-grids s and s + S/2, and both are written; adjoint: sub-grids s and
// s + S/2 are read and conj(lz_s) Z_s + conj(lz_{s+S/2}) Z_{s+S/2} is
// written to volume s.  Returns whether the pair form was used.
static bool z_pass(const Nufft& p, float sign, bool fm, const char* name)
{
	if (p.pair_zpass) {
		VolAxis va = VolAxis::make(p.n, 2, (long)(p.S / 2) * p.C);
		FftLaunchCfg cfg = fft_config(p.axis[2], va.transforms(), va.contig(), 1, sign, sign);

		if (cfg.fixed) {
			fft_config_pair(cfg);

			if (!cfg.L.contig && 0 == va.inner % (uint32_t)cfg.L.P) {
				cfg.name = name;
				LdGrid ld{ va, p.grid.ptr };
				StGrid st{ va, p.tabs, p.grid.ptr, fm };
				const long pair_off = (long)(p.S / 2) * p.C * p.N;
				if (sign < 0.f)
					launch_fft_axis_pair_io<false>(cfg, p.stream, ld, st, p.tabs.lin[2][0], p.tabs.lin[2][1],
								       pair_off);
				else
					launch_fft_axis_pair_io<true>(cfg, p.stream, ld, st, p.tabs.lin[2][0], p.tabs.lin[2][1],
								      pair_off);
				return true;
			}
		}
	}

	VolAxis va = VolAxis::make(p.n, 2, (long)p.S * p.C);
	FftLaunchCfg cfg = fft_config(p.axis[2], va.transforms(), va.contig(), 1, sign, sign);
	cfg.name = name;
	LdGrid ld{ va, p.grid.ptr };
	StGrid st{ va, p.tabs, p.grid.ptr, fm };
	launch_fft_axis(cfg, p.stream, ld, NoMid{}, st);
	return false;
}

// whether z_pass takes the pair form (same conditions)
static bool z_pass_pairs(const Nufft& p)
{
	if (!p.pair_zpass)
		return false;
	VolAxis va = VolAxis::make(p.n, 2, (long)(p.S / 2) * p.C);
	FftLaunchCfg cfg = fft_config(p.axis[2], va.transforms(), va.contig(), 1, -1.f, -1.f);
	if (!cfg.fixed)
		return false;
	fft_config_pair(cfg);
	return !cfg.L.contig && 0 == va.inner % (uint32_t)cfg.L.P;
}

void Nufft::forward(float2* ksp, const float2* src, const float2* sens)
{
	if (slice) {
		const bool pz = z_pass_pairs(*this);
		launch_slice_fwd(*this, src, sens, true, nullptr, "fft_fwd_slice", pz);
		const bool used = z_pass(*this, -1.f, true, "fft_fwd_axis");
		BAM_CHECK(used == pz, "z-pass form mismatch");
		grid_forward(gp, kb.ptr, ksp, grid.ptr, has_weights ? weights.ptr : nullptr, stream);
		return;
	}

	const int a0 = act[0];

	// first axis: image side -> grid (x roll x linphase), FFT
	{
		int cc = coil_group(S, C, n[a0], nullptr != sens);
		GeomS gs = GeomS::make(n, a0, S, C, cc);
		FftLaunchCfg cfg = fft_config(axis[0], gs.va.transforms(), true, S * cc, -1.f, -1.f);
		cfg.name = "fft_fwd_first";
		LdImgLin ld{ gs, tabs, src, sens, true };
		StGridS st{ gs, tabs, grid.ptr, 1 == nact };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
	}

	for (int j = 1; j < nact; j++) {

		VolAxis va = VolAxis::make(n, act[j], (long)S * C);
		FftLaunchCfg cfg = fft_config(axis[j], va.transforms(), va.contig(), 1, -1.f, -1.f);
		cfg.name = "fft_fwd_axis";
		LdGrid ld{ va, grid.ptr };
		StGrid st{ va, tabs, grid.ptr, j == nact - 1 };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
	}

	grid_forward(gp, kb.ptr, ksp, grid.ptr, has_weights ? weights.ptr : nullptr, stream);
}

void Nufft::adjoint(float2* dst, const float2* ksp, const float2* sens)
{
	// the output-stationary adjoint gridding writes every grid point
	grid_adjoint(gp, kb.ptr, grid.ptr, ksp, has_weights ? weights.ptr : nullptr, stream);

	if (slice) {

		const bool pz = z_pass(*this, +1.f, false, "fft_adj_axis");

		// with maps: one partial image per coil, summed in coil order
		const int G = C;
		float2* out = (nullptr != sens && G > 1) ? part.ptr : dst;

		launch_slice_inv(*this, out, sens, true, nullptr, "fft_adj_slice", pz);

		if (out != dst)
			sum_groups(dst, part.ptr, N, G, stream, nullptr);

		return;
	}

	for (int j = nact - 1; j >= 1; j--) {

		VolAxis va = VolAxis::make(n, act[j], (long)S * C);
		FftLaunchCfg cfg = fft_config(axis[j], va.transforms(), va.contig(), 1, +1.f, +1.f);
		cfg.name = "fft_adj_axis";
		LdGrid ld{ va, grid.ptr };
		StGrid st{ va, tabs, grid.ptr, false };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
	}

	const int cc = coil_group(S, C, n[act[0]], nullptr != sens);
	const int G = C / cc;
	GeomS gs = GeomS::make(n, act[0], S, C, cc);
	FftLaunchCfg cfg = fft_config(axis[0], gs.va.transforms(), true, S * cc, +1.f, +1.f);
	cfg.name = "fft_adj_last";
	LdGridS ld{ gs, grid.ptr };
	float2* out = dst;

	if (nullptr != sens && G > 1) {
		if ((long)part.n < (long)G * N)
			part.alloc((size_t)G * N);
		out = part.ptr;
	}

	StReduceS st{ gs, tabs, out, sens, true };
	launch_fft_axis(cfg, stream, ld, NoMid{}, st);

	if (out != dst)
		sum_groups(dst, part.ptr, N, G, stream, nullptr);
}

// the coil-group sum that ends a normal operator, or with fuse the CG step
// fused into it
static bool sum_groups_or_fuse(float2* dst, const float2* part, long N, int G, hipStream_t st, const int* skip,
			       const CgApFuse* fuse)
{
	if (nullptr != fuse) {
		BAM_CHECK((const float*)dst == fuse->Ap && fuse->len == 2 * N, "normal: fused CG step mismatch");
		cg_apdot_groups(*fuse, part, N, G, st);
		return true;
	}

	sum_groups(dst, part, N, G, st, skip);
	return false;
}

bool Nufft::normal(float2* dst, const float2* src, const float2* sens, const int* skip, const CgApFuse* fuse)
{
	BAM_CHECK(has_psf, "Toeplitz normal operator without PSF");
	BAM_CHECK(0 == nbasis, "normal: a plan with a subspace basis takes normal_basis");

	if (slice) {

		// sub-grids s and s + S/2 share the x-y transform H: the slice
		// kernels store and read H once per pair, and the z-pass applies both
		// members' z factors and PSFs to it (pair_z)
		VolAxis va = VolAxis::make(n, 2, (long)(S / 2) * C);
		FftLaunchCfg cfg = fft_config_skip(skip, axis[2], va.transforms(), va.contig(), 1, -1.f, +1.f);
		if (cfg.fixed)
			fft_config_pair(cfg);
		const bool pair_z = pair_zpass && cfg.fixed && !cfg.L.contig && 0 == va.inner % (uint32_t)cfg.L.P;

		if (!pair_z) {
			va = VolAxis::make(n, 2, (long)S * C);
			cfg = fft_config_skip(skip, axis[2], va.transforms(), va.contig(), 1, -1.f, +1.f);
		}

		launch_slice_fwd(*this, src, sens, false, skip, "fft_toeplitz_slice_fwd", pair_z);

		// the C coil volumes of a sub-grid share its PSF: launch order
		// column group -> volume, so that the PSF columns come from cache
		va.group_volumes(cfg.L.P);
		LdGrid ld{ va, grid.ptr };
		StGrid st{ va, tabs, grid.ptr, false };

		if (pair_z) {
			MdPsfPair md{ va, C, (long)(S / 2) * N, psf.ptr, tabs.lin[2][0], tabs.lin[2][1] };
			launch_fft_axis_pair(cfg, stream, ld, md, st);
		} else {
			MdPsf md{ va, C, psf.ptr };
			launch_fft_axis(cfg, stream, ld, md, st);
		}

		// with maps: one partial image per coil, summed in coil order
		const int G = C;
		float2* out = (nullptr != sens && G > 1) ? part.ptr : dst;

		launch_slice_inv(*this, out, sens, false, skip, "fft_toeplitz_slice_inv", pair_z);

		if (out != dst)
			return sum_groups_or_fuse(dst, part.ptr, N, G, stream, skip, fuse);

		return false;
	}

	const int cc = coil_group(S, C, n[act[0]], nullptr != sens);
	const int G = C / cc;
	GeomS gs = GeomS::make(n, act[0], S, C, cc);
	float2* out = dst;

	if (nullptr != sens && G > 1) {
		if ((long)part.n < (long)G * N)
			part.alloc((size_t)G * N);
		out = part.ptr;
	}

	if (1 == nact) {

		FftLaunchCfg cfg = fft_config_skip(skip, axis[0], gs.va.transforms(), true, S * cc, -1.f, +1.f);
		LdImgLin ld{ gs, tabs, src, sens, false };
		MdPsfS md{ gs, psf.ptr };
		StReduceS st{ gs, tabs, out, sens, false };
		launch_fft_axis(cfg, stream, ld, md, st);

		if (out != dst)
			return sum_groups_or_fuse(dst, part.ptr, N, G, stream, skip, fuse);
		return false;
	}

	{
		FftLaunchCfg cfg = fft_config_skip(skip, axis[0], gs.va.transforms(), true, S * cc, -1.f, -1.f);
		LdImgLin ld{ gs, tabs, src, sens, false };
		StGridS st{ gs, tabs, grid.ptr, false };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
	}

	for (int j = 1; j < nact; j++) {

		VolAxis va = VolAxis::make(n, act[j], (long)S * C);
		LdGrid ld{ va, grid.ptr };
		StGrid st{ va, tabs, grid.ptr, false };

		if (j == nact - 1) {
			FftLaunchCfg cfg = fft_config_skip(skip, axis[j], va.transforms(), va.contig(), 1, -1.f, +1.f);
			MdPsf md{ va, C, psf.ptr };
			launch_fft_axis(cfg, stream, ld, md, st);
		} else {
			FftLaunchCfg cfg = fft_config_skip(skip, axis[j], va.transforms(), va.contig(), 1, -1.f, -1.f);
			launch_fft_axis(cfg, stream, ld, NoMid{}, st);
		}
	}

	for (int j = nact - 2; j >= 1; j--) {

		VolAxis va = VolAxis::make(n, act[j], (long)S * C);
		FftLaunchCfg cfg = fft_config_skip(skip, axis[j], va.transforms(), va.contig(), 1, +1.f, +1.f);
		LdGrid ld{ va, grid.ptr };
		StGrid st{ va, tabs, grid.ptr, false };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
	}

	FftLaunchCfg cfg = fft_config_skip(skip, axis[0], gs.va.transforms(), true, S * cc, +1.f, +1.f);
	LdGridS ld{ gs, grid.ptr };
	StReduceS st{ gs, tabs, out, sens, false };
	launch_fft_axis(cfg, stream, ld, NoMid{}, st);

	if (out != dst)
		return sum_groups_or_fuse(dst, part.ptr, N, G, stream, skip, fuse);

	return false;
}

void Nufft::set_basis(int K, int T, long Mf, const float2* B_host)
{
	BAM_CHECK(K >= 1 && K <= BASIS_MAX, "basis: %d coefficients (at most %d)", K, BASIS_MAX);
	BAM_CHECK(T >= 1 && Mf >= 0, "basis: bad frame dims");
	nbasis = K;
	bT = T;
	bMf = Mf;
	bas.alloc((size_t)K * T);
	BAM_HIP(hipMemcpy(bas.ptr, B_host, sizeof(float2) * K * T, hipMemcpyHostToDevice));
}

// Toeplitz normal operator with a subspace basis (noncart/nufft.c:1546-1625
// with the K x K PSF of nufft_set_traj): per sub-grid s, FFT(linphase_s x_k)
// for every coil and coefficient, the PSF product over the coefficients,
// IFFT, conj(linphase_s), sum over s.  The last axis is transformed forward,
// multiplied (basis_psf_apply) and transformed back in three launches.
void Nufft::normal_basis(float2* dst, const float2* src, const int* skip)
{
	BAM_CHECK(has_psf && nbasis > 0, "normal_basis: no subspace PSF");
	BAM_CHECK(0 == C % nbasis, "normal_basis: channels %d not a multiple of %d coefficients", C, nbasis);
	const int Cc = C / nbasis;

	if (slice) {

		VolAxis va = VolAxis::make(n, 2, (long)S * C);
		launch_slice_fwd(*this, src, nullptr, false, skip, "fft_toeplitz_slice_fwd", false);

		FftLaunchCfg cf = fft_config_skip(skip, axis[2], va.transforms(), va.contig(), 1, -1.f, -1.f);
		cf.name = "fft_toeplitz_basis_fwd";
		LdGrid ld{ va, grid.ptr };
		StGrid st{ va, tabs, grid.ptr, false };
		launch_fft_axis(cf, stream, ld, NoMid{}, st);

		basis_psf_apply(grid.ptr, psf.ptr, S, Cc, nbasis, N, stream, skip);

		FftLaunchCfg ci = fft_config_skip(skip, axis[2], va.transforms(), va.contig(), 1, +1.f, +1.f);
		ci.name = "fft_toeplitz_basis_inv";
		launch_fft_axis(ci, stream, ld, NoMid{}, st);

		launch_slice_inv(*this, dst, nullptr, false, skip, "fft_toeplitz_slice_inv", false);
		return;
	}

	const int cc = coil_group(S, C, n[act[0]], false);
	GeomS gs = GeomS::make(n, act[0], S, C, cc);

	{
		FftLaunchCfg cfg = fft_config_skip(skip, axis[0], gs.va.transforms(), true, S * cc, -1.f, -1.f);
		cfg.name = "fft_toeplitz_basis_first";
		LdImgLin ld{ gs, tabs, src, nullptr, false };
		StGridS st{ gs, tabs, grid.ptr, false };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
	}

	for (int j = 1; j < nact; j++) {

		VolAxis va = VolAxis::make(n, act[j], (long)S * C);
		FftLaunchCfg cfg = fft_config_skip(skip, axis[j], va.transforms(), va.contig(), 1, -1.f, -1.f);
		cfg.name = "fft_toeplitz_basis_fwd";
		LdGrid ld{ va, grid.ptr };
		StGrid st{ va, tabs, grid.ptr, false };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
	}

	basis_psf_apply(grid.ptr, psf.ptr, S, Cc, nbasis, N, stream, skip);

	for (int j = nact - 1; j >= 1; j--) {

		VolAxis va = VolAxis::make(n, act[j], (long)S * C);
		FftLaunchCfg cfg = fft_config_skip(skip, axis[j], va.transforms(), va.contig(), 1, +1.f, +1.f);
		cfg.name = "fft_toeplitz_basis_inv";
		LdGrid ld{ va, grid.ptr };
		StGrid st{ va, tabs, grid.ptr, false };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
	}

	FftLaunchCfg cfg = fft_config_skip(skip, axis[0], gs.va.transforms(), true, S * cc, +1.f, +1.f);
	cfg.name = "fft_toeplitz_basis_last";
	LdGridS ld{ gs, grid.ptr };
	StReduceS st{ gs, tabs, dst, nullptr, false };
	launch_fft_axis(cfg, stream, ld, NoMid{}, st);
}

// ---------------------------------------------------------------------------
// Strang's circulant preconditioner (noncart/precond.c:46-71, 92-106)
// ---------------------------------------------------------------------------

__device__ inline float2 sep_fm(const SepTables& T, long i, const int n[3])
{
	const int x = (int)(i % n[0]);
	const int y = (int)((i / n[0]) % n[1]);
	const int z = (int)(i / ((long)n[0] * n[1]));
	return cmul(cmul(T.fm[0][x], T.fm[1][y]), T.fm[2][z]);
}

__device__ inline float2 sep_lin(const SepTables& T, int s, long i, const int n[3])
{
	const int c[3] = { (int)(i % n[0]), (int)((i / n[0]) % n[1]), (int)(i / ((long)n[0] * n[1])) };
	float2 v = make_float2(1.f, 0.f);
	for (int d = 0; d < 3; d++)
		if (T.bitpos[d] >= 0)
			v = cmul(v, T.lin[d][(s >> T.bitpos[d]) & 1][c[d]]);
	return v;
}

struct Int3 {
	int v[3];
};

// tmp[s][i] = psf[s][i] conj(fm(i)) (ifftmod before the inverse FFT)
__global__ void k_pre_load(SepTables T, Int3 n, long N, int S, const float2* __restrict__ psf, float2* __restrict__ tmp)
{
	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)S * N; e += (long)gridDim.x * blockDim.x)
		tmp[e] = cmulc(psf[e], sep_fm(T, e % N, n.v));
}

// acc(i) = sum_s conj(fm(i)) tmp[s](i) conj(lin_s(i)) (ifftmod after the
// inverse FFT, the cropped PSF of md_zfmacc2), then x fm(i) (fftmod before
// the forward FFT)
__global__ void k_pre_sum(SepTables T, Int3 n, long N, int S, const float2* __restrict__ tmp, float2* __restrict__ acc)
{
	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
		const float2 fm = sep_fm(T, i, n.v);
		float2 a = make_float2(0.f, 0.f);
		for (int s = 0; s < S; s++)
			a = cadd(a, cmulc(cmulc(tmp[(long)s * N + i], fm), sep_lin(T, s, i, n.v)));
		acc[i] = cmul(a, fm);
	}
}

// pre(i) = |fm(i) acc(i)| + 1e-3 (fftmod after the forward FFT, md_zabs,
// md_zsadd)
__global__ void k_pre_abs(SepTables T, Int3 n, long N, float2* __restrict__ acc)
{
	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
		const float2 v = cmul(acc[i], sep_fm(T, i, n.v));
		acc[i] = make_float2(sqrtf(v.x * v.x + v.y * v.y) + 1e-3f, 0.f);
	}
}

// dst[c][i] = src[c][i] / pre[i]
__global__ void k_pre_div(float2* __restrict__ dst, const float2* __restrict__ pre, long N, long total)
{
	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const float2 v = dst[e];
		const float p = pre[e % N].x;
		dst[e] = make_float2(v.x / p, v.y / p);
	}
}

void Nufft::compute_precond(DevBuf<float2>& pre) const
{
	BAM_CHECK(toeplitz && has_psf, "nufft_precond_create: needs a Toeplitz operator with its PSF");
	DevBuf<float2> tmp((size_t)S * N);
	pre.alloc(N);
	const Int3 nn{ { n[0], n[1], n[2] } };
	const unsigned blk = (unsigned)std::min<long>(((long)S * N + 255) / 256, 8192);
	const float sc = 1.f / sqrtf((float)N);

	hipLaunchKernelGGL(k_pre_load, dim3(blk), dim3(256), 0, stream, tabs, nn, N, S, psf.ptr, tmp.ptr);
	fft_nd(n, S, 7, +1.f, tmp.ptr, tmp.ptr, sc, stream);
	hipLaunchKernelGGL(k_pre_sum, dim3(blk), dim3(256), 0, stream, tabs, nn, N, S, tmp.ptr, pre.ptr);
	fft_nd(n, 1, 7, -1.f, pre.ptr, pre.ptr, sc, stream);
	hipLaunchKernelGGL(k_pre_abs, dim3(blk), dim3(256), 0, stream, tabs, nn, N, pre.ptr);
	BAM_HIP(hipGetLastError());
	BAM_HIP(hipStreamSynchronize(stream));
}

// dst = IFFT(FFT(src) / pre) per coil, unnormalised FFTs over the image
// dims (linop_fft_create, linops/someops.c:1732-1776): BART's
// nufft_precond_apply (noncart/precond.c:92-106)
void Nufft::apply_precond(const float2* pre, float2* dst, const float2* src, int ncoil) const
{
	fft_nd(n, ncoil, 7, -1.f, dst, src, 1.f, stream);
	const long total = (long)ncoil * N;
	const unsigned blk = (unsigned)std::min<long>((total + 255) / 256, 8192);
	hipLaunchKernelGGL(k_pre_div, dim3(blk), dim3(256), 0, stream, dst, pre, N, total);
	BAM_HIP(hipGetLastError());
	fft_nd(n, ncoil, 7, +1.f, dst, dst, 1.f, stream);
}

// ---------------------------------------------------------------------------
// Normal operators of many 2-D plans in one launch per pass (batched CG)
// ---------------------------------------------------------------------------

// The three passes of Nufft::normal's 2-D path (first axis with maps and
// linear phases; second axis, PSF, inverse; first axis back with the sums
// over sub-grids and coil groups) for every plan of a size class, each pass
// one launch of k_fft_axis_n_batch: the same work-groups, functors and
// arithmetic as the plans' own launches, so the results are the same bits.
struct NormalGroupSet {

	struct Group {
		int n0, n1;      // transform lengths of the first and second axis
		int nb[3] = { 0, 0, 0 };
		size_t lds[3] = { 0, 0, 0 };
		DevBuf<FftItem<LdImgLin, NoMid, StGridS>> it1;
		DevBuf<FftItem<LdGrid, MdPsf, StGrid>> it2;
		DevBuf<FftItem<LdGridS, NoMid, StReduceS>> it3;
		DevBuf<int2> bt[3];
	};

	std::vector<Group> groups;
};

bool normal_groupable(const Nufft& p, const float2* sens)
{
	if (2 != p.nact || p.slice || !p.toeplitz || !p.has_psf || nullptr == sens || p.nbasis > 0)
		return false;

	const int S = p.S, C = p.C;
	const int cc = coil_group(S, C, p.n[p.act[0]], true);
	GeomS gs = GeomS::make(p.n, p.act[0], S, C, cc);
	VolAxis va = VolAxis::make(p.n, p.act[1], (long)S * C);
	const FftLaunchCfg c1 = fft_config(p.axis[0], gs.va.transforms(), true, S * cc, -1.f, -1.f);
	const FftLaunchCfg c2 = fft_config(p.axis[1], va.transforms(), va.contig(), 1, -1.f, +1.f);
	return c1.fixed && c2.fixed && (C / cc <= 1 || (long)p.part.n >= (long)(C / cc) * p.N);
}

std::shared_ptr<NormalGroupSet> normal_groups_build(const std::vector<const Nufft*>& plans,
						    const std::vector<const float2*>& sens,
						    const std::vector<const float2*>& src, const std::vector<float2*>& dst,
						    const std::vector<const int*>& skip, std::vector<int>& ngroups_out)
{
	auto set = std::make_shared<NormalGroupSet>();
	std::vector<int> key_n0, key_n1;
	std::vector<std::vector<int>> members;

	for (size_t q = 0; q < plans.size(); q++) {
		const Nufft& p = *plans[q];
		const int n0 = p.n[p.act[0]], n1 = p.n[p.act[1]];
		size_t g = 0;
		while (g < members.size() && !(key_n0[g] == n0 && key_n1[g] == n1))
			g++;
		if (g == members.size()) {
			key_n0.push_back(n0);
			key_n1.push_back(n1);
			members.emplace_back();
		}
		members[g].push_back((int)q);
	}

	set->groups.resize(members.size());
	ngroups_out.assign(plans.size(), -1);

	for (size_t g = 0; g < members.size(); g++) {

		NormalGroupSet::Group& G = set->groups[g];
		G.n0 = key_n0[g];
		G.n1 = key_n1[g];
		std::vector<FftItem<LdImgLin, NoMid, StGridS>> h1;
		std::vector<FftItem<LdGrid, MdPsf, StGrid>> h2;
		std::vector<FftItem<LdGridS, NoMid, StReduceS>> h3;
		std::vector<int2> b1, b2, b3;

		for (int j = 0; j < (int)members[g].size(); j++) {

			const int q = members[g][j];
			const Nufft& p = *plans[q];
			const int S = p.S, C = p.C;
			const int cc = coil_group(S, C, p.n[p.act[0]], true);
			const int Gc = C / cc;
			GeomS gs = GeomS::make(p.n, p.act[0], S, C, cc);
			float2* out = (Gc > 1) ? p.part.ptr : dst[q];

			FftLaunchCfg c1 = fft_config_skip(skip[q], p.axis[0], gs.va.transforms(), true, S * cc, -1.f, -1.f);
			h1.push_back({ c1.L, LdImgLin{ gs, p.tabs, src[q], sens[q], false }, NoMid{},
				       StGridS{ gs, p.tabs, p.grid.ptr, false } });

			VolAxis va = VolAxis::make(p.n, p.act[1], (long)S * C);
			FftLaunchCfg c2 = fft_config_skip(skip[q], p.axis[1], va.transforms(), va.contig(), 1, -1.f, +1.f);
			h2.push_back({ c2.L, LdGrid{ va, p.grid.ptr }, MdPsf{ va, C, p.psf.ptr },
				       StGrid{ va, p.tabs, p.grid.ptr, false } });

			FftLaunchCfg c3 = fft_config_skip(skip[q], p.axis[0], gs.va.transforms(), true, S * cc, +1.f, +1.f);
			h3.push_back({ c3.L, LdGridS{ gs, p.grid.ptr }, NoMid{}, StReduceS{ gs, p.tabs, out, sens[q], false } });

			const FftLaunchCfg* cf[3] = { &c1, &c2, &c3 };
			std::vector<int2>* bt[3] = { &b1, &b2, &b3 };
			for (int ph = 0; ph < 3; ph++) {
				BAM_CHECK(cf[ph]->fixed, "grouped normal: pass without a fixed-length kernel");
				G.lds[ph] = std::max(G.lds[ph], cf[ph]->lds_bytes);
				for (unsigned b = 0; b < cf[ph]->blocks; b++)
					bt[ph]->push_back(make_int2(j, (int)b));
			}

			ngroups_out[q] = (int)g;
		}

		G.it1.alloc(h1.size());
		G.it2.alloc(h2.size());
		G.it3.alloc(h3.size());
		BAM_HIP(hipMemcpy(G.it1.ptr, h1.data(), sizeof(h1[0]) * h1.size(), hipMemcpyHostToDevice));
		BAM_HIP(hipMemcpy(G.it2.ptr, h2.data(), sizeof(h2[0]) * h2.size(), hipMemcpyHostToDevice));
		BAM_HIP(hipMemcpy(G.it3.ptr, h3.data(), sizeof(h3[0]) * h3.size(), hipMemcpyHostToDevice));
		const std::vector<int2>* bv[3] = { &b1, &b2, &b3 };
		for (int ph = 0; ph < 3; ph++) {
			G.nb[ph] = (int)bv[ph]->size();
			G.bt[ph].alloc(bv[ph]->size());
			BAM_HIP(hipMemcpy(G.bt[ph].ptr, bv[ph]->data(), sizeof(int2) * bv[ph]->size(), hipMemcpyHostToDevice));
		}
	}

	return set;
}

void normal_groups_launch(const NormalGroupSet& set, hipStream_t st)
{
	for (const NormalGroupSet::Group& G : set.groups) {
		launch_fft_axis_batch(G.n0, G.it1.ptr, G.bt[0].ptr, G.nb[0], G.lds[0], st, "fft_toeplitz_fwd");
		launch_fft_axis_batch(G.n1, G.it2.ptr, G.bt[1].ptr, G.nb[1], G.lds[1], st, "fft_toeplitz");
		launch_fft_axis_batch(G.n0, G.it3.ptr, G.bt[2].ptr, G.nb[2], G.lds[2], st, "fft_toeplitz_inv");
	}
}

// the partial images of a plan's grouped normal operator: (pointer, groups)
std::pair<const float2*, int> normal_group_partials(const Nufft& p, float2* dst)
{
	const int cc = coil_group(p.S, p.C, p.n[p.act[0]], true);
	const int Gc = p.C / cc;
	return (Gc > 1) ? std::make_pair((const float2*)p.part.ptr, Gc) : std::make_pair((const float2*)dst, 1);
}

// ---------------------------------------------------------------------------
// PSF (noncart/nufft.c:727-776 compute_psf2, :535-586 compute_psf_int)
// ---------------------------------------------------------------------------

// fftuc over the 2x grid: first pass loads x fftmod; the last pass stores
// x fftmod x 1/sqrt(N2) into the decomposed layout psf[s][n] (md_decompose,
// num/shuffle.c:84-130)
struct LdFm2 {
	VolAxis g;
	SepTables T2;
	const float2* src;
	bool fm;

	__device__ float2 operator()(long t, int k) const
	{
		uint32_t i, o, v;
		g.split(t, i, o, v);
		float2 x = src[g.offset_iov(i, o, v, k)];

		if (fm) {
			int px[3];
			g.coords(i, o, k, px);
			x = cmul(x, fm_val(T2, px));
		}

		return x;
	}
};

struct StDecomp {
	VolAxis g;
	SepTables T2;
	float2* dst;      // 2x grid (intermediate passes) or decomposed psf (last)
	bool last;
	bool fm_first;    // single-axis case: the load already applied the first fftmod
	float scale;
	int n1[3];        // sub-grid dims
	long N1;
	long vstride = 0; // output stride of the volumes (K x K basis PSFs)

	__device__ void store(const float2* lds, const FftLaunch& L, long t0) const
	{
		store_each(lds, L, t0, [&](long t, int k, float2 x) {
			uint32_t i, o, v;
			g.split(t, i, o, v);

			if (!last) {
				dst[g.offset_iov(i, o, v, k)] = x;
				return;
			}

			int px[3];
			g.coords(i, o, k, px);
			x = cmul(x, fm_val(T2, px));
			x = cscale(x, scale);

			int s = 0;
			int q[3];

			for (int d = 0; d < 3; d++) {
				if (T2.bitpos[d] >= 0) {
					s |= (px[d] & 1) << T2.bitpos[d];
					q[d] = px[d] >> 1;
				} else {
					q[d] = px[d];
				}
			}

			dst[(long)v * vstride + (long)s * N1 + q[0] + (long)n1[0] * (q[1] + (long)n1[1] * q[2])] = x;
		});
	}
};

void Nufft::compute_psf_from_traj(const float* traj, const float2* weights_host, bool periodic_psf)
{
	int n2[3];

	for (int d = 0; d < 3; d++)
		n2[d] = (1 == n[d]) ? 1 : 2 * n[d];

	std::vector<float> traj2(3 * M);

	for (long i = 0; i < 3 * M; i++)
		traj2[i] = traj[i] * 2.f;

	std::vector<float2> sqw;

	if (nullptr != weights_host) {

		sqw.resize(M);

		for (long m = 0; m < M; m++) {
			float2 w = weights_host[m];
			sqw[m] = make_float2(w.x * w.x + w.y * w.y, w.y * w.x - w.x * w.y);
		}
	}

	// with a basis, J = K^2 channels whose k-space is conj(B_k) B_k' in
	// place of the ones (compute_psf_int, noncart/nufft.c:535-586)
	const int J = (nbasis > 0) ? nbasis * nbasis : 1;
	BAM_CHECK(0 == nbasis || bMf * bT == M, "psf: basis frames (%d x %ld) do not match the samples (%ld)", bT, bMf,
		  M);

	Nufft p2;
	p2.stream = stream;
	p2.tune = false;  // one adjoint only
	p2.create(n2, J, M, traj2.data(), sqw.empty() ? nullptr : sqw.data(), false, periodic_psf, width, os, false);

	long N2 = p2.N;
	DevBuf<float2> ones(std::max<long>((long)J * M, 1));
	DevBuf<float2> psft((size_t)J * N2);

	if (nbasis > 0)
		basis_psf_vals(ones.ptr, bas.ptr, nbasis, bT, bMf, stream);
	else
		fill(ones.ptr, M, make_float2(1.f, 0.f), stream);

	p2.adjoint(psft.ptr, ones.ptr, nullptr);

	psf.alloc((size_t)J * S * N);

	const float scale = 1.f / sqrtf((float)N2);
	int last = p2.act[p2.nact - 1];

	for (int j = 0; j < p2.nact; j++) {

		int a = p2.act[j];
		VolAxis va = VolAxis::make(n2, a, J);
		FftLaunchCfg cfg = fft_config(p2.axis[j], va.transforms(), va.contig(), 1, -1.f, -1.f);
		cfg.name = "fft_psf";
		LdFm2 ld{ va, p2.tabs, psft.ptr, 0 == j };
		StDecomp st{ va, p2.tabs, (a == last) ? psf.ptr : psft.ptr, a == last, false, scale,
			     { n[0], n[1], n[2] }, N, (long)S * N };
		launch_fft_axis(cfg, stream, ld, NoMid{}, st);
	}

	apply_real_psf();
	BAM_HIP(hipStreamSynchronize(stream));
	has_psf = true;
}

__global__ void k_real_part(float2* p, long n)
{
	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
		p[i].y = 0.f;
}

void Nufft::apply_real_psf()
{
	if (!real_psf)
		return;

	const long n = (long)psf.n;
	const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 4096);
	hipLaunchKernelGGL(k_real_part, dim3(blocks), dim3(256), 0, stream, psf.ptr, n);
	BAM_HIP(hipGetLastError());
}

}  // namespace bam
