// fft.hpp -- LDS Stockham FFT along one axis of a stack of volumes (gfx950).
//
// This replaces the FFTW / cuFFT calls of BART (num/fft_plan.c:188-281,
// num/fft-cuda.c) with a hand-written kernel whose load, middle and store
// stages are functors, so that element-wise work before, between and after
// the transforms (deapodisation, linear phases, fftmod, coil maps, the
// Toeplitz PSF, the sum over sub-grids) is fused into the FFT passes.
//
// Convention (FFTW's, as used by BART): unnormalised, sign -1 forward,
// +1 backward.
//
// One work-group (512 threads = 8 wave64) holds P transforms of length n in
// LDS (P * n <= 8192 complex).  Each Stockham pass reads R strided values per
// butterfly into registers, applies twiddles from an FP64-built table, does
// a radix-R DFT and writes back in auto-sorted order.
#pragma once

#include <type_traits>
#include <utility>

#include "common.hpp"
#include "prof.hpp"

namespace bam {

constexpr int FFT_THREADS = 512;
constexpr int FFT_MAX_ELEMS = 8192;
constexpr int FFT_MAX_PASSES = 16;

struct AxisPlan {

	int n = 1;
	int nr = 0;
	int radix[FFT_MAX_PASSES] = {};
	const float2* tw = nullptr;  // exp(-2 pi i k / n), k < n
};

// host: radix plan + shared device twiddle table for length n
AxisPlan make_axis_plan(int n);
bool fft_size_supported(int n);

// Geometry of the transforms of one launch.
//   contig: transforms are contiguous rows (element stride 1); the P
//           transforms of a block are consecutive rows.
//   otherwise: element stride > 1; the P transforms of a block are adjacent
//           columns, so loads of the same k are coalesced.
struct FftLaunch {

	AxisPlan plan;
	int P = 1;
	int logP = 0;
	int pitch = 0;       // LDS distance of consecutive transforms (n or n + 1)
	bool contig = true;
	long total = 0;      // number of transforms
	float sign1 = -1.f;  // first transform
	float sign2 = +1.f;  // second transform (only when a middle stage is present)
	const int* skip = nullptr;  // device flag: when set (non-zero) the launch does nothing
};

// ---------------------------------------------------------------------------
// radix-R DFT in registers, w = exp(sgn 2 pi i / R)
// ---------------------------------------------------------------------------

__device__ inline float2 mul_i(float2 a, float sgn)  // a * (sgn i)
{
	return make_float2(-sgn * a.y, sgn * a.x);
}

template <int R> struct Dft;

template <> struct Dft<2> {
	__device__ static void run(float2* v, float) {
		float2 a = v[0], b = v[1];
		v[0] = cadd(a, b);
		v[1] = csub(a, b);
	}
};

template <> struct Dft<3> {
	__device__ static void run(float2* v, float sgn) {
		const float c = -0.5f;
		const float s = 0.86602540378443864676f;
		float2 t1 = cadd(v[1], v[2]);
		float2 t2 = make_float2(v[0].x + c * t1.x, v[0].y + c * t1.y);
		float2 t3 = mul_i(cscale(csub(v[1], v[2]), s), sgn);
		v[0] = cadd(v[0], t1);
		v[1] = cadd(t2, t3);
		v[2] = csub(t2, t3);
	}
};

template <> struct Dft<4> {
	__device__ static void run(float2* v, float sgn) {
		float2 a0 = cadd(v[0], v[2]);
		float2 a1 = csub(v[0], v[2]);
		float2 a2 = cadd(v[1], v[3]);
		float2 a3 = mul_i(csub(v[1], v[3]), sgn);
		v[0] = cadd(a0, a2);
		v[1] = cadd(a1, a3);
		v[2] = csub(a0, a2);
		v[3] = csub(a1, a3);
	}
};

template <> struct Dft<5> {
	__device__ static void run(float2* v, float sgn) {
		const float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
		const float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
		const float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
		const float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
		float2 a = v[0];
		float2 t1 = cadd(v[1], v[4]);
		float2 t2 = cadd(v[2], v[3]);
		float2 t3 = csub(v[1], v[4]);
		float2 t4 = csub(v[2], v[3]);
		float2 r1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
		float2 r2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
		float2 i1 = mul_i(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), sgn);
		float2 i2 = mul_i(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), sgn);
		v[0] = cadd(a, cadd(t1, t2));
		v[1] = cadd(r1, i1);
		v[4] = csub(r1, i1);
		v[2] = cadd(r2, i2);
		v[3] = csub(r2, i2);
	}
};

template <> struct Dft<8> {
	__device__ static void run(float2* v, float sgn) {
		float2 e[4] = { v[0], v[2], v[4], v[6] };
		float2 o[4] = { v[1], v[3], v[5], v[7] };
		Dft<4>::run(e, sgn);
		Dft<4>::run(o, sgn);
		const float r = 0.70710678118654752440f;
		// w8^1 = r (1 + sgn i), w8^2 = sgn i, w8^3 = r (-1 + sgn i)
		float2 o1 = make_float2(r * (o[1].x - sgn * o[1].y), r * (o[1].y + sgn * o[1].x));
		float2 o2 = mul_i(o[2], sgn);
		float2 o3 = make_float2(r * (-o[3].x - sgn * o[3].y), r * (-o[3].y + sgn * o[3].x));
		v[0] = cadd(e[0], o[0]);
		v[4] = csub(e[0], o[0]);
		v[1] = cadd(e[1], o1);
		v[5] = csub(e[1], o1);
		v[2] = cadd(e[2], o2);
		v[6] = csub(e[2], o2);
		v[3] = cadd(e[3], o3);
		v[7] = csub(e[3], o3);
	}
};

// radix 16: two radix-8 DFTs of the even and odd elements and the w16^k
// combination
template <> struct Dft<16> {
	__device__ static void run(float2* v, float sgn) {
		float2 e[8], o[8];
#pragma unroll
		for (int k = 0; k < 8; k++) {
			e[k] = v[2 * k];
			o[k] = v[2 * k + 1];
		}
		Dft<8>::run(e, sgn);
		Dft<8>::run(o, sgn);
		const float c = 0.92387953251128675613f;  // cos(pi / 8)
		const float s = 0.38268343236508977173f;  // sin(pi / 8)
		const float r = 0.70710678118654752440f;
		// w16^k = cos(2 pi k / 16) + sgn i sin(2 pi k / 16), k = 1 .. 7
		const float2 w[8] = { make_float2(1.f, 0.f),  make_float2(c, sgn * s),  make_float2(r, sgn * r),
				      make_float2(s, sgn * c),   make_float2(0.f, sgn),    make_float2(-s, sgn * c),
				      make_float2(-r, sgn * r), make_float2(-c, sgn * s) };
		v[0] = cadd(e[0], o[0]);
		v[8] = csub(e[0], o[0]);
		const float2 o4 = mul_i(o[4], sgn);
		v[4] = cadd(e[4], o4);
		v[12] = csub(e[4], o4);
#pragma unroll
		for (int k = 1; k < 8; k++) {
			if (4 == k)
				continue;
			const float2 t = cmul(o[k], w[k]);
			v[k] = cadd(e[k], t);
			v[k + 8] = csub(e[k], t);
		}
	}
};

// Generic odd prime radix (7, 11, 13): direct DFT with table twiddles.
template <int R> struct DftGeneric {
	__device__ static void run(float2* v, float sgn, const float2* tw, int n) {
		float2 o[R];
		const int step = n / R;
#pragma unroll
		for (int q = 0; q < R; q++) {
			float2 s = v[0];
#pragma unroll
			for (int r = 1; r < R; r++) {
				float2 w = tw[((r * q) % R) * step];
				if (sgn > 0.f)
					w.y = -w.y;
				s = cadd(s, cmul(v[r], w));
			}
			o[q] = s;
		}
#pragma unroll
		for (int q = 0; q < R; q++)
			v[q] = o[q];
	}
};

template <int R> __device__ inline void dft_any(float2* v, float sgn, const float2* tw, int n)
{
	if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16)
		Dft<R>::run(v, sgn);
	else
		DftGeneric<R>::run(v, sgn, tw, n);
}

// One Stockham pass of radix R over P transforms of length n in LDS.
// EPT = upper bound of elements per thread (P * n / FFT_THREADS).
template <int R, int EPT>
__device__ inline void stockham_pass(float2* lds, int n, int P, int Ns, const float2* tw, float sgn)
{
	constexpr int NB = (EPT + R - 1) / R;
	const int nR = n / R;
	const int nb = P * nR;
	const int tstep = n / (Ns * R);
	float2 v[NB][R];
	int dst[NB];

#pragma unroll
	for (int i = 0; i < NB; i++) {

		int b = threadIdx.x + i * FFT_THREADS;
		dst[i] = -1;

		if (b < nb) {

			int p = b / nR;
			int j = b - p * nR;
			int k = j % Ns;
			const float2* src = lds + p * n + j;

#pragma unroll
			for (int r = 0; r < R; r++) {

				float2 x = src[r * nR];

				if (r > 0 && k > 0) {

					float2 w = tw[r * k * tstep];
					if (sgn > 0.f)
						w.y = -w.y;
					x = cmul(x, w);
				}

				v[i][r] = x;
			}

			dft_any<R>(v[i], sgn, tw, n);
			dst[i] = p * n + (j / Ns) * Ns * R + k;
		}
	}

	lds_barrier();

#pragma unroll
	for (int i = 0; i < NB; i++)
		if (dst[i] >= 0)
#pragma unroll
			for (int q = 0; q < R; q++)
				lds[dst[i] + q * Ns] = v[i][q];

	lds_barrier();
}

// GEN: also instantiate the generic odd-prime radices (7, 11, 13), which
// cost registers; lengths of the form 2^a 3^b 5^c use GEN = false.
template <int EPT, bool GEN>
__device__ inline void run_passes(float2* lds, const AxisPlan& plan, int P, float sgn)
{
	int Ns = 1;

	for (int ip = 0; ip < plan.nr; ip++) {

		int R = plan.radix[ip];

		switch (R) {
		case 2: stockham_pass<2, EPT>(lds, plan.n, P, Ns, plan.tw, sgn); break;
		case 3: stockham_pass<3, EPT>(lds, plan.n, P, Ns, plan.tw, sgn); break;
		case 4: stockham_pass<4, EPT>(lds, plan.n, P, Ns, plan.tw, sgn); break;
		case 5: stockham_pass<5, EPT>(lds, plan.n, P, Ns, plan.tw, sgn); break;
		case 8: stockham_pass<8, EPT>(lds, plan.n, P, Ns, plan.tw, sgn); break;
		default:
			if constexpr (GEN) {
				switch (R) {
				case 7: stockham_pass<7, EPT>(lds, plan.n, P, Ns, plan.tw, sgn); break;
				case 11: stockham_pass<11, EPT>(lds, plan.n, P, Ns, plan.tw, sgn); break;
				case 13: stockham_pass<13, EPT>(lds, plan.n, P, Ns, plan.tw, sgn); break;
				default: break;
				}
			}
			break;
		}

		Ns *= R;
	}
}

// Identity middle stage.
struct NoMid {
	static constexpr bool active = false;
	static constexpr bool colblock = true;
	__device__ float2 operator()(long, int, float2 v) const { return v; }
	__device__ bool cb_ok(int) const { return true; }
	__device__ int base(long) const { return 0; }
	__device__ float2 at(int, int, int, float2 v) const { return v; }
};

// functors with column-block access (static constexpr bool colblock = true;
// cb_ok(P), base(t0), at(base, p, k [, v]))
template <class F, class = void>
struct colblock_of {
	static constexpr bool value = false;
};

template <class F>
struct colblock_of<F, std::void_t<decltype(F::colblock)>> {
	static constexpr bool value = F::colblock;
};

// ---------------------------------------------------------------------------
// The axis kernel.  Ld: float2 operator()(long t, int k) const.
// Md: float2 operator()(long t, int k, float2 v) const (when active).
// St: void store(const float2* lds, const FftLaunch&, long t0) const -- runs
//     the block's whole store loop so that it may reduce over transforms.
// ---------------------------------------------------------------------------

template <int EPT, bool GEN, class Ld, class Md, class St>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis(FftLaunch L, Ld ld, Md md, St st)
{
	extern __shared__ float2 lds[];

	if (nullptr != L.skip && 0 != *L.skip)
		return;

	const int n = L.plan.n;
	const int P = L.P;
	const long t0 = (long)blockIdx.x * P;
	const int tot = P * n;

	if (L.contig) {

		int p = threadIdx.x / n;
		int k = threadIdx.x - p * n;
		const int dp = FFT_THREADS / n;       // 0 when n > FFT_THREADS
		const int dk = FFT_THREADS - dp * n;

		for (int i = threadIdx.x; i < tot; i += FFT_THREADS) {

			long t = t0 + p;
			lds[p * n + k] = (t < L.total) ? ld(t, k) : make_float2(0.f, 0.f);
			p += dp;
			k += dk;
			while (k >= n) {
				k -= n;
				p++;
			}
		}

	} else {

		for (int i = threadIdx.x; i < tot; i += FFT_THREADS) {

			int p = i & (P - 1);
			int k = i >> L.logP;
			long t = t0 + p;
			lds[p * n + k] = (t < L.total) ? ld(t, k) : make_float2(0.f, 0.f);
		}
	}

	lds_barrier();

	run_passes<EPT, GEN>(lds, L.plan, P, L.sign1);

	if constexpr (Md::active) {

		for (int i = threadIdx.x; i < tot; i += FFT_THREADS) {

			int p, k;

			if (L.contig) {
				p = i / n;
				k = i - p * n;
			} else {
				p = i & (P - 1);
				k = i >> L.logP;
			}

			long t = t0 + p;

			if (t < L.total)
				lds[p * n + k] = md(t, k, lds[p * n + k]);
		}

		lds_barrier();

		run_passes<EPT, GEN>(lds, L.plan, P, L.sign2);
	}

	st.store(lds, L, t0);
}

// ---------------------------------------------------------------------------
// Compile-time lengths.  For the sizes of the image and oversampled grids the
// radix plan, the butterfly index arithmetic and the block shape are
// constants, and the twiddles are staged in LDS.
// ---------------------------------------------------------------------------

constexpr int ct_next_radix(int m)
{
	return (0 == m % 8) ? 8 : ((0 == m % 4) ? 4 : ((0 == m % 2) ? 2 : ((0 == m % 3) ? 3 : ((0 == m % 5) ? 5 : m))));
}

// radix 16 first in the slice kernels (two passes for 128 instead of
// three); the axis kernels keep radix 8 (radix 16 there: z-pass 4 % faster,
// 2-D passes slower)
constexpr bool SLICE_R16 = true;
constexpr bool AXIS_R16 = false;

// radix 16 first where it divides (fewer passes through LDS)
constexpr int ct_next_radix16(int m)
{
	return (0 == m % 16) ? 16 : ct_next_radix(m);
}

constexpr int ct_radix(int n, int idx, bool r16 = false)
{
	int m = n;
	for (int i = 0; i < idx; i++)
		m /= r16 ? ct_next_radix16(m) : ct_next_radix(m);
	return r16 ? ct_next_radix16(m) : ct_next_radix(m);
}

constexpr int ct_num_radices(int n, bool r16 = false)
{
	int m = n, c = 0;
	while (m > 1) {
		m /= r16 ? ct_next_radix16(m) : ct_next_radix(m);
		c++;
	}
	return c;
}

constexpr int ct_log2(int p)
{
	int l = 0;
	while ((1 << l) < p)
		l++;
	return l;
}

// transforms per block for a compile-time length (a power of two): up to
// 4096 elements, at least 16 columns (strided geometry) up to n = 512
constexpr int ct_block_p(int n)
{
	int p = 1;
	while (2 * p * n <= 4096)
		p *= 2;
	return (n <= 512 && p < 16) ? 16 : ((p < 8) ? 8 : p);
}

template <int N, int R, int NS, int P>
__device__ inline void ct_pass(float2* lds, const float2* tw, float sgn)
{
	constexpr int nR = N / R;
	constexpr int NB = P * nR;
	constexpr int PER = (NB + FFT_THREADS - 1) / FFT_THREADS;
	constexpr int TSTEP = N / (NS * R);
	float2 v[PER][R];
	int dst[PER];

	// see slice_pass: butterfly addresses are computed per pass
	int lane = threadIdx.x;
	asm volatile("" : "+v"(lane));

#pragma unroll
	for (int i = 0; i < PER; i++) {

		const int b = lane + i * FFT_THREADS;
		dst[i] = -1;

		if (0 == NB % FFT_THREADS || b < NB) {

			const int p = b / nR;
			const int j = b - p * nR;
			const int k = j % NS;
			const float2* src = lds + p * (N + 1) + j;

#pragma unroll
			for (int r = 0; r < R; r++) {

				float2 x = src[r * nR];

				if (r > 0 && NS > 1) {
					float2 w = tw[r * k * TSTEP];
					if (sgn > 0.f)
						w.y = -w.y;
					x = cmul(x, w);
				}

				v[i][r] = x;
			}

			dft_any<R>(v[i], sgn, tw, N);
			dst[i] = p * (N + 1) + (j / NS) * NS * R + k;
		}
	}

	lds_barrier();

#pragma unroll
	for (int i = 0; i < PER; i++)
		if (0 == NB % FFT_THREADS || dst[i] >= 0)
#pragma unroll
			for (int q = 0; q < R; q++)
				lds[dst[i] + q * NS] = v[i][q];

	lds_barrier();
}

template <int N, int P, int IDX, int NS>
__device__ inline void ct_passes(float2* lds, const float2* tw, float sgn)
{
	if constexpr (IDX < ct_num_radices(N, AXIS_R16)) {
		constexpr int R = ct_radix(N, IDX, AXIS_R16);
		ct_pass<N, R, NS, P>(lds, tw, sgn);
		ct_passes<N, P, IDX + 1, NS * R>(lds, tw, sgn);
	}
}

// the passes of a length-N transform after sub-transforms of length NS have
// been built in place, up to (not including) sub-transforms of length NSEND
template <int N, int P, int NS, int NSEND>
__device__ inline void ct_passes_between(float2* lds, const float2* tw, float sgn)
{
	if constexpr (NS < NSEND) {
		constexpr int R = ct_next_radix(NSEND / NS);
		ct_pass<N, R, NS, P>(lds, tw, sgn);
		ct_passes_between<N, P, NS * R, NSEND>(lds, tw, sgn);
	}
}

// Transform of a column block (strided axis) -- or forward transform,
// pointwise multiply and inverse transform -- with the first and last passes
// in registers: a thread's E = P N / FFT_THREADS points k0 + j N / E are the
// inputs of one butterfly of the first radix-E pass and the outputs of one
// butterfly of the last radix-E pass, so the loads, the stores and (with a
// multiply) the forward's last pass, the multiply and the inverse's first
// pass need no LDS round trip.
template <int N>
constexpr int ct_edge_radix()
{
	return ct_block_p(N) * N / FFT_THREADS;
}

template <int N>
constexpr bool ct_edges_in_regs()
{
	constexpr int E = ct_edge_radix<N>();
	return 0 == (ct_block_p(N) * N) % FFT_THREADS && (2 == E || 4 == E || 8 == E || 16 == E) &&
	       0 == N % (E * E) && 0 == FFT_THREADS % ct_block_p(N);
}

template <int N, int E, class Ld, class Md, class St>
__device__ inline void fft_mul_ifft_regs(const FftLaunch& L, const Ld& ld, const Md& md, const St& st, float2* lds,
					 const float2* tw, long t0)
{
	constexpr int P = ct_block_p(N);
	constexpr int LOGP = ct_log2(P);
	constexpr int PN = N + 1;
	constexpr int NE = N / E;

	const int p = threadIdx.x & (P - 1);
	const int k0 = threadIdx.x >> LOGP;
	const bool in = t0 + p < L.total;
	float2 v[E];

	// loads; first forward pass (radix E, no twiddles)
	{
		const auto lb = ld.base(t0);
#pragma unroll
		for (int j = 0; j < E; j++)
			v[j] = in ? ld.at(lb, p, k0 + j * NE) : make_float2(0.f, 0.f);
	}

	dft_any<E>(v, L.sign1, tw, N);

#pragma unroll
	for (int q = 0; q < E; q++)
		lds[p * PN + k0 * E + q] = v[q];

	lds_barrier();

	ct_passes_between<N, P, E, NE>(lds, tw, L.sign1);

	// last forward pass (radix E, sub-transforms of length N / E)
	auto last_pass = [&](float sgn) {
#pragma unroll
		for (int r = 0; r < E; r++) {
			float2 x = lds[p * PN + k0 + r * NE];
			if (r > 0) {
				float2 w = tw[r * k0];
				if (sgn > 0.f)
					w.y = -w.y;
				x = cmul(x, w);
			}
			v[r] = x;
		}
		dft_any<E>(v, sgn, tw, N);
	};

	last_pass(L.sign1);

	if constexpr (Md::active) {

		// pointwise multiply of the points k0 + q N / E; first inverse pass
		if (in) {
			const auto mb = md.base(t0);
#pragma unroll
			for (int q = 0; q < E; q++)
				v[q] = md.at(mb, p, k0 + q * NE, v[q]);
		}

		dft_any<E>(v, L.sign2, tw, N);

		lds_barrier();

#pragma unroll
		for (int q = 0; q < E; q++)
			lds[p * PN + k0 * E + q] = v[q];

		lds_barrier();

		ct_passes_between<N, P, E, NE>(lds, tw, L.sign2);

		last_pass(L.sign2);
	}

	if (in) {
		const auto sb = st.base(t0);
#pragma unroll
		for (int q = 0; q < E; q++)
			st.at(sb, p, k0 + q * NE, v[q]);
	}
}

// store functors that take single elements: put(t, k, value)
template <class F, class = void>
struct has_put : std::false_type {
};
template <class F>
struct has_put<F, std::void_t<decltype(std::declval<const F&>().put(0l, 0, float2{}))>> : std::true_type {
};

// Contiguous axis (transforms along rows): a thread takes the E points
// k0 + j N / E of one transform -- coalesced loads along k0 -- runs the
// first radix-E pass on them, the middle passes in LDS, and, when the store
// takes single elements, the last radix-E pass into registers and stores
// from there.
template <int N, int E, class Ld, class St>
__device__ inline void fft_contig_regs(const FftLaunch& L, const Ld& ld, const St& st, float2* lds, const float2* tw,
				       long t0)
{
	constexpr int P = ct_block_p(N);
	constexpr int PN = N + 1;
	constexpr int NE = N / E;

	const int p = threadIdx.x / NE;
	const int k0 = threadIdx.x - p * NE;
	const long t = t0 + p;
	const bool in = t < L.total;
	float2 v[E];

#pragma unroll
	for (int j = 0; j < E; j++)
		v[j] = in ? ld(t, k0 + j * NE) : make_float2(0.f, 0.f);

	dft_any<E>(v, L.sign1, tw, N);

#pragma unroll
	for (int q = 0; q < E; q++)
		lds[p * PN + k0 * E + q] = v[q];

	lds_barrier();

	if constexpr (has_put<St>::value) {

		ct_passes_between<N, P, E, NE>(lds, tw, L.sign1);

#pragma unroll
		for (int r = 0; r < E; r++) {
			float2 x = lds[p * PN + k0 + r * NE];
			if (r > 0) {
				float2 w = tw[r * k0];
				if (L.sign1 > 0.f)
					w.y = -w.y;
				x = cmul(x, w);
			}
			v[r] = x;
		}

		dft_any<E>(v, L.sign1, tw, N);

		if (in)
#pragma unroll
			for (int q = 0; q < E; q++)
				st.put(t, k0 + q * NE, v[q]);

	} else {

		ct_passes_between<N, P, E, N>(lds, tw, L.sign1);
		st.store(lds, L, t0);
	}
}

// The transforms [t0, t0 + P) of one launch (the body of k_fft_axis_n and of
// its batched form k_fft_axis_n_batch).
template <int N, class Ld, class Md, class St>
__device__ inline void fft_axis_block(const FftLaunch& L, const Ld& ld, const Md& md, const St& st, long t0,
				      float2* lds)
{
	constexpr int P = ct_block_p(N);
	constexpr int TOT = P * N;
	constexpr int LOGP = ct_log2(P);
	constexpr int PN = N + 1;  // transform pitch: strided accesses (p fastest) spread over the banks
	static_assert((1 << LOGP) == P, "transforms per block must be a power of two");

	float2* tw = lds + P * PN;

	// strided axis with column-block functors: per-block base addresses
	constexpr bool CBF = colblock_of<Ld>::value && colblock_of<Md>::value && colblock_of<St>::value;
	bool cb = false;

	if constexpr (CBF)
		cb = !L.contig && ld.cb_ok(P) && md.cb_ok(P) && st.cb_ok(P);

	for (int i = threadIdx.x; i < N; i += FFT_THREADS)
		tw[i] = L.plan.tw[i];

	if constexpr (CBF && ct_edges_in_regs<N>()) {
		if (cb) {
			fft_mul_ifft_regs<N, ct_edge_radix<N>()>(L, ld, md, st, lds, tw, t0);
			return;
		}
	}

	if constexpr (!Md::active && ct_edges_in_regs<N>()) {
		if (L.contig) {
			fft_contig_regs<N, ct_edge_radix<N>()>(L, ld, st, lds, tw, t0);
			return;
		}
	}

	if (L.contig) {

#pragma unroll 4
		for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) {
			const int p = i / N;
			const int k = i - p * N;
			const long t = t0 + p;
			lds[p * PN + k] = (t < L.total) ? ld(t, k) : make_float2(0.f, 0.f);
		}

	} else {

		bool done = false;

		if constexpr (CBF) {
			if (cb) {
				const auto lb = ld.base(t0);

#pragma unroll 4
				for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) {
					const int p = i & (P - 1);
					const int k = i >> LOGP;
					lds[p * PN + k] = (t0 + p < L.total) ? ld.at(lb, p, k) : make_float2(0.f, 0.f);
				}

				done = true;
			}
		}

		if (!done) {
#pragma unroll 4
			for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) {
				const int p = i & (P - 1);
				const int k = i >> LOGP;
				const long t = t0 + p;
				lds[p * PN + k] = (t < L.total) ? ld(t, k) : make_float2(0.f, 0.f);
			}
		}
	}

	lds_barrier();

	ct_passes<N, P, 0, 1>(lds, tw, L.sign1);

	if constexpr (Md::active) {

		bool done = false;

		if constexpr (CBF) {
			if (cb) {
				const auto mb = md.base(t0);

				for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) {
					const int p = i & (P - 1);
					const int k = i >> LOGP;
					if (t0 + p < L.total)
						lds[p * PN + k] = md.at(mb, p, k, lds[p * PN + k]);
				}

				done = true;
			}
		}

		if (!done) {

			for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) {

				int p, k;

				if (L.contig) {
					p = i / N;
					k = i - p * N;
				} else {
					p = i & (P - 1);
					k = i >> LOGP;
				}

				const long t = t0 + p;

				if (t < L.total)
					lds[p * PN + k] = md(t, k, lds[p * PN + k]);
			}
		}

		lds_barrier();

		ct_passes<N, P, 0, 1>(lds, tw, L.sign2);
	}

	if constexpr (CBF) {
		if (cb) {
			const auto sb = st.base(t0);

			for (int i = threadIdx.x; i < TOT; i += FFT_THREADS) {
				const int p = i & (P - 1);
				const int k = i >> LOGP;
				if (t0 + p < L.total)
					st.at(sb, p, k, lds[p * PN + k]);
			}

			return;
		}
	}

	st.store(lds, L, t0);
}

template <int N, class Ld, class Md, class St>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis_n(FftLaunch L, Ld ld, Md md, St st)
{
	extern __shared__ float2 lds[];

	if (nullptr != L.skip && 0 != *L.skip)
		return;

	fft_axis_block<N>(L, ld, md, st, (long)blockIdx.x * ct_block_p(N), lds);
}

// One pass of many problems of the same transform length in one launch:
// items[q] holds problem q's launch and functors, btab[b] = (q, block of
// problem q) for work-group b.
template <class Ld, class Md, class St>
struct FftItem {
	FftLaunch L;
	Ld ld;
	Md md;
	St st;
};

template <int N, class Ld, class Md, class St>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis_n_batch(const FftItem<Ld, Md, St>* __restrict__ items,
								  const int2* __restrict__ btab)
{
	extern __shared__ float2 lds[];
	const int2 e = btab[blockIdx.x];
	const FftItem<Ld, Md, St>& it = items[e.x];

	if (nullptr != it.L.skip && 0 != *it.L.skip)
		return;

	fft_axis_block<N>(it.L, it.ld, it.md, it.st, (long)e.y * ct_block_p(N), lds);
}

// Lengths whose pair kernel exists (k_fft_axis_pair needs the register-edge
// geometry of the column-block path).
inline bool fft_pair_len_ok(int n)
{
	return 64 == n || 128 == n || 256 == n || 512 == n;
}

// columns per block of the pair kernel: 2048 elements up to n = 128 (for
// n = 128, 16 columns of 4 points per thread need 63 registers and allow
// 8 waves per SIMD, against 112 and 4 for 32 columns: 0.94 against 1.05 ms
// per kooshball z-pass)
constexpr int pair_block_p(int n)
{
	return (n <= 128) ? 2048 / n : ct_block_p(n);
}

// the register-edge geometry for P columns of length N
template <int N, int P>
constexpr bool edges_in_regs_p()
{
	constexpr int E = P * N / FFT_THREADS;
	return 0 == (P * N) % FFT_THREADS && (2 == E || 4 == E || 8 == E || 16 == E) && 0 == N % (E * E) &&
	       0 == FFT_THREADS % P;
}

// Strided-axis transform, pointwise multiply and inverse transform for a
// pair of inputs that differ only by a factor along the axis: the column h
// (held in registers) is multiplied by zf(b, k), transformed, multiplied by
// the pair member's factor table at(b, ...), inverse-transformed, and
// conj(zf(b, k)) x result is summed over b = 0, 1 into the store.  One read
// and one write of the column serve both members.  Md provides base(t0),
// at(base, b, p, k, x), zf(b, k) and cb_ok(P); the host launches it only
// where the column-block geometry holds (cb_ok for every functor, strided).
template <int N, class Ld, class Md, class St>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis_pair(FftLaunch L, Ld ld, Md md, St st)
{
	constexpr int P = pair_block_p(N);
	static_assert(edges_in_regs_p<N, P>(), "pair kernel needs the register-edge geometry");
	constexpr int LOGP = ct_log2(P);
	constexpr int PN = N + 1;
	constexpr int E = P * N / FFT_THREADS;
	constexpr int NE = N / E;

	extern __shared__ float2 lds[];
	float2* tw = lds + P * PN;

	if (nullptr != L.skip && 0 != *L.skip)
		return;

	const long t0 = (long)blockIdx.x * P;

	for (int i = threadIdx.x; i < N; i += FFT_THREADS)
		tw[i] = L.plan.tw[i];

	const int p = threadIdx.x & (P - 1);
	const int k0 = threadIdx.x >> LOGP;
	const bool in = t0 + p < L.total;
	float2 h[E], acc[E], v[E];

	{
		const auto lb = ld.base(t0);
#pragma unroll
		for (int j = 0; j < E; j++) {
			h[j] = in ? ld.at(lb, p, k0 + j * NE) : make_float2(0.f, 0.f);
			acc[j] = make_float2(0.f, 0.f);
		}
	}

	const auto mb = md.base(t0);

	// last radix-E pass of a transform (sub-transforms of length N / E) from
	// LDS into registers
	auto last_pass = [&](float sgn) {
#pragma unroll
		for (int r = 0; r < E; r++) {
			float2 x = lds[p * PN + k0 + r * NE];
			if (r > 0) {
				float2 w = tw[r * k0];
				if (sgn > 0.f)
					w.y = -w.y;
				x = cmul(x, w);
			}
			v[r] = x;
		}
		dft_any<E>(v, sgn, tw, N);
	};

	for (int b = 0; b < 2; b++) {

		// factor along the axis; first forward pass in registers
#pragma unroll
		for (int j = 0; j < E; j++)
			v[j] = cmul(h[j], md.zf(b, k0 + j * NE));

		dft_any<E>(v, L.sign1, tw, N);

		if (b > 0)
			lds_barrier();  // the previous member's last pass has read LDS

#pragma unroll
		for (int q = 0; q < E; q++)
			lds[p * PN + k0 * E + q] = v[q];

		lds_barrier();
		ct_passes_between<N, P, E, NE>(lds, tw, L.sign1);
		last_pass(L.sign1);

		// pointwise multiply; first inverse pass
		if (in) {
#pragma unroll
			for (int q = 0; q < E; q++)
				v[q] = md.at(mb, b, p, k0 + q * NE, v[q]);
		}

		dft_any<E>(v, L.sign2, tw, N);

		lds_barrier();

#pragma unroll
		for (int q = 0; q < E; q++)
			lds[p * PN + k0 * E + q] = v[q];

		lds_barrier();
		ct_passes_between<N, P, E, NE>(lds, tw, L.sign2);
		last_pass(L.sign2);

#pragma unroll
		for (int q = 0; q < E; q++)
			acc[q] = cadd(acc[q], cmulc(v[q], md.zf(b, k0 + q * NE)));
	}

	if (in) {
		const auto sb = st.base(t0);
#pragma unroll
		for (int q = 0; q < E; q++)
			st.at(sb, p, k0 + q * NE, acc[q]);
	}
}

// Strided-axis transform for a pair of volumes that differ only by a
// factor along the axis, without a pointwise multiply.  Split (forward): the
// column h of volume v is multiplied by zf_b(k) and transformed, for b = 0, 1,
// into volumes v and v + pair_off.  Merge (adjoint): the columns of volumes
// v and v + pair_off are transformed and conj(zf_b(k)) x result is summed
// over b into volume v.  Same launch geometry as k_fft_axis_pair.
template <int N, bool MERGE, class Ld, class St>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis_pair_io(FftLaunch L, Ld ld, St st, const float2* zf0,
								  const float2* zf1, long pair_off)
{
	constexpr int P = pair_block_p(N);
	static_assert(edges_in_regs_p<N, P>(), "pair kernel needs the register-edge geometry");
	constexpr int LOGP = ct_log2(P);
	constexpr int PN = N + 1;
	constexpr int E = P * N / FFT_THREADS;
	constexpr int NE = N / E;

	extern __shared__ float2 lds[];
	float2* tw = lds + P * PN;

	if (nullptr != L.skip && 0 != *L.skip)
		return;

	const long t0 = (long)blockIdx.x * P;

	for (int i = threadIdx.x; i < N; i += FFT_THREADS)
		tw[i] = L.plan.tw[i];

	const int p = threadIdx.x & (P - 1);
	const int k0 = threadIdx.x >> LOGP;
	const bool in = t0 + p < L.total;
	float2 h[E], v[E];  // split: the input column; merge: the sum

	const auto lb = ld.base(t0);
	const auto sb = st.base(t0);

#pragma unroll
	for (int j = 0; j < E; j++)
		h[j] = (!MERGE && in) ? ld.at(lb, p, k0 + j * NE) : make_float2(0.f, 0.f);

	for (int b = 0; b < 2; b++) {

		const float2* zf = (0 == b) ? zf0 : zf1;

		if constexpr (MERGE) {
			auto lbb = lb;
			lbb.off += b * pair_off;
#pragma unroll
			for (int j = 0; j < E; j++)
				v[j] = in ? ld.at(lbb, p, k0 + j * NE) : make_float2(0.f, 0.f);
		} else {
#pragma unroll
			for (int j = 0; j < E; j++)
				v[j] = cmul(h[j], zf[k0 + j * NE]);
		}

		dft_any<E>(v, L.sign1, tw, N);

		if (b > 0)
			lds_barrier();  // the previous member's last pass has read LDS

#pragma unroll
		for (int q = 0; q < E; q++)
			lds[p * PN + k0 * E + q] = v[q];

		lds_barrier();
		ct_passes_between<N, P, E, NE>(lds, tw, L.sign1);

		// last radix-E pass into registers
#pragma unroll
		for (int r = 0; r < E; r++) {
			float2 x = lds[p * PN + k0 + r * NE];
			if (r > 0) {
				float2 w = tw[r * k0];
				if (L.sign1 > 0.f)
					w.y = -w.y;
				x = cmul(x, w);
			}
			v[r] = x;
		}
		dft_any<E>(v, L.sign1, tw, N);

		if constexpr (MERGE) {
#pragma unroll
			for (int q = 0; q < E; q++)
				h[q] = cadd(h[q], cmulc(v[q], zf[k0 + q * NE]));
		} else if (in) {
			auto sbb = sb;
			sbb.off += b * pair_off;
#pragma unroll
			for (int q = 0; q < E; q++)
				st.at(sbb, p, k0 + q * NE, v[q]);
		}
	}

	if (MERGE && in) {
#pragma unroll
		for (int q = 0; q < E; q++)
			st.at(sb, p, k0 + q * NE, h[q]);
	}
}

// ---------------------------------------------------------------------------
// Strided multi-transform passes in LDS for whole 2-D slices: P transforms
// of length N, element j of transform p at lds[p TS + j ES].  Butterflies
// are numbered with p fastest so that neighbouring lanes touch neighbouring
// transforms (consecutive addresses for the column transforms of a
// row-major slice, row pitch TS apart for the row transforms).  NTH threads.
// ---------------------------------------------------------------------------

template <int N, int R, int NS, int P, int ES, int TS, int NTH>
__device__ inline void slice_pass(float2* lds, const float2* tw, float sgn)
{
	constexpr int nR = N / R;
	constexpr int NB = P * nR;
	constexpr int PER = (NB + NTH - 1) / NTH;
	constexpr int TSTEP = N / (NS * R);
	float2 v[PER][R];
	int dst[PER];

	// the lane index passes through an empty asm so that the butterfly
	// addresses are computed here rather than hoisted out of the caller's
	// loops (they would occupy registers across all passes)
	int lane = threadIdx.x;
	asm volatile("" : "+v"(lane));

#pragma unroll
	for (int i = 0; i < PER; i++) {

		const int b = lane + i * NTH;
		dst[i] = -1;

		if (0 == NB % NTH || b < NB) {

			const int p = b % P;
			const int j = b / P;
			const int k = j % NS;
			const float2* src = lds + p * TS + j * ES;

#pragma unroll
			for (int r = 0; r < R; r++) {

				float2 x = src[r * nR * ES];

				if (r > 0 && NS > 1) {
					float2 w = tw[r * k * TSTEP];
					if (sgn > 0.f)
						w.y = -w.y;
					x = cmul(x, w);
				}

				v[i][r] = x;
			}

			dft_any<R>(v[i], sgn, tw, N);
			dst[i] = p * TS + ((j / NS) * NS * R + k) * ES;
		}

		// one butterfly's loads at a time (register pressure)
		__builtin_amdgcn_sched_barrier(0);
	}

	lds_barrier();

#pragma unroll
	for (int i = 0; i < PER; i++)
		if (0 == NB % NTH || dst[i] >= 0)
#pragma unroll
			for (int q = 0; q < R; q++)
				lds[dst[i] + q * NS * ES] = v[i][q];

	lds_barrier();
}

template <int N, int P, int ES, int TS, int NTH, int IDX = 0, int NS = 1>
__device__ inline void slice_passes(float2* lds, const float2* tw, float sgn)
{
	if constexpr (IDX < ct_num_radices(N, SLICE_R16)) {
		constexpr int R = ct_radix(N, IDX, SLICE_R16);
		slice_pass<N, R, NS, P, ES, TS, NTH>(lds, tw, sgn);
		slice_passes<N, P, ES, TS, NTH, IDX + 1, NS * R>(lds, tw, sgn);
	}
}

// The passes of a length-N transform after its first passes have built
// sub-transforms of length NS in place (M = N / NS still to combine).
template <int N, int M, int NS, int P, int ES, int TS, int NTH>
__device__ inline void slice_passes_rest(float2* lds, const float2* tw, float sgn)
{
	if constexpr (M > 1) {
		constexpr int R = SLICE_R16 ? ct_next_radix16(M) : ct_next_radix(M);
		slice_pass<N, R, NS, P, ES, TS, NTH>(lds, tw, sgn);
		slice_passes_rest<N, M / R, NS * R, P, ES, TS, NTH>(lds, tw, sgn);
	}
}

// the passes from sub-transforms of length NS up to (not including) length
// NSEND, radices as slice_passes chooses them
template <int N, int NS, int NSEND, int P, int ES, int TS, int NTH>
__device__ inline void slice_passes_range(float2* lds, const float2* tw, float sgn)
{
	if constexpr (NS < NSEND) {
		constexpr int R = SLICE_R16 ? ct_next_radix16(N / NS) : ct_next_radix(N / NS);
		slice_pass<N, R, NS, P, ES, TS, NTH>(lds, tw, sgn);
		slice_passes_range<N, NS * R, NSEND, P, ES, TS, NTH>(lds, tw, sgn);
	}
}

// radix of the last pass of slice_passes for length n
constexpr int slice_last_radix(int n)
{
	int m = n, r = n;
	while (m > 1) {
		r = SLICE_R16 ? ct_next_radix16(m) : ct_next_radix(m);
		m /= r;
	}
	return r;
}

// The last (radix R) pass of P length-N transforms stored as rows (element
// stride 1, row pitch LD) with the lanes along the butterfly index j of a
// row: butterfly (p, j) reads j + r N / R and yields the outputs x = j +
// q N / R of row p, handed to f(m, q, p, x, value) (m: the thread's
// butterfly count) instead of being written back.
template <int N, int R, int P, int LD, int NTH, class F>
__device__ inline void slice_last_pass_rows(const float2* lds, const float2* tw, float sgn, F&& f)
{
	constexpr int NS = N / R;
	constexpr int NB = P * NS;
	constexpr int PER = (NB + NTH - 1) / NTH;

#pragma unroll
	for (int m = 0; m < PER; m++) {

		const int b = threadIdx.x + m * NTH;

		if (0 != NB % NTH && b >= NB)
			break;

		const int p = b / NS;
		const int j = b - p * NS;
		float2 v[R];

#pragma unroll
		for (int r = 0; r < R; r++) {
			float2 x = lds[p * LD + j + r * NS];
			if (r > 0) {
				float2 w = tw[r * j];
				if (sgn > 0.f)
					w.y = -w.y;
				x = cmul(x, w);
			}
			v[r] = x;
		}

		dft_any<R>(v, sgn, tw, N);

#pragma unroll
		for (int q = 0; q < R; q++)
			f(m, q, p, j + q * NS, v[q]);
	}
}

// lengths with a compile-time kernel
inline bool ct_supported(int n)
{
	switch (n) {
	case 16: case 32: case 64: case 128: case 192: case 256: case 320: case 384: case 512: case 640: case 768:
		return true;
	default:
		return false;
	}
}

// Helper for element-wise stores: calls f(t, k, value) for every element of
// the block in an order that is coalesced for the launch's geometry.
template <class F>
__device__ inline void store_each(const float2* lds, const FftLaunch& L, long t0, const F& f)
{
	const int n = L.plan.n;
	const int P = L.P;
	const int tot = P * n;

	for (int i = threadIdx.x; i < tot; i += FFT_THREADS) {

		int p, k;

		if (L.contig) {
			p = i / n;
			k = i - p * n;
		} else {
			p = i & (P - 1);
			k = i >> L.logP;
		}

		long t = t0 + p;

		if (t < L.total)
			f(t, k, lds[p * L.pitch + k]);
	}
}

// Launch helper: picks P and EPT for the transform length and launches.
// P_min: minimum transforms per block (e.g. the group size of a reducing
// store); `contig` selects the geometry.
struct FftLaunchCfg {
	FftLaunch L;
	unsigned blocks;
	size_t lds_bytes;
	int ept;
	bool fixed = false;  // compile-time length kernel
	const char* name = "fft_axis";
};

FftLaunchCfg fft_config(const AxisPlan& plan, long total, bool contig, int P_min, float sign1, float sign2);

// the launch geometry of k_fft_axis_pair (its own columns per block)
inline void fft_config_pair(FftLaunchCfg& c)
{
	const int n = c.L.plan.n;
	const int P = pair_block_p(n);
	int logP = 0;
	while ((1 << logP) < P)
		logP++;
	c.L.P = P;
	c.L.logP = logP;
	c.blocks = (unsigned)((c.L.total + P - 1) / P);
	c.lds_bytes = (size_t)(P * (n + 1) + n) * sizeof(float2);
}

// k_fft_axis_n_batch over nblocks work-groups for transforms of length n
// (a compile-time length); items and btab device arrays
template <class Ld, class Md, class St>
void launch_fft_axis_batch(int n, const FftItem<Ld, Md, St>* items, const int2* btab, int nblocks, size_t lds_bytes,
			   hipStream_t stream, const char* name)
{
	if (nblocks <= 0)
		return;

	ProfScope prof(name, stream);
	const dim3 gb((unsigned)nblocks), tb(FFT_THREADS);

#define BAM_FFT_N(NN) case NN: hipLaunchKernelGGL((k_fft_axis_n_batch<NN, Ld, Md, St>), gb, tb, lds_bytes, stream, items, btab); break
	switch (n) {
		BAM_FFT_N(16);
		BAM_FFT_N(32);
		BAM_FFT_N(64);
		BAM_FFT_N(128);
		BAM_FFT_N(192);
		BAM_FFT_N(256);
		BAM_FFT_N(320);
		BAM_FFT_N(384);
		BAM_FFT_N(512);
		BAM_FFT_N(640);
		BAM_FFT_N(768);
	default:
		fail("batched FFT: no kernel for n = %d", n);
	}
#undef BAM_FFT_N

	BAM_HIP(hipGetLastError());
}

template <class Ld, class Md, class St>
void launch_fft_axis(const FftLaunchCfg& cfg, hipStream_t stream, const Ld& ld, const Md& md, const St& st)
{
	if (0 == cfg.blocks)
		return;

	ProfScope prof(cfg.name, stream);

	if (cfg.fixed) {

		const dim3 gb(cfg.blocks), tb(FFT_THREADS);
		const size_t lds = cfg.lds_bytes;

#define BAM_FFT_N(NN) case NN: hipLaunchKernelGGL((k_fft_axis_n<NN, Ld, Md, St>), gb, tb, lds, stream, cfg.L, ld, md, st); break
		switch (cfg.L.plan.n) {
			BAM_FFT_N(16);
			BAM_FFT_N(32);
			BAM_FFT_N(64);
			BAM_FFT_N(128);
			BAM_FFT_N(192);
			BAM_FFT_N(256);
			BAM_FFT_N(320);
			BAM_FFT_N(384);
			BAM_FFT_N(512);
			BAM_FFT_N(640);
			BAM_FFT_N(768);
		default:
			break;
		}
#undef BAM_FFT_N

		BAM_HIP(hipGetLastError());
		return;
	}

	bool gen = false;

	for (int i = 0; i < cfg.L.plan.nr; i++)
		gen = gen || (cfg.L.plan.radix[i] > 8 || 7 == cfg.L.plan.radix[i]);

	const dim3 gb(cfg.blocks), tb(FFT_THREADS);

	if (gen) {
		switch (cfg.ept) {
		case 8: hipLaunchKernelGGL((k_fft_axis<8, true, Ld, Md, St>), gb, tb, cfg.lds_bytes, stream, cfg.L, ld, md, st); break;
		case 16: hipLaunchKernelGGL((k_fft_axis<16, true, Ld, Md, St>), gb, tb, cfg.lds_bytes, stream, cfg.L, ld, md, st); break;
		default: hipLaunchKernelGGL((k_fft_axis<32, true, Ld, Md, St>), gb, tb, cfg.lds_bytes, stream, cfg.L, ld, md, st); break;
		}
	} else {
		switch (cfg.ept) {
		case 8: hipLaunchKernelGGL((k_fft_axis<8, false, Ld, Md, St>), gb, tb, cfg.lds_bytes, stream, cfg.L, ld, md, st); break;
		case 16: hipLaunchKernelGGL((k_fft_axis<16, false, Ld, Md, St>), gb, tb, cfg.lds_bytes, stream, cfg.L, ld, md, st); break;
		default: hipLaunchKernelGGL((k_fft_axis<32, false, Ld, Md, St>), gb, tb, cfg.lds_bytes, stream, cfg.L, ld, md, st); break;
		}
	}

	BAM_HIP(hipGetLastError());
}

template <class Ld, class Md, class St>
void launch_fft_axis_pair(const FftLaunchCfg& cfg, hipStream_t stream, const Ld& ld, const Md& md, const St& st)
{
	BAM_CHECK(cfg.fixed && !cfg.L.contig && fft_pair_len_ok(cfg.L.plan.n) &&
			  cfg.L.P == pair_block_p(cfg.L.plan.n),
		  "pair transform: unsupported geometry (n = %d, P = %d)", cfg.L.plan.n, cfg.L.P);

	if (0 == cfg.blocks)
		return;

	ProfScope prof(cfg.name, stream);
	const dim3 gb(cfg.blocks), tb(FFT_THREADS);

#define BAM_FFT_N(NN) case NN: hipLaunchKernelGGL((k_fft_axis_pair<NN, Ld, Md, St>), gb, tb, cfg.lds_bytes, stream, cfg.L, ld, md, st); break
	switch (cfg.L.plan.n) {
		BAM_FFT_N(64);
		BAM_FFT_N(128);
		BAM_FFT_N(256);
		BAM_FFT_N(512);
	default:
		fail("pair transform: no kernel for n = %d", cfg.L.plan.n);
	}
#undef BAM_FFT_N

	BAM_HIP(hipGetLastError());
}

template <bool MERGE, class Ld, class St>
void launch_fft_axis_pair_io(const FftLaunchCfg& cfg, hipStream_t stream, const Ld& ld, const St& st,
			     const float2* zf0, const float2* zf1, long pair_off)
{
	BAM_CHECK(cfg.fixed && !cfg.L.contig && fft_pair_len_ok(cfg.L.plan.n) &&
			  cfg.L.P == pair_block_p(cfg.L.plan.n),
		  "pair transform: unsupported geometry (n = %d, P = %d)", cfg.L.plan.n, cfg.L.P);

	if (0 == cfg.blocks)
		return;

	ProfScope prof(cfg.name, stream);
	const dim3 gb(cfg.blocks), tb(FFT_THREADS);

#define BAM_FFT_N(NN) case NN: hipLaunchKernelGGL((k_fft_axis_pair_io<NN, MERGE, Ld, St>), gb, tb, cfg.lds_bytes, stream, cfg.L, ld, st, zf0, zf1, pair_off); break
	switch (cfg.L.plan.n) {
		BAM_FFT_N(64);
		BAM_FFT_N(128);
		BAM_FFT_N(256);
		BAM_FFT_N(512);
	default:
		fail("pair transform: no kernel for n = %d", cfg.L.plan.n);
	}
#undef BAM_FFT_N

	BAM_HIP(hipGetLastError());
}

}  // namespace bam
