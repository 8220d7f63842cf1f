// chain.hip -- the chained NUFFT for oversampling factors other than 2
// (noncart/nufft_chain.c:192-244, selected by nufft_create2 for os != 2,
// noncart/nufft.c:1184-1199), and in the same form the non-decomposed
// NUFFT of decomp = false (nufft -1, :898-935: the image is the oversampled
// grid, no padding, rolloff with os = 1):
//   forward: x -> rolloff into the centre of the round(os n) grid
//            (linop_kb_rolloff_create, nufft_chain.c:60-118) -> centred
//            unitary FFT (fftmod(N2^-1/4) FFT fftmod(N2^-1/4),
//            linops/someops.c:1825-1879) -> KB interpolation on that grid
//            (grid2H with os, width, beta = calc_beta(os, width)) -> x
//            weights os^(d/2)
//   adjoint: the adjoint of every step, in reverse.
// The interpolation is a gather per sample; its adjoint is a gather per
// grid point over the samples of the nearby cells (samples sorted by cells
// of CHAIN_CS points), a tile of points per work-group walking the tile's
// samples in LDS batches, so no atomics are used and the result is
// deterministic.
#include "nufft_impl.hpp"
#include "prof.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace bam {

struct ChainGeom {
	int n[3], on[3], off[3], ncell[3], cs[3];
	long N, N2, M;
	int C;
	float os, width;
	bool periodic;
};

// tiles of the adjoint gather: extent t per dim (product <= 256) and
// number of tiles nt per dim
struct ChainTiles {
	int t[3], nt[3];
};

static ChainTiles tiles(const Chain& c)
{
	int act = 0;
	for (int d = 0; d < 3; d++)
		act += (c.on[d] > 1) ? 1 : 0;

	ChainTiles tl;
	int k = 0;

	for (int d = 0; d < 3; d++) {
		tl.t[d] = 1;
		if (c.on[d] > 1) {
			tl.t[d] = (3 == act) ? ((2 == k) ? 4 : 8) : ((2 == act) ? 16 : 256);
			k++;
		}
		tl.nt[d] = (c.on[d] + tl.t[d] - 1) / tl.t[d];
	}

	return tl;
}

// cells whose samples can reach some grid index of [a, b] along dim d: the
// cells within width / 2 of the range plus one on each side (rounding of
// the positions), wrapped (periodic) or clamped, every cell at most once;
// returns the count, fills the first cell
static int tile_span(const Chain& c, int d, int a, int b, int& first)
{
	if (1 == c.on[d]) {
		first = 0;
		return 1;
	}

	const int nc = c.ncell[d];
	int lo = (int)floorf(((float)a - 0.5f * c.width) / (float)c.cs[d]) - 1;
	int hi = (int)floorf(((float)b + 0.5f * c.width) / (float)c.cs[d]) + 1;

	if (hi - lo + 1 >= nc) {
		first = 0;
		return nc;
	}

	if (!c.periodic) {
		lo = std::max(lo, 0);
		hi = std::min(hi, nc - 1);
		first = lo;
		return std::max(hi - lo + 1, 0);
	}

	first = ((lo % nc) + nc) % nc;
	return hi - lo + 1;
}

static ChainGeom geom(const Chain& c)
{
	ChainGeom g;

	for (int d = 0; d < 3; d++) {
		g.n[d] = c.n[d];
		g.on[d] = c.on[d];
		g.off[d] = c.off[d];
		g.ncell[d] = c.ncell[d];
		g.cs[d] = c.cs[d];
	}

	g.N = c.N;
	g.N2 = c.N2;
	g.M = c.M;
	g.C = c.C;
	g.os = c.os;
	g.width = c.width;
	g.periodic = c.periodic;
	return g;
}

static unsigned nblocks(long n) { return (unsigned)std::min<long>(std::max<long>((n + 255) / 256, 1), 16384); }

// sample position on the grid of dim d: os k + on / 2 (noncart/grid.c:181-187)
__device__ inline float chain_pos(const ChainGeom& g, int d, float k)
{
	return grid_pos(g.os, k, 0.f, (g.on[d] > 1) ? (float)(g.on[d] / 2) : 0.f);
}

// first tap, tap count and the index offset of one dim (noncart/grid.c:415-441):
// cnt < 0: the sample has no taps (it is skipped)
struct TapRange {
	int sti, cnt, off;
};

__device__ inline TapRange tap_range(const ChainGeom& g, int d, float p)
{
	TapRange r;
	r.off = 0;

	if (1 == g.on[d]) {
		r.sti = 0;
		r.cnt = 1;
		return r;
	}

	int sti = (int)ceil((double)p - 0.5 * (double)g.width);
	int eni = (int)floor((double)p + 0.5 * (double)g.width);

	if (sti > eni) {
		r.sti = 0;
		r.cnt = -1;
		return r;
	}

	if (!g.periodic) {
		sti = max(sti, 0);
		eni = min(eni, g.on[d] - 1);
	} else {
		while (sti + r.off < 0)
			r.off += g.on[d];
	}

	r.sti = sti;
	r.cnt = max(eni - sti + 1, 0);
	return r;
}

// weight of tap j (grid.c:453-457); dims of size 1 weigh |0 - p|
__device__ inline float tap_weight(const ChainGeom& g, const float* kb, int d, int j, float p)
{
	return kb_lookup(kb, fabsf((float)((1 == g.on[d]) ? 0 : j) - p) / g.width);
}

// G[c][o] = roll(x) src[c][x] fma[o] inside the centred image, 0 outside
__global__ void k_chain_pad(ChainGeom g, const float* __restrict__ roll, const float2* __restrict__ fma,
			    const float2* __restrict__ src, float2* __restrict__ G)
{
	const long total = g.N2 * g.C;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long o = e % g.N2;
		const long c = e / g.N2;
		const int ox = (int)(o % g.on[0]);
		const int oy = (int)((o / g.on[0]) % g.on[1]);
		const int oz = (int)(o / ((long)g.on[0] * g.on[1]));
		const int x = ox - g.off[0], y = oy - g.off[1], z = oz - g.off[2];
		float2 v = make_float2(0.f, 0.f);

		if (x >= 0 && x < g.n[0] && y >= 0 && y < g.n[1] && z >= 0 && z < g.n[2]) {
			const long i = x + (long)g.n[0] * (y + (long)g.n[1] * z);
			v = cmul(cscale(src[c * g.N + i], roll[i]), fma[o]);
		}

		G[e] = v;
	}
}

// G[c][o] *= fma[o] (or conj(fma[o]))
__global__ void k_chain_mul(ChainGeom g, const float2* __restrict__ fma, float2* __restrict__ G, bool conj)
{
	const long total = g.N2 * g.C;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const float2 f = fma[e % g.N2];
		G[e] = conj ? cmulc(G[e], f) : cmul(G[e], f);
	}
}

// dst[c][x] = roll(x) (G[c][o] conj(fma[o])), o the centred position of x
__global__ void k_chain_crop(ChainGeom g, const float* __restrict__ roll, const float2* __restrict__ fma,
			     const float2* __restrict__ G, float2* __restrict__ dst)
{
	const long total = g.N * g.C;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long i = e % g.N;
		const long c = e / g.N;
		const int x = (int)(i % g.n[0]);
		const int y = (int)((i / g.n[0]) % g.n[1]);
		const int z = (int)(i / ((long)g.n[0] * g.n[1]));
		const long o = (x + g.off[0]) + (long)g.on[0] * ((y + g.off[1]) + (long)g.on[1] * (z + g.off[2]));
		dst[e] = cscale(cmulc(G[c * g.N2 + o], fma[o]), roll[i]);
	}
}

// ksp[c][m] = (sum over the taps of G[c]) x (weights x scale or scale)
// (grid2H, noncart/grid.c:153-199, grid_pointH :494-512: z outer, x inner)
__global__ void k_chain_gridH(ChainGeom g, const float* __restrict__ kbt, const float4* __restrict__ trj,
			      const float2* __restrict__ tmp, float scale, const float2* __restrict__ G,
			      float2* __restrict__ ksp)
{
	__shared__ float kb[KB_SIZE + 1];

	for (int i = threadIdx.x; i < KB_SIZE + 1; i += blockDim.x)
		kb[i] = kbt[i];

	__syncthreads();

	for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < g.M; m += (long)gridDim.x * blockDim.x) {

		const float4 k = trj[m];
		const float p[3] = { chain_pos(g, 0, k.x), chain_pos(g, 1, k.y), chain_pos(g, 2, k.z) };
		TapRange r[3];
		bool ok = true;

		for (int d = 0; d < 3; d++) {
			r[d] = tap_range(g, d, p[d]);
			ok = ok && (r[d].cnt >= 0);
		}

		const float2 f = (nullptr != tmp) ? tmp[m] : make_float2(scale, 0.f);

		for (int c = 0; c < g.C; c++) {

			float2 acc = make_float2(0.f, 0.f);

			if (ok) {
				const float2* Gc = G + (long)c * g.N2;

				for (int iz = 0; iz < r[2].cnt; iz++) {

					const int jz = r[2].sti + iz;
					const float dz = 1.f * tap_weight(g, kb, 2, jz, p[2]);
					const long oz = (1 == g.on[2]) ? 0 : (long)((jz + r[2].off) % g.on[2]);

					for (int iy = 0; iy < r[1].cnt; iy++) {

						const int jy = r[1].sti + iy;
						const float dy = dz * tap_weight(g, kb, 1, jy, p[1]);
						const long oy = (1 == g.on[1]) ? 0 : (long)((jy + r[1].off) % g.on[1]);

						for (int ix = 0; ix < r[0].cnt; ix++) {

							const int jx = r[0].sti + ix;
							const float dd = dy * tap_weight(g, kb, 0, jx, p[0]);
							const long ox = (1 == g.on[0]) ? 0 : (long)((jx + r[0].off) % g.on[0]);
							const float2 v = Gc[ox + g.on[0] * (oy + (long)g.on[1] * oz)];
							acc.x += v.x * dd;
							acc.y += v.y * dd;
						}
					}
				}
			}

			ksp[(long)c * g.M + m] = (nullptr != tmp) ? cmul(acc, f) : cscale(acc, scale);
		}
	}
}

constexpr int CHAIN_CC = 8;   // coils per work-group of the adjoint gather
constexpr int CHAIN_TB = 128; // samples per batch of the adjoint gather
constexpr int CHAIN_SMAX = 1024;  // samples per work-group of the adjoint gather

// kws[(ch M + s) CHAIN_CC + q] = ksp[c][m(s)] conj(weights x scale) (or
// x scale), c = ch CHAIN_CC + q, m(s) the original index of sorted sample s
// (0 past the last coil): the adjoint of the last diagonal, written in cell
// order with the coils of a chunk side by side
__global__ void k_chain_kscale(ChainGeom g, const int* __restrict__ sidx, const float2* __restrict__ tmp, float scale,
			       const float2* __restrict__ ksp, float2* __restrict__ kws)
{
	const int nch = (g.C + CHAIN_CC - 1) / CHAIN_CC;
	const long total = g.M * nch;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const int c0 = (int)(e / g.M) * CHAIN_CC;
		const long m = sidx[e % g.M];

		for (int q = 0; q < CHAIN_CC; q++) {
			float2 v = make_float2(0.f, 0.f);
			if (c0 + q < g.C) {
				const float2 x = ksp[(long)(c0 + q) * g.M + m];
				v = (nullptr != tmp) ? cmulc(x, tmp[m]) : cscale(x, scale);
			}
			kws[e * CHAIN_CC + q] = v;
		}
	}
}

// G[c][o] = sum over the samples of the nearby cells of w(o) kw[c][m]
// (the adjoint of k_chain_gridH: grid2, noncart/grid.c:202-253).
// A work-group owns a tile of grid points (one per thread), a chunk of
// CHAIN_CC coils and a slice [k CHAIN_SMAX, (k + 1) CHAIN_SMAX) of the
// samples of the cells that can reach the tile (taken in cell order: z, y,
// x, then sample order).  It walks its samples in batches: one thread per
// sample evaluates its taps (first grid index, count and KB weights per dim)
// into LDS and the coil values come in as one contiguous run; then every
// point adds the samples whose taps cover it, in order.  A tile with one
// slice writes the grid; the slices of a denser tile write partial tiles
// that k_chain_sum adds in slice order, so the result is deterministic and
// needs no atomics.
__global__ __launch_bounds__(256) void k_chain_grid(ChainGeom g, ChainTiles tl, const float* __restrict__ kbt,
						    const int4* __restrict__ work, const int* __restrict__ span,
						    const float4* __restrict__ spos, const int* __restrict__ cstart,
						    const float2* __restrict__ kws, float2* __restrict__ G,
						    float2* __restrict__ part, long nslots)
{
	__shared__ float kb[KB_SIZE + 1];
	__shared__ int sbase[CHAIN_TB][3];
	__shared__ int scnt[CHAIN_TB][3];
	__shared__ float sw[CHAIN_TB][3][CHAIN_TAPS];
	__shared__ float2 sv[CHAIN_TB * CHAIN_CC];

	for (int i = threadIdx.x; i < KB_SIZE + 1; i += blockDim.x)
		kb[i] = kbt[i];

	const int4 wi = work[blockIdx.x];  // tile, slice, slices of the tile, first partial slot
	const int ch = blockIdx.y;
	const int c0 = ch * CHAIN_CC;
	const long b = wi.x;
	const int tc[3] = { (int)(b % tl.nt[0]), (int)((b / tl.nt[0]) % tl.nt[1]), (int)(b / ((long)tl.nt[0] * tl.nt[1])) };
	const int l = threadIdx.x;
	const int lc[3] = { l % tl.t[0], (l / tl.t[0]) % tl.t[1], l / (tl.t[0] * tl.t[1]) };
	int o[3], first[3], count[3];
	bool inside = l < tl.t[0] * tl.t[1] * tl.t[2];

	for (int d = 0; d < 3; d++) {
		o[d] = tc[d] * tl.t[d] + lc[d];
		inside = inside && o[d] < g.on[d];
		first[d] = span[6 * b + d];
		count[d] = span[6 * b + 3 + d];
	}

	// this work-group's slice of the tile's samples
	const long v0 = (long)wi.y * CHAIN_SMAX;
	const long v1 = v0 + CHAIN_SMAX;
	long vo = 0;  // samples of the tile before the current run

	float2 acc[CHAIN_CC];

	for (int q = 0; q < CHAIN_CC; q++)
		acc[q] = make_float2(0.f, 0.f);

	for (int cz = 0; cz < count[2]; cz++)
		for (int cy = 0; cy < count[1]; cy++) {

			const int iz = (first[2] + cz) % g.ncell[2];
			const int iy = (first[1] + cy) % g.ncell[1];
			const long row = (long)g.ncell[0] * (iy + (long)g.ncell[1] * iz);
			// the x cells as up to two contiguous runs (wrapped)
			const int n0 = min(count[0], g.ncell[0] - first[0]);
			const int runs[2][2] = { { first[0], n0 }, { 0, count[0] - n0 } };

			for (int r = 0; r < 2; r++) {

				if (runs[r][1] <= 0)
					continue;

				const int r0 = cstart[row + runs[r][0]];
				const int r1 = cstart[row + runs[r][0] + runs[r][1]];
				// the part of [r0, r1) inside the slice
				const int s0 = (int)(r0 + std::min<long>(std::max<long>(v0 - vo, 0), r1 - r0));
				const int s1 = (int)(r0 + std::min<long>(std::max<long>(v1 - vo, 0), r1 - r0));
				vo += r1 - r0;

				for (int sb = s0; sb < s1; sb += CHAIN_TB) {

					const int nb = min(CHAIN_TB, s1 - sb);

					__syncthreads();  // the previous batch has been read

					if (threadIdx.x < nb) {
						const int j = threadIdx.x;
						const float4 pp = spos[sb + j];
						const float p[3] = { pp.x, pp.y, pp.z };

						for (int d = 0; d < 3; d++) {
							const TapRange tr = tap_range(g, d, p[d]);
							const int cnt = max(tr.cnt, 0);
							sbase[j][d] = (1 == g.on[d]) ? 0
								     : (g.periodic ? (tr.sti + tr.off) % g.on[d] : tr.sti);
							scnt[j][d] = cnt;
							for (int t = 0; t < cnt; t++)
								sw[j][d][t] = tap_weight(g, kb, d, tr.sti + t, p[d]);
						}
					}

					const float2* src = kws + ((long)ch * g.M + sb) * CHAIN_CC;

					for (int i = threadIdx.x; i < nb * CHAIN_CC; i += blockDim.x)
						sv[i] = src[i];

					__syncthreads();

					if (!inside)
						continue;

					for (int j = 0; j < nb; j++) {

						int dl[3];
						bool hit = true;

						for (int d = 2; d >= 0; d--) {
							int e = o[d] - sbase[j][d];
							if (g.periodic && e < 0)
								e += g.on[d];
							dl[d] = e;
							hit = hit && (unsigned)e < (unsigned)scnt[j][d];
						}

						if (!hit)
							continue;

						const float w = ((1.f * sw[j][2][dl[2]]) * sw[j][1][dl[1]]) * sw[j][0][dl[0]];

						for (int q = 0; q < CHAIN_CC; q++) {
							const float2 v = sv[j * CHAIN_CC + q];
							acc[q].x += v.x * w;
							acc[q].y += v.y * w;
						}
					}
				}
			}
		}

	if (1 == wi.z) {
		if (inside) {
			const long oi = o[0] + (long)g.on[0] * (o[1] + (long)g.on[1] * o[2]);
			for (int q = 0; q < CHAIN_CC; q++)
				if (c0 + q < g.C)
					G[(long)(c0 + q) * g.N2 + oi] = acc[q];
		}
	} else {
		float2* pp = part + ((long)ch * nslots + wi.w + wi.y) * (CHAIN_CC * 256);
		for (int q = 0; q < CHAIN_CC; q++)
			pp[q * 256 + l] = acc[q];
	}
}

// G[c][o] = sum over the slices k of the partial tiles, in order (tiles
// whose samples were split over several work-groups)
__global__ __launch_bounds__(256) void k_chain_sum(ChainGeom g, ChainTiles tl, const int4* __restrict__ split,
						   const float2* __restrict__ part, long nslots, float2* __restrict__ G)
{
	const int4 si = split[blockIdx.x];  // tile, slices, first partial slot
	const int ch = blockIdx.y;
	const int c0 = ch * CHAIN_CC;
	const long b = si.x;
	const int tc[3] = { (int)(b % tl.nt[0]), (int)((b / tl.nt[0]) % tl.nt[1]), (int)(b / ((long)tl.nt[0] * tl.nt[1])) };
	const int l = threadIdx.x;
	const int lc[3] = { l % tl.t[0], (l / tl.t[0]) % tl.t[1], l / (tl.t[0] * tl.t[1]) };
	int o[3];
	bool inside = l < tl.t[0] * tl.t[1] * tl.t[2];

	for (int d = 0; d < 3; d++) {
		o[d] = tc[d] * tl.t[d] + lc[d];
		inside = inside && o[d] < g.on[d];
	}

	if (!inside)
		return;

	const long oi = o[0] + (long)g.on[0] * (o[1] + (long)g.on[1] * o[2]);
	const float2* pp = part + ((long)ch * nslots + si.z) * (CHAIN_CC * 256) + l;

	for (int q = 0; q < CHAIN_CC; q++) {
		if (c0 + q >= g.C)
			break;
		float2 acc = pp[q * 256];
		for (int k = 1; k < si.y; k++) {
			const float2 v = pp[(long)k * (CHAIN_CC * 256) + q * 256];
			acc.x += v.x;
			acc.y += v.y;
		}
		G[(long)(c0 + q) * g.N2 + oi] = acc;
	}
}

void Chain::create(const int dims[3], int C_, long M_, const float* traj, const float2* weights_host, bool periodic_,
		   float width_, float os_, bool pad)
{
	BAM_CHECK(os_ >= 1.f, "nufft: oversampling %f below 1", os_);
	BAM_CHECK(width_ > 0.f && (int)floorf(width_) + 1 <= CHAIN_TAPS, "nufft: kernel width %f not supported",
		  width_);

	C = C_;
	M = M_;
	os = os_;
	width = width_;
	periodic = periodic_;
	N = N2 = 1;
	nd = 0;

	for (int d = 0; d < 3; d++) {

		n[d] = dims[d];
		on[d] = (n[d] > 1 && pad) ? (int)lround(os * n[d]) : n[d];
		off[d] = std::abs(on[d] / 2 - n[d] / 2);
		BAM_CHECK(1 == on[d] || (on[d] >= n[d] && on[d] > (int)floorf(width) + 1 && fft_size_supported(on[d])),
			  "nufft: oversampled size %d of dim %d not supported", on[d], d);
		// periodic grids wrap cells with the grid: cells of one point unless
		// CHAIN_CS divides the size
		cs[d] = (on[d] > 1 && (!periodic || 0 == on[d] % CHAIN_CS)) ? CHAIN_CS : 1;
		ncell[d] = (on[d] + cs[d] - 1) / cs[d];
		N *= n[d];
		N2 *= on[d];
		nd += (n[d] > 1);
	}

	// the table and the rolloff use beta rounded to float, as the first
	// kb_init call (the forward's rolloff) makes it (noncart/grid.c:45-60)
	beta = (double)(float)calc_beta(os, width);
	std::vector<float> kbh(KB_SIZE + 1);
	kb_table_host(beta, kbh.data());
	kb.alloc(KB_SIZE + 1);
	BAM_HIP(hipMemcpy(kb.ptr, kbh.data(), sizeof(float) * (KB_SIZE + 1), hipMemcpyHostToDevice));

	std::vector<float> rh(N);

	for (int z = 0; z < n[2]; z++)
		for (int y = 0; y < n[1]; y++)
			for (int x = 0; x < n[0]; x++) {
				const float ros = pad ? os : 1.f;
				const double rx = (n[0] > 1) ? rolloff_1d(n[0], x, ros, beta, width) : 1.;
				const double ry = (n[1] > 1) ? rolloff_1d(n[1], y, ros, beta, width) : 1.;
				const double rz = (n[2] > 1) ? rolloff_1d(n[2], z, ros, beta, width) : 1.;
				rh[x + (long)n[0] * (y + (long)n[1] * z)] = (float)(rx * ry * rz);
			}

	roll.alloc(N);
	BAM_HIP(hipMemcpy(roll.ptr, rh.data(), sizeof(float) * N, hipMemcpyHostToDevice));

	// fftmod of an array of N2^-1/4 (linop_fft_create_priv, someops.c:1835-1852)
	const float s4 = 1.f / sqrtf(sqrtf((float)N2));
	std::vector<float2> fh(N2);

	for (long o = 0; o < N2; o++) {
		const long c[3] = { o % on[0], (o / on[0]) % on[1], o / ((long)on[0] * on[1]) };
		double ph = 0.;
		for (int d = 0; d < 3; d++)
			if (on[d] > 1)
				ph += fftmod_phase(on[d], c[d]);
		const std::complex<double> v = fftmod_value(ph);
		fh[o] = make_float2(s4 * (float)v.real(), s4 * (float)v.imag());
	}

	fma.alloc(N2);
	BAM_HIP(hipMemcpy(fma.ptr, fh.data(), sizeof(float2) * N2, hipMemcpyHostToDevice));

	scale = powf(os, (float)nd / 2.f);
	has_weights = (nullptr != weights_host);

	if (has_weights) {
		std::vector<float2> th(M);
		for (long m = 0; m < M; m++)
			th[m] = cscale(weights_host[m], scale);
		tmp.alloc(std::max<long>(M, 1));
		if (M > 0)
			BAM_HIP(hipMemcpy(tmp.ptr, th.data(), sizeof(float2) * M, hipMemcpyHostToDevice));
	}

	// coordinates, and the samples sorted by cell (stable) for the adjoint
	std::vector<float4> kh(M), ph(M);
	std::vector<long> cell(M);
	const long ncells = (long)ncell[0] * ncell[1] * ncell[2];
	std::vector<int> cnt(ncells + 1, 0);

	for (long m = 0; m < M; m++) {

		float p[3];
		int ci[3];

		for (int d = 0; d < 3; d++) {
			const float k = traj[3 * m + d];
			float v = (float)((double)os * (double)k);
			if (on[d] > 1)
				v = v + (float)(on[d] / 2);
			p[d] = v;

			int c = 0;
			if (on[d] > 1) {
				int q = (int)floorf(v);
				q = periodic ? ((q % on[d]) + on[d]) % on[d] : std::min(std::max(q, 0), on[d] - 1);
				c = q / cs[d];
			}
			ci[d] = c;
		}

		kh[m] = make_float4(traj[3 * m], traj[3 * m + 1], traj[3 * m + 2], 0.f);
		ph[m] = make_float4(p[0], p[1], p[2], 0.f);
		cell[m] = ci[0] + (long)ncell[0] * (ci[1] + (long)ncell[1] * ci[2]);
		cnt[cell[m] + 1]++;
	}

	for (long c = 0; c < ncells; c++)
		cnt[c + 1] += cnt[c];

	std::vector<float4> sp(std::max<long>(M, 1));
	std::vector<int> si(std::max<long>(M, 1));
	std::vector<int> pos(cnt.begin(), cnt.end() - 1);

	for (long m = 0; m < M; m++) {
		const long j = pos[cell[m]]++;
		sp[j] = ph[m];
		si[j] = (int)m;
	}

	trj.alloc(std::max<long>(M, 1));
	spos.alloc(std::max<long>(M, 1));
	sidx.alloc(std::max<long>(M, 1));
	cstart.alloc(ncells + 1);

	if (M > 0) {
		BAM_HIP(hipMemcpy(trj.ptr, kh.data(), sizeof(float4) * M, hipMemcpyHostToDevice));
		BAM_HIP(hipMemcpy(spos.ptr, sp.data(), sizeof(float4) * M, hipMemcpyHostToDevice));
		BAM_HIP(hipMemcpy(sidx.ptr, si.data(), sizeof(int) * M, hipMemcpyHostToDevice));
	}

	BAM_HIP(hipMemcpy(cstart.ptr, cnt.data(), sizeof(int) * (ncells + 1), hipMemcpyHostToDevice));

	grid.alloc((size_t)C * N2);
	kwork.alloc(std::max<long>((long)((C + CHAIN_CC - 1) / CHAIN_CC) * CHAIN_CC * M, 1));

	// adjoint work: per tile its cell span and sample count, split into
	// slices of at most CHAIN_SMAX samples
	const ChainTiles tl = tiles(*this);
	const long ntiles = (long)tl.nt[0] * tl.nt[1] * tl.nt[2];
	BAM_CHECK(ntiles < (1L << 30), "nufft: grid too large for the adjoint gather");
	std::vector<int> sp6((size_t)ntiles * 6);
	std::vector<int4> wl, sl;
	long slots = 0;

	for (long t = 0; t < ntiles; t++) {

		const long tc[3] = { t % tl.nt[0], (t / tl.nt[0]) % tl.nt[1], t / ((long)tl.nt[0] * tl.nt[1]) };
		int f[3], n[3];

		for (int d = 0; d < 3; d++) {
			const int a = (int)tc[d] * tl.t[d];
			n[d] = tile_span(*this, d, a, std::min(a + tl.t[d], on[d]) - 1, f[d]);
			sp6[6 * t + d] = f[d];
			sp6[6 * t + 3 + d] = n[d];
		}

		long ns = 0;
		for (int cz = 0; cz < n[2]; cz++)
			for (int cy = 0; cy < n[1]; cy++) {
				const long row = (long)ncell[0] * ((f[1] + cy) % ncell[1] + (long)ncell[1] * ((f[2] + cz) % ncell[2]));
				const int n0 = std::min(n[0], ncell[0] - f[0]);
				ns += cnt[row + f[0] + n0] - cnt[row + f[0]];
				ns += cnt[row + (n[0] - n0)] - cnt[row];
			}

		const int K = (int)std::max<long>((ns + CHAIN_SMAX - 1) / CHAIN_SMAX, 1);

		for (int k = 0; k < K; k++)
			wl.push_back(make_int4((int)t, k, K, (K > 1) ? (int)slots : -1));

		if (K > 1) {
			sl.push_back(make_int4((int)t, K, (int)slots, 0));
			slots += K;
		}
	}

	BAM_CHECK(wl.size() < (1UL << 31) && slots < (1L << 30), "nufft: too many adjoint work items");
	nwork = (long)wl.size();
	nsplit = (long)sl.size();
	nslots = slots;
	cspan.alloc(sp6.size());
	BAM_HIP(hipMemcpy(cspan.ptr, sp6.data(), sizeof(int) * sp6.size(), hipMemcpyHostToDevice));
	cwork.alloc(wl.size());
	BAM_HIP(hipMemcpy(cwork.ptr, wl.data(), sizeof(int4) * wl.size(), hipMemcpyHostToDevice));
	csplit.alloc(std::max<size_t>(sl.size(), 1));
	if (!sl.empty())
		BAM_HIP(hipMemcpy(csplit.ptr, sl.data(), sizeof(int4) * sl.size(), hipMemcpyHostToDevice));
	cpart.alloc((size_t)std::max<long>(slots, 1) * ((C + CHAIN_CC - 1) / CHAIN_CC) * CHAIN_CC * 256);
}

void Chain::forward(float2* ksp, const float2* cim)
{
	const ChainGeom g = geom(*this);
	ProfScope prof("chain_forward", stream);

	hipLaunchKernelGGL(k_chain_pad, dim3(nblocks(N2 * C)), dim3(256), 0, stream, g, roll.ptr, fma.ptr, cim, grid.ptr);
	BAM_HIP(hipGetLastError());
	fft_nd(on, C, 7, -1.f, grid.ptr, grid.ptr, 1.f, stream);
	hipLaunchKernelGGL(k_chain_mul, dim3(nblocks(N2 * C)), dim3(256), 0, stream, g, fma.ptr, grid.ptr, false);
	BAM_HIP(hipGetLastError());

	if (M > 0) {
		hipLaunchKernelGGL(k_chain_gridH, dim3(nblocks(M)), dim3(256), 0, stream, g, kb.ptr, trj.ptr,
				   has_weights ? tmp.ptr : nullptr, scale, grid.ptr, ksp);
		BAM_HIP(hipGetLastError());
	}
}

void Chain::adjoint(float2* cim, const float2* ksp)
{
	const ChainGeom g = geom(*this);
	ProfScope prof("chain_adjoint", stream);

	const long nch = (C + CHAIN_CC - 1) / CHAIN_CC;

	if (M > 0) {
		hipLaunchKernelGGL(k_chain_kscale, dim3(nblocks(M * nch)), dim3(256), 0, stream, g, sidx.ptr,
				   has_weights ? tmp.ptr : nullptr, scale, ksp, kwork.ptr);
		BAM_HIP(hipGetLastError());
	}

	const ChainTiles tl = tiles(*this);
	BAM_CHECK(nch < 65536, "nufft: too many coils for the adjoint gather");
	hipLaunchKernelGGL(k_chain_grid, dim3((unsigned)nwork, (unsigned)nch), dim3(256), 0, stream, g, tl, kb.ptr,
			   cwork.ptr, cspan.ptr, spos.ptr, cstart.ptr, kwork.ptr, grid.ptr, cpart.ptr, nslots);
	BAM_HIP(hipGetLastError());

	if (nsplit > 0) {
		hipLaunchKernelGGL(k_chain_sum, dim3((unsigned)nsplit, (unsigned)nch), dim3(256), 0, stream, g, tl,
				   csplit.ptr, cpart.ptr, nslots, grid.ptr);
		BAM_HIP(hipGetLastError());
	}
	hipLaunchKernelGGL(k_chain_mul, dim3(nblocks(N2 * C)), dim3(256), 0, stream, g, fma.ptr, grid.ptr, true);
	BAM_HIP(hipGetLastError());
	fft_nd(on, C, 7, +1.f, grid.ptr, grid.ptr, 1.f, stream);
	hipLaunchKernelGGL(k_chain_crop, dim3(nblocks(N * C)), dim3(256), 0, stream, g, roll.ptr, fma.ptr, grid.ptr,
			   cim);
	BAM_HIP(hipGetLastError());
}

}  // namespace bam
