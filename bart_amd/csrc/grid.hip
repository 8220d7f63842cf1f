// grid.hip -- Kaiser-Bessel gridding for gfx950.
//
// Maths of noncart/grid.c:153-253 and :403-512 (tap range, clamp or wrap,
// separable weights from the 1001-entry KB table with linear interpolation,
// weight product in the order z, y, x).  Execution differs from
// gpu_grid.cu:
//   * samples are sorted once per trajectory by the cell (2^d grid points)
//     that holds their position; the plan lives with the operator;
//   * adjoint (k-space -> grid) is output-stationary: a work-group owns a tile
//     of all sub-grids for a group of coils, each wave a block of every
//     sub-grid (in 3-d 2 x 2 x 2 points = a 4 x 4 x 4 block of the 2x grid),
//     each lane one output in registers; a wave walks only the samples that
//     reach its block, with the coil values as wave-uniform loads.  There
//     are no atomics and no zero-fill of the grid;
//   * forward (grid -> k-space): a thread per sample gathers its taps in
//     cell-sorted order and stores k-space at the original positions.
#include "nufft_impl.hpp"
#include "prof.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <sys/stat.h>

namespace bam {

// ---------------------------------------------------------------------------
// Kaiser-Bessel table (noncart/grid.c:38-97, 516-519)
// ---------------------------------------------------------------------------

static double bessel_i0(double x)
{
	double q = 0.25 * x * x;
	double term = 1., sum = 1.;

	for (int k = 1; k < 500; k++) {
		term *= q / ((double)k * (double)k);
		sum += term;
		if (term < 1.e-17 * sum)
			break;
	}

	return sum;
}

double calc_beta(float os, float width)
{
	return M_PI * sqrt(pow((double)(width / os) * (os - 0.5), 2.) - 0.8);
}

void kb_table_host(double beta, float* table)
{
	double i0b = bessel_i0(beta);

	for (int i = 0; i < KB_SIZE + 1; i++) {

		double x = (double)i / (double)(KB_SIZE - 1) / 2.;
		table[i] = (fabs(x) >= 0.5) ? 0.f : (float)(bessel_i0(beta * sqrt(1. - pow(2. * x, 2.))) / i0b);
	}
}


// conj(fftmod) of grid point (x, y, z) applied to an adjoint gridding result
__device__ inline float2 apply_conj_fm(const GridGeom& g, float2 v, int x, int y, int z)
{
	if (nullptr == g.fm[0])
		return v;

	float2 f = g.fm[0][x];
	f = cmul(f, g.fm[1][y]);
	f = cmul(f, g.fm[2][z]);
	return cmulc(v, f);
}

// ---------------------------------------------------------------------------
// Plan: cell sort and tilings
// ---------------------------------------------------------------------------

constexpr int MAX_RANGES = 256;

// visits (samples of a tile's cells and their ring) per adjoint work item
constexpr int ADJ_MAXV = 2048;

// sample position on the grid, pos = os (k + shift) + n / 2, rounded as
// noncart/grid.c:181-187 rounds it (the device twin is grid_pos)
static float host_pos(const GridGeom& g, int s, int d, float k)
{
	const float b = (float)((double)g.os * ((double)k + (double)g.shift[s][d]));
	return b + (float)(g.n[d] / 2);
}

// Adjoint tiling.  A wave owns a block of every sub-grid: 2 x 2 x 2 points
// in 3-d (a 4 x 4 x 4 block of the 2x oversampled grid), 4 x 4 in 2-d, 32 in
// 1-d, one lane per (sub-grid, point), up to 16 coils per lane in registers.
// Tiles are 8 x 4 x 4 (16 waves), 16 x 16 (16 waves), 32 (1 wave); a tile
// dim is cut to the grid rounded up to the block.
static GridTiling choose_tiling(const GridGeom& g)
{
	int nd = 0;
	for (int d = 0; d < 3; d++)
		nd += (g.n[d] > 1);

	const int Tdef3[3] = { 8, 4, 4 };
	const int Tdef2[3] = { 16, 16, 1 };
	const int bdim = (3 == nd) ? 2 : ((2 == nd) ? 4 : 32);

	GridTiling t = {};
	int k = 0;
	int e = 0;
	t.ntiles = 1;
	t.Wtot = 1;
	t.nw = 1;

	for (int d = 0; d < 3; d++) {

		t.eoff[d] = -1;

		if (1 == g.n[d]) {
			t.T[d] = 1;
			t.b[d] = 1;
		} else {
			int want = (3 == nd) ? Tdef3[k] : ((2 == nd) ? Tdef2[k] : 32);
			k++;
			t.b[d] = bdim;
			t.T[d] = std::min(want, ((g.n[d] + bdim - 1) / bdim) * bdim);
			t.eoff[d] = e;
			e += 2 * t.T[d];
		}

		t.nt[d] = (g.n[d] + t.T[d] - 1) / t.T[d];
		t.ntiles *= t.nt[d];
		t.W[d] = t.T[d];
		t.Wtot *= t.T[d];
		t.nb[d] = t.T[d] / t.b[d];
		t.nw *= t.nb[d];
	}

	BAM_CHECK(e <= 64 && t.nw <= 16 && g.S * t.b[0] * t.b[1] * t.b[2] <= 64, "gridding: bad tiling");

	auto ilog2 = [](int v) {
		int l = 0;
		while ((1 << l) < v)
			l++;
		BAM_CHECK((1 << l) == v, "gridding: block dims must be powers of two");
		return l;
	};
	t.lb0 = ilog2(t.b[0]);
	t.lb01 = t.lb0 + ilog2(t.b[1]);
	t.lb = t.lb01 + ilog2(t.b[2]);
	t.cc = (g.C <= 4) ? 4 : ((g.C <= 8) ? 8 : 16);

	// cell rows visited by one work-group must fit the range list
	long rows = 2;
	for (int d = 1; d < 3; d++)
		if (g.n[d] > 1)
			rows *= (long)ceil((t.T[d] + g.reach_lo + g.reach_hi) / g.CS[d]) + 2;
	BAM_CHECK(rows <= MAX_RANGES, "gridding: tile too large for the cell range list");

	return t;
}

// cells [clo, chi] of one dim as up to two spans {a0, n0, a1, n1} (wrapped
// for periodic grids, clamped otherwise, every cell at most once); the
// device twin is cell_span()
static void host_cell_span(const GridGeom& g, int d, int clo, int chi, int sp[4])
{
	const int nc = g.ncell[d];
	sp[2] = 0;
	sp[3] = 0;

	if (chi - clo + 1 >= nc) {
		sp[0] = 0;
		sp[1] = nc;
		return;
	}

	if (!g.periodic) {
		clo = std::max(clo, 0);
		chi = std::min(chi, nc - 1);
		sp[0] = clo;
		sp[1] = std::max(0, chi - clo + 1);
		return;
	}

	int lo = ((clo % nc) + nc) % nc;
	int len = chi - clo + 1;
	sp[0] = lo;

	if (lo + len <= nc) {
		sp[1] = len;
	} else {
		sp[1] = nc - lo;
		sp[3] = len - sp[1];
	}
}

// Forward gather work list of 3-D grids (non-periodic, kernel width < 4):
// the samples in home-tile order (cell order within a tile) and work items of
// at most FW_MAXS samples of one tile.
constexpr int FW_T = 8;          // tile edge, grid points
constexpr int FW_SPT = 4;        // samples per thread
constexpr int FW_THREADS = 512;
constexpr int FW_MAXS = FW_SPT * FW_THREADS;
constexpr int FW_CC = 4;         // coils per staged region
constexpr int FW_E = 12;         // staged region edge: tile + halo of a kernel of width 3

static int taps_bound(const GridGeom& g);

void GridPlan::build_forward_tiles(const GridGeom& gg, const std::vector<int>& hsm, const std::vector<float4>& hsk,
				   const std::vector<int>& cell)
{
	ft = FwdTiling();

	// BART_AMD_FWD_TILES=off keeps the per-sample gather (for comparisons)
	const char* env = getenv("BART_AMD_FWD_TILES");

	if (nullptr != env && 0 == strcmp(env, "off"))
		return;

	bool all3 = gg.n[0] > 1 && gg.n[1] > 1 && gg.n[2] > 1;

	if (!all3 || gg.periodic || 4 != taps_bound(gg) || 0 == gg.M || gg.N * FW_CC >= (1l << 31))
		return;

	ft.T = FW_T;
	long ntiles = 1;

	for (int d = 0; d < 3; d++) {
		BAM_CHECK(0 == FW_T % gg.CS[d], "gridding: tile is not a multiple of the cell");
		ft.nt[d] = (gg.n[d] + FW_T - 1) / FW_T;
		ft.HL[d] = (int)ceilf(gg.reach_lo);
		ft.E[d] = FW_T + ft.HL[d] + (int)ceilf(gg.reach_hi);
		ntiles *= ft.nt[d];
	}

	ft.Etot = ft.E[0] * ft.E[1] * ft.E[2];

	// the kernel is built for one region edge; the region rows are loaded in
	// 16-byte pairs of points starting at even x
	if (FW_E != ft.E[0] || FW_E != ft.E[1] || FW_E != ft.E[2] || 0 != gg.n[0] % 2 || 0 != ft.HL[0] % 2)
		return;

	// the staged region of FW_CC coils and the KB table fit into 64 KiB
	if ((size_t)ft.Etot * FW_CC * sizeof(float2) + (KB_SIZE + 4) * sizeof(float) > 64 * 1024)
		return;

	const int cpt = FW_T / gg.CS[0];  // cells per tile edge (CS is equal in every dim)
	std::vector<int> tile_of(gg.M);
	std::vector<long> cnt(ntiles + 1, 0);

	// hsm/hsk are in cell order; cell[] is indexed by the original sample
	for (long i = 0; i < gg.M; i++) {
		int c = cell[hsm[i]];
		int cx = c % gg.ncell[0];
		int cy = (c / gg.ncell[0]) % gg.ncell[1];
		int cz = c / (gg.ncell[0] * gg.ncell[1]);
		int t = cx / cpt + ft.nt[0] * (cy / cpt + ft.nt[1] * (cz / cpt));
		tile_of[i] = t;
		cnt[t + 1]++;
	}

	for (long t = 0; t < ntiles; t++)
		cnt[t + 1] += cnt[t];

	std::vector<int> osm(gg.M);
	std::vector<float4> osk(gg.M);
	std::vector<long> pos(cnt.begin(), cnt.end() - 1);

	for (long i = 0; i < gg.M; i++) {
		long j = pos[tile_of[i]]++;
		osm[j] = hsm[i];
		osk[j] = hsk[i];
	}

	std::vector<int4> items;

	for (long t = 0; t < ntiles; t++)
		for (long a = cnt[t]; a < cnt[t + 1]; a += FW_MAXS)
			items.push_back(make_int4((int)t, (int)a, (int)std::min(cnt[t + 1], a + FW_MAXS), 0));

	std::stable_sort(items.begin(), items.end(), [](const int4& a, const int4& b) { return (a.z - a.y) > (b.z - b.y); });

	ft.nitems = (int)items.size();
	fsm.alloc(gg.M);
	fsk.alloc(gg.M);
	fwork.alloc(std::max(ft.nitems, 1));
	BAM_HIP(hipMemcpy(fsm.ptr, osm.data(), sizeof(int) * gg.M, hipMemcpyHostToDevice));
	BAM_HIP(hipMemcpy(fsk.ptr, osk.data(), sizeof(float4) * gg.M, hipMemcpyHostToDevice));
	BAM_HIP(hipMemcpy(fwork.ptr, items.data(), sizeof(int4) * ft.nitems, hipMemcpyHostToDevice));
	ft.on = true;
}

void GridPlan::build(const GridGeom& geom, const float* traj)
{
	g = geom;

	float maxshift = 0.f;
	for (int s = 0; s < g.S; s++)
		for (int d = 0; d < 3; d++)
			maxshift = std::max(maxshift, fabsf(g.shift[s][d] - g.shift[0][d]));

	g.reach_lo = 0.5f * g.width + g.os * maxshift;
	g.reach_hi = 0.5f * g.width;

	// distinct shift values per dim (the decomposition uses two: 0 and -1/2,
	// plus the odd-size correction) and the variant of every sub-grid
	for (int d = 0; d < 3; d++) {

		g.shv[d][0] = g.shift[0][d];
		g.shv[d][1] = g.shift[0][d];
		g.nshv[d] = 1;

		for (int s = 0; s < g.S; s++) {

			if (g.shift[s][d] == g.shv[d][0]) {
				g.sbit[s][d] = 0;
			} else if (1 == g.nshv[d] || g.shift[s][d] == g.shv[d][1]) {
				g.shv[d][1] = g.shift[s][d];
				g.nshv[d] = 2;
				g.sbit[s][d] = 1;
			} else {
				BAM_CHECK(false, "gridding: more than two shifts per dim");
			}
		}
	}

	long ncells = 1;

	for (int d = 0; d < 3; d++) {
		g.CS[d] = (g.n[d] > 1) ? 2 : 1;
		g.ncell[d] = (g.n[d] + g.CS[d] - 1) / g.CS[d];
		ncells *= g.ncell[d];
	}

	std::vector<int> cell(g.M);
	std::vector<int> wraps(g.M);
	std::vector<int> count(ncells + 1, 0);

	for (long m = 0; m < g.M; m++) {

		int cb[3];
		int code = 0;

		for (int d = 0; d < 3; d++) {

			if (1 == g.n[d]) {
				cb[d] = 0;
				code |= 8 << (4 * d);
				continue;
			}

			float p = host_pos(g, 0, d, traj[3 * m + d]);
			int q = 0;

			if (!std::isfinite(p))
				p = 0.f;

			if (g.periodic) {
				q = (int)floorf(p / (float)g.n[d]);
				q = std::min(std::max(q, -7), 7);
				p = p - (float)g.n[d] * (float)q;
			}

			code |= (q + 8) << (4 * d);

			int b = (int)floorf(p / (float)g.CS[d]);
			cb[d] = std::min(std::max(b, 0), g.ncell[d] - 1);
		}

		wraps[m] = code;
		cell[m] = cb[0] + g.ncell[0] * (cb[1] + g.ncell[1] * cb[2]);
		count[cell[m] + 1]++;
	}

	for (long c = 0; c < ncells; c++)
		count[c + 1] += count[c];

	std::vector<int> hsm(std::max<long>(g.M, 1));
	std::vector<float4> hsk(std::max<long>(g.M, 1));
	std::vector<int> pos(count.begin(), count.end() - 1);

	for (long m = 0; m < g.M; m++) {

		long i = pos[cell[m]]++;
		hsm[i] = (int)m;
		int code = wraps[m];
		float fcode;
		memcpy(&fcode, &code, sizeof(float));
		hsk[i] = make_float4(traj[3 * m + 0], traj[3 * m + 1], traj[3 * m + 2], fcode);
	}

	BAM_CHECK(g.M < (1l << 31), "too many samples");

	build_forward_tiles(g, hsm, hsk, cell);

	ta = choose_tiling(g);

	sm.alloc(std::max<long>(g.M, 1));
	si.alloc(std::max<long>(g.M, 1));
	sk.alloc(std::max<long>(g.M, 1));
	cstart.alloc(ncells + 1);

	BAM_HIP(hipMemcpy(sm.ptr, hsm.data(), sizeof(int) * std::max<long>(g.M, 1), hipMemcpyHostToDevice));
	{
		std::vector<int> hsi(std::max<long>(g.M, 1), 0);
		for (long i = 0; i < g.M; i++)
			hsi[hsm[i]] = (int)i;
		BAM_HIP(hipMemcpy(si.ptr, hsi.data(), sizeof(int) * std::max<long>(g.M, 1), hipMemcpyHostToDevice));
	}
	BAM_HIP(hipMemcpy(sk.ptr, hsk.data(), sizeof(float4) * std::max<long>(g.M, 1), hipMemcpyHostToDevice));

	BAM_HIP(hipMemcpy(cstart.ptr, count.data(), sizeof(int) * (ncells + 1), hipMemcpyHostToDevice));

	// room for every coils-per-work-group value the autotuner may pick
	// (4 .. the default): the coils rounded up to each
	cc_max = ta.cc;
	long kch = 0;
	for (int c = 4; c <= ta.cc; c *= 2)
		kch = std::max<long>(kch, (long)((g.C + c - 1) / c) * c);
	kspS.alloc((size_t)kch * std::max<long>(g.M, 1));

	// adjoint work list: the visits of a tile (its cells and their ring) in
	// chunks of at most MAXV; tiles with more than one chunk are split
	const long MAXV = ADJ_MAXV;
	std::vector<int4> hw;
	std::vector<int4> hsplit;
	const GridTiling& t = ta;
	int slots = 0;

	for (int tile = 0; tile < t.ntiles; tile++) {

		int tb[3] = { tile % t.nt[0], (tile / t.nt[0]) % t.nt[1], tile / (t.nt[0] * t.nt[1]) };
		int span[3][4];

		for (int d = 0; d < 3; d++) {

			int clo = 0, chi = 0;

			if (g.n[d] > 1) {
				int t0 = tb[d] * t.T[d];
				clo = (int)floorf(((float)t0 - g.reach_hi) / (float)g.CS[d]);
				chi = (int)floorf(((float)(t0 + t.T[d] - 1) + g.reach_lo) / (float)g.CS[d]);
			}

			host_cell_span(g, d, clo, chi, span[d]);
		}

		long F = 0;

		for (int zs = 0; zs < 2; zs++)
			for (int cz = span[2][2 * zs]; cz < span[2][2 * zs] + span[2][2 * zs + 1]; cz++)
				for (int ys = 0; ys < 2; ys++)
					for (int cy = span[1][2 * ys]; cy < span[1][2 * ys] + span[1][2 * ys + 1]; cy++)
						for (int xs = 0; xs < 2; xs++) {
							long row = (long)g.ncell[0] * (cy + (long)g.ncell[1] * cz);
							long a = row + span[0][2 * xs];
							F += count[a + span[0][2 * xs + 1]] - count[a];
						}

		if (F <= MAXV) {
			hw.push_back(make_int4(tile, 0, (int)F, 0));
		} else {
			const int first = slots;
			for (long b = 0; b < F; b += MAXV)
				hw.push_back(make_int4(tile, (int)b, (int)std::min(F, b + MAXV), 1 + slots++));
			hsplit.push_back(make_int4(tile, first, slots - first, 0));
		}
	}

	// largest items first: the persistent work-groups take them in this order
	std::stable_sort(hw.begin(), hw.end(), [](const int4& a, const int4& b) { return (a.z - a.y) > (b.z - b.y); });

	nwork = (int)hw.size();
	nsplit = (int)hsplit.size();
	ta.nwork = nwork;
	queue.alloc(1);
	{
		int dev = 0;
		BAM_HIP(hipGetDevice(&dev));
		BAM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	}
	nslot = slots;
	work.alloc(std::max(nwork, 1));
	split_tiles.alloc(std::max(nsplit, 1));
	BAM_HIP(hipMemcpy(work.ptr, hw.data(), sizeof(int4) * nwork, hipMemcpyHostToDevice));
	if (nsplit > 0)
		BAM_HIP(hipMemcpy(split_tiles.ptr, hsplit.data(), sizeof(int4) * nsplit, hipMemcpyHostToDevice));
	spart.alloc((size_t)std::max(nslot, 1) * g.S * g.C * t.Wtot);
}

// taps per dim: floor(p + w/2) - ceil(p - w/2) + 1 <= floor(w) + 1
static int taps_bound(const GridGeom& g)
{
	int tp = (int)floor(g.width) + 1;
	BAM_CHECK(tp <= MAX_TAPS, "gridding: kernel width %f too large", g.width);
	return (tp <= 4) ? 4 : 8;
}

// ---------------------------------------------------------------------------
// Adjoint input permutation: kspS[(ch M + i) cc + cl] = ksp[(ch cc + cl) M + m(i)] conj(w(m))
// (the weights of noncart/nufft.c:1474-1479 are applied here), written
// from the original sample order with si = the inverse of m(i)
// ---------------------------------------------------------------------------

// Thread per original sample m and coil chunk: the coil values are read
// along m (coalesced) and the cc values of the chunk are written as one
// contiguous run at the sample's sorted position si[m].
__global__ void k_permute_ksp_inv(long M, int C, int cc, const int* __restrict__ si, const float2* __restrict__ ksp,
				  const float2* __restrict__ weights, float2* __restrict__ kspS)
{
	const int nch = (C + cc - 1) / cc;
	const long tot = (long)nch * M;

	for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < tot; j += (long)gridDim.x * blockDim.x) {

		const long ch = j / M;
		const long m = j - ch * M;
		const long i = si[m];
		const float2 w = (nullptr != weights) ? weights[m] : make_float2(1.f, 0.f);
		float2* out = kspS + (ch * M + i) * cc;

		for (int cl = 0; cl < cc; cl++) {

			const int c = (int)ch * cc + cl;
			float2 v = make_float2(0.f, 0.f);

			if (c < C)
				v = (nullptr != weights) ? cmulc(ksp[(long)c * M + m], w) : ksp[(long)c * M + m];

			out[cl] = v;
		}
	}
}

// ---------------------------------------------------------------------------
// Cell ranges of a work-group: the sorted samples of the cells [clo, chi] per
// dim (wrapped for periodic grids, clamped otherwise; every cell at most
// once) as contiguous index ranges, one per row of cells along x (two when x
// wraps).
// ---------------------------------------------------------------------------

struct CellRanges {
	int start[MAX_RANGES];
	int pre[MAX_RANGES + 1];
	int n;
	int pad[2];
};

// dynamic LDS layout: KB table, cell ranges, then the float2 data (16-byte aligned)
constexpr size_t SMEM_TABLE = (KB_SIZE + 4) * sizeof(float);
constexpr size_t SMEM_DATA = ((SMEM_TABLE + sizeof(CellRanges) + 15) / 16) * 16;

__device__ inline void cell_span(const GridGeom& g, int d, int clo, int chi, int& a0, int& n0, int& a1, int& n1)
{
	const int nc = g.ncell[d];
	a1 = 0;
	n1 = 0;

	if (chi - clo + 1 >= nc) {
		a0 = 0;
		n0 = nc;
		return;
	}

	if (!g.periodic) {
		clo = max(clo, 0);
		chi = min(chi, nc - 1);
		a0 = clo;
		n0 = max(0, chi - clo + 1);
		return;
	}

	int lo = ((clo % nc) + nc) % nc;
	int len = chi - clo + 1;

	a0 = lo;

	if (lo + len <= nc) {
		n0 = len;
	} else {
		n0 = nc - lo;
		n1 = len - n0;
	}
}

// Builds the ranges in LDS, one per row of cells along x (and x span), in
// the order z, y, x; empty rows are kept (their prefix entries repeat).
// Returns the total sample count.
__device__ inline int build_ranges(const GridGeom& g, const int* __restrict__ cstart, const int clo[3],
				   const int chi[3], CellRanges& R)
{
	int xa[2], xn[2], ya[2], yn[2], za[2], zn[2];
	cell_span(g, 0, clo[0], chi[0], xa[0], xn[0], xa[1], xn[1]);
	cell_span(g, 1, clo[1], chi[1], ya[0], yn[0], ya[1], yn[1]);
	cell_span(g, 2, clo[2], chi[2], za[0], zn[0], za[1], zn[1]);

	const int nxs = (xn[0] > 0) + (xn[1] > 0);
	const int ny = yn[0] + yn[1];
	const int nrow = min(nxs * ny * (zn[0] + zn[1]), MAX_RANGES);  // the plan checks the bound

	for (int r = threadIdx.x; r < nrow; r += blockDim.x) {

		const int xs = (xn[0] > 0) ? (r % nxs) : 1;
		const int q = r / nxs;
		const int yi = q % ny;
		const int zi = q / ny;
		const int cy = (yi < yn[0]) ? (ya[0] + yi) : (ya[1] + yi - yn[0]);
		const int cz = (zi < zn[0]) ? (za[0] + zi) : (za[1] + zi - zn[0]);
		const long row = (long)g.ncell[0] * (cy + (long)g.ncell[1] * cz) + xa[xs];
		const int a = cstart[row];
		R.start[r] = a;
		R.pre[r + 1] = cstart[row + xn[xs]] - a;
	}

	__syncthreads();

	if (0 == threadIdx.x) {

		int tot = 0;
		R.pre[0] = 0;

		for (int r = 0; r < nrow; r++) {
			tot += R.pre[r + 1];
			R.pre[r + 1] = tot;
		}

		R.n = nrow;
	}

	__syncthreads();
	return R.pre[R.n];
}

// ---------------------------------------------------------------------------
// Taps of one sample along one dim (noncart/grid.c:415-457), computed once
// per sample and shift variant and reused by every sub-grid and coil.
// ---------------------------------------------------------------------------

__device__ inline float sample_pos_v(const GridGeom& g, int d, int v, float k)
{
	return grid_pos(g.os, k, g.shv[d][v], (float)(g.n[d] / 2));
}

template <int TP> struct Taps1 {
	int start;  // first tap (grid index before wrapping; clamped when not periodic)
	int cnt;    // number of taps
	float w[TP];
};

template <int TP>
__device__ inline void taps_raw(const GridGeom& g, int d, int v, float k, const float* table, Taps1<TP>& t)
{
	if (1 == g.n[d]) {
		t.start = 0;
		t.cnt = 1;
#pragma unroll
		for (int j = 0; j < TP; j++)
			t.w[j] = (0 == j) ? 1.f : 0.f;
		return;
	}

	const float pos = sample_pos_v(g, d, v, k);
	int sti = (int)ceil((double)pos - 0.5 * (double)g.width);
	int eni = (int)floor((double)pos + 0.5 * (double)g.width);

	if (!g.periodic) {
		sti = max(sti, 0);
		eni = min(eni, g.n[d] - 1);
	}

	t.start = sti;
	t.cnt = min(max(eni - sti + 1, 0), TP);

#pragma unroll
	for (int j = 0; j < TP; j++)
		t.w[j] = (j < t.cnt) ? kb_lookup(table, fabsf((float)(sti + j) - pos) / g.width) : 0.f;
}

// local index of tap j relative to a tile or window starting at org with
// extent range, or -1 (periodic grids: modulo n)
__device__ inline int tap_local(const GridGeom& g, int d, int w, int org, int range)
{
	int l;

	if (g.periodic) {
		const int nn = g.n[d];
		int gw = w % nn;
		if (gw < 0)
			gw += nn;
		l = gw - org;
		if (l < 0)
			l += nn;
		else if (l >= nn)
			l -= nn;
	} else {
		l = w - org;
	}

	return (l >= 0 && l < range) ? l : -1;
}

template <int TP>
__device__ inline void select_taps(const Taps1<TP>& a, const Taps1<TP>& b, bool useb, const GridGeom& g, int d,
				   int org, int range, int (&li)[TP], float (&wt)[TP], bool& any)
{
	const int start = useb ? b.start : a.start;
	const int cnt = useb ? b.cnt : a.cnt;
	bool an = false;

#pragma unroll
	for (int j = 0; j < TP; j++) {
		int l = (j < cnt) ? tap_local(g, d, start + j, org, range) : -1;
		li[j] = l;
		wt[j] = useb ? b.w[j] : a.w[j];
		an = an || (l >= 0);
	}

	any = an;
}

typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int ADJ_BATCH = 256;  // visits per batch (4 groups of 64)
constexpr int ADJ_WROW = 65;    // weight table row: 64 entries and a constant 1 (inactive dims)

template <typename X>
__device__ inline X sel3(int d, X a, X b, X c)
{
	return (0 == d) ? a : ((1 == d) ? b : c);
}

__global__ void k_reset_queue(int* q)
{
	if (0 == threadIdx.x)
		q[0] = 0;
}

// Adjoint gridding (k-space -> grid), output-stationary and atomic-free.
// grid layout [S][C][N].  Persistent work-groups (one per CU) take work
// items (tile, visit range, coil chunk) from a queue, largest first.
//
// A wave owns a block of b points per dim in every sub-grid (64 outputs:
// sub-grid s, point).  For each batch of visits (samples of the tile's cells
// and of the ring of neighbouring cells, in cell order):
//   1. one lane per (visit, dim, shift variant) evaluates the taps once and
//      writes the weight of every tile row (0 off the taps) and a bit mask
//      of the rows hit to LDS; the coil values of the batch, loaded during
//      the previous batch, go to LDS;
//   2. every wave tests from the masks which visits reach its block and
//      walks those two at a time: C[out][j] += sum_k A[out][k] B[k][j] on
//      the matrix cores (v_mfma_f32_32x32x2_f32), A = tap weight
//      wz * wy * wx of output out for visit k, B = value j (coil j / 2,
//      re / im) of visit k.
// Each product is accumulated in visit order, as in the sequential gridding
// of noncart/grid.c:240-250.  Tiles whose visits are split over several
// work items store partial tiles, summed in order by k_sum_split.
template <int TP, int CC>
__global__ __launch_bounds__(1024) void k_grid_adj(GridGeom g, GridTiling t, const float* __restrict__ kbt,
						   const float4* __restrict__ sk, const int* __restrict__ cstart,
						   const int4* __restrict__ work, int nitems, int* __restrict__ queue,
						   const float2* __restrict__ kspS, float2* __restrict__ grid,
						   float2* __restrict__ spart)
{
	extern __shared__ float4 smem4[];
	char* smem = (char*)smem4;
	float* table = (float*)smem;
	CellRanges& R = *(CellRanges*)(smem + SMEM_TABLE);
	float4* skall = (float4*)(smem + SMEM_DATA);          // [ADJ_MAXV] coordinates of each visit
	float2* vbuf = (float2*)(skall + ADJ_MAXV);           // [ADJ_BATCH][CC] coil values per visit
	int* sidx = (int*)(vbuf + ADJ_BATCH * CC);            // [ADJ_MAXV] sorted sample of each visit
	float* wtab = (float*)(sidx + ADJ_MAXV);              // [ADJ_BATCH][ADJ_WROW] tap weights per visit
	unsigned* rows = (unsigned*)(wtab + ADJ_WROW * ADJ_BATCH);  // [ADJ_BATCH][6] tile rows hit per dim, variant
	int* qitem = (int*)(rows + 6 * ADJ_BATCH);
	float2* fml = (float2*)(qitem + 4);                   // [n0 + n1 + n2] fftmod tables (when t.fmlds)

	const int nthr = blockDim.x;
	const int lane = threadIdx.x & 63;
	const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
	const int nwork = t.nwork;

	for (int i = threadIdx.x; i < KB_SIZE + 1; i += nthr)
		table[i] = kbt[i];

	if (0 == threadIdx.x)
		qitem[1] = 0;

	const bool fm = (nullptr != g.fm[0]);
	const bool fmlds = fm && (0 != t.fmlds);

	if (fmlds) {
		for (int i = threadIdx.x; i < g.n[0]; i += nthr)
			fml[i] = g.fm[0][i];
		for (int i = threadIdx.x; i < g.n[1]; i += nthr)
			fml[g.n[0] + i] = g.fm[1][i];
		for (int i = threadIdx.x; i < g.n[2]; i += nthr)
			fml[g.n[0] + g.n[1] + i] = g.fm[2][i];
	}

	// the wave's block; C rows this lane feeds with weights: l % 32 and 32 + l % 32
	const int B = t.b[0] * t.b[1] * t.b[2];
	const int wb[3] = { wid % t.nb[0], (wid / t.nb[0]) % t.nb[1], wid / (t.nb[0] * t.nb[1]) };
	const int jcol = lane & 31;
	int wo1[3], wo2[3];
	{
		const int r1 = lane & 31;
		const int r2 = 32 + (lane & 31);
		const int s1 = min(r1 / B, g.S - 1), s2 = min(r2 / B, g.S - 1);
		const int cb1 = r1 - (r1 / B) * B, cb2 = r2 - (r2 / B) * B;
		const int p1[3] = { wb[0] * t.b[0] + cb1 % t.b[0], wb[1] * t.b[1] + (cb1 / t.b[0]) % t.b[1],
				    wb[2] * t.b[2] + cb1 / (t.b[0] * t.b[1]) };
		const int p2[3] = { wb[0] * t.b[0] + cb2 % t.b[0], wb[1] * t.b[1] + (cb2 / t.b[0]) % t.b[1],
				    wb[2] * t.b[2] + cb2 / (t.b[0] * t.b[1]) };
#pragma unroll
		for (int d = 0; d < 3; d++) {
			wo1[d] = (t.eoff[d] < 0) ? 64 : (t.eoff[d] + g.sbit[s1][d] * t.T[d] + min(p1[d], t.T[d] - 1));
			wo2[d] = (t.eoff[d] < 0) ? 64 : (t.eoff[d] + g.sbit[s2][d] * t.T[d] + min(p2[d], t.T[d] - 1));
		}
	}

	constexpr int VPT = (ADJ_BATCH * CC + 1023) / 1024;  // coil values carried per lane (1024 lanes)

	while (true) {

		// next work item (the barrier also ends every wave's use of the
		// previous item's LDS data)
		__syncthreads();

		if (0 == threadIdx.x)
			qitem[0] = atomicAdd(queue, 1);

		__syncthreads();

		const int item = qitem[0];

		if (item >= nitems)
			break;

		// the constant-1 weight entry of every row (the output store of the
		// previous item reused this memory)
		for (int v = threadIdx.x; v < ADJ_BATCH; v += nthr)
			wtab[v * ADJ_WROW + 64] = 1.f;

		const int ch = item / nwork;
		const int4 wk = work[item - ch * nwork];  // (tile, first visit, end visit, partial slot + 1 or 0)
		const int tile = wk.x;
		const int c0 = ch * CC;
		const int ncc = min(CC, g.C - c0);

		int tb[3], t0[3], clo[3], chi[3];
		tb[0] = tile % t.nt[0];
		tb[1] = (tile / t.nt[0]) % t.nt[1];
		tb[2] = tile / (t.nt[0] * t.nt[1]);

		for (int d = 0; d < 3; d++) {
			t0[d] = tb[d] * t.T[d];
			if (1 == g.n[d]) {
				clo[d] = 0;
				chi[d] = 0;
			} else {
				clo[d] = (int)floorf(((float)t0[d] - g.reach_hi) / (float)g.CS[d]);
				chi[d] = (int)floorf(((float)(t0[d] + t.T[d] - 1) + g.reach_lo) / (float)g.CS[d]);
			}
		}

		const int Ftile = build_ranges(g, cstart, clo, chi, R);
		const int fb = min(wk.y, Ftile);
		const int fe = min(wk.z, Ftile);
		const int F = fe - fb;

		// sorted sample and coordinates of every visit of this item (F <= ADJ_MAXV)
		for (int f = fb + threadIdx.x; f < fe; f += nthr) {
			int lo = 0, hi = R.n;
			while (hi - lo > 1) {
				const int mid = (lo + hi) >> 1;
				if (R.pre[mid] <= f)
					lo = mid;
				else
					hi = mid;
			}
			const int i = R.start[lo] + (f - R.pre[lo]);
			sidx[f - fb] = i;
			skall[f - fb] = sk[i];
		}

		__syncthreads();

		// coil values: lane carries VPT float2 of the next batch (entries
		// threadIdx.x + 1024 q: visit e / CC, coil e % CC) through the current one
		const float2* kch = kspS + (long)ch * g.M * CC;
		float2 vnext[VPT];

#pragma unroll
		for (int q = 0; q < VPT; q++) {
			const int e = threadIdx.x + 1024 * q;
			vnext[q] = (e < ADJ_BATCH * CC && F > 0) ? kch[(long)sidx[min(e / CC, F - 1)] * CC + (e % CC)]
								 : make_float2(0.f, 0.f);
		}

		f32x16 C1 = {}, C2 = {};

		for (int base = 0; base < F; base += ADJ_BATCH) {

			const int nb = min(ADJ_BATCH, F - base);
			const float4* skc = skall + base;

			// the previous batch is done with the LDS buffers
			lds_barrier();

#pragma unroll
			for (int q = 0; q < VPT; q++) {
				const int e = threadIdx.x + 1024 * q;
				if (e < ADJ_BATCH * CC)
					vbuf[e] = vnext[q];
			}

			// work-groups of fewer than 1024 lanes load the rest now
			for (int e = threadIdx.x + nthr; e < ADJ_BATCH * CC; e += nthr)
				if (e % 1024 >= nthr)
					vbuf[e] = kch[(long)sidx[base + min(e / CC, nb - 1)] * CC + (e % CC)];

			// the next batch's values are loaded while this one is processed
			if (base + ADJ_BATCH < F) {
#pragma unroll
				for (int q = 0; q < VPT; q++) {
					const int e = threadIdx.x + 1024 * q;
					if (e < ADJ_BATCH * CC)
						vnext[q] = kch[(long)sidx[min(base + ADJ_BATCH + e / CC, F - 1)] * CC + (e % CC)];
				}
			}

			// 1. taps of the batch: lane (visit v, dim d, variant var)
			for (int jj = threadIdx.x; jj < 6 * nb; jj += nthr) {

				const int v = jj / 6;
				const int k = jj - 6 * v;
				const int d = k >> 1;
				const int var = k & 1;

				// per-dim parameters by selection (no indexed kernel-argument loads)
				const int eo = sel3(d, t.eoff[0], t.eoff[1], t.eoff[2]);

				if (eo < 0)
					continue;

				const float4 kc = skc[v];
				const float kd = sel3(d, kc.x, kc.y, kc.z);
				const int n = sel3(d, g.n[0], g.n[1], g.n[2]);
				const bool v1 = (1 == var) && (sel3(d, g.nshv[0], g.nshv[1], g.nshv[2]) > 1);
				const float sh = v1 ? sel3(d, g.shv[0][1], g.shv[1][1], g.shv[2][1])
						    : sel3(d, g.shv[0][0], g.shv[1][0], g.shv[2][0]);
				const float pos = grid_pos(g.os, kd, sh, (float)(n / 2));
				int sti = (int)ceil((double)pos - 0.5 * (double)g.width);
				int eni = (int)floor((double)pos + 0.5 * (double)g.width);

				if (!g.periodic) {
					sti = max(sti, 0);
					eni = min(eni, n - 1);
				}

				const int cnt = min(max(eni - sti + 1, 0), TP);
				float wq[TP];

#pragma unroll
				for (int q = 0; q < TP; q++)
					wq[q] = (q < cnt) ? kb_lookup(table, fabsf((float)(sti + q) - pos) / g.width) : 0.f;

				const int T = sel3(d, t.T[0], t.T[1], t.T[2]);
				const int td = sel3(d, t0[0], t0[1], t0[2]);
				float* wr = wtab + v * ADJ_WROW + eo + var * T;
				unsigned bits = 0u;

				for (int l = 0; l < T; l++) {

					const int gl = td + l;
					int j;

					if (g.periodic) {
						j = (gl - sti) % n;
						if (j < 0)
							j += n;
					} else {
						j = gl - sti;
					}

					const bool mem = (gl < n) && (j >= 0) && (j < cnt);
					float wl = 0.f;

#pragma unroll
					for (int q = 0; q < TP; q++)
						wl = (mem && j == q) ? wq[q] : wl;

					wr[l] = wl;
					bits |= mem ? (1u << l) : 0u;
				}

				rows[v * 6 + k] = bits;
			}

			lds_barrier();

			// 2. per group of 64 visits: those that reach this wave's block
			// (taps in its rows in every active dim, either variant), two at
			// a time
			for (int gb = 0; gb < nb; gb += 64) {

				const int vl = gb + min(lane, nb - gb - 1);
				bool reach = lane < nb - gb;

#pragma unroll
				for (int d = 0; d < 3; d++)
					if (t.eoff[d] >= 0) {
						const unsigned long long m = rows[vl * 6 + 2 * d] | rows[vl * 6 + 2 * d + 1];
						reach = reach && (0ull != ((m >> (wb[d] * t.b[d])) & ((1ull << t.b[d]) - 1ull)));
					}

				unsigned long long mh = __ballot(reach);

				while (0ull != mh) {

					const int vA = gb + __builtin_ctzll(mh);
					mh &= mh - 1ull;
					const bool hasB = 0ull != mh;
					const int vB = hasB ? (gb + __builtin_ctzll(mh)) : vA;
					if (hasB)
						mh &= mh - 1ull;

					const bool use = (lane < 32) || hasB;
					const int v = (lane < 32) ? vA : vB;
					const float* wv = wtab + v * ADJ_WROW;

					// z, y, x (BART order)
					float a1 = wv[wo1[2]];
					a1 = a1 * wv[wo1[1]];
					a1 = a1 * wv[wo1[0]];
					float a2 = wv[wo2[2]];
					a2 = a2 * wv[wo2[1]];
					a2 = a2 * wv[wo2[0]];
					float bv = (jcol < 2 * CC) ? ((const float*)vbuf)[v * 2 * CC + jcol] : 0.f;

					if (!use) {
						a1 = 0.f;
						a2 = 0.f;
						bv = 0.f;
					}

					C1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bv, C1, 0, 0, 0);
					C2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a2, bv, C2, 0, 0, 0);
				}
			}
		}

		// Store through LDS: the waves put their outputs into the tile
		// buffer otile[s][c][point] (plane pitch Wtot + 1, spread over the
		// banks; it reuses the batch buffers), then all lanes write the tile
		// out row by row.  C row i (of the wave's 64 outputs), column
		// j = 2 coil + (re, im): register r of lane l holds
		// i = 8 (r / 4) + 4 (l / 32) + r % 4 (+ 32 in C2), j = l % 32.
		const int OP = t.Wtot + 1;
		float* otile = (float*)skall;

		lds_barrier();

		{
			// (zero is 0, read from LDS here so that the row decoding is
			// not hoisted out of the item loop, where its 32 copies would
			// stay live)
			const int zero = qitem[1];
			const int c = (jcol + zero) >> 1;
			const int lhalf = (lane + zero) >> 5;

#pragma unroll
			for (int h = 0; h < 2; h++)
#pragma unroll
				for (int r = 0; r < 16; r++) {

					const int i = 32 * h + 8 * (r / 4) + 4 * lhalf + (r % 4);
					const int rs = i >> t.lb;
					const int rc = i & (B - 1);
					const int q0 = wb[0] * t.b[0] + (rc & (t.b[0] - 1));
					const int q1 = wb[1] * t.b[1] + ((rc >> t.lb0) & (t.b[1] - 1));
					const int q2 = wb[2] * t.b[2] + (rc >> t.lb01);

					if (i < g.S * B && c < CC)
						otile[(((long)rs * CC + c) * OP + q0 + t.W[0] * (q1 + t.W[1] * q2)) * 2 + (lane & 1)] =
							(0 == h) ? C1[r] : C2[r];
				}
		}

		lds_barrier();

		for (int e = threadIdx.x; e < g.S * CC * t.Wtot; e += nthr) {

			const int pq = e % t.Wtot;
			const int rest = e / t.Wtot;
			const int c = rest % CC;
			const int rs = rest / CC;
			const int x0 = t0[0] + pq % t.W[0];
			const int x1 = t0[1] + (pq / t.W[0]) % t.W[1];
			const int x2 = t0[2] + pq / (t.W[0] * t.W[1]);

			if (c >= ncc || x0 >= g.n[0] || x1 >= g.n[1] || x2 >= g.n[2])
				continue;

			const float* op = otile + (((long)rs * CC + c) * OP + pq) * 2;
			const float2 val = make_float2(op[0], op[1]);

			if (0 == wk.w) {
				const long idx = x0 + (long)g.n[0] * (x1 + (long)g.n[1] * x2);
				float2 o = val;
				if (fmlds) {
					float2 f = fml[x0];
					f = cmul(f, fml[g.n[0] + x1]);
					f = cmul(f, fml[g.n[0] + g.n[1] + x2]);
					o = cmulc(val, f);
				} else {
					o = apply_conj_fm(g, val, x0, x1, x2);
				}
				grid[((long)rs * g.C + c0 + c) * g.N + idx] = o;
			} else {
				// tile split over several work items (dense region): the
				// partial tile goes to its own slot, summed by k_sum_split
				spart[(((long)(wk.w - 1) * g.S + rs) * g.C + c0 + c) * t.Wtot + pq] = val;
			}
		}
	}
}

// tiles split over several work items: sum of their partial tiles in slot
// order (deterministic, no atomics); thread per (split tile, s, c, point)
__global__ void k_sum_split(GridGeom g, GridTiling t, const int4* __restrict__ tiles, int ntl,
			    const float2* __restrict__ spart, float2* __restrict__ grid)
{
	const int per = t.Wtot;
	const long SC = (long)g.S * g.C;
	const long tot = (long)ntl * SC * per;

	for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < tot; j += (long)gridDim.x * blockDim.x) {

		const long q = j / per;
		const int pnt = (int)(j - q * per);
		const long sc = q % SC;
		const int tl = (int)(q / SC);
		const int4 st = tiles[tl];
		const int tile = st.x;
		const int tb0 = tile % t.nt[0], tb1 = (tile / t.nt[0]) % t.nt[1], tb2 = tile / (t.nt[0] * t.nt[1]);
		const int gx = tb0 * t.T[0] + pnt % t.W[0];
		const int gy = tb1 * t.T[1] + (pnt / t.W[0]) % t.W[1];
		const int gz = tb2 * t.T[2] + pnt / (t.W[0] * t.W[1]);

		if (gx >= g.n[0] || gy >= g.n[1] || gz >= g.n[2])
			continue;

		float2 a = make_float2(0.f, 0.f);

		for (int k = 0; k < st.z; k++)
			a = cadd(a, spart[((long)(st.y + k) * SC + sc) * per + pnt]);

		grid[sc * g.N + gx + (long)g.n[0] * (gy + (long)g.n[1] * gz)] = apply_conj_fm(g, a, gx, gy, gz);
	}
}

void grid_adjoint(const GridPlan& gp, const float* kb, float2* grid, const float2* ksp, const float2* weights,
		  hipStream_t st)
{
	const GridGeom& g = gp.g;
	const GridTiling& t = gp.ta;

	if (0 == g.C)
		return;

	BAM_CHECK(t.cc <= 16, "gridding: at most 16 coils per work-group");

	{
		ProfScope prof("permute_ksp", st);
		int nch = (g.C + t.cc - 1) / t.cc;
		long tot = (long)nch * g.M;
		unsigned blocks = (unsigned)std::max<long>(1, std::min<long>((tot + 255) / 256, 16384));
		if (tot > 0)
			hipLaunchKernelGGL(k_permute_ksp_inv, dim3(blocks), dim3(256), 0, st, g.M, g.C, t.cc, gp.si.ptr, ksp,
					   weights, gp.kspS.ptr);
		BAM_HIP(hipGetLastError());
	}

	ProfScope prof("grid_adjoint", st);

	const int tp = taps_bound(g);
	const int nthr = 64 * t.nw;
	const int nch = (g.C + t.cc - 1) / t.cc;
	const int nitems = gp.nwork * nch;
	const unsigned nwg = (unsigned)std::max(1, std::min(nitems, gp.wpc * gp.ncu));
	size_t lds = SMEM_DATA + sizeof(float4) * ADJ_MAXV + sizeof(float2) * ADJ_BATCH * t.cc +
		     sizeof(int) * ADJ_MAXV + sizeof(float) * ADJ_WROW * ADJ_BATCH + sizeof(unsigned) * 6 * ADJ_BATCH +
		     sizeof(int) * 4;

	// the output tile (plane pitch Wtot + 1) reuses the batch buffers
	BAM_CHECK((size_t)g.S * t.cc * (t.Wtot + 1) * sizeof(float2) <=
			  sizeof(float4) * ADJ_MAXV + sizeof(float2) * ADJ_BATCH * t.cc + sizeof(int) * ADJ_MAXV +
				  sizeof(float) * ADJ_WROW * ADJ_BATCH,
		  "gridding: output tile larger than the batch buffers");

	// fftmod tables in LDS when they fit (160 KiB per work-group)
	GridTiling tl = t;
	const size_t fmb = sizeof(float2) * ((size_t)g.n[0] + g.n[1] + g.n[2]);
	tl.fmlds = (nullptr != g.fm[0] && lds + fmb <= 160 * 1024) ? 1 : 0;
	if (tl.fmlds)
		lds += fmb;

	hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, st, gp.queue.ptr);

#define BAM_ADJ(TP, CC)                                                                                     \
	hipLaunchKernelGGL((k_grid_adj<TP, CC>), dim3(nwg), dim3(nthr), lds, st, g, tl, kb, gp.sk.ptr, gp.cstart.ptr, \
			   gp.work.ptr, nitems, gp.queue.ptr, gp.kspS.ptr, grid, gp.spart.ptr)

	if (4 == tp) {
		if (t.cc <= 4) BAM_ADJ(4, 4);
		else if (t.cc <= 8) BAM_ADJ(4, 8);
		else BAM_ADJ(4, 16);
	} else {
		if (t.cc <= 4) BAM_ADJ(8, 4);
		else if (t.cc <= 8) BAM_ADJ(8, 8);
		else BAM_ADJ(8, 16);
	}
#undef BAM_ADJ

	BAM_HIP(hipGetLastError());

	if (gp.nsplit > 0) {
		ProfScope prof2("grid_adjoint_sum_split", st);
		long tot = (long)gp.nsplit * g.S * g.C * t.Wtot;
		unsigned zb = (unsigned)std::min<long>((tot + 255) / 256, 16384);
		hipLaunchKernelGGL(k_sum_split, dim3(zb), dim3(256), 0, st, g, t, gp.split_tiles.ptr, gp.nsplit,
				   gp.spart.ptr, grid);
		BAM_HIP(hipGetLastError());
	}
}

// ---------------------------------------------------------------------------
// Forward gridding (grid -> k-space), per-sample form for the grids the
// tiled kernel below does not take: one thread per sample and coil chunk
// gathers the taps of all sub-grids from the grid; sub-grids are summed in
// order s = 0 .. S-1 (noncart/nufft.c:1444-1445) and the result is
// multiplied by the weights (:1456-1457).  Samples are taken in cell-sorted
// order, so that the lanes of a wave read neighbouring grid lines (in the
// original order a 3-d readout crosses many lines: measured 8x slower on a
// 128^3 kooshball); the k-space stores go to the original positions.
// Launch (ceil(M / 256), ceil(C / CC)).
// ---------------------------------------------------------------------------

template <int TP, int CC>
__global__ __launch_bounds__(256) void k_grid_fwd(GridGeom g, const float* __restrict__ kbt,
						  const int* __restrict__ sm, const float4* __restrict__ sk,
						  const float2* __restrict__ grid, const float2* __restrict__ weights,
						  float2* __restrict__ ksp)
{
	__shared__ float table[KB_SIZE + 1];

	for (int i = threadIdx.x; i < KB_SIZE + 1; i += 256)
		table[i] = kbt[i];

	__syncthreads();

	const long i = (long)blockIdx.x * 256 + threadIdx.x;

	if (i >= g.M)
		return;

	const int c0 = blockIdx.y * CC;
	const int ncc = min(CC, g.C - c0);
	const float4 kc = sk[i];
	const float kk[3] = { kc.x, kc.y, kc.z };

	Taps1<TP> tv[3][2];

#pragma unroll
	for (int d = 0; d < 3; d++) {
		taps_raw<TP>(g, d, 0, kk[d], table, tv[d][0]);
		if (g.nshv[d] > 1)
			taps_raw<TP>(g, d, 1, kk[d], table, tv[d][1]);
		else
			tv[d][1] = tv[d][0];
	}

	float2 acc[CC];
#pragma unroll
	for (int cl = 0; cl < CC; cl++)
		acc[cl] = make_float2(0.f, 0.f);

	for (int s = 0; s < g.S; s++) {

		int li0[TP], li1[TP], li2[TP];
		float wt0[TP], wt1[TP], wt2[TP];
		bool a0, a1, a2;

		select_taps<TP>(tv[0][0], tv[0][1], 0 != g.sbit[s][0], g, 0, 0, g.n[0], li0, wt0, a0);
		select_taps<TP>(tv[1][0], tv[1][1], 0 != g.sbit[s][1], g, 1, 0, g.n[1], li1, wt1, a1);
		select_taps<TP>(tv[2][0], tv[2][1], 0 != g.sbit[s][2], g, 2, 0, g.n[2], li2, wt2, a2);

		if (!(a0 && a1 && a2))
			continue;

		float2 val[CC];
#pragma unroll
		for (int cl = 0; cl < CC; cl++)
			val[cl] = make_float2(0.f, 0.f);

		const float2* gs = grid + ((long)s * g.C + c0) * g.N;

#pragma unroll
		for (int iz = 0; iz < TP; iz++) {

			if (li2[iz] < 0)
				continue;

			float dz = 1.f * wt2[iz];

#pragma unroll
			for (int iy = 0; iy < TP; iy++) {

				if (li1[iy] < 0)
					continue;

				float dy = dz * wt1[iy];
				const float2* row = gs + (long)g.n[0] * (li1[iy] + (long)g.n[1] * li2[iz]);

#pragma unroll
				for (int ix = 0; ix < TP; ix++) {

					if (li0[ix] < 0)
						continue;

					float dd = dy * wt0[ix];
					const float2* p = row + li0[ix];

#pragma unroll
					for (int cl = 0; cl < CC; cl++) {
						if (cl < ncc) {
							float2 x = p[(long)cl * g.N];
							val[cl].x += x.x * dd;
							val[cl].y += x.y * dd;
						}
					}
				}
			}
		}

#pragma unroll
		for (int cl = 0; cl < CC; cl++)
			acc[cl] = cadd(acc[cl], val[cl]);
	}

	const int m = sm[i];
	const float2 w = (nullptr != weights) ? weights[m] : make_float2(1.f, 0.f);

#pragma unroll
	for (int cl = 0; cl < CC; cl++)
		if (cl < ncc)
			ksp[(long)(c0 + cl) * g.M + m] = (nullptr != weights) ? cmul(acc[cl], w) : acc[cl];
}

// ---------------------------------------------------------------------------
// Tiled forward gather (3-D grids): one work-group per work item, i.e. up to
// FW_MAXS samples of one home tile.  For every coil chunk and sub-grid the
// work-group stages the tile with its halo -- E^3 points of FW_CC coils, zero
// outside the grid -- in LDS with row-wise coalesced loads, and each thread
// gathers the taps of its samples from there.  The arithmetic is that of
// k_grid_fwd, term by term in the same order, so both give the same bits.
// ---------------------------------------------------------------------------

constexpr size_t FW_TABLE_BYTES = ((KB_SIZE + 4) * sizeof(float) + 15) / 16 * 16;

// E: staged region edge (tile + halo); LDS region layout [z][y][x][coil]
template <int CCH, int E>
__global__ __launch_bounds__(FW_THREADS) void k_grid_fwd_tile(GridGeom g, FwdTiling ft, const float* __restrict__ kbt,
							      const int4* __restrict__ work,
							      const int* __restrict__ fsm,
							      const float4* __restrict__ fsk,
							      const float2* __restrict__ grid,
							      const float2* __restrict__ weights,
							      float2* __restrict__ ksp)
{
	static_assert(CCH == 4 && E <= 16 && 0 == E % 2, "region rows are loaded by 8 lanes, 4 coils per point");

	extern __shared__ __align__(16) unsigned char fw_smem[];
	float* table = (float*)fw_smem;
	float2* reg = (float2*)(fw_smem + FW_TABLE_BYTES);

	for (int i = threadIdx.x; i < KB_SIZE + 1; i += FW_THREADS)
		table[i] = kbt[i];

	const int4 it = work[blockIdx.x];
	const int tile = it.x;
	const int o0 = (tile % ft.nt[0]) * ft.T - ft.HL[0];
	const int o1 = ((tile / ft.nt[0]) % ft.nt[1]) * ft.T - ft.HL[1];
	const int o2 = (tile / (ft.nt[0] * ft.nt[1])) * ft.T - ft.HL[2];
	const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
	const int Ni = (int)g.N;

	// region loads: 8 lanes per row of E points (two points per lane),
	// rows (coil, z, y)
	const int lx2 = 2 * (threadIdx.x & 7);
	const int gx = o0 + lx2;
	const bool xin = lx2 < E && gx >= 0 && gx < n0;
	constexpr int NROW = CCH * E * E;
	constexpr int RSTEP = FW_THREADS / 8;

	float4 kc[FW_SPT];
	bool valid[FW_SPT];

#pragma unroll
	for (int q = 0; q < FW_SPT; q++) {
		const int j = it.y + threadIdx.x + FW_THREADS * q;
		valid[q] = j < it.z;
		kc[q] = valid[q] ? fsk[j] : make_float4(0.f, 0.f, 0.f, 0.f);
	}

	for (int c0 = 0; c0 < g.C; c0 += CCH) {

		const int ncc = min(CCH, g.C - c0);
		float2 acc[FW_SPT][CCH];

#pragma unroll
		for (int q = 0; q < FW_SPT; q++)
#pragma unroll
			for (int cl = 0; cl < CCH; cl++)
				acc[q][cl] = make_float2(0.f, 0.f);

		for (int s = 0; s < g.S; s++) {

			const float2* gs = grid + ((long)s * g.C + c0) * g.N;

			// the previous region has been consumed (and the table is in)
			__syncthreads();

			for (int r = threadIdx.x >> 3; r < NROW; r += RSTEP) {

				const int cl = r / (E * E);
				const int zy = r - cl * (E * E);
				const int z = zy / E;
				const int gy = o1 + zy - z * E;
				const int gz = o2 + z;
				float4 v = make_float4(0.f, 0.f, 0.f, 0.f);

				// 32-bit offsets within the chunk (the plan bounds CCH N);
				// x pairs start at even points (even origin and row length)
				if (xin && cl < ncc && gy >= 0 && gy < n1 && gz >= 0 && gz < n2)
					v = *(const float4*)(gs + cl * Ni + gx + n0 * (gy + n1 * gz));

				if (lx2 < E) {
					float2* d = reg + (zy * E + lx2) * CCH + cl;
					d[0] = make_float2(v.x, v.y);
					d[CCH] = make_float2(v.z, v.w);
				}
			}

			__syncthreads();

			const int v0 = g.sbit[s][0], v1 = g.sbit[s][1], v2 = g.sbit[s][2];

#pragma unroll
			for (int q = 0; q < FW_SPT; q++) {

				if (!valid[q])
					continue;

				Taps1<4> t0, t1, t2;
				taps_raw<4>(g, 0, v0, kc[q].x, table, t0);
				taps_raw<4>(g, 1, v1, kc[q].y, table, t1);
				taps_raw<4>(g, 2, v2, kc[q].z, table, t2);

				// local tap indices (-1: no tap); taps outside the staged
				// region cannot occur for a sample of this tile
				int l0[4], l1[4], l2[4];
				bool a0 = false, a1 = false, a2 = false;

#pragma unroll
				for (int j = 0; j < 4; j++) {
					int w0 = t0.start + j, w1 = t1.start + j, w2 = t2.start + j;
					int q0 = w0 - o0, q1 = w1 - o1, q2 = w2 - o2;
					l0[j] = (j < t0.cnt && w0 >= 0 && w0 < n0 && q0 >= 0 && q0 < E) ? q0 : -1;
					l1[j] = (j < t1.cnt && w1 >= 0 && w1 < n1 && q1 >= 0 && q1 < E) ? q1 : -1;
					l2[j] = (j < t2.cnt && w2 >= 0 && w2 < n2 && q2 >= 0 && q2 < E) ? q2 : -1;
					a0 = a0 || (l0[j] >= 0);
					a1 = a1 || (l1[j] >= 0);
					a2 = a2 || (l2[j] >= 0);
				}

				if (!(a0 && a1 && a2))
					continue;

				float2 val[CCH];
#pragma unroll
				for (int cl = 0; cl < CCH; cl++)
					val[cl] = make_float2(0.f, 0.f);

#pragma unroll
				for (int iz = 0; iz < 4; iz++) {

					if (l2[iz] < 0)
						continue;

					float dz = 1.f * t2.w[iz];

#pragma unroll
					for (int iy = 0; iy < 4; iy++) {

						if (l1[iy] < 0)
							continue;

						float dy = dz * t1.w[iy];
						const float2* row = reg + (l2[iz] * E + l1[iy]) * E * CCH;

#pragma unroll
						for (int ix = 0; ix < 4; ix++) {

							if (l0[ix] < 0)
								continue;

							float dd = dy * t0.w[ix];
							const float4* pp = (const float4*)(row + l0[ix] * CCH);
							const float4 x01 = pp[0];
							const float4 x23 = pp[1];
							val[0].x += x01.x * dd;
							val[0].y += x01.y * dd;
							val[1].x += x01.z * dd;
							val[1].y += x01.w * dd;
							val[2].x += x23.x * dd;
							val[2].y += x23.y * dd;
							val[3].x += x23.z * dd;
							val[3].y += x23.w * dd;
						}
					}
				}

#pragma unroll
				for (int cl = 0; cl < CCH; cl++)
					acc[q][cl] = cadd(acc[q][cl], val[cl]);
			}
		}

#pragma unroll
		for (int q = 0; q < FW_SPT; q++) {

			if (!valid[q])
				continue;

			const int m = fsm[it.y + threadIdx.x + FW_THREADS * q];
			const float2 w = (nullptr != weights) ? weights[m] : make_float2(1.f, 0.f);

#pragma unroll
			for (int cl = 0; cl < CCH; cl++)
				if (cl < ncc)
					ksp[(long)(c0 + cl) * g.M + m] = (nullptr != weights) ? cmul(acc[q][cl], w) : acc[q][cl];
		}
	}
}

void grid_forward(const GridPlan& gp, const float* kb, float2* ksp, const float2* grid, const float2* weights,
		  hipStream_t st)
{
	const GridGeom& g = gp.g;

	if (0 == g.M || 0 == g.C)
		return;

	ProfScope prof("grid_forward", st);

	if (gp.ft.on) {

		const FwdTiling& ft = gp.ft;
		const size_t lds = FW_TABLE_BYTES + (size_t)FW_CC * ft.Etot * sizeof(float2);
		BAM_CHECK(4 == taps_bound(g) && !g.periodic && lds <= 64 * 1024 && FW_E == ft.E[0] &&
			  FW_E == ft.E[1] && FW_E == ft.E[2], "grid_forward: bad tiling");
		hipLaunchKernelGGL((k_grid_fwd_tile<FW_CC, FW_E>), dim3(ft.nitems), dim3(FW_THREADS), lds, st, g, ft,
				   kb, gp.fwork.ptr, gp.fsm.ptr, gp.fsk.ptr, grid, weights, ksp);
		BAM_HIP(hipGetLastError());
		return;
	}

	const int tp = taps_bound(g);
	// 4 coils per thread: the launch walks the samples chunk by chunk, so the
	// grid planes in use at a time are those of 4 coils (measured on a 128^3,
	// 16-coil kooshball: 7.7 ms at 4 coils, 12.4 ms at 16; 1 or 2 coils gain
	// nothing more)
	constexpr int CC = 4;
	dim3 blocks(ceil_div(g.M, 256), ceil_div(g.C, CC));

	if (4 == tp)
		hipLaunchKernelGGL((k_grid_fwd<4, CC>), blocks, dim3(256), 0, st, g, kb, gp.sm.ptr, gp.sk.ptr, grid,
				   weights, ksp);
	else
		hipLaunchKernelGGL((k_grid_fwd<8, CC>), blocks, dim3(256), 0, st, g, kb, gp.sm.ptr, gp.sk.ptr, grid,
				   weights, ksp);

	BAM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Autotuner of the adjoint gridding launch
// ---------------------------------------------------------------------------

struct TuneEntry {
	std::string key;
	int cc, wpc;
	float ms;
	int tried;
};

static std::mutex tune_mtx;
static std::vector<TuneEntry> tune_table;  // this process's choices (and the file's)
static bool tune_loaded = false;

static std::string tune_file()
{
	const char* f = getenv("BART_AMD_TUNE_FILE");
	if (nullptr != f)
		return f;
	const char* h = getenv("HOME");
	return std::string(nullptr != h ? h : "/tmp") + "/.cache/bart_amd_tune.txt";
}

static void tune_load()
{
	if (tune_loaded)
		return;
	tune_loaded = true;
	FILE* fp = fopen(tune_file().c_str(), "r");
	if (nullptr == fp)
		return;
	char key[256];
	int cc, wpc;
	float ms;
	while (4 == fscanf(fp, "%255s %d %d %f", key, &cc, &wpc, &ms))
		tune_table.push_back(TuneEntry{ key, cc, wpc, ms, 0 });
	fclose(fp);
}

static void tune_store(const TuneEntry& e)
{
	std::string f = tune_file();
	std::string dir = f.substr(0, f.find_last_of('/'));
	if (!dir.empty())
		(void)mkdir(dir.c_str(), 0755);
	FILE* fp = fopen(f.c_str(), "a");
	if (nullptr == fp)
		return;  // not writable: the choice holds for this process only
	fprintf(fp, "%s %d %d %.4f\n", e.key.c_str(), e.cc, e.wpc, e.ms);
	fclose(fp);
}

void tune_grid_adjoint(GridPlan& gp, const float* kb, float2* grid, hipStream_t st)
{
	const GridGeom& g = gp.g;

	static const bool on = [] {
		const char* e = getenv("BART_AMD_AUTOTUNE");
		return !(nullptr != e && 0 == strcmp(e, "off"));
	}();

	if (!on || 0 == g.C || 0 == g.M || 0 == gp.nwork)
		return;

	char kbuf[256];
	snprintf(kbuf, sizeof(kbuf), "adj:%dx%dx%d:S%d:C%d:M%ld:ccmax%d", g.n[0], g.n[1], g.n[2], g.S, g.C, g.M,
		 gp.cc_max);
	const std::string key = kbuf;

	{
		std::lock_guard<std::mutex> lock(tune_mtx);
		tune_load();
		for (const TuneEntry& e : tune_table)
			if (e.key == key && e.cc >= 4 && e.cc <= gp.cc_max && e.wpc >= 1 && e.wpc <= 4) {
				gp.ta.cc = e.cc;
				gp.wpc = e.wpc;
				return;
			}
	}

	// time every candidate on the plan with a zero k-space (two launches,
	// the second timed; the output is the same for all of them)
	DevBuf<float2> ksp((size_t)g.C * g.M);
	BAM_HIP(hipMemsetAsync(ksp.ptr, 0, sizeof(float2) * g.C * g.M, st));
	hipEvent_t e0, e1;
	BAM_HIP(hipEventCreate(&e0));
	BAM_HIP(hipEventCreate(&e1));
	const int cc0 = gp.ta.cc, wpc0 = gp.wpc;
	TuneEntry best{ key, cc0, wpc0, 1e30f, 0 };

	for (int cc = 4; cc <= gp.cc_max; cc *= 2) {
		for (int wpc = 1; wpc <= 4; wpc *= 2) {
			gp.ta.cc = cc;
			gp.wpc = wpc;
			grid_adjoint(gp, kb, grid, ksp.ptr, nullptr, st);
			BAM_HIP(hipEventRecord(e0, st));
			grid_adjoint(gp, kb, grid, ksp.ptr, nullptr, st);
			BAM_HIP(hipEventRecord(e1, st));
			BAM_HIP(hipEventSynchronize(e1));
			float ms = 0.f;
			BAM_HIP(hipEventElapsedTime(&ms, e0, e1));
			best.tried++;
			if (ms < best.ms) {
				best.ms = ms;
				best.cc = cc;
				best.wpc = wpc;
			}
		}
	}

	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	gp.ta.cc = best.cc;
	gp.wpc = best.wpc;

	std::lock_guard<std::mutex> lock(tune_mtx);
	tune_table.push_back(best);
	tune_store(best);
}

std::string tuning_report()
{
	std::lock_guard<std::mutex> lock(tune_mtx);
	std::string out = "[";
	for (size_t i = 0; i < tune_table.size(); i++) {
		const TuneEntry& e = tune_table[i];
		char b[512];
		snprintf(b, sizeof(b), "%s{\"shape\": \"%s\", \"coils_per_workgroup\": %d, \"workgroups_per_cu\": %d, "
			 "\"ms\": %.4f, \"candidates_timed\": %d}", i ? ", " : "", e.key.c_str(), e.cc, e.wpc, e.ms,
			 e.tried);
		out += b;
	}
	return out + "]";
}

}  // namespace bam
