// nufft_impl.hpp -- internal interface of the gfx950 NUFFT plan.
#pragma once

#include "common.hpp"
#include "fft.hpp"
#include "volume.hpp"

#include <complex>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace bam {

constexpr int KB_SIZE = 1000;
constexpr int MAX_TAPS = 8;  // taps per dimension (kernel width <= 7)

// sample position on a grid of n points, os (k + shift) + n / 2, rounded
// as noncart/grid.c:181-187 rounds it: os (k + shift) in double, rounded to
// float, then the float sum with n / 2
__device__ inline float grid_pos(float os, float k, float shift, float half)
{
	const float u = (float)((double)os * ((double)k + (double)shift));
	return u + half;
}

// intlookup (noncart/grid.c:126-149): x in [0, 0.5]; the interpolation
// (lerp, :120-123) in double.  No fused multiply-add: the products and
// sums are rounded one by one, as in the reference's build.
__device__ inline float kb_lookup(const float* table, float x)
{
#pragma clang fp contract(off)
	x *= 2.f;
	const float xs = x * (float)(KB_SIZE - 1);
	int index = (int)xs;
	index = min(max(index, 0), KB_SIZE - 1);
	const double f = (double)(xs - (float)index);
	return (float)((1. - f) * (double)table[index] + f * (double)table[index + 1]);
}

// host tables of nufft.hip: fftmod phase and value (num/fft.c:64-90) and the
// rolloff of one dim (noncart/grid.c:100-117, 522-531)
double fftmod_phase(long length, long j);
std::complex<double> fftmod_value(double phase);
double rolloff_1d(int d, int i, float os, double beta, double width);

// Gridding configuration of the decomposition sub-grids
// (noncart/grid.h:8-16 struct grid_conf_s, noncart/nufft.c:147-188).
struct GridGeom {

	int n[3];           // grid dims
	long N;             // grid points per coil
	int C;              // coils
	long M;             // samples per coil
	int S;              // sub-grids (shift sets)
	float os;
	float width;
	bool periodic;
	float shift[8][3];  // per sub-grid shift (including the odd-size correction)

	// the distinct shift values per dim (at most 2) and which one each
	// sub-grid uses
	float shv[3][2];
	int nshv[3];
	int sbit[8][3];

	// reach of a sample: taps w lie in [p - reach_lo, p + reach_hi] for its
	// shift-0 position p (union over the sub-grid shifts)
	float reach_lo, reach_hi;

	// cells: samples are sorted by the cell (CS^d grid points) that holds
	// their shift-0 position
	int CS[3];
	int ncell[3];

	// per-dim fftmod tables: the adjoint gridding stores grid x conj(fm)
	// (the first step of the adjoint chain, noncart/nufft.c:1462-1540);
	// null: no factor
	const float2* fm[3];
};

// Tiling of the adjoint gridding: a work-group owns a tile of T points per
// dim of all sub-grids for cc coils.  Each wave owns a block of b points per
// dim of every sub-grid (S x b0 b1 b2 = 64 lanes); the tile holds nb blocks
// per dim.  The per-visit weight table has 2 T[d] entries (two shift
// variants) for every active dim d, starting at eoff[d] (-1: inactive).
struct GridTiling {

	int T[3];
	int nt[3];
	int ntiles;
	int cc;    // coils per work-group
	int W[3];  // = T
	int Wtot;  // points per tile and sub-grid
	int b[3];  // wave block
	int nb[3]; // blocks per tile
	int nw;    // waves per work-group
	int eoff[3];
	int nwork; // work items per coil chunk
	int lb, lb0, lb01;  // log2 of b0 b1 b2, b0 and b0 b1 (all powers of two)
	int fmlds;          // fftmod tables staged in LDS by the adjoint (they fit)
};

// Tiling of the forward gather of 3-D grids: samples are grouped by the
// tile of T^3 grid points that holds their shift-0 position (their home
// tile); a work-group stages the tile and its halo (HL below, E - T - HL
// above, the reach of the kernel) of one sub-grid and a coil chunk in LDS.
struct FwdTiling {

	bool on = false;
	int T;
	int nt[3];
	int HL[3];
	int E[3];
	int Etot;
	int nitems;
};

struct GridPlan {

	GridGeom g;
	DevBuf<int> sm;          // sorted sample -> original sample index
	DevBuf<int> si;          // original sample -> sorted position
	DevBuf<float4> sk;       // sorted coordinates (kx, ky, kz) + wrap code
	DevBuf<int> cstart;      // cell -> first sorted sample (ncells + 1)
	GridTiling ta;           // adjoint tiling
	DevBuf<float2> kspS;     // adjoint input permuted to sorted order, coil chunks interleaved
	DevBuf<int4> work;       // adjoint work items (tile, first visit, end visit, partial slot + 1 or 0)
	DevBuf<int4> split_tiles; // tiles whose visits are spread over several work items: (tile, first slot, slots, 0)
	DevBuf<float2> spart;    // partial tiles of the split work items [slot][S][C][Wtot]
	int nslot = 0;

	int nwork = 0;
	int nsplit = 0;
	DevBuf<int> queue;       // adjoint work queue head
	int ncu = 1;             // compute units of the device
	int wpc = 2;             // persistent adjoint work-groups per compute unit (autotuned)
	int cc_max = 4;          // largest coils-per-work-group the k-space buffer holds

	FwdTiling ft;            // forward gather tiling (3-D)
	DevBuf<int> fsm;         // forward order -> original sample index
	DevBuf<float4> fsk;      // coordinates in forward order
	DevBuf<int4> fwork;      // forward work items (tile, first, end, 0), largest first

	// traj: host, M x 3 floats (real parts)
	void build(const GridGeom& geom, const float* traj);

private:
	void build_forward_tiles(const GridGeom& g, const std::vector<int>& hsm, const std::vector<float4>& hsk,
				 const std::vector<int>& cell);
};

// Separable per-dimension tables for the element-wise factors of the NUFFT:
// linear phase (with fftmod and 1/sqrt(n) folded in, one per shift bit),
// fftmod and rolloff.  Every factor of BART's tables is a product over the
// three dims (noncart/nufft.c:196-248, num/fft.c:64-157, grid.c:537-556).
struct SepTables {

	const float2* lin[3][2];
	const float2* fm[3];
	const float* roll[3];
	int bitpos[3];  // decomposition bit of each dim, -1 if the dim is trivial
	int n[3];

	// x-y products for 3-D grids (null otherwise), each [n1][n0]:
	// linxy[bx + 2 by] = lin[0][bx] x lin[1][by], fmxy = fm[0] x fm[1]
	const float2* linxy[4];
	const float2* fmxy;
};

struct CgApFuse;

struct Nufft {

	// geometry
	int n[3];
	long N;
	int S;
	int C;
	long M;
	int nact;
	int act[3];
	float width, os;
	double beta;
	bool periodic, toeplitz;

	// device data
	DevBuf<float> kb;
	DevBuf<float2> tab2;   // separable complex tables
	DevBuf<float> tabr;    // separable rolloff
	DevBuf<float2> tabxy;  // x-y product tables (3-D)
	SepTables tabs;
	DevBuf<float2> weights;
	bool has_weights = false;
	GridPlan gp;
	DevBuf<float2> psf;    // [S][N]
	bool has_psf = false;
	bool real_psf = false; // keep the real part of the PSF (conf.real, noncart/nufft.c:1128-1137)
	bool tune = true;      // autotune the adjoint gridding launch when the plan is built
	DevBuf<float2> grid;   // work [S][C][N]
	DevBuf<float2> part;   // per-coil-group partial images [G][N]
	AxisPlan axis[3];
	bool slice = false;    // 3-D with x-y slices transformed in LDS
	bool pair_zpass = false;  // Toeplitz z-pass over sub-grid pairs (slice path)

	hipStream_t stream = 0;

	// dims: image dims; traj: host M x 3 (real parts); weights: host M or null
	void create(const int dims[3], int C, long M, const float* traj, const float2* weights_host,
		    bool toeplitz, bool periodic, float width, float os, bool compute_psf);
	// geometry, tables and work buffers without a trajectory (the PSF-only
	// normal operator of nufft_create_normal)
	void create_geometry(const int dims[3], int C, bool toeplitz, bool periodic, float width, float os);
	// (re)sets trajectory (host, M x 3) and weights, builds the gridding
	// plan and, in Toeplitz mode with compute_psf, the PSF
	void set_traj(long M, const float* traj, const float2* weights_host, bool compute_psf);
	bool has_traj = false;

	// device pointers; sens == nullptr: coil images in/out, else single image
	// combined with coil maps
	void forward(float2* ksp, const float2* cim, const float2* sens);
	void adjoint(float2* cim, const float2* ksp, const float2* sens);
	// Toeplitz; skip: optional device flag that turns the launches into no-ops.
	// fuse (CG): where the operator ends with a sum of coil-group partial
	// images, that sum also takes the CG step dst += lambda src with its
	// partials; returns whether it did.
	bool normal(float2* dst, const float2* src, const float2* sens, const int* skip = nullptr,
		    const CgApFuse* fuse = nullptr);
	// PSF of the trajectory (periodic_psf: periodic gridding of the 2x
	// grid, as nufft_set_traj always uses)
	void compute_psf_from_traj(const float* traj, const float2* weights_host, bool periodic_psf);
	// Strang's circulant preconditioner of the Toeplitz operator
	// (noncart/precond.c:46-71): pre = |fftuc(sum_s ifftuc(psf_s)
	// conj(linphase_s))| + 1e-3 over the image dims; apply: dst = IFFT(FFT(src)
	// / pre) for ncoil coil images (unnormalised FFTs, precond.c:92-106)
	void compute_precond(DevBuf<float2>& pre) const;
	void apply_precond(const float2* pre, float2* dst, const float2* src, int ncoil) const;
	// with real_psf: the imaginary part of the PSF set to zero (a real PSF
	// multiplies both components of the grid alike, as BART's float PSF)
	void apply_real_psf();

	// Subspace basis (noncart/nufft.c:470-505, 1018-1036): nbasis = K
	// coefficients (0: none).  The plan's C channels are the coil-coefficient
	// pairs c + (C / K) k, its M samples the frames' samples m + Mf t (T
	// frames of Mf samples); bas = B [K][T] on the device (set before
	// create).  The PSF is then [K K][S][N], entry k K + k' the sub-grid PSFs
	// of sum_m |w_m|^2 conj(B[t_m][k]) B[t_m][k'].
	int nbasis = 0;
	int bT = 1;
	long bMf = 0;
	DevBuf<float2> bas;
	void set_basis(int K, int T, long Mf, const float2* B_host);
	// Toeplitz normal operator of the coefficient images (C channels in and
	// out, no maps): the K x K PSF product between the forward and inverse
	// transforms of the last axis
	void normal_basis(float2* dst, const float2* src, const int* skip = nullptr);
};

// Subspace basis kernels (basis.hip), layouts as described there; skip: an
// optional device flag that turns the launch into a no-op.
constexpr int BASIS_MAX = 16;
void basis_contract(float2* dst, const float2* src, const float2* B, int C, int K, int T, long Mf, hipStream_t st,
		    const int* skip = nullptr);
void basis_expand(float2* dst, const float2* src, const float2* B, int C, int K, int T, long Mf, hipStream_t st,
		  const int* skip = nullptr);
void maps_expand(float2* dst, const float2* sens, const float2* x, int C, int K, long N, hipStream_t st,
		 const int* skip = nullptr);
void maps_reduce(float2* x, const float2* sens, const float2* src, int C, int K, long N, hipStream_t st,
		 const int* skip = nullptr);
void basis_psf_apply(float2* grid, const float2* psf, int S, int C, int K, long N, hipStream_t st,
		     const int* skip = nullptr);
void basis_psf_vals(float2* v, const float2* B, int K, int T, long Mf, hipStream_t st);

// The chained NUFFT for os != 2 (noncart/nufft_chain.c:192-244): rolloff
// into the centre of the round(os n) grid, centred unitary FFT, KB
// interpolation on that grid with beta = calc_beta(os, width), weights times
// os^(d/2).  The adjoint gridding gathers per grid point from samples sorted
// by cells of CHAIN_CS points (1 along a periodic dim that CHAIN_CS does not
// divide).  No Toeplitz operator (BART has none here).
constexpr int CHAIN_CS = 4;
constexpr int CHAIN_TAPS = 16;
struct Chain {

	int n[3], on[3], off[3], ncell[3], cs[3];
	long N = 0, N2 = 0, M = 0;
	int C = 0;
	int nd = 0;
	float os = 2.f, width = 6.f;
	double beta = 0.;
	bool periodic = false;
	float scale = 1.f;
	bool has_weights = false;
	DevBuf<float> kb;
	DevBuf<float> roll;     // [N]
	DevBuf<float2> fma;     // [N2] fftmod x N2^-1/4
	DevBuf<float4> trj;     // [M] coordinates (original order)
	DevBuf<float2> tmp;     // [M] weights x scale
	DevBuf<float4> spos;    // sorted sample positions on the grid
	DevBuf<int> sidx;       // sorted sample -> original index
	DevBuf<int> cstart;     // [cells + 1]
	DevBuf<float2> grid;    // [C][N2]
	DevBuf<float2> kwork;   // [C / 8][M][8] scaled k-space in cell order
	DevBuf<int> cspan;      // adjoint: per tile its first cell and cell count per dim
	DevBuf<int4> cwork;     // adjoint work items (tile, slice, slices, first partial slot)
	DevBuf<int4> csplit;    // tiles split into slices (tile, slices, first partial slot)
	DevBuf<float2> cpart;   // partial tiles of the split tiles
	long nwork = 0, nsplit = 0, nslots = 0;
	hipStream_t stream = 0;

	// pad = false: the non-decomposed NUFFT of decomp = false
	// (noncart/nufft.c:898-935): the image is the oversampled grid itself,
	// rolloff with os = 1 over it, positions with os
	void create(const int dims[3], int C, long M, const float* traj, const float2* weights_host, bool periodic,
		    float width, float os, bool pad = true);
	void forward(float2* ksp, const float2* cim);
	void adjoint(float2* cim, const float2* ksp);
};

// Normal operators of several 2-D Toeplitz plans with coil maps, one launch
// per pass and size class (batched CG).  normal_groupable: whether plan p
// (with maps sens) can join; normal_groups_build: device tables for
// plans[q] applied to src[q] -> dst[q] (or its partial images), skip flags
// per plan, ngroups[q] = its size class; normal_groups_launch: the passes;
// normal_group_partials: the partial images a plan's passes leave and their
// count (dst itself and 1 without coil groups), to be summed by the caller.
struct NormalGroupSet;
bool normal_groupable(const Nufft& p, const float2* sens);
std::shared_ptr<NormalGroupSet> normal_groups_build(const std::vector<const Nufft*>& plans,
						    const std::vector<const float2*>& sens,
						    const std::vector<const float2*>& src, const std::vector<float2*>& dst,
						    const std::vector<const int*>& skip, std::vector<int>& ngroups);
void normal_groups_launch(const NormalGroupSet& set, hipStream_t st);
std::pair<const float2*, int> normal_group_partials(const Nufft& p, float2* dst);

void grid_adjoint(const GridPlan& gp, const float* kb, float2* grid, const float2* ksp, const float2* weights,
		  hipStream_t st);

// Autotuner of the adjoint gridding launch: coils per work-group (4, 8, 16
// up to the default) and persistent work-groups per compute unit (1, 2, 4)
// are timed on the plan itself (the output does not depend on them) and
// the fastest is kept, per shape class (dims, coils, samples) for the
// process and in a cache file (BART_AMD_TUNE_FILE, default
// $HOME/.cache/bart_amd_tune.txt; BART_AMD_AUTOTUNE=off: defaults).
void tune_grid_adjoint(GridPlan& gp, const float* kb, float2* grid, hipStream_t st);
// JSON list of the tuned shape classes of this process
std::string tuning_report();
void grid_forward(const GridPlan& gp, const float* kb, float2* ksp, const float2* grid, const float2* weights,
		  hipStream_t st);

void fft_nd(const int dims[3], long batch, unsigned flags, float sign, float2* dst, const float2* src,
	    float scale, hipStream_t stream);

void kb_table_host(double beta, float* table);
void fill(float2* p, long n, float2 v, hipStream_t st, const int* skip = nullptr);
// d += a x (vops->axpy)
void axpy(float2* d, const float2* x, float a, long n, hipStream_t st, const int* skip = nullptr);
// dst = sum_g part[g N ..], g = 0 .. G-1 in order (partial images)
void sum_images(float2* dst, const float2* part, long N, int G, hipStream_t st, const int* skip);
double calc_beta(float os, float width);

}  // namespace bam
