// espirit.hip -- ESPIRiT coil-sensitivity calibration (src/ecalib.c,
// calib/calib.c:715-1094) on gfx950:
//   1. the Gram matrix A^H A of the calibration (Casorati) matrix, A[p][k]
//      the k-space value at block position p and kernel entry k (kx fastest,
//      then ky, kz, coil) -- one 16 x 16 tile per work-group, the blocks
//      staged in LDS, FP64 sums (covariance_function_fft, calmat.c:181);
//   2. its eigenvectors (rocSOLVER cheevd); the kernels are the eigenvectors
//      of the conjugate, as BART's row-major matrix reads in LAPACK, in
//      descending order, kept while sv_i / sv_0 > sqrt(threshold);
//   3. the kernels zero-padded to twice their size, ifftmod, inverse FFT,
//      and per point of that grid the C x C matrix sum_k conj(K_c1) K_c2 /
//      (kx ky kz xh yh zh) (compute_imgcov);
//   4. its sinc interpolation to the image, one axis at a time
//      (misc/resize.c:234-292);
//   5. per image point the orthogonal iteration of num/linalg.c:1296-1314
//      for the maps, the crop by eigenvalue and the phase of the channel
//      principal component (eigenmaps, crop_sens, fixphase2).
#include "espirit.hpp"
#include "nufft_impl.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <complex>

namespace bam {

constexpr int ESP_MAXC = 32;  // coils
constexpr int ESP_MAXM = 4;   // maps

static unsigned nblk(long n) { return (unsigned)std::min<long>(std::max<long>((n + 255) / 256, 1), 16384); }

// ---------------------------------------------------------------------------
// Calibration region (misc/mri2.c:101-217)
// ---------------------------------------------------------------------------

std::vector<float2> extract_calib_host(const long calsize[3], const long dims[3], int C, const float2* ksp,
				       long caldims[3])
{
	const long X = dims[0], Y = dims[1], Z = dims[2];
	const long td[3] = { X, std::min(calsize[1], Y), std::min(calsize[2], Z) };
	const long tp[3] = { (X - td[0]) / 2, (Y - td[1]) / 2, (Z - td[2]) / 2 };

	auto kval = [&](int c, long x, long y, long z) {
		return ksp[((c * Z + (z + tp[2])) * Y + (y + tp[1])) * X + (x + tp[0])];
	};

	// pattern: a sample is present if any coil holds a non-zero value
	std::vector<char> pat(td[0] * td[1] * td[2], 0);

	for (long z = 0; z < td[2]; z++)
		for (long y = 0; y < td[1]; y++)
			for (long x = 0; x < td[0]; x++)
				for (int c = 0; c < C; c++) {
					const float2 v = kval(c, x, y, z);
					if (0.f != v.x || 0.f != v.y) {
						pat[(z * td[1] + y) * td[0] + x] = 1;
						break;
					}
				}

	long cal[3] = { 1, 1, 1 };
	long pos[3] = { (td[0] - 1) / 2, (td[1] - 1) / 2, (td[2] - 1) / 2 };
	bool stop[3] = { false, false, false };

	while (!(stop[0] && stop[1] && stop[2])) {

		for (int i = 0; i < 3; i++) {

			if (cal[i] == td[i] || cal[i] >= calsize[i])
				stop[i] = true;

			if (stop[i])
				continue;

			cal[i]++;
			pos[i] = (td[i] - cal[i]) / 2;

			long cnt = 0;

			for (long z = pos[2]; z < pos[2] + cal[2]; z++)
				for (long y = pos[1]; y < pos[1] + cal[1]; y++)
					for (long x = pos[0]; x < pos[0] + cal[0]; x++)
						cnt += pat[(z * td[1] + y) * td[0] + x];

			if (cnt != cal[0] * cal[1] * cal[2]) {
				cal[i]--;
				pos[i] = (td[i] - cal[i]) / 2;
				stop[i] = true;
			}
		}
	}

	// move along the readout to the position of maximal energy
	long best = 0;
	double beng = 0.;

	for (long r = 0; r < td[0] - cal[0] + 1; r++) {

		double e = 0.;

		for (int c = 0; c < C; c++)
			for (long z = pos[2]; z < pos[2] + cal[2]; z++)
				for (long y = pos[1]; y < pos[1] + cal[1]; y++)
					for (long x = r; x < r + cal[0]; x++) {
						const float2 v = kval(c, x, y, z);
						e += (double)v.x * v.x + (double)v.y * v.y;
					}

		if (e > beng) {
			best = r;
			beng = e;
		}
	}

	pos[0] = best;

	for (int i = 0; i < 3; i++)
		caldims[i] = cal[i];

	std::vector<float2> out((size_t)C * cal[0] * cal[1] * cal[2]);

	for (int c = 0; c < C; c++)
		for (long z = 0; z < cal[2]; z++)
			for (long y = 0; y < cal[1]; y++)
				for (long x = 0; x < cal[0]; x++)
					out[((c * cal[2] + z) * cal[1] + y) * cal[0] + x] =
						kval(c, x + pos[0], y + pos[1], z + pos[2]);

	return out;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

struct CasGeom {
	int cx, cy, cz, kx, ky, kz, C;
	int px, py, pz;
	long P;
	int N;
};

__device__ inline float2 cas_entry(const CasGeom& g, const float2* cal, long p, int k)
{
	const int ox = k % g.kx;
	int t = k / g.kx;
	const int oy = t % g.ky;
	t /= g.ky;
	const int oz = t % g.kz;
	const int c = t / g.kz;
	const int qx = (int)(p % g.px);
	const long r = p / g.px;
	const int qy = (int)(r % g.py);
	const int qz = (int)(r / g.py);
	return cal[(((long)c * g.cz + qz + oz) * g.cy + qy + oy) * g.cx + qx + ox];
}

// G[i][j] = sum_p conj(A[p][i]) A[p][j] over the rows p of the calibration
// matrix, in double.  A work-group computes a 64 x 64 tile of the upper
// block triangle (each thread 4 x 4 entries, rows ti + 16 a, columns
// tj + 16 b) from 16 rows of A at a time staged in LDS, and writes the
// mirrored tile as its conjugate, so G is Hermitian bit for bit.
constexpr int CG_T = 64;   // tile edge
constexpr int CG_KP = 16;  // rows of A per LDS stage

__global__ __launch_bounds__(256) void k_cas_gram(CasGeom g, const float2* __restrict__ cal, float2* __restrict__ G)
{
	__shared__ float2 As[CG_KP][CG_T], Bs[CG_KP][CG_T];

	// block -> (bi, bj), bj >= bi, row by row of the upper block triangle
	const int nb = (g.N + CG_T - 1) / CG_T;
	int bi = 0;
	long rem = blockIdx.x;
	while (rem >= nb - bi) {
		rem -= nb - bi;
		bi++;
	}
	const int bj = bi + (int)rem;

	const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;
	double re[4][4], im[4][4];

#pragma unroll
	for (int a = 0; a < 4; a++)
#pragma unroll
		for (int c = 0; c < 4; c++) {
			re[a][c] = 0.;
			im[a][c] = 0.;
		}

	for (long p0 = 0; p0 < g.P; p0 += CG_KP) {

		for (int q = threadIdx.x; q < CG_KP * CG_T; q += 256) {
			const int pp = q / CG_T, kk = q % CG_T;
			const long p = p0 + pp;
			const int ki = bi * CG_T + kk, kj = bj * CG_T + kk;
			As[pp][kk] = (p < g.P && ki < g.N) ? cas_entry(g, cal, p, ki) : make_float2(0.f, 0.f);
			Bs[pp][kk] = (p < g.P && kj < g.N) ? cas_entry(g, cal, p, kj) : make_float2(0.f, 0.f);
		}

		__syncthreads();

		for (int pp = 0; pp < CG_KP; pp++) {

			float2 x[4], y[4];

#pragma unroll
			for (int a = 0; a < 4; a++) {
				x[a] = As[pp][ti + 16 * a];
				y[a] = Bs[pp][tj + 16 * a];
			}

#pragma unroll
			for (int a = 0; a < 4; a++)
#pragma unroll
				for (int c = 0; c < 4; c++) {
					re[a][c] += (double)x[a].x * y[c].x + (double)x[a].y * y[c].y;
					im[a][c] += (double)x[a].x * y[c].y - (double)x[a].y * y[c].x;
				}
		}

		__syncthreads();
	}

#pragma unroll
	for (int a = 0; a < 4; a++)
#pragma unroll
		for (int c = 0; c < 4; c++) {
			const int i = bi * CG_T + ti + 16 * a, j = bj * CG_T + tj + 16 * c;
			if (i < g.N && j < g.N) {
				G[(long)i * g.N + j] = make_float2((float)re[a][c], (float)im[a][c]);
				if (bi != bj)
					G[(long)j * g.N + i] = make_float2((float)re[a][c], -(float)im[a][c]);
			}
		}
}

struct PadGeom {
	int kx, ky, kz, C, n;
	int xh, yh, zh;
	long N, Nh;
};

// K[k][c][z][y][x] = kernel k (column N - 1 - k of the eigenvectors) zero-padded
// about the centre, times ifftmod
__global__ void k_pad_kernels(PadGeom g, const float2* __restrict__ vec, const float2* __restrict__ ifm,
			      float2* __restrict__ K)
{
	const long total = (long)g.n * g.C * g.Nh;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long r = e % g.Nh;
		const long t = e / g.Nh;
		const int c = (int)(t % g.C);
		const int k = (int)(t / g.C);
		const int x = (int)(r % g.xh), y = (int)((r / g.xh) % g.yh), z = (int)(r / ((long)g.xh * g.yh));
		const int ox = x - (g.xh / 2 - g.kx / 2), oy = y - (g.yh / 2 - g.ky / 2), oz = z - (g.zh / 2 - g.kz / 2);
		float2 v = make_float2(0.f, 0.f);

		if (ox >= 0 && ox < g.kx && oy >= 0 && oy < g.ky && oz >= 0 && oz < g.kz) {
			const float2* ker = vec + (g.N - 1 - k) * g.N;
			v = cmul(ker[(((long)c * g.kz + oz) * g.ky + oy) * g.kx + ox], ifm[r]);
		}

		K[e] = v;
	}
}

// cov[l][r] = sum_k conj(K[k][i][r]) K[k][j][r] / scale, l = i (i + 1) / 2 + j, j <= i
__global__ void k_imgcov(PadGeom g, const float2* __restrict__ K, float scale, float2* __restrict__ cov)
{
	const long cosize = (long)g.C * (g.C + 1) / 2;
	const long total = cosize * g.Nh;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long r = e % g.Nh;
		const int l = (int)(e / g.Nh);
		int i = 0;
		while ((i + 1) * (i + 2) / 2 <= l)
			i++;
		const int j = l - i * (i + 1) / 2;
		double re = 0., im = 0.;

		for (int k = 0; k < g.n; k++) {
			const float2 a = K[((long)k * g.C + i) * g.Nh + r];
			const float2 b = K[((long)k * g.C + j) * g.Nh + r];
			re += (double)a.x * b.x + (double)a.y * b.y;
			im += (double)a.x * b.y - (double)a.y * b.x;
		}

		cov[e] = make_float2((float)(re / scale), (float)(im / scale));
	}
}

struct AxisGeom {
	int d[3];     // current dims (x, y, z)
	int axis;
	long nb;      // batch
};

__device__ inline int axis_coord(const AxisGeom& g, long r)
{
	if (0 == g.axis)
		return (int)(r % g.d[0]);
	if (1 == g.axis)
		return (int)((r / g.d[0]) % g.d[1]);
	return (int)(r / ((long)g.d[0] * g.d[1]));
}

// A *= tab[coordinate along the axis]
__global__ void k_mod_axis(AxisGeom g, const float2* __restrict__ tab, float2* __restrict__ A)
{
	const long np = (long)g.d[0] * g.d[1] * g.d[2];
	const long total = np * g.nb;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
		A[e] = cmul(A[e], tab[axis_coord(g, e % np)]);
}

// dst (dims o) = src (dims g) zero-padded about the centre along the axis,
// times fm_in (source position) and ifm_out (destination position)
__global__ void k_pad_axis(AxisGeom g, AxisGeom o, const float2* __restrict__ fm_in,
			   const float2* __restrict__ ifm_out, const float2* __restrict__ src, float2* __restrict__ dst)
{
	const long npo = (long)o.d[0] * o.d[1] * o.d[2];
	const long npi = (long)g.d[0] * g.d[1] * g.d[2];
	const long total = npo * o.nb;
	const int off = o.d[o.axis] / 2 - g.d[g.axis] / 2;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long r = e % npo;
		const long b = e / npo;
		int c[3] = { (int)(r % o.d[0]), (int)((r / o.d[0]) % o.d[1]), (int)(r / ((long)o.d[0] * o.d[1])) };
		const int oa = c[o.axis];
		const int q = oa - off;
		float2 v = make_float2(0.f, 0.f);

		if (q >= 0 && q < g.d[g.axis]) {
			c[o.axis] = q;
			const long ri = c[0] + (long)g.d[0] * (c[1] + (long)g.d[1] * c[2]);
			v = cmul(cmul(src[b * npi + ri], fm_in[q]), ifm_out[oa]);
		}

		dst[e] = v;
	}
}

struct MapsGeom {
	long np;
	int C, M, iters;
	float crop;
	bool rotphase;
};

__device__ inline float2 hval(const MapsGeom& g, const float2* cov, long p, int j, int k)
{
	if (j >= k)
		return cov[((long)j * (j + 1) / 2 + k) * g.np + p];
	return cconj(cov[((long)k * (k + 1) / 2 + j) * g.np + p]);
}

// sum over the G lanes of a group (butterfly, fixed order)
template <int G>
__device__ inline double group_sum(double v)
{
#pragma unroll
	for (int m = G / 2; m >= 1; m /= 2)
		v += __shfl_xor(v, m, G);
	return v;
}

// per point: orthogonal iteration (rows times H, gram_schmidt from the last
// row, num/linalg.c:1296-1314), maps u = row M - 1 - u with eigenvalue val,
// crop, phase of the principal component.  A group of G >= C lanes serves
// one point: lane k holds column k of H in registers and component k of
// each row; a row times H takes the row's components from the other lanes
// by shuffles, the dot products of the Gram-Schmidt step are group sums.
// Products and sums in double, as the single-lane form.
template <int G>
__global__ __launch_bounds__(256) void k_eigenmaps(MapsGeom g, const float2* __restrict__ cov,
						   const float2* __restrict__ rot, float2* __restrict__ maps,
						   float2* __restrict__ emaps)
{
	const int k = threadIdx.x % G;
	const long p0 = ((long)blockIdx.x * blockDim.x + threadIdx.x) / G;
	const bool live = p0 < g.np;
	const long p = live ? p0 : g.np - 1;  // idle groups run on the last point and store nothing
	const bool col = k < g.C;

	float2 h[G];

#pragma unroll
	for (int j = 0; j < G; j++)
		h[j] = (col && j < g.C) ? hval(g, cov, p, j, k) : make_float2(0.f, 0.f);

	float2 row[ESP_MAXM];
	float val[ESP_MAXM];

#pragma unroll
	for (int u = 0; u < ESP_MAXM; u++) {
		row[u] = make_float2((u == k) ? 1.f : 0.f, 0.f);
		val[u] = 0.f;
	}

	for (int it = 0; it < g.iters; it++) {

#pragma unroll
		for (int u = 0; u < ESP_MAXM; u++) {

			if (u >= g.M)
				break;

			const float2 t = row[u];
			double re = 0., im = 0.;

#pragma unroll
			for (int j = 0; j < G; j++) {
				const float2 a = make_float2(__shfl(t.x, j, G), __shfl(t.y, j, G));
				re += (double)a.x * h[j].x - (double)a.y * h[j].y;
				im += (double)a.x * h[j].y + (double)a.y * h[j].x;
			}

			row[u] = make_float2((float)re, (float)im);
		}

#pragma unroll
		for (int r = ESP_MAXM - 1; r >= 0; r--) {

			if (r >= g.M)
				continue;

#pragma unroll
			for (int j = r + 1; j < ESP_MAXM; j++) {

				if (j >= g.M)
					break;

				const float2 a = row[r], b = row[j];
				const float sr = (float)group_sum<G>((double)a.x * b.x + (double)a.y * b.y);
				const float si = (float)group_sum<G>((double)a.y * b.x - (double)a.x * b.y);
				row[r] = csub(row[r], cmul(make_float2(sr, si), row[j]));
			}

			const double nn = group_sum<G>((double)row[r].x * row[r].x + (double)row[r].y * row[r].y);
			val[r] = sqrtf((float)nn);
			row[r] = make_float2(row[r].x / val[r], row[r].y / val[r]);
		}
	}

	const float2 rc = col ? rot[k] : make_float2(0.f, 0.f);

#pragma unroll
	for (int u = 0; u < ESP_MAXM; u++) {

		if (u >= g.M)
			break;

		// map u is row M - 1 - u
		float2 a = make_float2(0.f, 0.f);
		float e = 0.f;

#pragma unroll
		for (int r = 0; r < ESP_MAXM; r++)
			if (r == g.M - 1 - u) {
				a = row[r];
				e = val[r];
			}

		const float w = (fabsf(e) >= g.crop) ? 1.f : 0.f;
		a = cscale(a, w);
		float2 ph = make_float2(1.f, 0.f);

		if (g.rotphase) {
			const double tr = group_sum<G>((double)a.x * rc.x + (double)a.y * rc.y);
			const double ti = group_sum<G>((double)a.y * rc.x - (double)a.x * rc.y);
			const double m = sqrt(tr * tr + ti * ti);
			if (m > 0.)
				ph = make_float2((float)(tr / m), (float)(ti / m));
		}

		if (live && col)
			maps[((long)u * g.C + k) * g.np + p] = cmulc(a, ph);

		if (live && 0 == k && nullptr != emaps)
			emaps[(long)u * g.np + p] = make_float2(e, 0.f);
	}
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------

static std::vector<float2> mod_table(int n, bool inv)
{
	std::vector<float2> t(n);
	for (int j = 0; j < n; j++) {
		std::complex<double> v = (n > 1) ? fftmod_value(fftmod_phase(n, j)) : 1.;
		if (inv)
			v = std::conj(v);
		t[j] = make_float2((float)v.real(), (float)v.imag());
	}
	return t;
}

struct Rocsolver {
	rocblas_handle h = nullptr;
	explicit Rocsolver(hipStream_t st)
	{
		BAM_CHECK(rocblas_status_success == rocblas_create_handle(&h), "ecalib: rocBLAS handle");
		rocblas_set_stream(h, st);
	}
	~Rocsolver() { (void)rocblas_destroy_handle(h); }

	// eigenvalues (ascending) and eigenvectors (columns, column-major) of the
	// n x n Hermitian matrix A as the column-major routine reads it
	void heevd(int n, float2* A, float* D)
	{
		DevBuf<float> E(std::max(n, 1));
		DevBuf<int> info(1);
		const rocblas_status s = rocsolver_cheevd(h, rocblas_evect_original, rocblas_fill_lower, n,
							  (rocblas_float_complex*)A, n, D, E.ptr, info.ptr);
		BAM_CHECK(rocblas_status_success == s, "ecalib: eigendecomposition failed (status %d)", (int)s);
		int ih = 0;
		BAM_HIP(hipMemcpy(&ih, info.ptr, sizeof(int), hipMemcpyDeviceToHost));
		BAM_CHECK(0 == ih, "ecalib: eigendecomposition did not converge (info %d)", ih);
	}
};

int espirit_calib(const EspiritConf& conf, const long img[3], int C, int maps, const long caldims[3],
		  const float2* cal, float2* maps_out, float2* emaps, float* svals, hipStream_t st)
{
	BAM_CHECK(C >= 1 && C <= ESP_MAXC, "ecalib: %d coils (at most %d)", C, ESP_MAXC);
	BAM_CHECK(maps >= 1 && maps <= std::min(C, ESP_MAXM), "ecalib: %d maps (at most %d)", maps,
		  std::min(C, ESP_MAXM));

	int kd[3], hd[3];
	for (int d = 0; d < 3; d++) {
		kd[d] = (img[d] > 1) ? (int)conf.kdims[d] : 1;
		BAM_CHECK(kd[d] >= 1 && caldims[d] >= kd[d], "ecalib: calibration region smaller than the kernel");
		hd[d] = (1 == kd[d]) ? 1 : 2 * kd[d];
		BAM_CHECK(1 == hd[d] || fft_size_supported(hd[d]), "ecalib: kernel size %d not supported", kd[d]);
	}

	CasGeom cg;
	cg.cx = (int)caldims[0];
	cg.cy = (int)caldims[1];
	cg.cz = (int)caldims[2];
	cg.kx = kd[0];
	cg.ky = kd[1];
	cg.kz = kd[2];
	cg.C = C;
	cg.px = cg.cx - kd[0] + 1;
	cg.py = cg.cy - kd[1] + 1;
	cg.pz = cg.cz - kd[2] + 1;
	cg.P = (long)cg.px * cg.py * cg.pz;
	cg.N = kd[0] * kd[1] * kd[2] * C;
	const int N = cg.N;

	// 1. Gram matrix of the calibration matrix
	const long ncal = (long)C * cg.cx * cg.cy * cg.cz;
	DevBuf<float2> dcal(ncal);
	BAM_HIP(hipMemcpyAsync(dcal.ptr, cal, sizeof(float2) * ncal, hipMemcpyHostToDevice, st));
	DevBuf<float2> G((size_t)N * N);
	const long nbt = (N + CG_T - 1) / CG_T;
	hipLaunchKernelGGL(k_cas_gram, dim3((unsigned)(nbt * (nbt + 1) / 2)), dim3(256), 0, st, cg, dcal.ptr, G.ptr);
	BAM_HIP(hipGetLastError());

	// 2. eigenvectors: the column-major routine reads the row-major Gram
	// matrix as its conjugate (as lapack_eig does in BART)
	Rocsolver rs(st);
	DevBuf<float> D(N);
	rs.heevd(N, G.ptr, D.ptr);

	std::vector<float> dh(N);
	BAM_HIP(hipMemcpyAsync(dh.data(), D.ptr, sizeof(float) * N, hipMemcpyDeviceToHost, st));
	BAM_HIP(hipStreamSynchronize(st));

	std::vector<float> val(N);
	for (int i = 0; i < N; i++)
		val[i] = (dh[N - 1 - i] < 0.f) ? 0.f : sqrtf(dh[N - 1 - i]);

	if (nullptr != svals)
		for (int i = 0; i < N; i++)
			svals[i] = val[i];

	BAM_CHECK(val[0] > 0.f, "No signal.");

	int nk = 0;
	for (int i = 0; i < N; i++)
		if (val[i] / val[0] > sqrtf(conf.threshold))
			nk++;

	BAM_CHECK(nk >= 1, "ecalib: no kernels above the threshold");

	// 3. image-space covariance on the grid of twice the kernel size
	PadGeom pg;
	pg.kx = kd[0];
	pg.ky = kd[1];
	pg.kz = kd[2];
	pg.C = C;
	pg.n = nk;
	pg.xh = hd[0];
	pg.yh = hd[1];
	pg.zh = hd[2];
	pg.N = N;
	pg.Nh = (long)hd[0] * hd[1] * hd[2];

	std::vector<float2> ifm(pg.Nh);
	for (long r = 0; r < pg.Nh; r++) {
		const long c[3] = { r % hd[0], (r / hd[0]) % hd[1], r / ((long)hd[0] * hd[1]) };
		double ph = 0.;
		for (int d = 0; d < 3; d++)
			if (hd[d] > 1)
				ph += fftmod_phase(hd[d], c[d]);
		const std::complex<double> v = std::conj(fftmod_value(ph));
		ifm[r] = make_float2((float)v.real(), (float)v.imag());
	}

	DevBuf<float2> difm(pg.Nh);
	BAM_HIP(hipMemcpyAsync(difm.ptr, ifm.data(), sizeof(float2) * pg.Nh, hipMemcpyHostToDevice, st));
	DevBuf<float2> K((size_t)nk * C * pg.Nh);
	hipLaunchKernelGGL(k_pad_kernels, dim3(nblk((long)nk * C * pg.Nh)), dim3(256), 0, st, pg, G.ptr, difm.ptr,
			   K.ptr);
	BAM_HIP(hipGetLastError());
	fft_nd(hd, (long)nk * C, 7, +1.f, K.ptr, K.ptr, 1.f, st);

	const long cosize = (long)C * (C + 1) / 2;
	DevBuf<float2> cov(cosize * pg.Nh);
	const float scale = (float)((double)kd[0] * kd[1] * kd[2] * (double)pg.Nh);
	hipLaunchKernelGGL(k_imgcov, dim3(nblk(cosize * pg.Nh)), dim3(256), 0, st, pg, K.ptr, scale, cov.ptr);
	BAM_HIP(hipGetLastError());

	// 4. sinc interpolation to the image, z first
	int cur[3] = { hd[0], hd[1], hd[2] };
	DevBuf<float2> a = std::move(cov);

	for (int ax = 2; ax >= 0; ax--) {

		if (cur[ax] == img[ax])
			continue;

		BAM_CHECK(cur[ax] < img[ax] && fft_size_supported((int)img[ax]), "ecalib: image dim %ld not supported",
			  img[ax]);

		int nxt[3] = { cur[0], cur[1], cur[2] };
		nxt[ax] = (int)img[ax];
		const std::vector<float2> fin = mod_table(cur[ax], false), iout = mod_table(nxt[ax], true);
		DevBuf<float2> dfin(fin.size()), diout(iout.size());
		BAM_HIP(hipMemcpyAsync(dfin.ptr, fin.data(), sizeof(float2) * fin.size(), hipMemcpyHostToDevice, st));
		BAM_HIP(hipMemcpyAsync(diout.ptr, iout.data(), sizeof(float2) * iout.size(), hipMemcpyHostToDevice, st));

		AxisGeom gi = { { cur[0], cur[1], cur[2] }, ax, cosize };
		AxisGeom go = { { nxt[0], nxt[1], nxt[2] }, ax, cosize };
		const long ni = (long)cur[0] * cur[1] * cur[2] * cosize;
		const long no = (long)nxt[0] * nxt[1] * nxt[2] * cosize;

		hipLaunchKernelGGL(k_mod_axis, dim3(nblk(ni)), dim3(256), 0, st, gi, dfin.ptr, a.ptr);
		BAM_HIP(hipGetLastError());
		fft_nd(cur, cosize, 1u << ax, -1.f, a.ptr, a.ptr, 1.f, st);
		DevBuf<float2> b(no);
		hipLaunchKernelGGL(k_pad_axis, dim3(nblk(no)), dim3(256), 0, st, gi, go, dfin.ptr, diout.ptr, a.ptr, b.ptr);
		BAM_HIP(hipGetLastError());
		fft_nd(nxt, cosize, 1u << ax, +1.f, b.ptr, b.ptr, 1.f, st);
		hipLaunchKernelGGL(k_mod_axis, dim3(nblk(no)), dim3(256), 0, st, go, diout.ptr, b.ptr);
		BAM_HIP(hipGetLastError());
		BAM_HIP(hipStreamSynchronize(st));

		a = std::move(b);
		for (int d = 0; d < 3; d++)
			cur[d] = nxt[d];
	}

	// 5. channel principal component (scc, calib/cc.c:45-63): eigenvectors
	// of the C x C covariance as the column-major routine reads it
	std::vector<float2> rot(C, make_float2(0.f, 0.f));

	if (conf.rotphase) {

		const long cs = (long)cg.cx * cg.cy * cg.cz;
		std::vector<float2> T((size_t)C * C);

		for (int i = 0; i < C; i++)
			for (int j = 0; j < C; j++) {
				double re = 0., im = 0.;
				for (long q = 0; q < cs; q++) {
					const float2 x = cal[(long)i * cs + q], y = cal[(long)j * cs + q];
					re += (double)x.x * y.x + (double)x.y * y.y;
					im += (double)x.x * y.y - (double)x.y * y.x;
				}
				T[(size_t)i * C + j] = make_float2((float)re, (float)im);
			}

		DevBuf<float2> dT((size_t)C * C);
		DevBuf<float> dD(C);
		BAM_HIP(hipMemcpyAsync(dT.ptr, T.data(), sizeof(float2) * C * C, hipMemcpyHostToDevice, st));
		rs.heevd(C, dT.ptr, dD.ptr);
		BAM_HIP(hipMemcpyAsync(T.data(), dT.ptr, sizeof(float2) * C * C, hipMemcpyDeviceToHost, st));
		BAM_HIP(hipStreamSynchronize(st));

		// largest eigenvalue: the last column; its first entry made real
		const float2* v = T.data() + (size_t)(C - 1) * C;
		const double m0 = sqrt((double)v[0].x * v[0].x + (double)v[0].y * v[0].y);
		const float2 ph = (m0 > 0.) ? make_float2((float)(v[0].x / m0), (float)(v[0].y / m0)) : make_float2(1.f, 0.f);

		for (int c = 0; c < C; c++)
			rot[c] = cmulc(v[c], ph);
	}

	DevBuf<float2> drot(C);
	BAM_HIP(hipMemcpyAsync(drot.ptr, rot.data(), sizeof(float2) * C, hipMemcpyHostToDevice, st));

	MapsGeom mg;
	mg.np = img[0] * img[1] * img[2];
	mg.C = C;
	mg.M = maps;
	mg.iters = conf.num_orthiter;
	mg.crop = conf.crop;
	mg.rotphase = conf.rotphase;
	// lanes per point: the smallest group of 4, 8, 16 or 32 that holds C
	const int lanes = (C <= 4) ? 4 : (C <= 8) ? 8 : (C <= 16) ? 16 : 32;
	const long nb = (mg.np * lanes + 255) / 256;
	BAM_CHECK(nb < (1L << 31), "ecalib: image too large");
	auto kern = (4 == lanes) ? k_eigenmaps<4> : (8 == lanes) ? k_eigenmaps<8> : (16 == lanes) ? k_eigenmaps<16>
											 : k_eigenmaps<32>;
	hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), 0, st, mg, a.ptr, drot.ptr, maps_out, emaps);
	BAM_HIP(hipGetLastError());
	BAM_HIP(hipStreamSynchronize(st));

	return nk;
}

}  // namespace bam
