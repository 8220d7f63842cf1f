// cnn.hip -- 3x3 convolutions of the CG warm-start network on the matrix
// cores (gfx950): BF16 operands, FP32 accumulation.
//
// No counterpart in BART: the network that supplies a CG initial guess (the
// x0 of pics -W, iter/itop.c:54-66) is new.  Layout: activations
// [B][C][H][W] float, weights [Cout][Cin][3][3] float (k = ci 9 + 3 u + v),
// zero padding, cross-correlation:
//   y[b][co][i][j] = bias[co] + sum_{ci,u,v} w[co][ci][u][v] x[b][ci][i+u-1][j+v-1]
// Every layer is a GEMM of M = output channels, N = pixels, K = Cin 9 taps
// (forward), or M = output channels, N = Cin 9 taps, K = pixels (weight
// gradient), on v_mfma_f32_32x32x8_bf16: lane l supplies A[l % 32][4 (l / 32)
// + t] and B[4 (l / 32) + t][l % 32], t < 4; accumulator register r of lane
// l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  Operands are rounded to
// bfloat16 (nearest, ties to even) when they are staged into LDS.
#include "cnn.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <random>
#include <vector>

namespace bam {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ inline unsigned short f2bf(float f)
{
	unsigned u = __float_as_uint(f);
	u += 0x7FFFu + ((u >> 16) & 1u);  // nearest, ties to even (finite values)
	return (unsigned short)(u >> 16);
}

constexpr int CV_MAXC = 64;    // channels (in and out) per layer at most
constexpr int CV_PIX = 64;     // forward: pixels (row segment) per work-group
constexpr int CV_ROWS = 4;     // forward: rows per work-group of a large launch
constexpr int CV_FWD_THREADS = 256;
constexpr int CV_WG_THREADS = 512;
constexpr int CV_WG_ROWS = 8;  // weight gradient: image rows per work-group
constexpr int CV_WG_MAXW = 256;

__host__ __device__ inline int round8(int v) { return (v + 7) / 8 * 8; }

// ---------------------------------------------------------------------------
// Forward: work-group = 64 pixels of one row x 64 output channels, 4 waves of
// one 32 x 32 tile each.  K is ordered (tap, channel) with the channels
// padded to CP (a multiple of 8): the weights are first converted to
// bfloat16 in that order ([64][9 CP], k_conv_wprep) and read from there
// (cached), and the 3 x 66 input patch is staged in LDS channel-fastest, so
// that a lane's 4 consecutive k of one matrix-core step -- 4 channels of one
// tap -- are one 8-byte read for A and one LDS read for B.
// ---------------------------------------------------------------------------

__host__ __device__ inline int chan_pad(int c) { return (c + 7) / 8 * 8; }

// wb[co][t CP + ci] = bf16(w[co][ci][t]), zero for padded co / ci
__global__ void k_conv_wprep(const float* __restrict__ w, int Cin, int Cout, int CP,
			     unsigned short* __restrict__ wb)
{
	const int KP = 9 * CP;
	const int n = CV_MAXC * KP;

	for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
		const int co = e / KP;
		const int k = e - co * KP;
		const int t = k / CP;
		const int ci = k - t * CP;
		wb[e] = f2bf((co < Cout && ci < Cin) ? w[(co * Cin + ci) * 9 + t] : 0.f);
	}
}

template <int ROWS>
__global__ __launch_bounds__(CV_FWD_THREADS) void k_conv3x3_fwd(const float* __restrict__ x, int Cin, int H, int W,
								const unsigned short* __restrict__ wb,
								const float* __restrict__ bias,
								const float* __restrict__ res, int Cout, int relu,
								float* __restrict__ y)
{
	extern __shared__ __align__(16) unsigned char cv_smem[];
	const int CP = chan_pad(Cin);
	unsigned short* xl = (unsigned short*)cv_smem;          // [ROWS + 2][66][CP]
	constexpr int PR = ROWS + 2;                          // patch rows

	const int x0 = blockIdx.x * CV_PIX;
	const int y0 = blockIdx.y * ROWS;
	const int b = blockIdx.z;

	// input patch, channel-fastest
	for (int e = threadIdx.x; e < CP * PR * 66; e += CV_FWD_THREADS) {
		const int ci = e / (PR * 66);
		const int r = e - ci * (PR * 66);
		const int u = r / 66;
		const int xx = r - u * 66;
		const int gx = x0 + xx - 1;
		const int gy = y0 + u - 1;
		float v = 0.f;
		if (ci < Cin && gy >= 0 && gy < H && gx >= 0 && gx < W)
			v = x[((long)(b * Cin + ci) * H + gy) * W + gx];
		xl[(u * 66 + xx) * CP + ci] = f2bf(v);
	}

	__syncthreads();

	const int lane = threadIdx.x & 63;
	const int wv = threadIdx.x >> 6;
	const int cot = wv & 1;
	const int pt = wv >> 1;
	const int i = lane & 31;
	const int hk = lane >> 5;
	// A fragments straight from the prepared weights (a few KiB, cached),
	// each used for the ROWS rows of the block
	const unsigned short* wrow = wb + (cot * 32 + i) * (9 * CP) + 4 * hk;
	const int px = pt * 32 + i;

	f32x16 acc[ROWS];
#pragma unroll
	for (int q = 0; q < ROWS; q++)
#pragma unroll
		for (int r = 0; r < 16; r++)
			acc[q][r] = 0.f;

	for (int t = 0; t < 9; t++) {

		const int u = t / 3;
		const int v = t - 3 * u;
		const unsigned short* xrow = xl + (u * 66 + px + v) * CP + 4 * hk;
		const unsigned short* wt = wrow + t * CP;

		for (int c0 = 0; c0 < CP; c0 += 8) {
			const bf16x4 a = *(const bf16x4*)(wt + c0);
#pragma unroll
			for (int q = 0; q < ROWS; q++) {
				const bf16x4 bv = *(const bf16x4*)(xrow + q * 66 * CP + c0);
				acc[q] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bv, acc[q], 0, 0, 0);
			}
		}
	}

	const int gx = x0 + px;

	if (gx < W) {
#pragma unroll
		for (int q = 0; q < ROWS; q++) {
			const int yy = y0 + q;
			if (yy >= H)
				break;
#pragma unroll
			for (int r = 0; r < 16; r++) {
				const int co = cot * 32 + 8 * (r / 4) + 4 * hk + (r % 4);
				if (co < Cout) {
					float v = acc[q][r] + ((nullptr != bias) ? bias[co] : 0.f);
					if (relu)
						v = fmaxf(v, 0.f);
					const long o = ((long)(b * Cout + co) * H + yy) * W + gx;
					if (nullptr != res)
						v += res[o];
					y[o] = v;
				}
			}
		}
	}
}

static size_t fwd_lds(int Cin, int rows)
{
	return sizeof(unsigned short) * (size_t)(rows + 2) * 66 * chan_pad(Cin);
}

// w'[ci][co][u][v] = w[co][ci][2 - u][2 - v]: the weights of the transposed
// convolution (the data gradient)
__global__ void k_conv_flip(const float* __restrict__ w, int Cout, int Cin, float* __restrict__ wt)
{
	const int n = Cout * Cin * 9;

	for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
		const int co = e / (Cin * 9);
		const int r = e - co * Cin * 9;
		const int ci = r / 9;
		const int t = r - ci * 9;
		wt[(ci * Cout + co) * 9 + (8 - t)] = w[e];
	}
}

// ---------------------------------------------------------------------------
// Weight gradient: work-group = CV_WG_ROWS rows of one image; per row the
// output-gradient row (M = Cout x K = W pixels) and the 3 input rows around
// it (N = Cin 9 taps) are staged in LDS; 8 waves share the
// ceil(Cout / 32) x ceil(Cin 9 / 32) tiles.  Each work-group writes its
// partial [64][NP] gradient; k_conv_wgrad_sum adds the partials in order.
// ---------------------------------------------------------------------------

constexpr int CV_WG_TILES = 4;  // tiles per wave at most (2 x 14 tiles over 8 waves)

__global__ __launch_bounds__(CV_WG_THREADS) void k_conv3x3_wgrad(const float* __restrict__ x,
								  const float* __restrict__ dy, int Cin, int Cout,
								  int H, int W, float* __restrict__ part)
{
	extern __shared__ __align__(16) unsigned char cv_smem[];
	const int K = Cin * 9;
	const int NP = (K + 31) / 32 * 32;
	const int WX = W + 2;
	unsigned short* dyl = (unsigned short*)cv_smem;         // [64][W]
	unsigned short* xl = dyl + CV_MAXC * W;                 // [Cin][3][W + 2]
	int* koff = (int*)(xl + round8(Cin * 3 * WX));          // [NP]

	const int nrg = (H + CV_WG_ROWS - 1) / CV_WG_ROWS;
	const int b = blockIdx.x / nrg;
	const int y0 = (blockIdx.x - b * nrg) * CV_WG_ROWS;

	for (int k = threadIdx.x; k < NP; k += CV_WG_THREADS) {
		const int ci = k / 9;
		const int t = k - ci * 9;
		koff[k] = (k < K) ? (ci * 3 + t / 3) * WX + (t % 3) : -1;
	}

	const int lane = threadIdx.x & 63;
	const int wv = threadIdx.x >> 6;
	const int i = lane & 31;
	const int hk = lane >> 5;
	const int ncot = (Cout + 31) / 32;
	const int nkt = NP / 32;
	const int ntiles = ncot * nkt;

	f32x16 acc[CV_WG_TILES];
#pragma unroll
	for (int q = 0; q < CV_WG_TILES; q++)
#pragma unroll
		for (int r = 0; r < 16; r++)
			acc[q][r] = 0.f;

	for (int yy = y0; yy < min(H, y0 + CV_WG_ROWS); yy++) {

		__syncthreads();

		for (int e = threadIdx.x; e < CV_MAXC * W; e += CV_WG_THREADS) {
			const int co = e / W;
			const int xx = e - co * W;
			dyl[e] = f2bf((co < Cout) ? dy[((long)(b * Cout + co) * H + yy) * W + xx] : 0.f);
		}

		for (int e = threadIdx.x; e < Cin * 3 * WX; e += CV_WG_THREADS) {
			const int cu = e / WX;
			const int ci = cu / 3;
			const int u = cu - ci * 3;
			const int gx = e - cu * WX - 1;
			const int gy = yy + u - 1;
			float v = 0.f;
			if (gy >= 0 && gy < H && gx >= 0 && gx < W)
				v = x[((long)(b * Cin + ci) * H + gy) * W + gx];
			xl[e] = f2bf(v);
		}

		__syncthreads();

#pragma unroll
		for (int q = 0; q < CV_WG_TILES; q++) {

			const int tile = wv + 8 * q;

			if (tile >= ntiles)
				break;

			const int cot = tile / nkt;
			const int kt = tile - cot * nkt;
			const unsigned short* arow = dyl + (cot * 32 + i) * W;
			const int o = koff[kt * 32 + i];

			for (int p0 = 0; p0 < W; p0 += 8) {

				const int pb = p0 + 4 * hk;
				const bf16x4 a = *(const bf16x4*)(arow + pb);
				bf16x4 bv;
#pragma unroll
				for (int t = 0; t < 4; t++)
					bv[t] = (o >= 0) ? (short)xl[o + pb + t] : (short)0;
				acc[q] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bv, acc[q], 0, 0, 0);
			}
		}
	}

	// partial [64][NP] of this work-group
	float* pp = part + (long)blockIdx.x * CV_MAXC * NP;

#pragma unroll
	for (int q = 0; q < CV_WG_TILES; q++) {

		const int tile = wv + 8 * q;

		if (tile >= ntiles)
			break;

		const int cot = tile / nkt;
		const int kt = tile - cot * nkt;

#pragma unroll
		for (int r = 0; r < 16; r++) {
			const int co = cot * 32 + 8 * (r / 4) + 4 * hk + (r % 4);
			pp[co * NP + kt * 32 + i] = acc[q][r];
		}
	}
}

// dw[co][k] = sum over the partials in work-group order
__global__ void k_conv_wgrad_sum(const float* __restrict__ part, int nparts, int Cout, int K, int NP,
				 float* __restrict__ dw)
{
	const int n = Cout * K;

	for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
		const int co = e / K;
		const int k = e - co * K;
		float s = 0.f;
		for (int p = 0; p < nparts; p++)
			s += part[((long)p * CV_MAXC + co) * NP + k];
		dw[e] = s;
	}
}

// db[co] = sum over b, i, j of dy (one work-group per channel, fixed order)
__global__ __launch_bounds__(256) void k_conv_bgrad(const float* __restrict__ dy, int B, int Cout, long HW,
						     float* __restrict__ db)
{
	__shared__ float sh[256];
	const int co = blockIdx.x;
	float s = 0.f;

	for (int b = 0; b < B; b++)
		for (long e = threadIdx.x; e < HW; e += 256)
			s += dy[((long)b * Cout + co) * HW + e];

	sh[threadIdx.x] = s;
	__syncthreads();

	for (int h = 128; h > 0; h >>= 1) {
		if (threadIdx.x < h)
			sh[threadIdx.x] += sh[threadIdx.x + h];
		__syncthreads();
	}

	if (0 == threadIdx.x)
		db[co] = sh[0];
}

static size_t wgrad_lds(int Cin, int W)
{
	const int NP = (Cin * 9 + 31) / 32 * 32;
	return sizeof(unsigned short) * ((size_t)CV_MAXC * W + round8(Cin * 3 * (W + 2))) + sizeof(int) * NP;
}

// work buffers of the weight gradient and the flipped weights
struct CnnWork {
	std::mutex mtx;
	DevBuf<float> part;
	DevBuf<float> wt;
	DevBuf<unsigned short> wb;
};

static CnnWork& cnn_work()
{
	static CnnWork w;
	return w;
}

static void check_shape(int B, int Cin, int Cout, int H, int W)
{
	BAM_CHECK(B >= 1 && H >= 1 && W >= 1, "conv3x3: bad shape");
	BAM_CHECK(Cin >= 1 && Cin <= CV_MAXC && Cout >= 1 && Cout <= CV_MAXC,
		  "conv3x3: 1 to %d channels in and out (got %d, %d)", CV_MAXC, Cin, Cout);
}

// the bfloat16 weights of the forward kernel, prepared per call
static void fwd_launch(const float* x, int B, int Cin, int H, int W, const float* w, const float* bias,
		       const float* res, int Cout, bool relu, float* y, hipStream_t st, DevBuf<unsigned short>& wb)
{
	check_shape(B, Cin, Cout, H, W);
	BAM_CHECK(H <= 65535 && B <= 65535, "conv3x3: too many rows or images");
	const int CP = chan_pad(Cin);
	const long nw = (long)CV_MAXC * 9 * CP;
	if ((long)wb.n < nw)
		wb.alloc(nw);
	hipLaunchKernelGGL(k_conv_wprep, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, w, Cin, Cout, CP, wb.ptr);
	// CV_ROWS rows per work-group reuse each weight fragment CV_ROWS times;
	// a single image keeps one row per work-group so that every CU has work
	const long big = (long)((W + CV_PIX - 1) / CV_PIX) * ((H + CV_ROWS - 1) / CV_ROWS) * B;
	const int rows = (big >= 2048) ? CV_ROWS : 1;
	dim3 grid((W + CV_PIX - 1) / CV_PIX, (H + rows - 1) / rows, B);
	if (CV_ROWS == rows)
		hipLaunchKernelGGL(k_conv3x3_fwd<CV_ROWS>, grid, dim3(CV_FWD_THREADS), fwd_lds(Cin, rows), st, x, Cin, H, W,
				   wb.ptr, bias, res, Cout, relu ? 1 : 0, y);
	else
		hipLaunchKernelGGL(k_conv3x3_fwd<1>, grid, dim3(CV_FWD_THREADS), fwd_lds(Cin, rows), st, x, Cin, H, W,
				   wb.ptr, bias, res, Cout, relu ? 1 : 0, y);
	BAM_HIP(hipGetLastError());
}

void conv3x3_fwd(const float* x, int B, int Cin, int H, int W, const float* w, const float* bias, const float* res,
		 int Cout, bool relu, float* y, hipStream_t st)
{
	// the prepared weights are rewritten by the next call: calls are ordered
	// on the one stream every layer call uses
	CnnWork& cw = cnn_work();
	std::lock_guard<std::mutex> lock(cw.mtx);
	fwd_launch(x, B, Cin, H, W, w, bias, res, Cout, relu, y, st, cw.wb);
}

void conv3x3_bwd_data(const float* dy, int B, int Cout, int H, int W, const float* w, int Cin, float* dx,
		      hipStream_t st)
{
	check_shape(B, Cin, Cout, H, W);
	CnnWork& cw = cnn_work();
	std::lock_guard<std::mutex> lock(cw.mtx);
	const long n = (long)Cout * Cin * 9;
	if ((long)cw.wt.n < n)
		cw.wt.alloc(n);
	hipLaunchKernelGGL(k_conv_flip, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, Cout, Cin, cw.wt.ptr);
	BAM_HIP(hipGetLastError());
	fwd_launch(dy, B, Cout, H, W, cw.wt.ptr, nullptr, nullptr, Cin, false, dx, st, cw.wb);
	BAM_HIP(hipStreamSynchronize(st));  // the flipped weights are reused by the next call
}

void conv3x3_bwd_weights(const float* x, const float* dy, int B, int Cin, int Cout, int H, int W, float* dw,
			 float* db, hipStream_t st)
{
	check_shape(B, Cin, Cout, H, W);
	BAM_CHECK(W <= CV_WG_MAXW && 0 == W % 8, "conv3x3 weight gradient: width a multiple of 8, at most %d",
		  CV_WG_MAXW);
	CnnWork& cw = cnn_work();
	std::lock_guard<std::mutex> lock(cw.mtx);
	const int NP = (Cin * 9 + 31) / 32 * 32;
	BAM_CHECK(((Cout + 31) / 32) * (NP / 32) <= 8 * CV_WG_TILES, "conv3x3 weight gradient: too many tiles");
	const int nparts = B * ((H + CV_WG_ROWS - 1) / CV_WG_ROWS);
	const long np = (long)nparts * CV_MAXC * NP;
	if ((long)cw.part.n < np)
		cw.part.alloc(np);
	hipLaunchKernelGGL(k_conv3x3_wgrad, dim3(nparts), dim3(CV_WG_THREADS), wgrad_lds(Cin, W), st, x, dy, Cin, Cout,
			   H, W, cw.part.ptr);
	const int n = Cout * Cin * 9;
	hipLaunchKernelGGL(k_conv_wgrad_sum, dim3((n + 255) / 256), dim3(256), 0, st, cw.part.ptr, nparts, Cout,
			   Cin * 9, NP, dw);
	if (nullptr != db)
		hipLaunchKernelGGL(k_conv_bgrad, dim3(Cout), dim3(256), 0, st, dy, B, Cout, (long)H * W, db);
	BAM_HIP(hipGetLastError());
	BAM_HIP(hipStreamSynchronize(st));  // the partials are reused by the next call
}

}  // namespace bam

// ---------------------------------------------------------------------------
// The warm-start network: 6 convolution layers 2 -> C -> C -> C -> C -> C -> 2
// with ReLU after the first five and the input added to the output
// (residual), trained on mean squared error with Adam.
// ---------------------------------------------------------------------------

namespace bam {

// g[i] = 0 where the layer's output a[i] was clipped by the ReLU
__global__ void k_relu_bwd(float* __restrict__ g, const float* __restrict__ a, long n)
{
	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
		if (!(a[i] > 0.f))
			g[i] = 0.f;
}

// d = 2 (out - t) / n; per-work-group partial sums of (out - t)^2
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ out, const float* __restrict__ t, long n,
					      float* __restrict__ d, double* __restrict__ part)
{
	__shared__ double sh[256];
	double s = 0.;

	for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
		const float e = out[i] - t[i];
		d[i] = 2.f * e / (float)n;
		s += (double)e * (double)e;
	}

	sh[threadIdx.x] = s;
	__syncthreads();

	for (int h = 128; h > 0; h >>= 1) {
		if (threadIdx.x < h)
			sh[threadIdx.x] += sh[threadIdx.x + h];
		__syncthreads();
	}

	if (0 == threadIdx.x)
		part[blockIdx.x] = sh[0];
}

// Adam (Kingma & Ba, ICLR 2015) with bias correction
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
		       float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
		const float gi = g[i];
		const float mi = b1 * m[i] + (1.f - b1) * gi;
		const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
		m[i] = mi;
		v[i] = vi;
		p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
	}
}

constexpr int WN_LAYERS = 6;
constexpr int WN_MSE_BLOCKS = 256;

WarmNet::WarmNet(int channels, unsigned long seed)
{
	BAM_CHECK(channels >= 1 && channels <= CV_MAXC, "warm-start network: 1 to %d channels", CV_MAXC);
	ch[0] = 2;
	for (int l = 1; l < WN_LAYERS; l++)
		ch[l] = channels;
	ch[WN_LAYERS] = 2;

	long n = 0;
	for (int l = 0; l < WN_LAYERS; l++) {
		woff[l] = n;
		n += (long)ch[l + 1] * ch[l] * 9;
		boff[l] = n;
		n += ch[l + 1];
	}
	nparams = n;

	// He-normal weights (std sqrt(2 / fan-in)); the last layer x 0.1 so that
	// the untrained network is close to the identity; zero biases
	std::vector<float> h(n, 0.f);
	std::mt19937_64 rng(seed);
	std::normal_distribution<double> nd(0., 1.);
	for (int l = 0; l < WN_LAYERS; l++) {
		const double sd = sqrt(2. / (9. * ch[l])) * ((WN_LAYERS - 1 == l) ? 0.1 : 1.);
		for (long i = 0; i < (long)ch[l + 1] * ch[l] * 9; i++)
			h[woff[l] + i] = (float)(sd * nd(rng));
	}

	params.alloc(n);
	grads.alloc(n);
	m.alloc(n);
	v.alloc(n);
	BAM_HIP(hipMemcpy(params.ptr, h.data(), sizeof(float) * n, hipMemcpyHostToDevice));
	BAM_HIP(hipMemset(m.ptr, 0, sizeof(float) * n));
	BAM_HIP(hipMemset(v.ptr, 0, sizeof(float) * n));
	lpart.alloc(WN_MSE_BLOCKS);
}

void WarmNet::ensure(int B, int H, int W)
{
	const long hid = (long)B * ch[1] * H * W;
	if ((long)act.n < (WN_LAYERS - 1) * hid)
		act.alloc((WN_LAYERS - 1) * hid);
	if ((long)g0.n < hid) {
		g0.alloc(hid);
		g1.alloc(hid);
	}
	const long io = (long)B * 2 * H * W;
	if ((long)out.n < io)
		out.alloc(io);
}

// a_l: input of layer l (a_0 = x)
const float* WarmNet::a(int l, const float* x, int B, int H, int W) const
{
	return (0 == l) ? x : act.ptr + (long)(l - 1) * B * ch[1] * H * W;
}

void WarmNet::forward(const float* x, int B, int H, int W, float* y, hipStream_t st)
{
	ensure(B, H, W);

	for (int l = 0; l < WN_LAYERS; l++) {
		const bool last = WN_LAYERS - 1 == l;
		float* dst = last ? y : (float*)a(l + 1, x, B, H, W);
		conv3x3_fwd(a(l, x, B, H, W), B, ch[l], H, W, params.ptr + woff[l], params.ptr + boff[l],
			    last ? x : nullptr, ch[l + 1], !last, dst, st);
	}
}

// Backpropagation of the output gradient in g0 through the layer outputs of
// the last forward (input x), then one Adam step.
void WarmNet::backward_update(const float* x, int B, int H, int W, float lr, hipStream_t st)
{
	float* g = g0.ptr;      // gradient at the output of layer l
	float* gn = g1.ptr;

	for (int l = WN_LAYERS - 1; l >= 0; l--) {

		const float* al = a(l, x, B, H, W);
		conv3x3_bwd_weights(al, g, B, ch[l], ch[l + 1], H, W, grads.ptr + woff[l], grads.ptr + boff[l], st);

		if (l > 0) {
			conv3x3_bwd_data(g, B, ch[l + 1], H, W, params.ptr + woff[l], ch[l], gn, st);
			const long na = (long)B * ch[l] * H * W;
			hipLaunchKernelGGL(k_relu_bwd, dim3(1024), dim3(256), 0, st, gn, al, na);
			BAM_HIP(hipGetLastError());
			std::swap(g, gn);
		}
	}

	step++;
	const float b1 = 0.9f, b2 = 0.999f;
	const float c1 = 1.f - (float)pow(b1, step), c2 = 1.f - (float)pow(b2, step);
	hipLaunchKernelGGL(k_adam, dim3(256), dim3(256), 0, st, params.ptr, grads.ptr, m.ptr, v.ptr, nparams, lr, b1, b2,
			   1e-8f, c1, c2);
	BAM_HIP(hipGetLastError());
}

double WarmNet::train_step(const float* x, const float* t, int B, int H, int W, float lr, hipStream_t st)
{
	ensure(B, H, W);
	forward(x, B, H, W, out.ptr, st);

	const long n = (long)B * 2 * H * W;
	hipLaunchKernelGGL(k_mse, dim3(WN_MSE_BLOCKS), dim3(256), 0, st, out.ptr, t, n, g0.ptr, lpart.ptr);
	BAM_HIP(hipGetLastError());

	backward_update(x, B, H, W, lr, st);

	std::vector<double> hp(WN_MSE_BLOCKS);
	BAM_HIP(hipMemcpyAsync(hp.data(), lpart.ptr, sizeof(double) * WN_MSE_BLOCKS, hipMemcpyDeviceToHost, st));
	BAM_HIP(hipStreamSynchronize(st));
	double s = 0.;
	for (double v : hp)
		s += v;
	return s / (double)n;
}

void WarmNet::train_step_grad(const float* x, const float* dout, int B, int H, int W, float lr, hipStream_t st)
{
	ensure(B, H, W);
	const long n = (long)B * 2 * H * W;
	BAM_HIP(hipMemcpyAsync(g0.ptr, dout, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
	backward_update(x, B, H, W, lr, st);
	BAM_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// The network's input and the guess it gives (first CG iterate from zero,
// scale, channels)
// ---------------------------------------------------------------------------

constexpr int GS_BLOCKS = 256;
constexpr int GS_THREADS = 256;

// partial sums of the first step along a direction d for the right-hand
// side b: d^H b (re, im), d^H (A^H A d + lambda d) and max |d|^2, per block
// (FP64; d = b gives CG's first step from zero)
__global__ __launch_bounds__(GS_THREADS) void k_guess_partials(const float* __restrict__ d, const float* __restrict__ b,
							       const float* __restrict__ Ad, float lambda, long n,
							       double* __restrict__ part)
{
	__shared__ double sh[4][GS_THREADS / 64];
	double db = 0., dbi = 0., dad = 0., mx = 0.;

	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
		const float2 dv = ((const float2*)d)[i];
		const float2 bv = ((const float2*)b)[i];
		const float2 av = ((const float2*)Ad)[i];
		const float a0 = av.x + lambda * dv.x;
		const float a1 = av.y + lambda * dv.y;
		db += (double)(dv.x * bv.x) + (double)(dv.y * bv.y);
		dbi += (double)(dv.x * bv.y) - (double)(dv.y * bv.x);
		dad += (double)(dv.x * a0) + (double)(dv.y * a1);
		mx = fmax(mx, (double)dv.x * dv.x + (double)dv.y * dv.y);
	}

	for (int off = 32; off > 0; off >>= 1) {
		db += __shfl_down(db, off, 64);
		dbi += __shfl_down(dbi, off, 64);
		dad += __shfl_down(dad, off, 64);
		mx = fmax(mx, __shfl_down(mx, off, 64));
	}

	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;

	if (0 == lane) {
		sh[0][wid] = db;
		sh[1][wid] = dad;
		sh[2][wid] = mx;
		sh[3][wid] = dbi;
	}

	__syncthreads();

	if (0 == threadIdx.x) {
		double a = 0., c = 0., m = 0., ai = 0.;
		for (int w = 0; w < GS_THREADS / 64; w++) {
			a += sh[0][w];
			c += sh[1][w];
			m = fmax(m, sh[2][w]);
			ai += sh[3][w];
		}
		part[blockIdx.x] = a;
		part[GS_BLOCKS + blockIdx.x] = c;
		part[2 * GS_BLOCKS + blockIdx.x] = m;
		part[3 * GS_BLOCKS + blockIdx.x] = ai;
	}
}

// sc = { re alpha, s, 1 / s, im alpha }: x1 = alpha d with alpha =
// d^H b / d^H (A^H A + lambda I) d, s = max |x1| (1 when x1 = 0)
__global__ void k_guess_scalars(const double* __restrict__ part, float* __restrict__ sc)
{
	if (0 != threadIdx.x)
		return;

	double a = 0., c = 0., m = 0., ai = 0.;
	for (int i = 0; i < GS_BLOCKS; i++) {
		a += part[i];
		c += part[GS_BLOCKS + i];
		m = fmax(m, part[2 * GS_BLOCKS + i]);
		ai += part[3 * GS_BLOCKS + i];
	}

	const float rs = (float)a;
	const float rsi = (float)ai;
	const float pAp = (float)c;
	const float alpha = (0.f != pAp) ? rs / pAp : 0.f;
	const float alpha_i = (0.f != pAp) ? rsi / pAp : 0.f;
	float s = sqrtf(alpha * alpha + alpha_i * alpha_i) * (float)sqrt(m);
	if (!(s > 0.f))
		s = 1.f;
	sc[0] = alpha;
	sc[1] = s;
	sc[2] = 1.f / s;
	sc[3] = alpha_i;
}

// ch[0][i] = re(v_i) f, ch[1][i] = im(v_i) f with v = alpha d (mode 0) or
// v = img (mode 1; f = 1 / scale given)
__global__ void k_guess_pack(const float2* __restrict__ v, const float* __restrict__ sc, float f, long npix,
			     int mode, float* __restrict__ ch)
{
	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
		float2 x = v[i];
		float g = f;
		if (0 == mode) {
			x = make_float2(sc[0] * x.x - sc[3] * x.y, sc[0] * x.y + sc[3] * x.x);
			g = sc[2];
		}
		ch[i] = x.x * g;
		ch[npix + i] = x.y * g;
	}
}

// x0 = s (ch[0] + i ch[1])
__global__ void k_guess_unpack(const float* __restrict__ ch, const float* __restrict__ sc, long npix,
			       float2* __restrict__ x0)
{
	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x)
		x0[i] = make_float2(sc[1] * ch[i], sc[1] * ch[npix + i]);
}

void guess_scalars(const float2* d, const float2* b, const float2* Ad, float lambda, long npix, double* part,
		   float* sc, hipStream_t st)
{
	hipLaunchKernelGGL(k_guess_partials, dim3(GS_BLOCKS), dim3(GS_THREADS), 0, st, (const float*)d,
			   (const float*)b, (const float*)Ad, lambda, npix, part);
	hipLaunchKernelGGL(k_guess_scalars, dim3(1), dim3(64), 0, st, part, sc);
	BAM_HIP(hipGetLastError());
}

void guess_pack(const float2* v, const float* sc, float f, long npix, bool scaled_first_iterate, float* ch,
		hipStream_t st)
{
	const unsigned blocks = (unsigned)std::min<long>((npix + 255) / 256, 2048);
	hipLaunchKernelGGL(k_guess_pack, dim3(blocks), dim3(256), 0, st, v, sc, f, npix, scaled_first_iterate ? 0 : 1,
			   ch);
	BAM_HIP(hipGetLastError());
}

void guess_unpack(const float* ch, const float* sc, long npix, float2* x0, hipStream_t st)
{
	const unsigned blocks = (unsigned)std::min<long>((npix + 255) / 256, 2048);
	hipLaunchKernelGGL(k_guess_unpack, dim3(blocks), dim3(256), 0, st, ch, sc, npix, x0);
	BAM_HIP(hipGetLastError());
}

}  // namespace bam
