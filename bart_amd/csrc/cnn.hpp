// cnn.hpp -- 3x3 convolution layers of the CG warm-start network (cnn.hip).
#pragma once

#include "common.hpp"

namespace bam {

// y = [relu](conv3x3(x, w) + bias) [+ res]; device pointers, [B][C][H][W]
void conv3x3_fwd(const float* x, int B, int Cin, int H, int W, const float* w, const float* bias, const float* res,
		 int Cout, bool relu, float* y, hipStream_t st);
// dx = data gradient of conv3x3 for the output gradient dy
void conv3x3_bwd_data(const float* dy, int B, int Cout, int H, int W, const float* w, int Cin, float* dx,
		      hipStream_t st);
// dw, db = weight and bias gradients (db may be null)
void conv3x3_bwd_weights(const float* x, const float* dy, int B, int Cin, int Cout, int H, int W, float* dw,
			 float* db, hipStream_t st);

// Input and guess of the warm-start network.  guess_scalars: sc = { re
// alpha, s, 1 / s, im alpha } of the step along a direction d for the
// right-hand side b, x1 = alpha d with alpha = d^H b / d^H (A^H A d +
// lambda d) (d = b: CG's first iterate from zero, alpha as CG computes it),
// given Ad = A^H A d, and its scale s = max |x1|; part: 4 x 256 doubles.
// guess_pack: channels [2][npix] (real, imaginary) of x1 / s
// (scaled_first_iterate, v = d) or of v / scale (f = 1 / scale).
// guess_unpack: x0 = s (ch[0] + i ch[1]).
constexpr int GUESS_PARTS = 4 * 256;
void guess_scalars(const float2* d, const float2* b, const float2* Ad, float lambda, long npix, double* part,
		   float* sc, hipStream_t st);
void guess_pack(const float2* v, const float* sc, float f, long npix, bool scaled_first_iterate, float* ch,
		hipStream_t st);
void guess_unpack(const float* ch, const float* sc, long npix, float2* x0, hipStream_t st);

// The warm-start network (2 -> C x 5 -> 2 channels, ReLU, residual), its
// parameters packed [W0 b0 W1 b1 ...] in one buffer, Adam state beside them.
struct WarmNet {

	int ch[7];
	long woff[6], boff[6];
	long nparams = 0;
	int step = 0;
	DevBuf<float> params, grads, m, v;
	DevBuf<float> act, g0, g1, out;  // layer outputs 1..5, gradient ping-pong, output
	DevBuf<double> lpart;

	WarmNet(int channels, unsigned long seed);
	void ensure(int B, int H, int W);
	const float* a(int l, const float* x, int B, int H, int W) const;
	// y = x + net(x); x, y [B][2][H][W] device
	void forward(const float* x, int B, int H, int W, float* y, hipStream_t st);
	// one Adam step on mean((x + net(x) - t)^2); returns the loss before the step
	double train_step(const float* x, const float* t, int B, int H, int W, float lr, hipStream_t st);
	// one Adam step for a loss whose gradient with respect to the output of
	// the last forward(x) is dout (the layer outputs of that forward are used)
	void train_step_grad(const float* x, const float* dout, int B, int H, int W, float lr, hipStream_t st);
	void backward_update(const float* x, int B, int H, int W, float lr, hipStream_t st);
};

}  // namespace bam
