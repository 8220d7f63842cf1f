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

}  // namespace bam
