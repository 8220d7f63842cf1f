// nudft.hpp -- direct non-uniform DFT operator (noncart/nudft.c).
#pragma once

#include "common.hpp"

namespace bam {

struct NudftGeom {
	int n[3];
	double h[3];   // n_d / 2
	long N;        // pixels per coil
	int C;
	long M;
	float scale;   // 1 / sqrt(N C)
};

struct Nudft {

	NudftGeom g = {};
	DevBuf<double4> grad;  // per sample: -2 pi k_d / n_d
	hipStream_t stream = 0;

	// dims: image dims; traj: host M x 3 (real parts)
	void create(const int dims[3], int ncoils, long nsamples, const float* traj);
	void forward(float2* ksp, const float2* img);
	void adjoint(float2* img, const float2* ksp);
};

}  // namespace bam
