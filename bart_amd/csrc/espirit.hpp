// espirit.hpp -- ESPIRiT coil-sensitivity calibration on gfx950 (espirit.hip).
#pragma once

#include "common.hpp"

#include <vector>

namespace bam {

// The options of struct ecalib_conf (calib/calib.h:12-34) this path uses.
struct EspiritConf {
	long kdims[3];
	float threshold;
	float crop;
	bool softcrop;
	int num_orthiter;
	bool rotphase;
};

// Calibration region of a Cartesian k-space [C][Z][Y][X] (host): the
// centred box that grows while fully sampled, up to calsize, then moved
// along the readout to maximal energy (misc/mri2.c:101-217).  caldims[3]
// (x, y, z) out; returns [C][cz][cy][cx].
std::vector<float2> extract_calib_host(const long calsize[3], const long dims[3], int C, const float2* ksp,
				       long caldims[3]);

// calib2 (calib/calib.c:715-798) without the options that change the
// algorithm (weighting, intensity, automatic thresholds, Nystroem): the
// Gram matrix of the calibration matrix and its eigenvectors, the image-space
// covariance of the kept kernels, its sinc interpolation to the image, the
// point-wise orthogonal iteration for `maps` maps, cropping, and the phase
// of the channel principal component.  cal: host [C][cz][cy][cx]; maps,
// emaps: device [maps][C][Z][Y][X] and [maps][Z][Y][X] (emaps may be null);
// svals: host, N = kx ky kz C singular values (may be null).  Returns the
// number of kernels kept.
int espirit_calib(const EspiritConf& conf, const long img[3], int C, int maps, const long caldims[3],
		  const float2* cal, float2* maps_out, float2* emaps, float* svals, hipStream_t st);

}  // namespace bam
