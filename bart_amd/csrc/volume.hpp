// volume.hpp -- index geometry of a stack of 3-d volumes (BART column-major
// layout: x fastest, then y, z, then the stack index), viewed as the set of
// 1-d transforms along one axis.
#pragma once

#include "common.hpp"

namespace bam {

// Division by a run-time constant via multiply-high (round-up method).
struct FastDiv {

	uint32_t d = 1, m = 1, s = 0;

	FastDiv() = default;

	explicit FastDiv(uint32_t div) : d(div)
	{
		s = 0;
		while ((1ull << s) < div)
			s++;
		m = (uint32_t)(((((1ull << s) - div) << 32) / div) + 1);
	}

	__host__ __device__ inline uint32_t div(uint32_t n) const
	{
#ifdef __HIP_DEVICE_COMPILE__
		uint32_t t = __umulhi(n, m);
#else
		uint32_t t = (uint32_t)(((uint64_t)n * m) >> 32);
#endif
		return (uint32_t)(((uint64_t)t + n) >> s);
	}

	__host__ __device__ inline void divmod(uint32_t n, uint32_t& q, uint32_t& r) const
	{
		q = div(n);
		r = n - q * d;
	}
};

// Transforms along `axis` of `nvol` volumes of dims n[3].
// Transform t decomposes as t = i + inner * (o + outer * v), with i the
// index over the dims below the axis, o over the dims above, v the volume.
struct VolAxis {

	int n[3];
	long N;
	int axis;
	uint32_t inner, outer;
	long nvol;
	FastDiv f_inner, f_outer, f_n0, f_n1;

	// volume-interleaved order (set by group_volumes): consecutive groups of
	// vgrp transforms run through the volumes before the next columns,
	// t = il + vgrp (v + nvol ih) with i = il + vgrp ih (outer == 1 only)
	uint32_t vgrp = 0;
	FastDiv f_vgrp, f_nvol;

	static VolAxis make(const int dims[3], int axis, long nvol)
	{
		VolAxis g;

		for (int i = 0; i < 3; i++)
			g.n[i] = dims[i];

		g.N = (long)dims[0] * dims[1] * dims[2];
		g.axis = axis;
		long in = 1, out = 1;

		for (int d = 0; d < axis; d++)
			in *= dims[d];
		for (int d = axis + 1; d < 3; d++)
			out *= dims[d];

		g.inner = (uint32_t)in;
		g.outer = (uint32_t)out;
		g.nvol = nvol;
		g.f_inner = FastDiv((uint32_t)in);
		g.f_outer = FastDiv((uint32_t)out);
		g.f_n0 = FastDiv((uint32_t)dims[0]);
		g.f_n1 = FastDiv((uint32_t)dims[1]);

		BAM_CHECK(in * out * nvol < (1l << 32), "too many transforms (%ld)", in * out * nvol);
		return g;
	}

	long transforms() const { return (long)inner * outer * nvol; }
	bool contig() const { return 1 == inner; }

	// Launch blocks of P transforms then take the same P columns of every
	// volume one after the other (data shared by the volumes, such as a
	// PSF, is reused from cache).  Returns false (order unchanged) when the
	// geometry does not allow it.
	bool group_volumes(int P)
	{
		if (1 != outer || P <= 1 || 0 != inner % (uint32_t)P)
			return false;
		vgrp = (uint32_t)P;
		f_vgrp = FastDiv((uint32_t)P);
		f_nvol = FastDiv((uint32_t)nvol);
		return true;
	}

	// (i, o, v) of transform t
	__host__ __device__ inline void split(long t, uint32_t& i, uint32_t& o, uint32_t& v) const
	{
		if (0 != vgrp) {
			uint32_t r, il, ih;
			f_vgrp.divmod((uint32_t)t, r, il);
			f_nvol.divmod(r, ih, v);
			i = il + vgrp * ih;
			o = 0;
			return;
		}

		uint32_t rest;
		f_inner.divmod((uint32_t)t, rest, i);
		f_outer.divmod(rest, v, o);
	}

	__host__ __device__ inline long offset_iov(uint32_t i, uint32_t o, uint32_t v, int k) const
	{
		return (long)v * N + i + (long)inner * (k + (long)n[axis] * o);
	}

	__host__ __device__ inline long offset(long t, int k) const
	{
		uint32_t i, o, v;
		split(t, i, o, v);
		return offset_iov(i, o, v, k);
	}

	// pixel coordinates (x, y, z) of element k of transform (i, o)
	__host__ __device__ inline void coords(uint32_t i, uint32_t o, int k, int c[3]) const
	{
		uint32_t q, r;

		if (0 == axis) {
			c[0] = k;
			f_n1.divmod(o, q, r);
			c[1] = (int)r;
			c[2] = (int)q;
		} else if (1 == axis) {
			c[0] = (int)i;
			c[1] = k;
			c[2] = (int)o;
		} else {
			f_n0.divmod(i, q, r);
			c[0] = (int)r;
			c[1] = (int)q;
			c[2] = k;
		}
	}
};

}  // namespace bam
