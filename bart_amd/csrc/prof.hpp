// prof.hpp -- optional per-kernel timing with HIP events.
//
// When enabled (bart_amd_profile(1)), every kernel launch of the library is
// bracketed by two events on the stream it is launched on; the report sums
// the elapsed times per kernel class.  Launches inside stream capture are not
// recorded.  Disabled (the default) it costs one branch per launch.
#pragma once

#include "common.hpp"

namespace bam {

bool prof_enabled();
// set while a stream capture is in progress on this thread (launches inside a
// captured graph are not timed)
void prof_set_capturing(bool on);
bool prof_capturing();
void prof_record(const char* name, hipEvent_t e0, hipEvent_t e1);
hipEvent_t prof_event();

struct ProfScope {

	const char* name;
	hipStream_t st;
	hipEvent_t e0 = nullptr;

	ProfScope(const char* n, hipStream_t s) : name(n), st(s)
	{
		if (!prof_enabled() || prof_capturing())
			return;

		e0 = prof_event();
		(void)hipEventRecord(e0, st);
	}

	~ProfScope()
	{
		if (nullptr == e0)
			return;

		hipEvent_t e1 = prof_event();
		(void)hipEventRecord(e1, st);
		prof_record(name, e0, e1);
	}
};

}  // namespace bam
