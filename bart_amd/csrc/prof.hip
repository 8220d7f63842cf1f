// prof.hip -- per-kernel HIP-event timing (see prof.hpp) and its C ABI.
#include "../../include/bart_amd.h"
#include "prof.hpp"

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace bam {

struct ProfRec {
	const char* name;
	hipEvent_t e0, e1;
};

static std::atomic<bool> prof_on{ false };
static std::mutex prof_mtx;
static std::vector<ProfRec> prof_recs;
static std::vector<hipEvent_t> prof_pool;

static thread_local bool capturing = false;

bool prof_enabled() { return prof_on.load(std::memory_order_relaxed); }
void prof_set_capturing(bool on) { capturing = on; }
bool prof_capturing() { return capturing; }

hipEvent_t prof_event()
{
	std::lock_guard<std::mutex> lock(prof_mtx);

	if (!prof_pool.empty()) {
		hipEvent_t e = prof_pool.back();
		prof_pool.pop_back();
		return e;
	}

	hipEvent_t e;
	BAM_HIP(hipEventCreate(&e));
	return e;
}

void prof_record(const char* name, hipEvent_t e0, hipEvent_t e1)
{
	std::lock_guard<std::mutex> lock(prof_mtx);
	prof_recs.push_back({ name, e0, e1 });
}

}  // namespace bam

using namespace bam;

extern "C" {

void bart_amd_profile(int enable)
{
	prof_on.store(0 != enable);
}

// Writes {"name": [count, total_ms, min_ms, max_ms], ...}.  Returns the
// length needed (including the terminating zero); the records are cleared
// only when the report fits into buf.
int bart_amd_profile_report(char* buf, int len)
{
	std::vector<ProfRec> recs;
	{
		std::lock_guard<std::mutex> lock(prof_mtx);
		recs = prof_recs;
	}

	struct Acc { long n = 0; double tot = 0., mn = 1e30, mx = 0.; };
	std::map<std::string, Acc> acc;
	long failed = 0;
	int last_err = 0;

	for (auto& r : recs) {

		float ms = 0.f;
		hipError_t e = hipEventSynchronize(r.e1);

		if (hipSuccess == e)
			e = hipEventElapsedTime(&ms, r.e0, r.e1);

		if (hipSuccess != e) {
			failed++;
			last_err = (int)e;
			(void)hipGetLastError();
		} else {
			Acc& a = acc[r.name];
			a.n++;
			a.tot += ms;
			a.mn = std::min(a.mn, (double)ms);
			a.mx = std::max(a.mx, (double)ms);
		}
	}

	std::string s = "{";
	bool first = true;

	if (failed > 0) {
		char tmp[128];
		snprintf(tmp, sizeof(tmp), "\"_failed\": [%ld, %d, 0, 0]", failed, last_err);
		s += tmp;
		first = false;
	}

	for (auto& kv : acc) {
		char tmp[256];
		snprintf(tmp, sizeof(tmp), "%s\"%s\": [%ld, %.6f, %.6f, %.6f]", first ? "" : ", ", kv.first.c_str(),
			 kv.second.n, kv.second.tot, kv.second.mn, kv.second.mx);
		s += tmp;
		first = false;
	}

	s += "}";

	if (nullptr != buf && len > (int)s.size()) {

		snprintf(buf, len, "%s", s.c_str());

		std::lock_guard<std::mutex> lock(prof_mtx);
		// drop the reported records (new ones may have been appended meanwhile)
		prof_recs.erase(prof_recs.begin(), prof_recs.begin() + recs.size());
		for (auto& r : recs) {
			prof_pool.push_back(r.e0);
			prof_pool.push_back(r.e1);
		}
	}

	return (int)s.size() + 1;
}

}  // extern "C"
