// vj_pair_plan.hpp -- capacity rule of the device grouping's pair buffer
// (host code, no HIP types, so the CPU tests compile it on its own).
#ifndef VJ_PAIR_PLAN_HPP
#define VJ_PAIR_PLAN_HPP

#include <algorithm>
#include <cstdint>

namespace vj {

// The largest pair buffer the engine keeps (entries of 4 bytes: 4 GiB).
constexpr int64_t kPairCapMax = int64_t{1} << 30;

// After a detect whose images hold nPairs similar pairs in all, with a
// buffer of cap entries: the device grouping result is usable when every
// pair was stored (nPairs <= cap); otherwise the call is grouped on the host
// and the buffer grows to nPairs plus a quarter plus 64K entries, but never
// past kPairCapMax (beyond it, such calls stay on the host).  Returns whether
// the device result is usable; *nextCap is the capacity for the next call.
inline bool pair_plan(int64_t nPairs, int64_t cap, int64_t* nextCap) {
    if (nPairs >= 0 && nPairs <= cap) {
        *nextCap = cap;
        return true;
    }
    const int64_t want = nPairs + nPairs / 4 + 65536;
    *nextCap = std::max(cap, std::min(want, kPairCapMax));
    return false;
}

}  // namespace vj

#endif
