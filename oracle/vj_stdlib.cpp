// vj_stdlib.cpp -- TEST INFRASTRUCTURE: the C++ standard-library steps of
// the reference's training loop, for the oracle's cascade driver
// (oracle/pyoracle.py train_cascade).  The reference draws its stage
// positives, its mining-image order and its top-up picks with std::mt19937
// through std::shuffle and std::uniform_int_distribution, and keeps each
// image's best boxes with std::partial_sort (src/main.cu:1589-1591,
// 1848-1849, 1866-1877, 1915-1918).  What those calls return is the
// standard library's behaviour, so the oracle calls the same algorithms on
// the same inputs; only tests load this library.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <random>
#include <vector>

extern "C" {

// The permutation std::shuffle applies to n elements with std::mt19937(seed):
// out[i] is the original position of the element that ends at position i.
void or_std_shuffle(int n, uint32_t seed, int32_t* out) {
    std::vector<int32_t> v(static_cast<size_t>(n));
    std::iota(v.begin(), v.end(), 0);
    std::mt19937 rng(seed);
    std::shuffle(v.begin(), v.end(), rng);
    std::copy(v.begin(), v.end(), out);
}

// count successive draws of std::uniform_int_distribution<int>(lo, hi) from
// one std::mt19937(seed).
void or_std_uniform_ints(int count, int lo, int hi, uint32_t seed, int32_t* out) {
    std::mt19937 rng(seed);
    std::uniform_int_distribution<int> pick(lo, hi);
    for (int i = 0; i < count; ++i) out[i] = pick(rng);
}

// std::partial_sort of n records by score, descending, for the first keep
// positions: order[i] is the original position of the record at position i
// (the same comparisons and moves as sorting the records themselves).
void or_std_partial_sort_desc(const float* score, int n, int keep, int32_t* order) {
    std::vector<int32_t> v(static_cast<size_t>(n));
    std::iota(v.begin(), v.end(), 0);
    std::partial_sort(v.begin(), v.begin() + keep, v.end(),
                      [score](int32_t a, int32_t b) { return score[a] > score[b]; });
    std::copy(v.begin(), v.end(), order);
}

}  // extern "C"
