// host_sampling.h — the sequential, RNG-driven steps of clustering stay on the
// host (they are a few hundred draws).  MT19937 with the 1998 Knuth-69069
// seeding reproduces the stream Renzoku's Random draws from (random.h:17-71,
// mtwist.c:378-406: reversed state, identical outputs), and DiscretePdf keeps
// its sequential normalisation and strict-< CDF search (pdf.cc:18-98).
#pragma once
#pragma clang fp contract(off)

#include <cmath>
#include <cstdint>
#include <vector>

namespace rzk {

class Mt19937 {
public:
    explicit Mt19937(uint32_t seed = 2013) { seed32(seed); }
    void seed32(uint32_t seed) {
        if (seed == 0) seed = 4357;
        x_[0] = seed;
        for (int i = 1; i < kN; ++i) x_[i] = 69069u * x_[i - 1];
        i_ = kN;
    }
    uint32_t next_u32() {
        if (i_ >= kN) twist();
        uint32_t y = x_[i_++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
    // Random::operator(): 24-bit float in [0, 1)
    float next_float() { return (float)(next_u32() & 0x00ffffffu) * (1.0f / 16777216.0f); }

private:
    static constexpr int kN = 624, kM = 397;
    void twist() {
        for (int k = 0; k < kN; ++k) {
            uint32_t y = (x_[k] & 0x80000000u) | (x_[(k + 1) % kN] & 0x7fffffffu);
            x_[k] = x_[(k + kM) % kN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        i_ = 0;
    }
    uint32_t x_[kN];
    int i_ = kN;
};

// DiscretePdf with an O(log n) search that returns exactly what the linear
// scan "first i with u < cdf[i]" returns: that index is also the first one
// where the running maximum of the cdf exceeds u, and the running maximum is
// monotone, so a binary search on it finds the same index even when rounding
// or negative weights make the cdf itself non-monotone.
class DiscretePdf {
public:
    bool build(const float* arr, size_t n) {
        // every entry is written below whenever the pdf is valid
        pdf_.resize(n);
        cdf_.resize(n);
        runmax_.resize(n);
        valid_ = false;
        if (n == 0) return false;
        float sum = 0.f;
        for (size_t i = 0; i < n; ++i) sum += arr[i];
        if (sum == 0.0f) return false;
        for (size_t i = 0; i < n; ++i) pdf_[i] = arr[i] / sum;
        // cdf[i] = cdf[i - 1] + pdf[i] (the last one then forced to 1) and its
        // running maximum, in one pass
        float acc = 0.f, m = -INFINITY;
        for (size_t i = 0; i < n; ++i) {
            acc = i == 0 ? pdf_[0] : acc + pdf_[i];
            const float ci = i + 1 == n ? 1.0f : acc;
            cdf_[i] = ci;
            if (ci > m) m = ci;              // NaN entries never satisfy u < cdf
            runmax_[i] = m;
        }
        valid_ = true;
        return true;
    }
    // returns index or -1; pdf_k = 0 when -1
    int sample(float u, float& pdf_k) const {
        size_t n = cdf_.size();
        if (!valid_ || n == 0) { pdf_k = 0.f; return -1; }
        // first i with u < runmax[i]
        if (!(u < runmax_[n - 1])) { pdf_k = 0.f; return -1; }
        size_t lo = 0, hi = n - 1;
        while (lo < hi) {
            size_t mid = lo + (hi - lo) / 2;
            if (u < runmax_[mid]) hi = mid; else lo = mid + 1;
        }
        pdf_k = pdf_[lo];
        return (int)lo;
    }
    const std::vector<float>& pdf() const { return pdf_; }
    const std::vector<float>& cdf() const { return cdf_; }
    bool valid() const { return valid_; }

private:
    std::vector<float> pdf_, cdf_, runmax_;
    bool valid_ = false;
};

// DiscretePdf over w[idx[0]], ..., w[idx[n - 1]] built and sampled once with u:
// the same sums, quotients and strict-< search, walked only up to the index
// sampled.  Returns it (pdf_k its probability) or -1 (pdf_k = 0) when the
// weights sum to zero or there are none.
inline int sample_members(const float* w, const int* idx, int n, float u, float& pdf_k) {
    pdf_k = 0.f;
    if (n <= 0) return -1;
    float sum = 0.f;
    for (int i = 0; i < n; ++i) sum += w[idx[i]];
    if (sum == 0.0f) return -1;
    float acc = 0.f, m = -INFINITY;
    for (int i = 0; i < n; ++i) {
        const float p = w[idx[i]] / sum;
        acc = i == 0 ? p : acc + p;
        const float ci = i + 1 == n ? 1.0f : acc;
        if (ci > m) m = ci;
        if (u < m) { pdf_k = p; return i; }
    }
    return -1;
}

}  // namespace rzk
