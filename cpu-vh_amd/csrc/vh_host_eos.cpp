#include "vh_host_eos.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

namespace vh {

namespace {
constexpr double kHbarc = 0.197326938;
constexpr double kEosFactor = 13.8997;
constexpr int kLen = 100000;
}  // namespace

void HostEos::load(const double *raw) {
    kind = 1;
    p_tb.assign(kLen, 0.0);
    T_tb.assign(kLen, 0.0);
    for (int ii = 0; ii < kLen; ++ii) {
        const double ev = raw[4 * ii] / kHbarc;
        if (ii == 0) e_min = ev;
        if (ii == 1) e_spacing = ev - e_min;
        if (ii == kLen - 1) e_max = ev;
        p_tb[ii] = raw[4 * ii + 1] / kHbarc;
        T_tb[ii] = raw[4 * ii + 3] / kHbarc;
    }
}

double HostEos::interp(double e, const std::vector<double> &tb) const {
    int idx = static_cast<int>((e - e_min) / e_spacing);
    idx = std::min(kLen - 2, idx);
    idx = std::max(0, idx);
    const double frac = (e - (idx * e_spacing + e_min)) / e_spacing;
    return tb[idx] * (1. - frac) + tb[idx + 1] * frac;
}

double HostEos::pressure(double e) const {
    if (kind == 1) return std::max(1.e-15, interp(e, p_tb));
    return e / 3.;
}

double HostEos::cs2(double e) const {
    if (kind == 1) {
        const double p = pressure(e), p_r = pressure(e + e_spacing);
        double v = e_spacing / (p_r - p);
        return std::max(0.01, std::min(1. / 3., v));
    }
    return 1. / 3.;
}

double HostEos::temperature(double e) const {
    if (kind == 1) return std::max(1.e-15, interp(e, T_tb));
    return powf((float)(e / kEosFactor), 0.25f);
}

double HostEos::energy_density(double T) const {
    if (kind == 1) {
        double lo = 1e-15, hi = e_max, mid = (hi + lo) / 2.;
        if (T < temperature(lo)) return lo;
        int iter = 0;
        while (((hi - lo) / mid > 1e-8 && (hi - lo) > 1e-15) && iter < 1000) {
            if (T < temperature(mid)) hi = mid;
            else lo = mid;
            mid = (hi + lo) / 2.;
            ++iter;
        }
        return mid;
    }
    return kEosFactor * powf((float)T, 4.0f);
}

bool HostEos::cs2_is_one_third() const {
    if (kind != 1) return true;
    for (int i = 0; i + 1 < kLen; ++i) {
        const double d = p_tb[i + 1] - p_tb[i];
        if (!(d > 1e-9 * std::fabs(p_tb[i + 1])) || !(d < 3 * e_spacing)) return false;
    }
    return true;
}

const QuarterRootExceptions &quarter_root_exceptions() {
    static QuarterRootExceptions X;
    static std::once_flag once;
    std::call_once(once, [] {
        const float flo = std::ldexp(1.0f, -32), fhi = std::ldexp(1.0f, 16);
        std::memcpy(&X.lo, &flo, 4);
        std::memcpy(&X.hi, &fhi, 4);
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const uint32_t span = X.hi - X.lo;
        std::vector<std::vector<uint32_t>> keys(nt), vals(nt);
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nt; ++w)
            th.emplace_back([&, w] {
                const uint32_t a = X.lo + (uint32_t)((uint64_t)span * w / nt);
                const uint32_t b = X.lo + (uint32_t)((uint64_t)span * (w + 1) / nt);
                for (uint32_t u = a; u < b; ++u) {
                    float x;
                    std::memcpy(&x, &u, 4);
                    const float g = powf(x, 0.25f);
                    const float c = (float)std::sqrt(std::sqrt((double)x));
                    uint32_t gb, cb;
                    std::memcpy(&gb, &g, 4);
                    std::memcpy(&cb, &c, 4);
                    if (gb != cb) {
                        keys[w].push_back(u);
                        vals[w].push_back(gb);
                    }
                }
            });
        for (auto &t : th) t.join();
        for (unsigned w = 0; w < nt; ++w) {
            X.key.insert(X.key.end(), keys[w].begin(), keys[w].end());
            X.val.insert(X.val.end(), vals[w].begin(), vals[w].end());
        }
        const uint32_t nb = (span >> X.shift) + 1;
        X.bucket.assign(nb + 1, 0);
        size_t pos = 0;
        for (uint32_t b = 0; b <= nb; ++b) {
            const uint64_t start = (uint64_t)X.lo + ((uint64_t)b << X.shift);
            while (pos < X.key.size() && X.key[pos] < start) ++pos;
            X.bucket[b] = (uint32_t)pos;
        }
    });
    return X;
}

}  // namespace vh
