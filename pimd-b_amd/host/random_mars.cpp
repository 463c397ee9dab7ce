// random_mars.cpp -- see random_mars.hpp.  The steps follow the published
// algorithm (James 1990, section 3.4): seed split into two lag-table seeds,
// each table entry built bit by bit from a 3-lag multiplicative generator mod
// 179 and a linear congruential generator mod 169, then
// u = (table[97-lag] - table[33-lag]) mod 1 minus a Weyl term mod 1.
#include "random_mars.hpp"

#include <cmath>
#include <stdexcept>

namespace pimdb_host {

RanMars::RanMars(int seed) {
    if (seed <= 0 || seed > 900000000) throw std::invalid_argument("Invalid seed for Marsaglia random # generator");
    const int ij = (seed - 1) / 30082;
    const int kl = (seed - 1) - 30082 * ij;
    // three-lag generator state (a, b, c) mod 179 and the congruential d mod 169
    int a = (ij / 177) % 177 + 2;
    int b = ij % 177 + 2;
    int c = (kl / 169) % 178 + 1;
    int d = kl % 169;
    for (int e = 1; e <= 97; ++e) {
        double value = 0.0, bit = 0.5;
        for (int n = 0; n < 24; ++n) {
            const int next = ((a * b) % 179) * c % 179;
            a = b;
            b = c;
            c = next;
            d = (53 * d + 1) % 169;
            if ((d * next) % 64 >= 32) value = value + bit;
            bit = 0.5 * bit;
        }
        lag_[e] = value;
    }
    weyl_ = 362436.0 / 16777216.0;
    weyl_step_ = 7654321.0 / 16777216.0;
    weyl_mod_ = 16777213.0 / 16777216.0;
    hi_ = 97;
    lo_ = 33;
    uniform();   // the reference discards the first draw
}

double RanMars::uniform() {
    double u = lag_[hi_] - lag_[lo_];
    if (u < 0.0) u += 1.0;
    lag_[hi_] = u;
    if (--hi_ == 0) hi_ = 97;
    if (--lo_ == 0) lo_ = 97;
    weyl_ -= weyl_step_;
    if (weyl_ < 0.0) weyl_ += weyl_mod_;
    u -= weyl_;
    if (u < 0.0) u += 1.0;
    return u;
}

// Polar (Marsaglia) method; returns the second of each pair first and keeps
// the first for the next call, as random_mars.cpp:83-103.
double RanMars::gaussian() {
    if (has_spare_) {
        has_spare_ = false;
        return spare_;
    }
    double v1, v2, r2;
    do {
        v1 = 2.0 * uniform() - 1.0;
        v2 = 2.0 * uniform() - 1.0;
        r2 = v1 * v1 + v2 * v2;
    } while (r2 >= 1.0 || r2 == 0.0);
    const double scale = std::sqrt(-2.0 * std::log(r2) / r2);
    spare_ = v1 * scale;
    has_spare_ = true;
    return v2 * scale;
}

}  // namespace pimdb_host
