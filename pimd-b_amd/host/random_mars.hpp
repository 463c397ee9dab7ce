// random_mars.hpp -- RANMAR, the Marsaglia-Zaman-Tsang lagged-Fibonacci
// generator with a Weyl sequence (F. James, Comput. Phys. Commun. 60 (1990)
// 329), and the polar Gaussian on top of it: the per-bead noise streams of
// the reference's Langevin thermostat (libs/random_mars.cpp:11-103, one
// stream RanMars(seed + bead), src/simulation.cpp:59).
//
// The host side draws these streams in the reference's order and hands the
// deviates to pimdb_step, so a GPU run reproduces a reference run's noise.
#pragma once

#include <array>

namespace pimdb_host {

class RanMars {
public:
    // 0 < seed <= 900000000, else std::invalid_argument (random_mars.cpp:16-17).
    explicit RanMars(int seed);
    double uniform();
    double gaussian();

private:
    std::array<double, 98> lag_{};   // entries 1..97
    int hi_ = 97, lo_ = 33;          // the two lag positions
    double weyl_ = 0.0, weyl_step_ = 0.0, weyl_mod_ = 0.0;
    bool has_spare_ = false;
    double spare_ = 0.0;
};

}  // namespace pimdb_host
