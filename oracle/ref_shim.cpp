// ref_shim.cpp -- C entry points over the reference's own RANMAR and unit
// table, compiled together with libs/random_mars.cpp and src/units.cpp of
// higj/pimd-b where they lie (see Makefile, target "ref").  TEST
// INFRASTRUCTURE ONLY: used by tests/ to check the oracle's restatement of
// those two files.  The rest of the reference path includes <format>, which
// this toolchain lacks, so it is not built.
#include <string>
#include "random_mars.h"
#include "units.h"

extern "C" {
void* ref_ranmars_new(int seed) { return new RanMars(seed); }
void ref_ranmars_free(void* r) { delete static_cast<RanMars*>(r); }
double ref_ranmars_uniform(void* r) { return static_cast<RanMars*>(r)->uniform(); }
double ref_ranmars_gaussian(void* r) { return static_cast<RanMars*>(r)->gaussian(); }
double ref_units_to_internal(const char* family, const char* unit, double v) {
    return Units::convertToInternal(family, unit, v);
}
}
