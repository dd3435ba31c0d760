// One translation unit per radial half-width: compiled once for each
// MGBF_RADIAL_H in 1..10 so the variants build in parallel.
#include "mgbf_radial.hpp"
#include "mgbf_radial_mx.hpp"

#ifndef MGBF_RADIAL_H
#error "MGBF_RADIAL_H must be defined"
#endif

#define MGBF_CAT2(a, b) a##b
#define MGBF_CAT(a, b) MGBF_CAT2(a, b)

namespace mgbf {
hipError_t MGBF_CAT(launch_radial_H, MGBF_RADIAL_H)(const RadialArgs& a, hipStream_t st) {
    if (a.mfma) return launch_radial_mx_h<MGBF_RADIAL_H>(a, st);
    return launch_radial_h<MGBF_RADIAL_H>(a, st);
}
}  // namespace mgbf
