// Test-only C entry points over the host-side line decompositions of the
// product (mgbf::line_triad / mgbf::line_hexad in mgbf_coef.cpp), built by
// the CPU tests with g++ together with mgbf_coef.cpp.
#include "mgbf_coef.hpp"

extern "C" int probe_triad(const double* a, int* l, double* w) {
    int ll[3][2];
    if (!mgbf::line_triad(a, ll, w)) return 0;
    for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 2; ++e) l[2 * c + e] = ll[c][e];
    return 1;
}

extern "C" int probe_hexad(const double* a, int* l, double* w) {
    int ll[7][3];
    if (!mgbf::line_hexad(a, ll, w)) return 0;
    for (int c = 0; c < 7; ++c)
        for (int e = 0; e < 3; ++e) l[3 * c + e] = ll[c][e];
    return 1;
}
