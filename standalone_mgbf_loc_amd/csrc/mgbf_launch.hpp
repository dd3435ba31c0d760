// Launch interface between the C ABI layer and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace mgbf {

constexpr int kMaxHaloDev = 10;
constexpr int kRWidthDev = 8;
constexpr int kPWidthDev = 4;

// w(|a|,|b|) of the radial filter as [b][a]
struct RadialCoef {
    float w[kMaxHaloDev + 1][kMaxHaloDev + 1];
};

struct RadialArgs {
    const float* in;
    float* out;
    const float* wfield;  // per-point scale, or nullptr
    float wconst;         // constant scale when wfield is nullptr
    int K, n, m;
    int H;                // max(hx, hy)
    int TJ;               // output rows per wavefront
    bool exact_disc;
    RadialCoef coef;
};

struct TransferArgs {
    const float* src;
    float* dst;
    int K, nf, mf;        // fine generation size
    const int* xs;        // window starts along x
    const float* xw;      // window weights along x
    const int* ys;
    const float* yw;
};

hipError_t launch_radial(const RadialArgs& a, hipStream_t st);
hipError_t launch_vertical(const float* in, float* out, const float* coef, int K, int lm,
                           int nfilt, int hz, long long npts, hipStream_t st);
hipError_t launch_restrict(const TransferArgs& a, hipStream_t st);
hipError_t launch_prolong(const TransferArgs& a, const float* addend, hipStream_t st);
hipError_t launch_scale(const float* in, float* out, const float* wfield, float wconst,
                        int use_sqrt, long long count, hipStream_t st);
hipError_t launch_copy(const float* in, float* out, long long count, hipStream_t st);
hipError_t launch_dot(const float* a, const float* b, long long count, double* acc,
                      hipStream_t st);

}  // namespace mgbf
