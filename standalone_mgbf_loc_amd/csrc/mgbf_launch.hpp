// Launch interface between the C ABI layer and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

struct mgbf_ctx;

namespace mgbf {

// Static device scratch of mgbf_apply_noalloc (mgbf_noalloc.hip): 32 Mi
// floats (128 MiB) per device, from the load of the code object to the end
// of the process.
constexpr size_t kNoallocScratchFloats = size_t(32) << 20;
// Looks up the scratch's device address on `device` (which must be current)
// once per device; MGBF_OK or an error status with a message in *err.
int noalloc_setup(int device, std::string* err);
// The scratch's device address on `device`, or nullptr before its setup.
float* noalloc_scratch(int device);
// mgbf_apply with the context's workspaces replaced, for this one call, by
// slices of the `nfloats` floats at scratch_of(context's device)
// (mgbf_capi.cpp).
int apply_in_scratch(mgbf_ctx* c, int op, const float* in, float* out, void* stream,
                     float* (*scratch_of)(int device), size_t nfloats);

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
    const float* coarse;  // next coarser generation to prolong and add, or nullptr
    const float* band;    // vertical band (lm x (2 vhb + 1)) applied at the end, or nullptr
    int vhb;              // its half-width
    int lm, nfilt;        // column length and vertically filtered part of the stack
    int K, n, m;
    int klo;              // first k filtered (k < klo is left alone): 0, or K3 for
                          // the 2-D variables after the 3-D ones of procedures 4, 5
    int H;                // max(hx, hy)
    int TI;               // output columns per lane (2 or 4)
    // matrix-core engine (mgbf_radial_mx.hpp)
    bool mfma;
    const void* bfrag;      // FP16 hi/lo B fragments of the Toeplitz weight bands T_b
    const void* rfrag;      // FP16 B fragments of the x restriction (3 tile classes), or nullptr
    int wexp;               // power-of-two exponent applied to the weights
    int TJ;               // output rows per wavefront
    // restriction to the next generation fused into the pass (TI = 4, H >= 3)
    float* rout;            // coarse output, or nullptr
    const float4* rytab;    // per fine row: weights into the 6 open coarse rows (2 x float4)
    const int2* rymeta;     // per fine row: lowest open coarse row, rows completed
    bool exact_disc;
    bool separable;       // coef.w[0][a] = fx(a), coef.w[1][b] = fy(b): w(a,b) = fx(a) fy(b)
    bool quad;            // the prolongation epilogue uses the quadratic transfer (lquart)
    RadialCoef coef;
};

// kernel argument record of the radial kernel
struct RadialParams {
    const float* in;
    float* out;
    const float* wfield;
    const float* coarse;
    const float* band;
    float wconst;
    int K, n, m, TJ, lm, nfilt;
    int klo;                  // first k of the pass
    int ntile_i, ntile_j, nitems;
    // restriction to the next generation fused into the pass (RES)
    float* rout;              // coarse output (n/2 x m/2 x K), or nullptr
    const float4* rytab;      // per fine row: weights into the 6 open coarse rows (2 x float4)
    const int2* rymeta;       // per fine row: lowest open coarse row, rows completed
    RadialCoef W;
};

// the matrix-core engine takes the vertical band epilogue below this
// half-width (at 8 the band's registers would spill beside the ring)
constexpr int kMxMaxBandH = 8;

// kernel argument record of the matrix-core radial kernel (constant
// weights only: a per-point weight field follows as a scale pass)
struct RadialMxParams {
    const float* in;
    float* out;
    const float* coarse;
    const float* band;
    const void* bfrag;
    const void* rfrag;
    float* rout;
    const float4* rytab;
    const int2* rymeta;
    float wconst;
    int wexp;
    int K, n, m, TJ, lm, nfilt;
    int ntile_i, ntile_j;
};

#define MGBF_DECLARE_RADIAL(h) hipError_t launch_radial_H##h(const RadialArgs& a, hipStream_t st);
MGBF_DECLARE_RADIAL(1) MGBF_DECLARE_RADIAL(2) MGBF_DECLARE_RADIAL(3) MGBF_DECLARE_RADIAL(4)
MGBF_DECLARE_RADIAL(5) MGBF_DECLARE_RADIAL(6) MGBF_DECLARE_RADIAL(7) MGBF_DECLARE_RADIAL(8)
MGBF_DECLARE_RADIAL(9) MGBF_DECLARE_RADIAL(10)
#undef MGBF_DECLARE_RADIAL

struct TransferArgs {
    const float* src;
    float* dst;
    int K, nf, mf;        // fine generation size
    const int* xs;        // window starts along x
    const float* xw;      // window weights along x
    const int* ys;
    const float* yw;
    bool quad;            // quadratic stencil (direct2 / adjoint2) instead of the 4-point one
};

constexpr int kMaxVBand = 20;   // half-band of Vf Vf^T for hz <= 10

struct VBandW {
    float w[2 * kMaxVBand + 1];
};

struct VBandArgs {
    const float* in;
    float* out;
    const float* band;   // lm x (2HB+1) rows
    VBandW wint;         // interior row
    int ilo, ihi;        // rows using wint
    int HB;
    int K, lm, nfilt;
    long long npts;
};

// The radial filter of a generation with a per-point aspect tensor field:
// w holds, per point p = i + n j, the weights of the (2H+1)^2 offsets
// |gx|, |gy| <= H (t = (gy+H)(2H+1) + gx+H), zero off the point's ellipse.
struct AspectArgs {
    const float* in;
    float* out;
    const float* w;
    int K, n, m, H;
};
hipError_t launch_aspect(const AspectArgs& a, bool transpose, hipStream_t st);

hipError_t launch_radial(const RadialArgs& a, hipStream_t st);
hipError_t launch_vband(const VBandArgs& a, hipStream_t st);
hipError_t launch_restrict(const TransferArgs& a, hipStream_t st);
hipError_t launch_prolong(const TransferArgs& a, const float* addend, hipStream_t st);
hipError_t launch_scale(const float* in, float* out, const float* wfield, float wconst,
                        int use_sqrt, long long count, hipStream_t st);
hipError_t launch_copy(const float* in, float* out, long long count, hipStream_t st);
hipError_t launch_fill(float* p, float v, long long count, hipStream_t st);

// Separable window transfer between two grids of K-long columns:
// out(k,i,j) = sum_ty yw[j][ty] sum_tx xw[i][tx] in(k, xs[i]+tx, ys[j]+ty)
struct XferArgs {
    const float* in;
    float* out;
    int K;
    int nin_x, nin_y, nout_x, nout_y;
    const int* xs;
    const float* xw;
    int wx;
    const int* ys;
    const float* yw;
    int wy;
};
hipError_t launch_xfer2d(const XferArgs& a, hipStream_t st);
// The same separable transfer by output tiles of 8 x 8 points: tile (p, q)
// reads the input block x0[p] .. x0[p]+SX-1 by y0[q] .. y0[q]+SY-1 and
// applies the dense tile weights wx[p][t][c] (8 x SX) and wy[q][t][r]
// (8 x SY), zero outside each output's window. SX in {8, 12, 16, 20, 24, 32}.
struct XferTileArgs {
    const float* in;
    float* out;
    int K;
    int nin_x, nin_y, nout_x, nout_y;
    const int* x0;
    const float* wx;
    int SX;
    const int* y0;
    const float* wy;
    int SY;
};
hipError_t launch_xfer_tile(const XferTileArgs& a, hipStream_t st);
// The vertical map of a transfer, fused into the tile kernel, in windows of
// at most 4 consecutive levels per 64-level chunk of the filter stack (the
// wave's lanes). Direct (filter -> analysis, after the horizontal pass):
// chunk c writes the analysis levels lo[c] .. lo[c]+cnt[c]-1 (cnt <= 128),
// output level o = sum_u w[4 o + u] h(64 c + r0[o] + u), h the chunk's
// horizontally interpolated levels. Adjoint (analysis -> filter, before the
// horizontal pass): chunk c reads the analysis levels lo[c] ..
// lo[c]+cnt[c]-1 (cnt <= 128) of every input point, and filter level k
// takes sum_u w[4 k + u] in(lo[c] + r0[k] + u). kv: the stacked levels of
// the analysis side.
struct XferVArgs {
    const int* lo;
    const int* cnt;
    const int* r0;
    const float* w;
    int kv;
};
hipError_t launch_xfer_tile_v(const XferTileArgs& a, const XferVArgs& v, bool adjoint, hipStream_t st);
// Window transfer along the stacked k of every point:
// out(ko,p) = sum_u kw[ko][u] in(ks[ko]+u, p)
hipError_t launch_kwin(const float* in, float* out, int kin, int kout, long long npts, const int* ks,
                       const float* kw, int w, hipStream_t st);
// acc[0] = sum_e a[e] b[e] accumulated in FP64; acc holds 1 + kDotBlocks
// doubles (the partial sums after the first).
constexpr int kDotBlocks = 1024;
hipError_t launch_dot(const float* a, const float* b, long long count, double* acc,
                      hipStream_t st);

// The filter of procedures 4 and 5 (filtering_rad3 / _lin3) on one generation: the first
// K3 = nv*lm entries of the stack are nv columns of lm levels filtered over
// (i, j, L) by the 3-D taps, the rest by the 2-D taps; half-sample mirror
// in i and j, whole-sample mirror in L. transpose: the adjoint (the 2-D and
// horizontal parts are symmetric; the level mirror is transposed).
struct BallArgs {
    const float* in;
    float* out;
    const int4* t3;      // 3-D taps (a, b, c, 0)
    const float* w3;
    int n3;
    const int4* t2;      // 2-D taps (a, b, 0, 0) of the 2-D variables
    const float* w2;
    int n2;
    int K, n, m, lm, K3;
    int transpose;
    int zero_levels;     // zeros beyond the column ends instead of the level mirror
    int klo;             // only k >= klo is written
};
hipError_t launch_ball(const BallArgs& a, hipStream_t st);

// The same 3-D filter on the 3-D variables in the streaming form of the
// radial kernel (mgbf_ball3.hip): sign-symmetric taps with |a|,|b| <= H and
// |c| <= HZ, a wave holding whole columns (lm divides 64, lm >= 2 HZ + 2).
// ball: procedure 4's taps, w = wr[a^2+b^2+c^2] inside the lattice ball of
// radius H+1 (H = HZ, level mirror); otherwise procedure 5's box of line
// products w = fx[|a|] fy[|b|] fz[|c|] under zero_levels.
constexpr int kBall3MaxH = 4;
constexpr int kBall3MaxHZ = 4;
struct Ball3Coef {
    float wr[(kBall3MaxH + 1) * (kBall3MaxH + 1)];
    float fx[kBall3MaxH + 1], fy[kBall3MaxH + 1], fz[kBall3MaxHZ + 1];
};
struct Ball3Args {
    const float* in;
    float* out;
    const float* coarse;   // next coarser generation to prolong and add (forward only), or nullptr
    float* rout;           // the input restricted to the next generation (transposed ball pass), or nullptr
    const float4* rytab;   // its per-fine-row weights and closing rows (as RadialArgs)
    const int2* rymeta;
    const float* wfield;   // per-point scale of the output, or nullptr
    float wconst;          // constant scale when wfield is nullptr
    int K, n, m, lm, K3;   // only k < K3 is written
    int H, HZ, TJ;
    int ball, transpose, zero_levels;
    int quad;              // the restriction is the quadratic one (lquart)
    Ball3Coef coef;
};
struct Ball3Params {
    const float* in;
    float* out;
    const float* coarse;
    float* rout;
    const float4* rytab;
    const int2* rymeta;
    const float* wfield;
    float wconst;
    int K, n, m, TJ, lm, K3;
    int ntile_i, ntile_j, nitems;
    Ball3Coef W;
};
hipError_t launch_ball3(const Ball3Args& a, hipStream_t st);
hipError_t launch_ball3_H1(const Ball3Args& a, hipStream_t st);
hipError_t launch_ball3_H2(const Ball3Args& a, hipStream_t st);
hipError_t launch_ball3_H3(const Ball3Args& a, hipStream_t st);
hipError_t launch_ball3_H4(const Ball3Args& a, hipStream_t st);

}  // namespace mgbf
