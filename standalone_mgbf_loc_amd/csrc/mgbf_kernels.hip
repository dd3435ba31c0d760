// HIP kernels of the MGBF operator for CDNA4 (gfx950): vertical band,
// restriction, prolongation, scaling and reductions; the radial filter kernel
// is in mgbf_radial.hpp (one translation unit per half-width).
//
// Field layout: k fastest, then x (i), then y (j); element (k,i,j) at
// k + K*(i + n*j). One wavefront (64 lanes) owns 64 consecutive k; every
// stage acts on each k independently except the vertical filter, which mixes
// levels of one column. Halo fills of the reference (boco_2d, half-sample
// mirror at the domain edge) become wave-uniform index remapping, so no
// padded copies are stored; adjoint folds are never formed because the
// radial filter is symmetric under that mirror (F^T = F) and the restriction
// uses precomputed transposed tables.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mgbf_device.hpp"
#include "mgbf_launch.hpp"

namespace mgbf {
namespace {

// ---------------------------------------------------------------------------
// Banded vertical operator along the stack (Vf, Vf^T or the product Vf Vf^T):
// out(k) = sum_{t=-HB..HB} c(L,t) in(k+t) for k < nfilt, L = k mod lm, with the
// whole-sample mirror folded into the band; entries k >= nfilt pass through.
// Rows ilo <= L < ihi (clear of the mirror folds) share the uniform interior
// weights; the rows near the ends of a column read their row of the table. A block stages a
// contiguous span of up to kVSpan elements (one chunk of a point's column, or
// several whole points) plus the halo in LDS, so HBM is streamed contiguously.
// ---------------------------------------------------------------------------
constexpr int kVPer = 8;                    // elements per thread
constexpr int kVSpan = kBlock * kVPer;      // elements per block

template <int HB, bool FIXL>
__global__ void __launch_bounds__(kBlock)
vband_kernel(const float* __restrict__ in, float* __restrict__ out,
             const float* __restrict__ band, VBandW wint, int ilo, int ihi, int K, int lm,
             int nfilt, long long N, int chunks_per_point, int pspan) {
    constexpr int NB = 2 * HB + 1;
    __shared__ float s[kVSpan + 2 * HB];
    long long e0;
    int kbase, span;
    if (chunks_per_point > 0) {            // K > kVSpan: one chunk of one point
        const long long pt = blockIdx.x / chunks_per_point;
        kbase = (int)(blockIdx.x % chunks_per_point) * kVSpan;
        e0 = pt * K + kbase;
        span = K - kbase < kVSpan ? K - kbase : kVSpan;
    } else {                               // whole points, pspan = ppb * K elements
        kbase = 0;
        e0 = (long long)blockIdx.x * pspan;
        span = (int)(N - e0 < pspan ? N - e0 : pspan);
    }
    const int tid = threadIdx.x;
    float tmp[kVPer];
#pragma unroll
    for (int r = 0; r < kVPer; ++r) {
        const int i = tid + r * kBlock;
        const long long e = e0 - HB + i;
        tmp[r] = (i < span + 2 * HB && e >= 0 && e < N) ? in[e] : 0.f;
    }
    float tail = 0.f;
    if (tid < 2 * HB) {
        const int i = kVSpan + tid;
        const long long e = e0 - HB + i;
        tail = (i < span + 2 * HB && e >= 0 && e < N) ? in[e] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < kVPer; ++r) s[tid + r * kBlock] = tmp[r];
    if (tid < 2 * HB) s[kVSpan + tid] = tail;
    __syncthreads();
    if constexpr (FIXL) {
        // every column is filtered and lm divides the block: the level of
        // each element of this thread is tid mod lm, its band row is fixed
        float vc[NB];
        const float* c = band + (tid % lm) * NB;
#pragma unroll
        for (int t = 0; t < NB; ++t) vc[t] = c[t];
#pragma unroll
        for (int r = 0; r < kVPer; ++r) {
            const int l = tid + r * kBlock;
            if (l < span) {
                const float* row = s + l + HB;
                float v = 0.f;
#pragma unroll
                for (int t = 0; t < NB; ++t) v = __builtin_fmaf(vc[t], row[t - HB], v);
                out[e0 + l] = v;
            }
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < kVPer; ++r) {
        const int l = tid + r * kBlock;
        if (l < span) {
            const int k = chunks_per_point > 0 ? kbase + l : l % K;
            const float* row = s + l + HB;
            float v;
            if (k < nfilt) {
                const int L = k % lm;
                v = 0.f;
                if (L >= ilo && L < ihi) {
#pragma unroll
                    for (int t = 0; t < NB; ++t) v = __builtin_fmaf(wint.w[t], row[t - HB], v);
                } else {
                    const float* c = band + L * NB;
#pragma unroll
                    for (int t = 0; t < NB; ++t) v = __builtin_fmaf(c[t], row[t - HB], v);
                }
            } else {
                v = row[0];
            }
            out[e0 + l] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Prolongation P (coarse -> fine), optionally accumulated: the coarse field is
// read through the half-sample mirror (boco_2d) and interpolated with the
// stencil S, x then y (direct1, or direct2 for the quadratic stencil). One
// lane computes TJ fine rows of one fine column for its k; the NR
// x-interpolated coarse rows they need stay in registers.
// ---------------------------------------------------------------------------
template <int TJ, bool Q>
__global__ void __launch_bounds__(kBlock)
prolong2_kernel(const float* __restrict__ coarse, const float* addend, float* out, int K, int nf,
                int mf, int nkc, int nstrip, int nitems) {
    using S = Stencil<Q>;
    constexpr int NP = S::NP;
    constexpr int CO = S::OFF[0] < S::OFF[1] ? S::OFF[0] : S::OFF[1];  // lowest coarse offset
    constexpr int NR = TJ / 2 + NP - 1 + (S::OFF[1] - CO);
    const int lane = threadIdx.x & (kWave - 1);
    const int item = (int)xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
    if (item >= nitems) return;
    const int kc = item % nkc;
    const int rest = item / nkc;
    const int i = rest % nf;
    const int js = rest / nf;
    const int k = kc * kWave + lane;
    const bool kvalid = k < K;
    const int kk = kvalid ? k : K - 1;
    const int nc = nf / 2, mc = mf / 2;
    const int u = i >> 1;
    const int par = i & 1;
    const int cs = u + (par ? S::OFF[1] : S::OFF[0]);
    float cw[NP];
    int xo[NP];
#pragma unroll
    for (int t = 0; t < NP; ++t) {
        cw[t] = par ? S::C[1][t] : S::C[0][t];
        xo[t] = mirror_half(cs + t, nc) * K;
    }
    const int j0 = js * TJ;
    const int u0 = j0 >> 1;
    const size_t crow = (size_t)nc * K;
    float xv[NR];   // coarse rows u0 + CO .. u0 + CO + NR - 1, interpolated along x
#pragma unroll
    for (int t = 0; t < NR; ++t) {
        const int r = mirror_half(u0 + CO + t, mc);
        const float* row = coarse + kk + (size_t)r * crow;
        float a = 0.f;
#pragma unroll
        for (int c = 0; c < NP; ++c) a = __builtin_fmaf(cw[c], row[xo[c]], a);
        xv[t] = a;
    }
    if (!kvalid) return;
    const size_t frow = (size_t)nf * K;
#pragma unroll
    for (int q = 0; q < TJ; ++q) {
        const int j = j0 + q;
        if (j < mf) {
            const int pq = q & 1;
            const int b = (q >> 1) + (pq ? S::OFF[1] : S::OFF[0]) - CO;
            float a = 0.f;
#pragma unroll
            for (int t = 0; t < NP; ++t) a = __builtin_fmaf(pq ? S::C[1][t] : S::C[0][t], xv[b + t], a);
            const size_t e = kk + (size_t)j * frow + (size_t)i * K;
            out[e] = (addend ? addend[e] : 0.f) + a;
        }
    }
}


// ---------------------------------------------------------------------------
// Restriction R = Fold o P^T (fine -> coarse). One lane computes TI2 coarse
// columns x TJ2 coarse rows for its k, coarse row by coarse row. An interior
// coarse row J is the NR-point stencil S::R over the x-restricted fine rows
// 2J+ROFF .. 2J+ROFF+NR-1 (2J-3 .. 2J+4 for the 4-point stencil, 2J-2 ..
// 2J+3 for the quadratic one), which a sliding window keeps in registers: the
// next row needs two new fine rows. Rows and columns within S::EDGE of an
// edge carry the folded mirror contributions and use the window tables.
// ---------------------------------------------------------------------------
// Edge tile of the restriction: window tables in x and y, one coarse point
// at a time (only the tiles within two columns of an edge come here).
template <int TI2>
__device__ void restrict_edge_tile(const float* __restrict__ fine, float* __restrict__ coarse,
                                   int kk, bool kvalid, int K, int nf, int mf, int I0, int J0,
                                   int J1, const int* __restrict__ rxs,
                                   const float* __restrict__ rxw, const int* __restrict__ rys,
                                   const float* __restrict__ ryw) {
    constexpr int WR = kRWidthDev;
    const int nc = nf / 2;
    const size_t frow = (size_t)nf * K;
    for (int c = 0; c < TI2; ++c) {
        const int I = I0 + c;
        if (I >= nc) break;
        const int xs = rxs[I];
        for (int J = J0; J < J1; ++J) {
            const int ys = rys[J];
            float acc = 0.f;
            for (int ty = 0; ty < WR; ++ty) {
                const int jj = ys + ty < mf ? ys + ty : mf - 1;
                const float wy = ys + ty < mf ? ryw[J * WR + ty] : 0.f;
                const float* row = fine + kk + (size_t)jj * frow;
                float a = 0.f;
                for (int tx = 0; tx < WR; ++tx) {
                    const int ii = xs + tx < nf ? xs + tx : nf - 1;
                    const float wx = xs + tx < nf ? rxw[I * WR + tx] : 0.f;
                    const float fv = row[(size_t)ii * K];
                    a = __builtin_fmaf(wx, fv, a);
                }
                acc = __builtin_fmaf(wy, a, acc);
            }
            if (kvalid) coarse[kk + (size_t)K * ((size_t)I + (size_t)nc * J)] = acc;
        }
    }
}

template <int TI2, int TJ2, bool Q>
__global__ void __launch_bounds__(kBlock)
restrict3_kernel(const float* __restrict__ fine, float* __restrict__ coarse, int K, int nf,
                 int mf, const int* __restrict__ rxs, const float* __restrict__ rxw,
                 const int* __restrict__ rys, const float* __restrict__ ryw, int nkc, int ntile_i,
                 int nstrip, int nitems) {
    constexpr int WR = kRWidthDev;
    using S = Stencil<Q>;
    constexpr int NRW = S::NR, E = S::EDGE;
    constexpr int NCF = 2 * TI2 + NRW - 2;
    const int lane = threadIdx.x & (kWave - 1);
    const int item = (int)xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock +
                     __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= nitems) return;
    // k chunks fastest: consecutive waves read the whole k extent of a column
    // (K x 4 contiguous bytes) together, which keeps DRAM pages open; the
    // halo columns this re-reads mostly hit the caches
    const int kc = item % nkc;
    const int rest = item / nkc;
    const int ti = rest % ntile_i;
    const int js = rest / ntile_i;
    const int k = kc * kWave + lane;
    const bool kvalid = k < K;
    const int kk = kvalid ? k : K - 1;
    const int nc = nf / 2, mc = mf / 2;
    const int I0 = ti * TI2, J0 = js * TJ2;
    const int J1 = J0 + TJ2 < mc ? J0 + TJ2 : mc;
    if (!(I0 >= E && I0 + TI2 <= nc - E)) {
        restrict_edge_tile<TI2>(fine, coarse, kk, kvalid, K, nf, mf, I0, J0, J1, rxs, rxw, rys, ryw);
        return;
    }
    const size_t frow = (size_t)nf * K;
    // interior columns: fine columns 2*I0+ROFF .. 2*I0+ROFF+NCF-1, addressed
    // as the row's buffer + lane k + wave-uniform column offset
    const float* base = fine + (size_t)(2 * I0 + S::ROFF) * K;
    const int vk = kk * 4;
    auto xrow = [&](int fr, float* out) {
        const __amdgpu_buffer_rsrc_t rs = buf_rsrc(base + (size_t)fr * frow);
        float v[NCF];
#pragma unroll
        for (int c = 0; c < NCF; ++c) v[c] = buf_ld(rs, vk, c * K * 4);
#pragma unroll
        for (int c = 0; c < TI2; ++c) {
            float a = 0.f;
#pragma unroll
            for (int t = 0; t < NRW; ++t) a = __builtin_fmaf(S::R[t], v[2 * c + t], a);
            out[c] = a;
        }
    };
    float win[NRW][TI2];   // x-restricted fine rows 2J+ROFF .. 2J+ROFF+NRW-1
    int wbase = -1000;     // fine row held in win[0]
    for (int J = J0; J < J1; ++J) {
        float acc[TI2];
        if (J >= E && J <= mc - 1 - E) {
            const int fb = 2 * J + S::ROFF;
            if (fb == wbase + 2) {
#pragma unroll
                for (int t = 0; t < NRW - 2; ++t)
#pragma unroll
                    for (int c = 0; c < TI2; ++c) win[t][c] = win[t + 2][c];
                xrow(fb + NRW - 2, win[NRW - 2]);
                xrow(fb + NRW - 1, win[NRW - 1]);
            } else {
#pragma unroll
                for (int t = 0; t < NRW; ++t) xrow(fb + t, win[t]);
            }
            wbase = fb;
#pragma unroll
            for (int c = 0; c < TI2; ++c) {
                float a = 0.f;
#pragma unroll
                for (int t = 0; t < NRW; ++t) a = __builtin_fmaf(S::R[t], win[t][c], a);
                acc[c] = a;
            }
        } else {
            // fold rows: the y window table of coarse row J
            const int ys = rys[J];
#pragma unroll
            for (int c = 0; c < TI2; ++c) acc[c] = 0.f;
            for (int t = 0; t < WR; ++t) {
                const int jj = ys + t < mf ? ys + t : mf - 1;
                const float wy = ys + t < mf ? ryw[J * WR + t] : 0.f;
                float xv[TI2];
                xrow(jj, xv);
#pragma unroll
                for (int c = 0; c < TI2; ++c) acc[c] = __builtin_fmaf(wy, xv[c], acc[c]);
            }
            wbase = -1000;
        }
        if (kvalid) {
            const __amdgpu_buffer_rsrc_t ro = buf_rsrc(coarse + (size_t)K * ((size_t)I0 + (size_t)nc * J));
#pragma unroll
            for (int c = 0; c < TI2; ++c) buf_st(acc[c], ro, vk, c * K * 4);
        }
    }
}

// out = in * (wfield ? f(wfield) : f(wconst)), f = identity or sqrt
__global__ void __launch_bounds__(kBlock)
scale_kernel(const float* __restrict__ in, float* __restrict__ out,
             const float* __restrict__ wfield, float wconst, int use_sqrt, long long count) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += stride) {
        float w = wfield ? wfield[e] : wconst;
        if (use_sqrt) w = __builtin_sqrtf(w);
        out[e] = w * in[e];
    }
}

__global__ void __launch_bounds__(kBlock)
copy_kernel(const float* __restrict__ in, float* __restrict__ out, long long count) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += stride)
        out[e] = in[e];
}

// FP64-accumulated inner product in two launches with a fixed order (so
// the result does not depend on scheduling): each of kDotBlocks blocks sums
// its grid-stride share into partial[block], 16-byte loads when both arrays
// are 16-byte aligned; one block then sums the partials.
__device__ __forceinline__ double block_sum(double s) {
    __shared__ double ws[kWavesPerBlock];
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kWavesPerBlock; ++w) t += ws[w];
    return t;
}

template <bool VEC>
__global__ void __launch_bounds__(kBlock)
dot_partial_kernel(const float* __restrict__ a, const float* __restrict__ b, long long count,
                   double* __restrict__ partial) {
    double s = 0.0;
    const long long tid = (long long)blockIdx.x * kBlock + threadIdx.x;
    const long long stride = (long long)gridDim.x * kBlock;
    long long head = 0;
    if constexpr (VEC) {
        const long long n4 = count >> 2;
        const float4* a4 = reinterpret_cast<const float4*>(a);
        const float4* b4 = reinterpret_cast<const float4*>(b);
        for (long long e = tid; e < n4; e += stride) {
            const float4 x = a4[e], y = b4[e];
            s = __builtin_fma((double)x.x, (double)y.x, s);
            s = __builtin_fma((double)x.y, (double)y.y, s);
            s = __builtin_fma((double)x.z, (double)y.z, s);
            s = __builtin_fma((double)x.w, (double)y.w, s);
        }
        head = n4 << 2;
    }
    for (long long e = head + tid; e < count; e += stride) s = __builtin_fma((double)a[e], (double)b[e], s);
    const double t = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kBlock) dot_final_kernel(const double* __restrict__ partial, int n,
                                                           double* __restrict__ acc) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) s += partial[i];
    const double t = block_sum(s);
    if (threadIdx.x == 0) *acc = t;
}

__global__ void __launch_bounds__(kBlock) fill_kernel(float* p, float v, long long count) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += stride)
        p[e] = v;
}

unsigned grid_for(long long count) {
    long long g = (count + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

hipError_t launch_radial(const RadialArgs& a, hipStream_t st) {
    // a k offset takes no band (whole columns are not in the wave) and no
    // matrix-core engine
    if (a.klo < 0 || a.klo >= a.K || (a.klo && (a.mfma || a.band))) return hipErrorInvalidValue;
    switch (a.H) {
        case 1: return launch_radial_H1(a, st);
        case 2: return launch_radial_H2(a, st);
        case 3: return launch_radial_H3(a, st);
        case 4: return launch_radial_H4(a, st);
        case 5: return launch_radial_H5(a, st);
        case 6: return launch_radial_H6(a, st);
        case 7: return launch_radial_H7(a, st);
        case 8: return launch_radial_H8(a, st);
        case 9: return launch_radial_H9(a, st);
        case 10: return launch_radial_H10(a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ball3(const Ball3Args& a, hipStream_t st) {
    if (a.HZ < 0 || a.HZ > kBall3MaxHZ || a.lm < 2 * a.HZ + 2 || kWave % a.lm != 0 || a.TJ < 1 ||
        a.K3 > a.K || a.K3 < 1)
        return hipErrorInvalidValue;
    switch (a.H) {
        case 1: return launch_ball3_H1(a, st);
        case 2: return launch_ball3_H2(a, st);
        case 3: return launch_ball3_H3(a, st);
        case 4: return launch_ball3_H4(a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_restrict(const TransferArgs& a, hipStream_t st) {
    constexpr int TI2 = 4, TJ2 = 16;
    const int nc = a.nf / 2, mc = a.mf / 2;
    const int ntile_i = (nc + TI2 - 1) / TI2;
    const int nstrip = (mc + TJ2 - 1) / TJ2;
    const int nkc = (a.K + kWave - 1) / kWave;
    const long long nitems = (long long)ntile_i * nstrip * nkc;
    const unsigned nblocks = (unsigned)((nitems + kWavesPerBlock - 1) / kWavesPerBlock);
    if (a.quad)
        hipLaunchKernelGGL((restrict3_kernel<TI2, TJ2, true>), dim3(nblocks), dim3(kBlock), 0, st,
                           a.src, a.dst, a.K, a.nf, a.mf, a.xs, a.xw, a.ys, a.yw, nkc,
                           ntile_i, nstrip, (int)nitems);
    else
        hipLaunchKernelGGL((restrict3_kernel<TI2, TJ2, false>), dim3(nblocks), dim3(kBlock), 0, st,
                           a.src, a.dst, a.K, a.nf, a.mf, a.xs, a.xw, a.ys, a.yw, nkc,
                           ntile_i, nstrip, (int)nitems);
    return hipGetLastError();
}

hipError_t launch_prolong(const TransferArgs& a, const float* addend, hipStream_t st) {
    constexpr int TJ = 16;
    const int nstrip = (a.mf + TJ - 1) / TJ;
    const int nkc = (a.K + kWave - 1) / kWave;
    const long long nitems = (long long)a.nf * nstrip * nkc;
    const unsigned nblocks = (unsigned)((nitems + kWavesPerBlock - 1) / kWavesPerBlock);
    if (a.quad)
        hipLaunchKernelGGL((prolong2_kernel<TJ, true>), dim3(nblocks), dim3(kBlock), 0, st, a.src,
                           addend, a.dst, a.K, a.nf, a.mf, nkc, nstrip, (int)nitems);
    else
        hipLaunchKernelGGL((prolong2_kernel<TJ, false>), dim3(nblocks), dim3(kBlock), 0, st, a.src,
                           addend, a.dst, a.K, a.nf, a.mf, nkc, nstrip, (int)nitems);
    return hipGetLastError();
}

template <int HB>
hipError_t launch_vband_h(const VBandArgs& a, hipStream_t st) {
    const bool fixl = a.nfilt == a.K && kBlock % a.lm == 0 && a.K % a.lm == 0;
    const long long N = (long long)a.K * a.npts;
    int cpp = 0, pspan = 0;
    long long nblocks;
    if (a.K > kVSpan) {
        cpp = (a.K + kVSpan - 1) / kVSpan;
        nblocks = a.npts * cpp;
    } else {
        const int ppb = kVSpan / a.K;
        pspan = ppb * a.K;
        nblocks = (N + pspan - 1) / pspan;
    }
    if (fixl)
        hipLaunchKernelGGL((vband_kernel<HB, true>), dim3((unsigned)nblocks), dim3(kBlock), 0, st,
                           a.in, a.out, a.band, a.wint, a.ilo, a.ihi, a.K, a.lm, a.nfilt, N, cpp,
                           pspan);
    else
        hipLaunchKernelGGL((vband_kernel<HB, false>), dim3((unsigned)nblocks), dim3(kBlock), 0, st,
                           a.in, a.out, a.band, a.wint, a.ilo, a.ihi, a.K, a.lm, a.nfilt, N, cpp,
                           pspan);
    return hipGetLastError();
}

hipError_t launch_vband(const VBandArgs& a, hipStream_t st) {
    switch (a.HB) {
#define MGBF_VB(h) case h: return launch_vband_h<h>(a, st);
        MGBF_VB(1) MGBF_VB(2) MGBF_VB(3) MGBF_VB(4) MGBF_VB(5) MGBF_VB(6) MGBF_VB(7)
        MGBF_VB(8) MGBF_VB(9) MGBF_VB(10) MGBF_VB(11) MGBF_VB(12) MGBF_VB(13) MGBF_VB(14)
        MGBF_VB(15) MGBF_VB(16) MGBF_VB(17) MGBF_VB(18) MGBF_VB(19) MGBF_VB(20)
#undef MGBF_VB
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_scale(const float* in, float* out, const float* wfield, float wconst,
                        int use_sqrt, long long count, hipStream_t st) {
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(count)), dim3(kBlock), 0, st, in, out, wfield,
                       wconst, use_sqrt, count);
    return hipGetLastError();
}

hipError_t launch_copy(const float* in, float* out, long long count, hipStream_t st) {
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(count)), dim3(kBlock), 0, st, in, out, count);
    return hipGetLastError();
}

// One wave per output point and 64 consecutive k: the window weights are
// wave-uniform (scalar loads), every tap a 256-byte coalesced read.
__global__ void __launch_bounds__(kBlock) xfer2d_kernel(XferArgs a, long long nitems) {
    const long long item = (long long)blockIdx.x * kWavesPerBlock +
                           __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= nitems) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int nkc = (a.K + kWave - 1) / kWave;
    const int kc = (int)(item % nkc);
    const long long pt = item / nkc;
    const int i = (int)(pt % a.nout_x), j = (int)(pt / a.nout_x);
    const int k = kc * kWave + lane;
    if (k >= a.K) return;
    const int x0 = a.xs[i], y0 = a.ys[j];
    float acc = 0.f;
    for (int ty = 0; ty < a.wy; ++ty) {
        const float wyv = a.yw[(size_t)j * a.wy + ty];
        const float* row = a.in + k + (size_t)a.K * ((size_t)a.nin_x * (y0 + ty) + x0);
        float r = 0.f;
        for (int tx = 0; tx < a.wx; ++tx) r = __builtin_fmaf(a.xw[(size_t)i * a.wx + tx], row[(size_t)tx * a.K], r);
        acc = __builtin_fmaf(wyv, r, acc);
    }
    a.out[k + (size_t)a.K * ((size_t)a.nout_x * j + i)] = acc;
}

// One wave per 8 x 8 output points and 64 consecutive k: the input block of
// the tile streams through registers row by row; each row is interpolated in
// x for the 8 output columns (dense tile weights, wave-uniform) and added
// into the 8 output rows with their y weights.
// MODE 1 (direct transfer with its vertical map): the 64 k are filter
// levels; each finished output row goes through a per-wave LDS tile, and
// each lane forms up to two analysis levels of every point from a window
// of 4 of the chunk's levels. MODE 2 (adjoint transfer): every input point
// brings the chunk's analysis levels (up to 128, two coalesced loads per
// lane) through a per-wave LDS tile, and each lane maps its window of 4 to
// its filter level before the horizontal pass.
template <int SX, int MODE>
__global__ void __launch_bounds__(kBlock) xfer_tile_kernel(XferTileArgs a, XferVArgs v, int ntx, int nty,
                                                         long long nitems) {
    constexpr int TB = 8;
    constexpr int DP = kWave + 4;        // LDS pitch of an output point's levels (MODE 1)
    constexpr int AP = 2 * kWave + 4;    // LDS pitch of an input point's levels (MODE 2)
    constexpr int LDSW = MODE == 1 ? TB * DP : (MODE == 2 ? SX * AP : 1);
    __shared__ float lds_v[kWavesPerBlock][LDSW];
    // the tile's x weights, read by every lane at the same address (held in
    // scalar registers they spill: TB x SX of them beside the column offsets)
    __shared__ __align__(16) float lds_wx[kWavesPerBlock][TB * SX];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long long item = (long long)xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wv;
    if (item >= nitems) return;
    const int lane = threadIdx.x & (kWave - 1);
    // With a vertical map the chunks of one tile are neighbours in the
    // launch: their level runs on the analysis grid (lm_a levels a column)
    // share cache lines at the chunk boundaries. The adjoint, which reads
    // the analysis grid, takes the tiles in strips of kStrip columns from
    // top to bottom: a row of tiles starts as the row above it finishes and
    // finds the input rows they share still cached.
    int p, q, kc;
    if constexpr (MODE == 0) {
        p = (int)(item % ntx);
        const long long rest = item / ntx;
        q = (int)(rest % nty);
        kc = (int)(rest / nty);
    } else if constexpr (MODE == 1) {
        const int nkc = (a.K + kWave - 1) / kWave;
        kc = (int)(item % nkc);
        const long long rest = item / nkc;
        p = (int)(rest % ntx);
        q = (int)(rest / ntx);
    } else {
        constexpr int kStrip = 16;
        const int nkc = (a.K + kWave - 1) / kWave;
        kc = (int)(item % nkc);
        const long long t = item / nkc;
        const int nfull = ntx / kStrip;
        const long long fullsz = (long long)kStrip * nty;
        if (t < nfull * fullsz) {
            const long long r = t % fullsz;
            q = (int)(r / kStrip);
            p = (int)(t / fullsz) * kStrip + (int)(r % kStrip);
        } else {
            const long long r = t - nfull * fullsz;
            const int wl = ntx - nfull * kStrip;
            q = (int)(r / wl);
            p = nfull * kStrip + (int)(r % wl);
        }
    }
    const int k = kc * kWave + lane;
    const bool kvalid = k < a.K;
    const int vk = (kvalid ? k : a.K - 1) * 4;
    const int x0 = a.x0[p], y0 = a.y0[q];
    const float* WX = a.wx + (size_t)p * TB * SX;
    for (int e = lane; e < TB * SX; e += kWave) lds_wx[wv][e] = WX[e];
    const float* WXL = lds_wx[wv];
    const float* WY = a.wy + (size_t)q * TB * a.SY;
    float* const col = lds_v[wv];
    // levels per input point: the filter stack, or (MODE 2) the analysis one
    const int kin = MODE == 2 ? v.kv : a.K;
    int soff[SX];
#pragma unroll
    for (int c = 0; c < SX; ++c) soff[c] = (x0 + c) * kin * 4;
    // MODE 2: the chunk's analysis levels lo .. lo+cnt-1, loaded as lo + lane
    // and lo + 64 + lane (clamped into the span, zero beyond it), and the
    // lane's window r0 .. r0+3 of them
    int voff[2] = {vk, vk};
    bool vin[2] = {true, true};
    int r0 = 0;
    float vw[4] = {0.f, 0.f, 0.f, 0.f};
    if constexpr (MODE == 2) {
        const int lo = v.lo[kc], cnt = v.cnt[kc];
        const int last = cnt > 0 ? cnt - 1 : 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = lane + h * kWave;
            vin[h] = e < cnt;
            voff[h] = (lo + (e < last ? e : last)) * 4;
        }
        const int kk = kvalid ? k : a.K - 1;
        r0 = v.r0[kk];
#pragma unroll
        for (int u = 0; u < 4; ++u) vw[u] = v.w[(size_t)kk * 4 + u];
        settle(r0);
#pragma unroll
        for (int u = 0; u < 4; ++u) settle(vw[u]);
        // the pads past the last level: read by zero taps only
        for (int e = lane; e < 4 * SX; e += kWave) col[(e >> 2) * AP + 2 * kWave + (e & 3)] = 0.f;
    }
    float acc[TB][TB];
#pragma unroll
    for (int u = 0; u < TB; ++u)
#pragma unroll
        for (int t = 0; t < TB; ++t) acc[u][t] = 0.f;
    const size_t rowstride = (size_t)a.nin_x * kin;
    constexpr int NH = MODE == 2 ? 2 : 1;
    // the next input row is loaded while the current one is combined
    float vn[SX][NH];
    auto load_row = [&](int y) __attribute__((always_inline)) {
        const __amdgpu_buffer_rsrc_t rs = buf_rsrc(a.in + (size_t)y * rowstride);
#pragma unroll
        for (int c = 0; c < SX; ++c)
#pragma unroll
            for (int h = 0; h < NH; ++h) vn[c][h] = buf_ld(rs, voff[h], soff[c]);
    };
    load_row(y0);
    for (int r = 0; r < a.SY; ++r) {
        float vc[SX];
        if constexpr (MODE == 2) {
            // the row's points through LDS, then each lane's window
#pragma unroll
            for (int c = 0; c < SX; ++c) {
                col[c * AP + lane] = vin[0] ? vn[c][0] : 0.f;
                col[c * AP + kWave + lane] = vin[1] ? vn[c][1] : 0.f;
            }
            if (r + 1 < a.SY) load_row(y0 + r + 1);
#pragma unroll
            for (int c = 0; c < SX; ++c) {
                const float* cv = col + c * AP + r0;
                vc[c] = __builtin_fmaf(vw[0], cv[0],
                                       __builtin_fmaf(vw[1], cv[1], __builtin_fmaf(vw[2], cv[2], vw[3] * cv[3])));
            }
        } else {
#pragma unroll
            for (int c = 0; c < SX; ++c) vc[c] = vn[c][0];
            if (r + 1 < a.SY) load_row(y0 + r + 1);
        }
        float xr[TB];
        // the x weights again at every row (an offset the compiler cannot
        // hoist): held in registers across the loop they would take TB x SX
        // VGPRs
        int wo = 0;
        asm("" : "+s"(wo) : "s"(r));
        const float* wl = WXL + wo;
#pragma unroll
        for (int t = 0; t < TB; ++t) {
            float z = 0.f;
#pragma unroll
            for (int c = 0; c < SX; ++c) z = __builtin_fmaf(wl[t * SX + c], vc[c], z);
            xr[t] = z;
        }
#pragma unroll
        for (int u = 0; u < TB; ++u) {
            const float wy = WY[u * a.SY + r];
#pragma unroll
            for (int t = 0; t < TB; ++t) acc[u][t] = __builtin_fmaf(wy, xr[t], acc[u][t]);
        }
    }
    if constexpr (MODE == 1) {
        // every lane (also those past K, whose levels no window reads)
        // writes its level of each point; a lane forms analysis levels
        // lo + lane and lo + 64 + lane of the chunk
        const int lo = v.lo[kc], cnt = v.cnt[kc];
        int o0[2];
        bool ov[2];
        float ow[2][4];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int o = s * kWave + lane;
            ov[s] = o < cnt;
            // a lane past cnt reads output lo's entry and drops its result
            const int ko = lo + (ov[s] ? o : 0);
            o0[s] = v.r0[ko];
#pragma unroll
            for (int u = 0; u < 4; ++u) ow[s][u] = v.w[(size_t)ko * 4 + u];
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            settle(o0[s]);
#pragma unroll
            for (int u = 0; u < 4; ++u) settle(ow[s][u]);
        }
        if (lane < 4 * TB) col[(lane >> 2) * DP + kWave + (lane & 3)] = 0.f;
        const size_t orow = (size_t)a.nout_x * v.kv;
#pragma unroll
        for (int u = 0; u < TB; ++u) {
            const int j = q * TB + u;
            if (j >= a.nout_y) break;
#pragma unroll
            for (int t = 0; t < TB; ++t) col[t * DP + lane] = acc[u][t];
            const __amdgpu_buffer_rsrc_t ro = buf_rsrc(a.out + (size_t)j * orow);
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                const int i = p * TB + t;
                if (i >= a.nout_x) break;
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const float* cv = col + t * DP + o0[s];
                    const float y = __builtin_fmaf(
                        ow[s][0], cv[0], __builtin_fmaf(ow[s][1], cv[1], __builtin_fmaf(ow[s][2], cv[2], ow[s][3] * cv[3])));
                    if (ov[s]) buf_st(y, ro, (lo + s * kWave + lane) * 4, i * v.kv * 4);
                }
            }
        }
        return;
    }
    if (!kvalid) return;
#pragma unroll
    for (int u = 0; u < TB; ++u) {
        const int j = q * TB + u;
        if (j < a.nout_y) {
            const __amdgpu_buffer_rsrc_t ro = buf_rsrc(a.out + (size_t)j * a.nout_x * a.K + (size_t)p * TB * a.K);
#pragma unroll
            for (int t = 0; t < TB; ++t)
                if (p * TB + t < a.nout_x) buf_st(acc[u][t], ro, vk, t * a.K * 4);
        }
    }
}

// One wave per 64 consecutive output k and kWinPoints consecutive points:
// each lane reads its window (start and weights) once and applies it to
// every point of the run.
constexpr int kWinPoints = 8;
constexpr int kWinMaxWidth = 8;

template <int W>
__global__ void __launch_bounds__(kBlock)
kwin_kernel(const float* __restrict__ in, float* __restrict__ out, int kin, int kout, long long npts,
            const int* __restrict__ ks, const float* __restrict__ kw, int w, int nkc, long long nitems) {
    const long long item = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (item >= nitems) return;
    const int kc = (int)(item % nkc);
    const long long p0 = (item / nkc) * kWinPoints;
    const int ko = kc * kWave + (threadIdx.x & (kWave - 1));
    if (ko >= kout) return;
    const int s0 = ks[ko];
    if constexpr (W == 0) {   // any width: the weights are read per point
        for (int q = 0; q < kWinPoints; ++q) {
            const long long p = p0 + q;
            if (p < npts) {
                const float* col = in + p * kin + s0;
                float acc = 0.f;
                for (int u = 0; u < w; ++u) acc = __builtin_fmaf(kw[(size_t)ko * w + u], col[u], acc);
                out[p * kout + ko] = acc;
            }
        }
    } else {
        float c[W];
#pragma unroll
        for (int u = 0; u < W; ++u) c[u] = u < w ? kw[(size_t)ko * w + u] : 0.f;
        // every load of the run is issued before the first store
        float x[kWinPoints][W];
#pragma unroll
        for (int q = 0; q < kWinPoints; ++q) {
            const long long p = p0 + q < npts ? p0 + q : npts - 1;
            const float* col = in + p * kin + s0;
#pragma unroll
            for (int u = 0; u < W; ++u) x[q][u] = u < w ? col[u] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < kWinPoints; ++q) {
            float acc = 0.f;
#pragma unroll
            for (int u = 0; u < W; ++u) acc = __builtin_fmaf(c[u], x[q][u], acc);
            if (p0 + q < npts) out[(p0 + q) * kout + ko] = acc;
        }
    }
}

// 3-D filter of procedures 4 and 5 (vrbeta3 / vrbeta3T with sup_vrbeta3's
// level mirror; the hexad lines with zeros beyond the column ends) and the
// 2-D filter of the 2-D variables, as a gather: one
// lane per stacked k, one wave per 64 k of a grid point, taps wave-uniform.
// The transpose gathers every level L whose mirrored shift L + c lands on
// the output level (at most one direct and one mirrored preimage).
template <bool T>
__global__ void __launch_bounds__(kBlock) ball_kernel(BallArgs a, int nkc, long long nitems) {
    const long long item = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (item >= nitems) return;
    const int kc = (int)(item % nkc) + a.klo / kWave;
    const long long rest = item / nkc;
    const int i = (int)(rest % a.n), j = (int)(rest / a.n);
    const int k = kc * kWave + (threadIdx.x & (kWave - 1));
    if (k >= a.K || k < a.klo) return;
    const int lm = a.lm;
    float acc = 0.f;
    if (k < a.K3) {
        const int v0 = (k / lm) * lm, L = k - v0;
        for (int t = 0; t < a.n3; ++t) {
            const int4 g = a.t3[t];
            const int ii = mirror_half(i + g.x, a.n), jj = mirror_half(j + g.y, a.m);
            const float* col = a.in + (size_t)a.K * ((size_t)ii + (size_t)a.n * jj) + v0;
            float s;
            if (!T) {
                const int Ls = L + g.z;
                s = (a.zero_levels && (Ls < 0 || Ls >= lm)) ? 0.f : col[mirror_whole(Ls, lm)];
            } else {
                const int L1 = L - g.z, L2 = -L - g.z, L3 = 2 * (lm - 1) - L - g.z;
                s = (L1 >= 0 && L1 < lm) ? col[L1] : 0.f;
                if (!a.zero_levels) {
                    if (L > 0 && L2 >= 0 && L2 < lm) s += col[L2];
                    if (L < lm - 1 && L3 >= 0 && L3 < lm) s += col[L3];
                }
            }
            acc = __builtin_fmaf(a.w3[t], s, acc);
        }
    } else {
        for (int t = 0; t < a.n2; ++t) {
            const int4 g = a.t2[t];
            const int ii = mirror_half(i + g.x, a.n), jj = mirror_half(j + g.y, a.m);
            acc = __builtin_fmaf(a.w2[t], a.in[k + (size_t)a.K * ((size_t)ii + (size_t)a.n * jj)], acc);
        }
    }
    a.out[k + (size_t)a.K * ((size_t)i + (size_t)a.n * j)] = acc;
}

hipError_t launch_ball(const BallArgs& a, hipStream_t st) {
    if (a.klo < 0 || a.klo >= a.K) return hipErrorInvalidValue;
    const int nkc = (a.K + kWave - 1) / kWave - a.klo / kWave;
    const long long nitems = (long long)nkc * a.n * a.m;
    const long long nblocks = (nitems + kWavesPerBlock - 1) / kWavesPerBlock;
    if (nblocks >= (1ll << 31)) return hipErrorInvalidValue;
    if (a.transpose)
        hipLaunchKernelGGL(ball_kernel<true>, dim3((unsigned)nblocks), dim3(kBlock), 0, st, a, nkc, nitems);
    else
        hipLaunchKernelGGL(ball_kernel<false>, dim3((unsigned)nblocks), dim3(kBlock), 0, st, a, nkc, nitems);
    return hipGetLastError();
}

// The radial filter of a generation whose aspect tensor varies per point
// (vrbeta2 / vrbeta2T with el(:,:,ix,iy) and ss(ix,iy), jp_pbfil.f90:
// 888-928, 1011-1050). Point p = i + n j carries the weights of its taps
// over the box |gx|, |gy| <= H, w[p][t] with t = (gy+H)(2H+1) + gx+H
// (ss included, zero off its ellipse). One wave per point and 64 k:
//   F   (T = false): out(p) = sum_t w[p][t] in(mirror(p + g_t)), the gather
//                    of vrbeta2 over the mirrored halo of boco_2d;
//   F^T (T = true):  the scatter of vrbeta2T followed by the fold of
//                    bocoT_2d, as a gather: out(q) = sum over the mirror
//                    preimages e of q (per axis q, -1-q within H of the low
//                    edge, 2n-1-q within H of the high one) and the taps t
//                    whose source s = e - g_t lies in the domain, of
//                    w[s][t] in(s).
template <bool T>
__global__ void __launch_bounds__(kBlock) aspect_kernel(AspectArgs a, int nkc, long long nitems) {
    const long long item = __builtin_amdgcn_readfirstlane(
        (int)((long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)));
    if (item >= nitems) return;
    const int kc = (int)(item % nkc);
    const long long pt = item / nkc;
    const int i = (int)(pt % a.n), j = (int)(pt / a.n);
    const int k = kc * kWave + (threadIdx.x & (kWave - 1));
    const bool kv = k < a.K;
    const int kk = kv ? k : a.K - 1;
    const int H = a.H, NB = 2 * H + 1, nt = NB * NB, n = a.n, m = a.m;
    const size_t K = (size_t)a.K;
    float acc = 0.f;
    if (!T) {
        const float* wp = a.w + (size_t)(i + (size_t)n * j) * nt;
        for (int gy = -H; gy <= H; ++gy) {
            const int jj = mirror_half(j + gy, m);
            for (int gx = -H; gx <= H; ++gx) {
                const float wv = wp[(gy + H) * NB + gx + H];
                if (wv != 0.f)
                    acc = __builtin_fmaf(wv, a.in[kk + K * ((size_t)mirror_half(i + gx, n) + (size_t)n * jj)], acc);
            }
        }
    } else {
        int ex[3], ey[3], nx = 0, ny = 0;
        ex[nx++] = i;
        if (i <= H - 1) ex[nx++] = -1 - i;
        if (i >= n - H) ex[nx++] = 2 * n - 1 - i;
        ey[ny++] = j;
        if (j <= H - 1) ey[ny++] = -1 - j;
        if (j >= m - H) ey[ny++] = 2 * m - 1 - j;
        for (int b = 0; b < ny; ++b)
            for (int gy = -H; gy <= H; ++gy) {
                const int sy = ey[b] - gy;
                if (sy < 0 || sy >= m) continue;
                for (int c = 0; c < nx; ++c)
                    for (int gx = -H; gx <= H; ++gx) {
                        const int sx = ex[c] - gx;
                        if (sx < 0 || sx >= n) continue;
                        const size_t sp = (size_t)sx + (size_t)n * sy;
                        const float wv = a.w[sp * nt + (gy + H) * NB + gx + H];
                        if (wv != 0.f) acc = __builtin_fmaf(wv, a.in[kk + K * sp], acc);
                    }
            }
    }
    if (kv) a.out[k + K * ((size_t)i + (size_t)n * j)] = acc;
}

hipError_t launch_aspect(const AspectArgs& a, bool transpose, hipStream_t st) {
    if (a.H < 1 || a.n < a.H || a.m < a.H || a.K < 1) return hipErrorInvalidValue;
    const int nkc = (a.K + kWave - 1) / kWave;
    const long long nitems = (long long)nkc * a.n * a.m;
    const long long nblocks = (nitems + kWavesPerBlock - 1) / kWavesPerBlock;
    if (nitems >= (1ll << 31)) return hipErrorInvalidValue;
    if (transpose)
        hipLaunchKernelGGL(aspect_kernel<true>, dim3((unsigned)nblocks), dim3(kBlock), 0, st, a, nkc, nitems);
    else
        hipLaunchKernelGGL(aspect_kernel<false>, dim3((unsigned)nblocks), dim3(kBlock), 0, st, a, nkc, nitems);
    return hipGetLastError();
}

hipError_t launch_xfer2d(const XferArgs& a, hipStream_t st) {
    const long long nkc = (a.K + kWave - 1) / kWave;
    const long long nitems = nkc * a.nout_x * a.nout_y;
    const long long nblocks = (nitems + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL(xfer2d_kernel, dim3((unsigned)nblocks), dim3(kBlock), 0, st, a, nitems);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_xfer_tile_mode(const XferTileArgs& a, const XferVArgs& v, hipStream_t st) {
    const int ntx = (a.nout_x + 7) / 8, nty = (a.nout_y + 7) / 8;
    const long long nkc = (a.K + kWave - 1) / kWave;
    const long long nitems = nkc * ntx * nty;
    const long long nblocks = (nitems + kWavesPerBlock - 1) / kWavesPerBlock;
    if (nblocks >= (1ll << 31) || a.SY < 1 || a.SY > a.nin_y || a.SX > a.nin_x) return hipErrorInvalidValue;
    const dim3 g((unsigned)nblocks), b(kBlock);
    switch (a.SX) {
        case 8: hipLaunchKernelGGL((xfer_tile_kernel<8, MODE>), g, b, 0, st, a, v, ntx, nty, nitems); break;
        case 12: hipLaunchKernelGGL((xfer_tile_kernel<12, MODE>), g, b, 0, st, a, v, ntx, nty, nitems); break;
        case 16: hipLaunchKernelGGL((xfer_tile_kernel<16, MODE>), g, b, 0, st, a, v, ntx, nty, nitems); break;
        case 20: hipLaunchKernelGGL((xfer_tile_kernel<20, MODE>), g, b, 0, st, a, v, ntx, nty, nitems); break;
        case 24: hipLaunchKernelGGL((xfer_tile_kernel<24, MODE>), g, b, 0, st, a, v, ntx, nty, nitems); break;
        case 32: hipLaunchKernelGGL((xfer_tile_kernel<32, MODE>), g, b, 0, st, a, v, ntx, nty, nitems); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_xfer_tile(const XferTileArgs& a, hipStream_t st) {
    return launch_xfer_tile_mode<0>(a, XferVArgs{}, st);
}

hipError_t launch_xfer_tile_v(const XferTileArgs& a, const XferVArgs& v, bool adjoint, hipStream_t st) {
    if (!v.lo || !v.cnt || !v.r0 || !v.w || v.kv < 1) return hipErrorInvalidValue;
    // 32-bit offsets inside a row of analysis points
    const long long npts = adjoint ? a.nin_x : a.nout_x;
    if (npts * v.kv * 4 >= (1ll << 31)) return hipErrorInvalidValue;
    return adjoint ? launch_xfer_tile_mode<2>(a, v, st) : launch_xfer_tile_mode<1>(a, v, st);
}

hipError_t launch_kwin(const float* in, float* out, int kin, int kout, long long npts, const int* ks,
                       const float* kw, int w, hipStream_t st) {
    if (w < 1) return hipErrorInvalidValue;
    const int nkc = (kout + kWave - 1) / kWave;
    const long long nitems = (long long)nkc * ((npts + kWinPoints - 1) / kWinPoints);
    const long long nblocks = (nitems + kWavesPerBlock - 1) / kWavesPerBlock;
    if (nblocks >= (1ll << 31)) return hipErrorInvalidValue;
    const dim3 g((unsigned)nblocks), b(kBlock);
    if (w <= 4)
        hipLaunchKernelGGL(kwin_kernel<4>, g, b, 0, st, in, out, kin, kout, npts, ks, kw, w, nkc, nitems);
    else if (w <= kWinMaxWidth)
        hipLaunchKernelGGL(kwin_kernel<kWinMaxWidth>, g, b, 0, st, in, out, kin, kout, npts, ks, kw, w, nkc,
                           nitems);
    else
        hipLaunchKernelGGL(kwin_kernel<0>, g, b, 0, st, in, out, kin, kout, npts, ks, kw, w, nkc, nitems);
    return hipGetLastError();
}

hipError_t launch_fill(float* p, float v, long long count, hipStream_t st) {
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(count)), dim3(kBlock), 0, st, p, v, count);
    return hipGetLastError();
}

hipError_t launch_dot(const float* a, const float* b, long long count, double* acc,
                      hipStream_t st) {
    double* partial = acc + 1;
    const bool vec = ((reinterpret_cast<size_t>(a) | reinterpret_cast<size_t>(b)) & 15) == 0;
    if (vec)
        hipLaunchKernelGGL(dot_partial_kernel<true>, dim3(kDotBlocks), dim3(kBlock), 0, st, a, b, count, partial);
    else
        hipLaunchKernelGGL(dot_partial_kernel<false>, dim3(kDotBlocks), dim3(kBlock), 0, st, a, b, count, partial);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kBlock), 0, st, partial, kDotBlocks, acc);
    return hipGetLastError();
}

}  // namespace mgbf
