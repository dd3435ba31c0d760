// HIP kernels of the MGBF operator for CDNA4 (gfx950).
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

#include "mgbf_launch.hpp"

namespace mgbf {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;

// Half-sample mirror (X(1-t)=X(t), X(n+t)=X(n+1-t), 1-based), clamped into
// the domain: a clamped index only feeds outputs outside the domain.
__device__ __forceinline__ int mirror_half(int q, int n) {
    q = q < 0 ? -1 - q : q;
    q = q >= n ? 2 * n - 1 - q : q;
    return q < 0 ? 0 : (q >= n ? n - 1 : q);
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of logical blocks so neighbouring tiles share its L2.
__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nblocks) {
    constexpr unsigned kXcd = 8;
    if (nblocks % kXcd) return bid;
    return (bid % kXcd) * (nblocks / kXcd) + bid / kXcd;
}

template <int H>
__host__ __device__ constexpr int disc_extent(int b) {
    int a = H;
    while (a >= 0 && a * a + b * b >= (H + 1) * (H + 1)) --a;
    return a;
}

// ---------------------------------------------------------------------------
// Radial 2-D beta filter F (gather form, half-sample mirror halo):
//   out(k,i,j) = scale(k,i,j) * sum_{a,b} w(|a|,|b|) in(k, mx(i+a), my(j+b))
// scale = wfield (per point) or wconst. One wave computes TI x TJ outputs for
// its 64 k: input rows stream through registers; for every input row the
// x-pair sums P_a = v(i+a)+v(i-a) give the per-row partial sums
// Z_b = sum_a w(a,b) P_a, which are added to the output rows j = r-b and r+b
// held in a ring of 2H+1 accumulator rows.
// ---------------------------------------------------------------------------
template <int H, int TI, bool DISC>
__global__ void __launch_bounds__(kBlock)
radial_kernel(const float* __restrict__ in, float* __restrict__ out,
              const float* __restrict__ wfield, float wconst,
              int K, int n, int m, int TJ, int ntile_i, int ntile_j, int nitems,
              RadialCoef W) {
    constexpr int NR = 2 * H + 1;
    constexpr int NC = TI + 2 * H;
    const int lane = threadIdx.x & (kWave - 1);
    const int item = (int)xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
    if (item >= nitems) return;
    const int ti = item % ntile_i;
    const int rest = item / ntile_i;
    const int tj = rest % ntile_j;
    const int kc = rest / ntile_j;
    const int k = kc * kWave + lane;
    const bool kvalid = k < K;
    const int kk = kvalid ? k : K - 1;
    const int i0 = ti * TI, j0 = tj * TJ;
    const size_t rowstride = (size_t)n * K;

    int coff[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) coff[c] = mirror_half(i0 - H + c, n) * K;

    float acc[NR][TI];
#pragma unroll
    for (int s = 0; s < NR; ++s)
#pragma unroll
        for (int t = 0; t < TI; ++t) acc[s][t] = 0.f;

    const int tj_eff = TJ < m - j0 ? TJ : m - j0;
    const int nsteps = tj_eff + 2 * H;
    const float* base_in = in + kk;
    float* base_out = out + kk;
    const float* base_w = wfield ? wfield + kk : nullptr;

    for (int rb = 0; rb < nsteps; rb += NR) {
#pragma unroll
        for (int s = 0; s < NR; ++s) {
            const int rr = rb + s;
            if (rr < nsteps) {
                const int r = mirror_half(j0 - H + rr, m);
                const float* row = base_in + (size_t)r * rowstride;
                float v[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) v[c] = row[coff[c]];
#pragma unroll
                for (int t = 0; t < TI; ++t) {
                    float P[H + 1];
                    P[0] = v[t + H];
#pragma unroll
                    for (int a = 1; a <= H; ++a) P[a] = v[t + H + a] + v[t + H - a];
#pragma unroll
                    for (int b = 0; b <= H; ++b) {
                        const int A = DISC ? disc_extent<H>(b) : H;
                        float z = W.w[b][0] * P[0];
#pragma unroll
                        for (int a = 1; a <= H; ++a)
                            if (a <= A) z = __builtin_fmaf(W.w[b][a], P[a], z);
                        acc[(s + b) % NR][t] += z;
                        if (b > 0) acc[(s + NR - b) % NR][t] += z;
                    }
                }
                // output row j = j0 + rr - 2H is complete
                const int jrel = rr - 2 * H;
                constexpr int slot_off = NR - H;   // (s - H) mod NR
                if (jrel >= 0) {
                    const int j = j0 + jrel;
                    if (jrel < tj_eff && kvalid) {
#pragma unroll
                        for (int t = 0; t < TI; ++t) {
                            const int i = i0 + t;
                            if (i < n) {
                                const size_t e = (size_t)j * rowstride + (size_t)i * K;
                                const float sc = base_w ? base_w[e] : wconst;
                                base_out[e] = sc * acc[(s + slot_off) % NR][t];
                            }
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < TI; ++t) acc[(s + slot_off) % NR][t] = 0.f;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Vertical 1-D beta filter on a generation-1 field: for k < nfilt (the
// vertical columns of the stack, lm levels each) out(k) = sum_t c(L,t)
// in(k - L + L + t), L = k mod lm, with the whole-sample mirror folded into
// the band table c (lm x (2hz+1)); entries k >= nfilt pass through.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
vertical_kernel(const float* __restrict__ in, float* __restrict__ out,
                const float* __restrict__ coef, int K, int lm, int nfilt, int hz,
                long long npts) {
    const int k = blockIdx.y * kWave + (threadIdx.x & (kWave - 1));
    const long long p = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (k >= K || p >= npts) return;
    const float* col = in + p * K;
    float r;
    if (k < nfilt) {
        const int L = k % lm;
        const int c0 = k - L;
        const float* cf = coef + L * (2 * hz + 1);
        r = 0.f;
        for (int t = -hz; t <= hz; ++t) {
            int q = L + t;
            q = q < 0 ? 0 : (q >= lm ? lm - 1 : q);   // zero-weight taps stay in range
            r = __builtin_fmaf(cf[t + hz], col[c0 + q], r);
        }
    } else {
        r = col[k];
    }
    out[p * K + k] = r;
}

// ---------------------------------------------------------------------------
// Restriction R (fine n x m -> coarse n/2 x m/2) from window tables: one lane
// computes TJ2 coarse rows of one coarse column; every fine row of the strip
// is x-restricted once and added to the coarse rows whose window holds it.
// ---------------------------------------------------------------------------
template <int TJ2>
__global__ void __launch_bounds__(kBlock)
restrict_kernel(const float* __restrict__ fine, float* __restrict__ coarse, int K, int nf,
                int mf, const int* __restrict__ rxs, const float* __restrict__ rxw,
                const int* __restrict__ rys, const float* __restrict__ ryw, int nstrip,
                int nitems) {
    constexpr int WR = kRWidthDev;
    const int lane = threadIdx.x & (kWave - 1);
    const int item = (int)xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
    if (item >= nitems) return;
    const int nc = nf / 2, mc = mf / 2;
    const int I = item % nc;
    const int rest = item / nc;
    const int js = rest % nstrip;
    const int kc = rest / nstrip;
    const int k = kc * kWave + lane;
    const bool kvalid = k < K;
    const int kk = kvalid ? k : K - 1;
    const int J0 = js * TJ2;
    const int J1 = J0 + TJ2 < mc ? J0 + TJ2 : mc;
    const int xs = rxs[I];
    float xw[WR];
    int xo[WR];
#pragma unroll
    for (int t = 0; t < WR; ++t) {
        xw[t] = rxw[I * WR + t];
        const int ii = xs + t < nf ? xs + t : nf - 1;
        xo[t] = ii * K;
    }
    const int jlo = rys[J0];
    int jhi = rys[J1 - 1] + WR;
    if (jhi > mf) jhi = mf;
    float acc[TJ2];
#pragma unroll
    for (int q = 0; q < TJ2; ++q) acc[q] = 0.f;
    const size_t rowstride = (size_t)nf * K;
    for (int jj = jlo; jj < jhi; ++jj) {
        const float* row = fine + kk + (size_t)jj * rowstride;
        float xr = 0.f;
#pragma unroll
        for (int t = 0; t < WR; ++t) xr = __builtin_fmaf(xw[t], row[xo[t]], xr);
#pragma unroll
        for (int q = 0; q < TJ2; ++q) {
            const int J = J0 + q;
            if (J < J1) {
                const int d = jj - rys[J];
                if (d >= 0 && d < WR) acc[q] = __builtin_fmaf(ryw[J * WR + d], xr, acc[q]);
            }
        }
    }
    if (!kvalid) return;
#pragma unroll
    for (int q = 0; q < TJ2; ++q) {
        const int J = J0 + q;
        if (J < J1) coarse[kk + (size_t)K * ((size_t)I + (size_t)nc * J)] = acc[q];
    }
}

// ---------------------------------------------------------------------------
// Prolongation P (coarse n/2 x m/2 -> fine n x m), optionally accumulated:
// out = (addend ? addend : 0) + P coarse. One lane computes TJ fine rows of
// one fine column.
// ---------------------------------------------------------------------------
template <int TJ>
__global__ void __launch_bounds__(kBlock)
prolong_kernel(const float* __restrict__ coarse, const float* addend, float* out, int K,
               int nf, int mf, const int* __restrict__ pxs, const float* __restrict__ pxw,
               const int* __restrict__ pys, const float* __restrict__ pyw, int nstrip,
               int nitems) {
    constexpr int WP = kPWidthDev;
    const int lane = threadIdx.x & (kWave - 1);
    const int item = (int)xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
    if (item >= nitems) return;
    const int nc = nf / 2;
    const int i = item % nf;
    const int rest = item / nf;
    const int js = rest % nstrip;
    const int kc = rest / nstrip;
    const int k = kc * kWave + lane;
    const bool kvalid = k < K;
    const int kk = kvalid ? k : K - 1;
    const int j0 = js * TJ;
    const int j1 = j0 + TJ < mf ? j0 + TJ : mf;
    const int xs = pxs[i];
    float xw[WP];
    int xo[WP];
#pragma unroll
    for (int u = 0; u < WP; ++u) {
        xw[u] = pxw[i * WP + u];
        const int ii = xs + u < nc ? xs + u : nc - 1;
        xo[u] = ii * K;
    }
    const int mc = mf / 2;
    const int clo = pys[j0];
    int chi = pys[j1 - 1] + WP;
    if (chi > mc) chi = mc;
    float acc[TJ];
#pragma unroll
    for (int q = 0; q < TJ; ++q) acc[q] = 0.f;
    const size_t crow = (size_t)nc * K;
    for (int jc = clo; jc < chi; ++jc) {
        const float* row = coarse + kk + (size_t)jc * crow;
        float xv = 0.f;
#pragma unroll
        for (int u = 0; u < WP; ++u) xv = __builtin_fmaf(xw[u], row[xo[u]], xv);
#pragma unroll
        for (int q = 0; q < TJ; ++q) {
            const int j = j0 + q;
            if (j < j1) {
                const int d = jc - pys[j];
                if (d >= 0 && d < WP) acc[q] = __builtin_fmaf(pyw[j * WP + d], xv, acc[q]);
            }
        }
    }
    if (!kvalid) return;
    const size_t frow = (size_t)nf * K;
#pragma unroll
    for (int q = 0; q < TJ; ++q) {
        const int j = j0 + q;
        if (j < j1) {
            const size_t e = kk + (size_t)j * frow + (size_t)i * K;
            out[e] = (addend ? addend[e] : 0.f) + acc[q];
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

// FP64-accumulated inner product; one double atomic per block into *acc.
__global__ void __launch_bounds__(kBlock)
dot_kernel(const float* __restrict__ a, const float* __restrict__ b, long long count,
           double* acc) {
    __shared__ double part[kBlock];
    double s = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += stride)
        s += (double)a[e] * (double)b[e];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(acc, part[0]);
}

__global__ void zero_double_kernel(double* p) { *p = 0.0; }

unsigned grid_for(long long count) {
    long long g = (count + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

template <int H, int TI>
hipError_t launch_radial_h(const RadialArgs& a, hipStream_t st) {
    const int ntile_i = (a.n + TI - 1) / TI;
    const int ntile_j = (a.m + a.TJ - 1) / a.TJ;
    const int nkc = (a.K + kWave - 1) / kWave;
    const long long nitems = (long long)ntile_i * ntile_j * nkc;
    const unsigned nblocks = (unsigned)((nitems + kWavesPerBlock - 1) / kWavesPerBlock);
    if (a.exact_disc)
        hipLaunchKernelGGL((radial_kernel<H, TI, true>), dim3(nblocks), dim3(kBlock), 0, st, a.in,
                           a.out, a.wfield, a.wconst, a.K, a.n, a.m, a.TJ, ntile_i, ntile_j,
                           (int)nitems, a.coef);
    else
        hipLaunchKernelGGL((radial_kernel<H, TI, false>), dim3(nblocks), dim3(kBlock), 0, st, a.in,
                           a.out, a.wfield, a.wconst, a.K, a.n, a.m, a.TJ, ntile_i, ntile_j,
                           (int)nitems, a.coef);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_radial(const RadialArgs& a, hipStream_t st) {
    switch (a.H) {
        case 1: return launch_radial_h<1, 4>(a, st);
        case 2: return launch_radial_h<2, 4>(a, st);
        case 3: return launch_radial_h<3, 4>(a, st);
        case 4: return launch_radial_h<4, 4>(a, st);
        case 5: return launch_radial_h<5, 4>(a, st);
        case 6: return launch_radial_h<6, 4>(a, st);
        case 7: return launch_radial_h<7, 4>(a, st);
        case 8: return launch_radial_h<8, 4>(a, st);
        case 9: return launch_radial_h<9, 4>(a, st);
        case 10: return launch_radial_h<10, 4>(a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_vertical(const float* in, float* out, const float* coef, int K, int lm,
                           int nfilt, int hz, long long npts, hipStream_t st) {
    dim3 grid((unsigned)((npts + kWavesPerBlock - 1) / kWavesPerBlock),
              (unsigned)((K + kWave - 1) / kWave));
    hipLaunchKernelGGL(vertical_kernel, grid, dim3(kBlock), 0, st, in, out, coef, K, lm, nfilt,
                       hz, npts);
    return hipGetLastError();
}

hipError_t launch_restrict(const TransferArgs& a, hipStream_t st) {
    constexpr int TJ2 = 8;
    const int nc = a.nf / 2, mc = a.mf / 2;
    const int nstrip = (mc + TJ2 - 1) / TJ2;
    const int nkc = (a.K + kWave - 1) / kWave;
    const long long nitems = (long long)nc * nstrip * nkc;
    const unsigned nblocks = (unsigned)((nitems + kWavesPerBlock - 1) / kWavesPerBlock);
    hipLaunchKernelGGL(restrict_kernel<TJ2>, dim3(nblocks), dim3(kBlock), 0, st, a.src, a.dst,
                       a.K, a.nf, a.mf, a.xs, a.xw, a.ys, a.yw, nstrip, (int)nitems);
    return hipGetLastError();
}

hipError_t launch_prolong(const TransferArgs& a, const float* addend, hipStream_t st) {
    constexpr int TJ = 16;
    const int nstrip = (a.mf + TJ - 1) / TJ;
    const int nkc = (a.K + kWave - 1) / kWave;
    const long long nitems = (long long)a.nf * nstrip * nkc;
    const unsigned nblocks = (unsigned)((nitems + kWavesPerBlock - 1) / kWavesPerBlock);
    hipLaunchKernelGGL(prolong_kernel<TJ>, dim3(nblocks), dim3(kBlock), 0, st, a.src, addend,
                       a.dst, a.K, a.nf, a.mf, a.xs, a.xw, a.ys, a.yw, nstrip, (int)nitems);
    return hipGetLastError();
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

hipError_t launch_dot(const float* a, const float* b, long long count, double* acc,
                      hipStream_t st) {
    hipLaunchKernelGGL(zero_double_kernel, dim3(1), dim3(1), 0, st, acc);
    hipLaunchKernelGGL(dot_kernel, dim3(grid_for(count)), dim3(kBlock), 0, st, a, b, count, acc);
    return hipGetLastError();
}

}  // namespace mgbf
