// Radial 2-D beta filter on the matrix cores (gfx950 v_mfma_f32_16x16x32_f16).
//
// For one input row r the filter's contribution to output row r+-b is a
// banded Toeplitz product along x: Z_b(i, k) = sum_c T_b(i, c) v(c, k) with
// T_b(i, c) = w(|c - H - i|, b) over a window of 16 + 2H <= 32 columns. A
// block of 4 wavefronts owns 16 output columns x 64 k x TJ output rows; wave
// w owns k0 + 16w .. +15 (the N dimension of the MFMA). Each input row is
// staged once in LDS as FP16 hi/lo pairs (v = hi + lo to 22 bits) and every
// wave forms Z_b with three MFMAs (hi*hi + hi*lo + lo*hi, FP32 accumulate),
// then adds Z_b into the ring of 2H+1 output-row accumulators (the same ring
// as the vector kernel). Exactness of the split relies on range: every k is
// scaled by a power of two fixed for the launch from max|v| over the input
// field (written by the restriction that read it), the weights by a power of
// two fixed per filter; both are undone exactly at the store.
#pragma once

#include "mgbf_device.hpp"
#include "mgbf_launch.hpp"

namespace mgbf {

typedef _Float16 mgbf_half8 __attribute__((ext_vector_type(8)));
typedef float mgbf_float4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaCols = 16;      // output columns per block
constexpr int kMfmaWin = 32;       // window columns (MFMA K)
constexpr int kLdsPitch = 40;      // halves per k row in LDS (80 B, staggers banks)

// power-of-two exponent that brings |v| <= m below 2^15 (FP16 keeps headroom)
__device__ __forceinline__ int mfma_exp(float m) {
    if (!(m > 0.f)) return 0;
    const int e = 14 - ilogbf(m);
    return e < -100 ? -100 : (e > 100 ? 100 : e);
}

template <int H>
__global__ void __launch_bounds__(kBlock) radial_mfma_kernel(RadialMfmaParams p) {
    static_assert(kMfmaCols + 2 * H <= kMfmaWin, "window exceeds one MFMA K step");
    constexpr int NR = 2 * H + 1;
    __shared__ __align__(16) _Float16 sh[2][2][kWave][kLdsPitch];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = tid >> 6;
    const int K = p.K, n = p.n, m = p.m;
    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int ti = tile % p.ntile_i;
    const int rest = tile / p.ntile_i;
    const int tj = rest % p.ntile_j;
    const int kc = rest / p.ntile_j;
    const int i0 = ti * kMfmaCols, j0 = tj * p.TJ, k0 = kc * kWave;
    const size_t rowstride = (size_t)n * K;

    // staging role: k = k0 + lane, window columns w*8 .. w*8+7
    const int ks = k0 + lane < K ? k0 + lane : K - 1;
    int coff[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) coff[q] = mirror_half(i0 - H + w * 8 + q, n) * K;
    const float sc_in = __builtin_ldexpf(1.f, mfma_exp(p.kmax_in[ks]));

    // compute role: k = k0 + 16w + (lane & 15), output columns (lane>>4)*4 .. +3
    const int kq = k0 + 16 * w + (lane & 15);
    const bool kvalid = kq < K;
    const int kqc = kvalid ? kq : K - 1;
    const float unscale = __builtin_ldexpf(1.f, -(mfma_exp(p.kmax_in[kqc]) + p.wexp));

    mgbf_half8 Ah[H + 1], Al[H + 1];
    const mgbf_half8* af = reinterpret_cast<const mgbf_half8*>(p.afrag);
#pragma unroll
    for (int b = 0; b <= H; ++b) {
        Ah[b] = af[(b * 2 + 0) * kWave + lane];
        Al[b] = af[(b * 2 + 1) * kWave + lane];
    }

    mgbf_float4 acc[NR];
#pragma unroll
    for (int s = 0; s < NR; ++s) acc[s] = mgbf_float4{0.f, 0.f, 0.f, 0.f};

    const int tj_eff = p.TJ < m - j0 ? p.TJ : m - j0;
    const int nsteps = tj_eff + 2 * H;
    const float* base_in = p.in + ks;
    float v[8];
    {
        const float* row = base_in + (size_t)mirror_half(j0 - H, m) * rowstride;
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = row[coff[q]];
    }
    const int ocol0 = i0 + (lane >> 4) * 4;
    float tmax = 0.f;   // max |out| of this lane's k (for the next pass's scale)

    for (int rb = 0; rb < nsteps; rb += NR) {
#pragma unroll
        for (int s = 0; s < NR; ++s) {
            const int rr = rb + s;
            if (rr < nsteps) {
                const int buf = rr & 1;
                // stage row rr: FP16 hi/lo of the scaled values
                mgbf_half8 hi, lo;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float x = v[q] * sc_in;
                    const _Float16 h = (_Float16)x;
                    hi[q] = h;
                    lo[q] = (_Float16)(x - (float)h);
                }
                *reinterpret_cast<mgbf_half8*>(&sh[buf][0][lane][w * 8]) = hi;
                *reinterpret_cast<mgbf_half8*>(&sh[buf][1][lane][w * 8]) = lo;
                if (rr + 1 < nsteps) {
                    const float* row = base_in + (size_t)mirror_half(j0 - H + rr + 1, m) * rowstride;
#pragma unroll
                    for (int q = 0; q < 8; ++q) v[q] = row[coff[q]];
                }
                __syncthreads();
                const mgbf_half8 Bh = *reinterpret_cast<const mgbf_half8*>(
                    &sh[buf][0][16 * w + (lane & 15)][(lane >> 4) * 8]);
                const mgbf_half8 Bl = *reinterpret_cast<const mgbf_half8*>(
                    &sh[buf][1][16 * w + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
                for (int b = 0; b <= H; ++b) {
                    mgbf_float4 z = mgbf_float4{0.f, 0.f, 0.f, 0.f};
                    z = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[b], Bh, z, 0, 0, 0);
                    z = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[b], Bl, z, 0, 0, 0);
                    z = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[b], Bh, z, 0, 0, 0);
                    acc[(s + b) % NR] += z;
                    if (b > 0) acc[(s + NR - b) % NR] += z;
                }
                const int jrel = rr - 2 * H;
                constexpr int slot_off = NR - H;
                if (jrel >= 0 && jrel < tj_eff) {
                    const int j = j0 + jrel;
                    const mgbf_float4 o = acc[(s + slot_off) % NR];
                    float mx = 0.f;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int col = ocol0 + q;
                        if (col < n && kvalid) {
                            const size_t e = (size_t)j * rowstride + (size_t)col * K + kq;
                            const float sc = p.wfield ? p.wfield[e] : p.wconst;
                            const float val = sc * unscale * o[q];
                            p.out[e] = val;
                            mx = fmaxf(mx, fabsf(val));
                        }
                    }
                    tmax = fmaxf(tmax, mx);
                }
                acc[(s + slot_off) % NR] = mgbf_float4{0.f, 0.f, 0.f, 0.f};
            }
        }
    }
    if (p.kmax_out && kvalid)
        atomicMax(reinterpret_cast<unsigned int*>(p.kmax_out) + kq, __float_as_uint(tmax));
}

template <int H>
hipError_t launch_radial_mfma_h(const RadialArgs& a, hipStream_t st) {
    if constexpr (kMfmaCols + 2 * H > kMfmaWin) {
        return hipErrorInvalidValue;
    } else {
        RadialMfmaParams p{};
        p.in = a.in;
        p.out = a.out;
        p.wfield = a.wfield;
        p.wconst = a.wconst;
        p.kmax_in = a.kmax_in;
        p.kmax_out = a.kmax_out;
        p.afrag = a.afrag;
        p.wexp = a.wexp;
        p.K = a.K;
        p.n = a.n;
        p.m = a.m;
        p.TJ = a.TJ;
        p.ntile_i = (a.n + kMfmaCols - 1) / kMfmaCols;
        p.ntile_j = (a.m + a.TJ - 1) / a.TJ;
        const int nkc = (a.K + kWave - 1) / kWave;
        const long long nblocks = (long long)p.ntile_i * p.ntile_j * nkc;
        hipLaunchKernelGGL((radial_mfma_kernel<H>), dim3((unsigned)nblocks), dim3(kBlock), 0, st, p);
        return hipGetLastError();
    }
}

}  // namespace mgbf
