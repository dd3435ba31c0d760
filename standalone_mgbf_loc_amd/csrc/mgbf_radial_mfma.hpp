// Radial 2-D beta filter on the matrix cores (gfx950 v_mfma_f32_16x16x32_f16).
//
// For one input row r the filter's contribution to output row r+-b is a
// banded Toeplitz product along x: Z_b(i, k) = sum_c T_b(i, c) v(c, k) with
// T_b(i, c) = w(|c - H - i|, b) over a window of 16 + 2H <= 32 columns. A
// block of 4 wavefronts owns 16 output columns x 64 k x TJ output rows; wave
// w owns k0 + 16w .. +15 (the N dimension of the MFMA). Each input row is
// loaded two rows ahead, staged once in LDS as FP16 hi/lo pairs (v = hi + lo
// to 22 bits), and every wave forms Z_b with three MFMAs (hi*hi + hi*lo +
// lo*hi, FP32 accumulate), then adds Z_b into the ring of 2H+1 output-row
// accumulators (the same ring as the vector kernel). Exactness of the split
// relies on range: every k is scaled by a power of two fixed for the launch
// from max|v| over the input field (written by the kernel that produced it),
// the weights by a power of two fixed per filter; both are undone exactly at
// the store.
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
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int K = p.K, n = p.n, m = p.m;
    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int ti = tile % p.ntile_i;
    const int rest = tile / p.ntile_i;
    const int tj = rest % p.ntile_j;
    const int kc = rest / p.ntile_j;
    const int i0 = ti * kMfmaCols, j0 = tj * p.TJ, k0 = kc * kWave;
    const size_t rowstride = (size_t)n * K;

    // staging role: k = k0 + lane, window columns w*8 .. w*8+7 (byte offsets
    // from the row start: per lane k, per column wave-uniform)
    const int ks = k0 + lane < K ? k0 + lane : K - 1;
    const int vks = ks * 4;
    int soff[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) soff[q] = mirror_half(i0 - H + w * 8 + q, n) * K * 4;
    const float sc_in = __builtin_ldexpf(1.f, mfma_exp(p.kmax_in[ks]));

    // compute role: k = k0 + 16w + (lane & 15), output columns (lane>>4)*4 .. +3
    const int kq = k0 + 16 * w + (lane & 15);
    const bool kvalid = kq < K;
    const int kqc = kvalid ? kq : K - 1;
    const float unscale = __builtin_ldexpf(1.f, -(mfma_exp(p.kmax_in[kqc]) + p.wexp));
    const float cscale = p.wconst * unscale;
    const int ocol0 = i0 + (lane >> 4) * 4;
    const int vout = ((lane >> 4) * 4 * K + kqc) * 4;
    const bool track = p.kmax_out != nullptr;

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
    // two input rows in flight: row r sits in vb[r & 1] from two steps
    // before it is staged
    float vb[2][8];
    auto load_row = [&](int rr, float* dst) {
        const __amdgpu_buffer_rsrc_t rs = buf_rsrc(p.in + (size_t)mirror_half(j0 - H + rr, m) * rowstride);
#pragma unroll
        for (int q = 0; q < 8; ++q) dst[q] = buf_ld(rs, vks, soff[q]);
    };
    load_row(0, vb[0]);
    if (nsteps > 1) load_row(1, vb[1]);
    // max |out| of this lane's k (for the next pass's scale), as float bits:
    // for non-negative floats the integer order is the float order
    unsigned tmax = 0u;

    // unrolled over 2*NR rows so that the ring slot (s % NR), the register
    // buffer and the LDS buffer (s & 1) are all compile-time
    for (int rb = 0; rb < nsteps; rb += 2 * NR) {
#pragma unroll
        for (int s = 0; s < 2 * NR; ++s) {
            const int rr = rb + s;
            if (rr < nsteps) {
                const int buf = s & 1;
                const int slot = s % NR;
                // stage row rr: FP16 hi/lo of the scaled values
                mgbf_half8 hi, lo;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float x = vb[buf][q] * sc_in;
                    const _Float16 h = (_Float16)x;
                    hi[q] = h;
                    lo[q] = (_Float16)(x - (float)h);
                }
                *reinterpret_cast<mgbf_half8*>(&sh[buf][0][lane][w * 8]) = hi;
                *reinterpret_cast<mgbf_half8*>(&sh[buf][1][lane][w * 8]) = lo;
                if (rr + 2 < nsteps) load_row(rr + 2, vb[buf]);
                __syncthreads();
                const mgbf_half8 Bh = *reinterpret_cast<const mgbf_half8*>(
                    &sh[buf][0][16 * w + (lane & 15)][(lane >> 4) * 8]);
                const mgbf_half8 Bl = *reinterpret_cast<const mgbf_half8*>(
                    &sh[buf][1][16 * w + (lane & 15)][(lane >> 4) * 8]);
                // b = 0 accumulates in place; b > 0 feeds two output rows
                acc[slot] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[0], Bh, acc[slot], 0, 0, 0);
                acc[slot] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[0], Bl, acc[slot], 0, 0, 0);
                acc[slot] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[0], Bh, acc[slot], 0, 0, 0);
#pragma unroll
                for (int b = 1; b <= H; ++b) {
                    mgbf_float4 z = mgbf_float4{0.f, 0.f, 0.f, 0.f};
                    z = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[b], Bh, z, 0, 0, 0);
                    z = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[b], Bl, z, 0, 0, 0);
                    z = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[b], Bh, z, 0, 0, 0);
                    acc[(slot + b) % NR] += z;
                    acc[(slot + NR - b) % NR] += z;
                }
                const int jrel = rr - 2 * H;
                const int oslot = (slot + NR - H) % NR;
                if (jrel >= 0 && jrel < tj_eff) {
                    const size_t rowoff = (size_t)(j0 + jrel) * rowstride + (size_t)i0 * K;
                    const __amdgpu_buffer_rsrc_t ro = buf_rsrc(p.out + rowoff);
                    const mgbf_float4 o = acc[oslot];
                    const __amdgpu_buffer_rsrc_t rw = buf_rsrc(p.wfield ? p.wfield + rowoff : p.out);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (ocol0 + q < n && kvalid) {
                            const float sc = p.wfield ? buf_ld(rw, vout, q * K * 4) * unscale : cscale;
                            const float val = sc * o[q];
                            buf_st(val, ro, vout, q * K * 4);
                            tmax = umax(tmax, __float_as_uint(val) & 0x7fffffffu);
                        }
                    }
                }
                acc[oslot] = mgbf_float4{0.f, 0.f, 0.f, 0.f};
            }
        }
    }
    if (track && kvalid)
        atomicMax(reinterpret_cast<unsigned int*>(p.kmax_out) + kq, tmax);
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
        // byte offsets within one row (and the row's last tile) must fit the
        // 32-bit buffer offsets
        if ((long long)a.n * a.K * 4 + (long long)kMfmaCols * a.K * 4 >= (1ll << 31))
            return hipErrorInvalidValue;
        p.ntile_i = (a.n + kMfmaCols - 1) / kMfmaCols;
        p.ntile_j = (a.m + a.TJ - 1) / a.TJ;
        const int nkc = (a.K + kWave - 1) / kWave;
        const long long nblocks = (long long)p.ntile_i * p.ntile_j * nkc;
        hipLaunchKernelGGL((radial_mfma_kernel<H>), dim3((unsigned)nblocks), dim3(kBlock), 0, st, p);
        return hipGetLastError();
    }
}

}  // namespace mgbf
