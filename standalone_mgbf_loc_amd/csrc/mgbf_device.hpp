// Device helpers shared by the HIP kernels of the MGBF operator.
#pragma once

#include <type_traits>
#include <utility>

#include <hip/hip_runtime.h>

namespace mgbf {

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

// Whole-sample mirror (X(1-t)=X(1+t), X(n+t)=X(n-t), 1-based), clamped
// into the domain.
__device__ __forceinline__ int mirror_whole(int q, int n) {
    q = q < 0 ? -q : q;
    q = q >= n ? 2 * (n - 1) - q : q;
    return q < 0 ? 0 : (q >= n ? n - 1 : q);
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of logical blocks so neighbouring tiles share its L2.
__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nblocks) {
    constexpr unsigned kXcd = 8;
    if (nblocks % kXcd) return bid;
    return (bid % kXcd) * (nblocks / kXcd) + bid / kXcd;
}

// Buffer resource over 4 GiB from `base`: raw dword accesses address
// base + voffset (per lane) + soffset (wave-uniform), both in bytes, so a row
// access needs no per-lane 64-bit address arithmetic. Callers keep every
// offset below 2^31 (one generation row is limited to that size).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const float* base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)0xffffffff,
                                             0x00020000);
}

// f(integral_constant<int, I>) for I in the sequence, unrolled at compile
// time whatever the size of f's body
template <class F, int... I>
__device__ __forceinline__ void static_for(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}

// Makes a value loaded before a loop count as consumed here: the empty asm
// reads and "rewrites" its registers, so the compiler waits for the load
// once, at this point. A load still pending where a loop first uses it is
// otherwise waited for inside the loop, with a count that also drains the
// loop's own prefetches.
template <class T>
__device__ __forceinline__ void settle(T& x) {
    asm("" : "+v"(x));
}

__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, voff, soff, 0);
}

__device__ __forceinline__ unsigned umax(unsigned a, unsigned b) { return a > b ? a : b; }

// Whole-wave single-lane shifts (DPP wave_shl:1 / wave_shr:1): lane l reads
// lane l+1 (from above) or lane l-1 (from below); the lane with no source
// reads 0.
// bound_ctrl: the lane without a source lane writes 0 (no old value to set up)
__device__ __forceinline__ float lane_from_above(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xf,
                                                              0xf, true));
}
__device__ __forceinline__ float lane_from_below(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xf,
                                                              0xf, true));
}

// Interpolation constants of the offset grids (def_offset_coef): fine index
// 2u (even, 0-based) reads coarse u-2..u+1 with kQc, fine 2u+1 reads coarse
// u-1..u+2 with kPc.
constexpr float kPc[4] = {-9.f / 128.f, 111.f / 128.f, 29.f / 128.f, -3.f / 128.f};
constexpr float kQc[4] = {-3.f / 128.f, 29.f / 128.f, 111.f / 128.f, -9.f / 128.f};

// Restriction weights of fine rows 2J-3 .. 2J+4 onto an interior coarse row J
// (the transpose of the two interpolation stencils above).
constexpr float kW8[8] = {-3.f / 128.f, -9.f / 128.f, 29.f / 128.f, 111.f / 128.f,
                          111.f / 128.f, 29.f / 128.f, -9.f / 128.f, -3.f / 128.f};

// Inter-generation transfer along one axis, as a compile-time stencil.
// Fine index i (0-based, u = i >> 1) reads coarse u + OFF[i & 1] + t,
// t < NP, with weight C[i & 1][t], the coarse indices mirrored (half-sample)
// into the domain. The interior restriction onto coarse I reads fine
// 2I + ROFF + t, t < NR, with weight R[t]; coarse indices within EDGE of a
// domain edge also receive mirrored contributions.
//   Stencil<false>: the 4-point offset constants (direct1 / adjoint)
//   Stencil<true>:  the quadratic a_coef / b_coef (direct2 / adjoint2)
template <bool Q>
struct Stencil;

template <>
struct Stencil<false> {
    static constexpr int NP = 4, NR = 8, ROFF = -3, EDGE = 2;
    static constexpr int OFF[2] = {-2, -1};
    static constexpr float C[2][4] = {{kQc[0], kQc[1], kQc[2], kQc[3]},
                                      {kPc[0], kPc[1], kPc[2], kPc[3]}};
    static constexpr float R[8] = {kW8[0], kW8[1], kW8[2], kW8[3],
                                   kW8[4], kW8[5], kW8[6], kW8[7]};
};

template <>
struct Stencil<true> {
    static constexpr int NP = 3, NR = 6, ROFF = -2, EDGE = 1;
    static constexpr int OFF[2] = {-1, -1};
    static constexpr float C[2][3] = {{5.f / 32.f, 30.f / 32.f, -3.f / 32.f},
                                      {-3.f / 32.f, 30.f / 32.f, 5.f / 32.f}};
    static constexpr float R[6] = {-3.f / 32.f, 5.f / 32.f, 30.f / 32.f,
                                   30.f / 32.f, 5.f / 32.f, -3.f / 32.f};
};

// Restriction weights of the two coarse indices next to a domain edge over
// the first 8 fine indices: R = P^T of the stencil S (fine i reads coarse
// (i >> 1) + S::OFF[i & 1] + t with weight S::C[i & 1][t]), the coarse
// indices mirrored (half-sample) into the domain. The far edge is the mirror
// image: R[nc-1-I][nf-1-t] = R[I][t] (both stencils are symmetric).
struct EdgeRW {
    float w[2][8];
};
template <bool Q>
constexpr EdgeRW make_edge_rw() {
    EdgeRW e{};
    for (int i = 0; i < 8; ++i) {
        const int par = i & 1, u = i >> 1;
        for (int t = 0; t < Stencil<Q>::NP; ++t) {
            int I = u + Stencil<Q>::OFF[par] + t;
            if (I < 0) I = -1 - I;
            if (I < 2) e.w[I][i] += Stencil<Q>::C[par][t];
        }
    }
    return e;
}
constexpr EdgeRW kEdgeRW = make_edge_rw<false>();    // 4-point (direct1 / adjoint)
constexpr EdgeRW kEdgeRWQ = make_edge_rw<true>();    // quadratic (direct2 / adjoint2)

}  // namespace mgbf
