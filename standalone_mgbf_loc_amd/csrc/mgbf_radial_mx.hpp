// Radial 2-D beta filter on the matrix cores (gfx950 v_mfma_f32_16x16x32_f16),
// with the epilogues of the vector kernel (weighting, prolongation, vertical
// band, restriction to the next generation).
//
// For one input row r the filter's contribution to output row r+-b is a
// banded Toeplitz product along x: Z_b(k, i) = sum_c v(k, c) T_b(c, i),
// T_b(c, i) = w(|c - H - i|, b) over a window of 16 + 2H <= 32 columns. A
// block of 4 wavefronts owns 16 output columns x 64 consecutive k x TJ
// output rows.
//
// Memory reaches the waves through LDS: each step's input row window (and,
// with PRO, the coarse row the prolongation window takes in) is fetched by
// direct 16-byte buffer-to-LDS loads NB - 1 steps ahead into a per-wave
// slot, and a step waits with an explicit vector-memory count for its own
// slot only. Loads complete in order, so every load the loop waits for is
// issued in step order; weight fields are therefore left to the vector
// kernel (a per-point weight would be one more load per step).
//
// Phase A (matrix cores), wave w: k0 + 16w .. +15 is the MFMA's M, the
// window columns its K, the 16 output columns its N. Lane l reads k =
// l % 16 at window columns 8(l/16) .. +7 (the A operand layout), splits the
// values, scaled by a power of two 2^e(k), into FP16 hi + lo (22 bits), and
// forms Z_b = A_hi T_hi + A_lo T_hi + A_hi T_lo with FP32 accumulation; T_b
// (hi/lo, scaled by 2^wexp) stays in registers for the whole launch. Z_b is
// added into a ring of 2H+1 output-row accumulators in the MFMA's C layout
// (lane l: k = 4(l/16) .. +3 of the tile at output column l % 16). The
// exponent e(k) starts high and is lowered, with the ring rescaled by the
// exact power of two, whenever a loaded row would exceed 2^14 after
// scaling; it is undone exactly at the output (ldexp).
// With RES the same A operand meets the restriction's x weights (multiples
// of 1/128, or 1/32 for the quadratic transfer, exact in FP16) in two more
// MFMAs: the input row restricted in
// x over the block's 8 coarse columns.
// A finished output row (and the x-restricted row) is unscaled, takes the
// constant weight and the prolongation of the next coarser generation
// (PRO) in the C layout, and goes to LDS. QP selects the quadratic
// transfer between generations (lquart) for the prolongation and the
// restriction.
//
// Phase B (after one barrier per row), wave w owns output columns
// i0 + 4w .. +3 for all 64 k, lane = k: the vertical band (VHB: the levels
// k+d read from the LDS rows of the block, which holds whole columns), the
// y-restriction of the x-restricted rows into the open coarse rows
// (coarse columns i0/2 + 2w, +1), and stores of 256 contiguous bytes per
// column. The LDS tiles are double-buffered, so one barrier per row orders
// both phases.
#pragma once

#include <type_traits>
#include <utility>

#include "mgbf_device.hpp"
#include "mgbf_launch.hpp"

namespace mgbf {

typedef _Float16 mx_half8 __attribute__((ext_vector_type(8)));
typedef float mx_float4 __attribute__((ext_vector_type(4)));

constexpr int kMxCols = 16;        // output columns per block
constexpr int kMxWin = 32;         // window columns (the MFMA's K)
constexpr int kMxPitch = 20;       // floats per k row of the LDS output tile (80 B)
constexpr int kMxRPitch = 68;      // floats per coarse column of the LDS x-restricted tile
constexpr int kMxSlots = 6;        // LDS slots per wave (5 steps in flight)
constexpr int kMxCCols = 12;       // coarse columns a wave's prolongation reads
constexpr int kMxExInit = 100;     // exponent before the first nonzero row
constexpr int kMxExMin = -120;

__device__ __forceinline__ mx_float4 mx_ld4(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(mx_float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// 16 bytes per lane from the buffer straight into LDS: lane l's bytes land
// at byte lds + 16 l. Completion is counted by the vector memory counter
// only, and the caller waits for it explicitly. The LDS address arrives as
// an integer: the compiler then treats the load as writing any LDS (so it
// keeps LDS accesses in order around it) but does not track it as a store
// to one variable, which would make it wait for every load in flight
// before later LDS reads.
__device__ __forceinline__ void mx_ld_lds16(__amdgpu_buffer_rsrc_t r, unsigned lds, int voff) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        r, (__attribute__((address_space(3))) void*)(size_t)__builtin_amdgcn_readfirstlane(lds), 16, voff, 0, 0,
        0);
}

// byte address of an LDS location
__device__ __forceinline__ unsigned mx_lds_addr(const float* x) {
    return (unsigned)(size_t)(const __attribute__((address_space(3))) float*)x;
}

// f(integral_constant<int, I>) for I = 0 .. N-1, unrolled at compile time
template <class F, int... I>
__device__ __forceinline__ void mx_static_for(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}

// Makes a value loaded before the row loop count as consumed there: the
// empty asm reads (and "rewrites") its registers, so the compiler waits for
// the load once, here. A load still pending at the loop's back edge is
// otherwise waited for inside every row step, with a count that also drains
// the row prefetch issued after it.
template <class T>
__device__ __forceinline__ void mx_settle(T& x) {
    asm volatile("" : "+v"(x));
}

// Waits until at most n vector memory operations of the wave are
// outstanding (n rounded down to a multiple of 4, at most 60; n is
// wave-uniform). Loads, direct-to-LDS loads and stores leave the vector
// memory counter in the order they were issued.
__device__ __forceinline__ void mx_wait_vmem(int n) {
    switch (n < 60 ? n >> 2 : 15) {
        case 15: asm volatile("s_waitcnt vmcnt(60)" ::: "memory"); break;
        case 14: asm volatile("s_waitcnt vmcnt(56)" ::: "memory"); break;
        case 13: asm volatile("s_waitcnt vmcnt(52)" ::: "memory"); break;
        case 12: asm volatile("s_waitcnt vmcnt(48)" ::: "memory"); break;
        case 11: asm volatile("s_waitcnt vmcnt(44)" ::: "memory"); break;
        case 10: asm volatile("s_waitcnt vmcnt(40)" ::: "memory"); break;
        case 9: asm volatile("s_waitcnt vmcnt(36)" ::: "memory"); break;
        case 8: asm volatile("s_waitcnt vmcnt(32)" ::: "memory"); break;
        case 7: asm volatile("s_waitcnt vmcnt(28)" ::: "memory"); break;
        case 6: asm volatile("s_waitcnt vmcnt(24)" ::: "memory"); break;
        case 5: asm volatile("s_waitcnt vmcnt(20)" ::: "memory"); break;
        case 4: asm volatile("s_waitcnt vmcnt(16)" ::: "memory"); break;
        case 3: asm volatile("s_waitcnt vmcnt(12)" ::: "memory"); break;
        case 2: asm volatile("s_waitcnt vmcnt(8)" ::: "memory"); break;
        case 1: asm volatile("s_waitcnt vmcnt(4)" ::: "memory"); break;
        default: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    }
}

template <int H, bool PRO, int VHB, bool RES, int NB, bool QP>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) radial_mx_kernel(RadialMxParams p) {
    static_assert(kMxCols + 2 * H <= kMxWin, "window exceeds one MFMA K step");
    static_assert(!RES || (H >= 3 && !PRO && VHB == 0), "restriction epilogue shape");
    // at most 3 (NB - 2) row and coarse-row loads and 4 (NB - 1) stores are
    // behind a row, and the waits count up to 60 of the counter's 63
    static_assert(NB >= 2 && 3 * (NB - 2) + 4 * (NB - 1) <= 60, "row slots");
    constexpr int NR = 2 * H + 1;
    constexpr int NORow = kWave + 2 * VHB;    // LDS output rows: 64 k plus zero pads
    __shared__ __align__(16) float lds_o[2][NORow * kMxPitch];
    __shared__ __align__(16) float lds_r[RES ? 2 : 1][RES ? 8 * kMxRPitch : 1];
    // per wave, NB slots: an input row window (32 columns x 16 k) and, with
    // PRO, the coarse row the step's window takes in (12 columns x 16 k)
    __shared__ __align__(16) float lds_a[kBlock / kWave][NB][kMxWin * 16];
    __shared__ __align__(16) float lds_c[PRO ? kBlock / kWave : 1][PRO ? NB : 1][PRO ? kMxCCols * 16 : 4];

    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g4 = lane >> 4, ln = lane & 15;
    const int K = p.K, n = p.n, m = p.m;
    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int ti = tile % p.ntile_i;
    const int rest = tile / p.ntile_i;
    const int tj = rest % p.ntile_j;
    const int kc = rest / p.ntile_j;
    const int i0 = ti * kMxCols, j0 = tj * p.TJ, k0 = kc * kWave;
    const size_t rowstride = (size_t)n * K;

    // the zero pads of the band rows (never written by the rows themselves)
    if constexpr (VHB > 0) {
        for (int q = threadIdx.x; q < 2 * VHB * kMxPitch; q += kBlock) {
            const int r = q / kMxPitch, c = q % kMxPitch;
            const int row = r < VHB ? r : kWave + r;
            lds_o[0][row * kMxPitch + c] = 0.f;
            lds_o[1][row * kMxPitch + c] = 0.f;
        }
    }

    // ---- phase A roles ----
    // The input window of a row reaches LDS by two direct 16-byte loads:
    // window column c = 8 g + e (g = lane / 16 of the A layout, e its value)
    // sits at position 4 e + g, 16 k each, so lane l reads value e at
    // 64 e + l (64 consecutive floats per read, no bank conflict). Load q,
    // lane l: the 4 k from 4 (l % 4) at the column of position 16 q + l / 4.
    int vr[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int pos = 16 * q + (lane >> 2);
        const int c = 8 * (pos & 3) + (pos >> 2);
        vr[q] = (k0 + 16 * w + 4 * (lane & 3)) * 4 + mirror_half(i0 - H + c, n) * K * 4;
    }
    const int kC = k0 + 16 * w + 4 * g4;      // first of the 4 k of the C layout
    const int iC = i0 + ln;                   // output column of the C layout

    mx_half8 bh[H + 1], bl[H + 1];
    {
        const mx_half8* bf = reinterpret_cast<const mx_half8*>(p.bfrag);
#pragma unroll
        for (int b = 0; b <= H; ++b) {
            bh[b] = bf[(b * 2 + 0) * kWave + lane];
            bl[b] = bf[(b * 2 + 1) * kWave + lane];
        }
#pragma unroll
        for (int b = 0; b <= H; ++b) {
            mx_settle(bh[b]);
            mx_settle(bl[b]);
        }
    }
    mx_half8 rx;
    if constexpr (RES) {
        const int cls = i0 == 0 ? 1 : (i0 + kMxCols >= n ? 2 : 0);
        rx = reinterpret_cast<const mx_half8*>(p.rfrag)[cls * kWave + lane];
        mx_settle(rx);
    }

    // scaling state: A lane exponent, and per C row the output exponents
    int ex = kMxExInit;
    float sc = __builtin_amdgcn_ldexpf(1.f, ex);
    float thr = __builtin_amdgcn_ldexpf(1.f, 14 - ex);
    int eC[4];   // output exponent of the C row: -e(k) - wexp
#pragma unroll
    for (int q = 0; q < 4; ++q) eC[q] = -ex - p.wexp;

    mx_float4 acc[NR];
#pragma unroll
    for (int s = 0; s < NR; ++s) acc[s] = mx_float4{0.f, 0.f, 0.f, 0.f};

    // prolongation window (C layout): NP x-interpolated coarse rows of the
    // lane's 4 k at its fine column (fine i reads coarse (i >> 1) + OFF[i & 1]
    // .. +NP-1 with the stencil's weights C[i & 1]). The lane's coarse
    // columns cb .. cb+NP-1 are columns cci .. of the wave's 12 (i0/2 - 2 ..
    // i0/2 + 9), which a step's coarse row brings to LDS in one 16-byte
    // load: lane f < 48 takes the 4 k from 4 (f % 4) at column f / 4.
    using PS = Stencil<QP>;
    constexpr int NP = PS::NP;
    const int nc = n / 2, mc = m / 2;
    const size_t crow = (size_t)nc * K;
    mx_float4 XV[NP];
    float cxw[NP];
    int cci = 0, voc = 0;
    // the window's first coarse row for output row j: (j >> 1) + OFF[j & 1]
    auto wbase = [](int j) { return QP ? (j >> 1) - 1 : ((j + 1) >> 1) - 2; };
    static_assert(PS::OFF[0] == (QP ? -1 : -2) && PS::OFF[1] == -1, "prolongation window offsets");
    if constexpr (PRO) {
        const int par = iC & 1;
        const int cb = (iC >> 1) + PS::OFF[par];
        cci = cb - (i0 / 2 - 2);
        int ccol[NP];
#pragma unroll
        for (int t = 0; t < NP; ++t) {
            ccol[t] = kC * 4 + mirror_half(cb + t, nc) * K * 4;
            cxw[t] = par ? PS::C[1][t] : PS::C[0][t];
        }
        voc = (k0 + 16 * w + 4 * (lane & 3)) * 4 + mirror_half(i0 / 2 - 2 + (lane >> 2), nc) * K * 4;
        const int cbase = wbase(j0);
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const __amdgpu_buffer_rsrc_t rc = buf_rsrc(p.coarse + (size_t)mirror_half(cbase + u, mc) * crow);
            mx_float4 x = mx_float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int t = 0; t < NP; ++t) x += cxw[t] * mx_ld4(rc, ccol[t], 0);
            XV[u] = x;
        }
    }

    // ---- phase B roles: lane = k, columns i0 + 4w .. +3 ----
    const int vkB = (k0 + lane) * 4;
    constexpr int NVB = VHB > 0 ? 2 * VHB + 1 : 1;
    float vc[NVB];
    if constexpr (VHB > 0) {
        const int k = k0 + lane;
        if (k < p.nfilt) {
            const float* row = p.band + (k % p.lm) * NVB;
#pragma unroll
            for (int d = 0; d < NVB; ++d) vc[d] = row[d];
        } else {
#pragma unroll
            for (int d = 0; d < NVB; ++d) vc[d] = d == VHB ? 1.f : 0.f;
        }
#pragma unroll
        for (int d = 0; d < NVB; ++d) mx_settle(vc[d]);
    }

    const int tj_eff = p.TJ < m - j0 ? p.TJ : m - j0;
    const int nsteps = tj_eff + 2 * H;

    // fused restriction state (phase B layout): coarse columns i0/2 + 2w + c
    constexpr int NQ = 6;
    float racc[NQ][2];
    const int Jlo = j0 / 2, Jhi = (j0 + tj_eff) / 2;
    if constexpr (RES) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) racc[q][0] = racc[q][1] = 0.f;
    }

    // Step rr reads input row rr (j0 - H + rr, mirrored) and has an output
    // row j0 + rr - 2H when 2H <= rr < 2H + tj_eff. With PRO the window of 4
    // coarse rows moves down one row at output rows whose base wbase(j)
    // differs from the previous row's.
    const int olo = 2 * H, ohi = 2 * H + tj_eff;
    const unsigned lds_a0 = mx_lds_addr(&lds_a[w][0][0]);
    const unsigned lds_c0 = PRO ? mx_lds_addr(&lds_c[w][0][0]) : 0u;
    auto cmove_at = [&](int rr) __attribute__((always_inline)) {
        const int j = j0 + rr - 2 * H;
        return rr > olo && rr < ohi && wbase(j) != wbase(j - 1);
    };
    // Everything a step reads from memory reaches LDS NB - 1 steps ahead,
    // the coarse row before the input row, so that the wait for the input
    // row covers both: the loads complete in order, and waiting for a load
    // issued later would also wait for every row issued before it.
    auto issue_step = [&](int rr, int sl) __attribute__((always_inline)) {
        if constexpr (PRO) {
            if (cmove_at(rr)) {
                const int j = j0 + rr - 2 * H;
                const __amdgpu_buffer_rsrc_t rc =
                    buf_rsrc(p.coarse + (size_t)mirror_half(wbase(j) + NP - 1, mc) * crow);
                if (lane < kMxCCols * 4) mx_ld_lds16(rc, lds_c0 + sl * (kMxCCols * 16 * 4), voc);
            }
        }
        const __amdgpu_buffer_rsrc_t rs = buf_rsrc(p.in + (size_t)mirror_half(j0 - H + rr, m) * rowstride);
        const unsigned la = lds_a0 + sl * (kMxWin * 16 * 4);
        mx_ld_lds16(rs, la, vr[0]);
        mx_ld_lds16(rs, la + kWave * 16, vr[1]);
    };
    // Vector memory operations issued after input row rr's loads, counted
    // when step rr waits for it (a lower bound): the later rows' loads and
    // the 4 output stores of each step with an output row from the step
    // that issued row rr on.
    auto outputs_in = [&](int a, int b) __attribute__((always_inline)) {   // steps with an output row in [a, b)
        const int lo = a > olo ? a : olo, hi = b < ohi ? b : ohi;
        return hi > lo ? hi - lo : 0;
    };
    auto behind = [&](int rr) __attribute__((always_inline)) {
        const int last = rr + NB - 2 < nsteps - 1 ? rr + NB - 2 : nsteps - 1;
        const int a = rr - NB + 1 > 0 ? rr - NB + 1 : 0;
        return 2 * (last - rr) + 4 * outputs_in(a, rr);
    };
    for (int d = 0; d < NB - 1 && d < nsteps; ++d) issue_step(d, d);
    int rslot = 0;   // slot of the step

    // One row step: s is the step's place in the ring (rr % NR), FAST a
    // block of NR steps inside the tile where every step has an output row,
    // issues the row NB - 1 steps ahead, waits behind the steady count of
    // later operations and (RES) reads a row away from the domain edges.
    constexpr int kSteady = 2 * (NB - 2) + 4 * (NB - 1);
    auto step = [&](const int rr, auto S, auto F) __attribute__((always_inline)) {
        constexpr int s = decltype(S)::value;
        constexpr bool FAST = decltype(F)::value;
        const int buf = rr & 1;
        const int jrel = rr - 2 * H;
        const bool orow = FAST || (jrel >= 0 && jrel < tj_eff);
        const int j = j0 + jrel;
        const int slot = rslot;
        // steps are scheduled one by one: interleaving them would hold two
        // steps' operands at once
        __builtin_amdgcn_sched_barrier(0);
        // ================= phase A =================
        mx_wait_vmem(FAST ? kSteady : behind(rr));
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = lds_a[w][slot][e * kWave + lane];
        // into the slot of the previous step, read there
        if (FAST || rr + NB - 1 < nsteps) issue_step(rr + NB - 1, slot == 0 ? NB - 1 : slot - 1);
        rslot = slot == NB - 1 ? 0 : slot + 1;

        float mx = fabsf(v[0]);
#pragma unroll
        for (int e = 1; e < 8; ++e) mx = fmaxf(mx, fabsf(v[e]));
        if (__any(mx > thr)) {
            // a k of this wave leaves the FP16 range: lower its exponent so
            // that the row's max scales below 2^14 (the 4 lanes of a k agree
            // on the max), rescale its ring rows exactly. The ring is
            // rescaled as whole C tuples (two power-of-two factors, each a
            // normal float), which keeps the ring in its registers.
            float mk = fmaxf(mx, __shfl_xor(mx, 16));
            mk = fmaxf(mk, __shfl_xor(mk, 32));
            int exn = ex;
            if (mk > thr) {
                exn = 14 - __builtin_amdgcn_frexp_expf(mk);
                exn = exn < kMxExMin ? kMxExMin : exn;
            }
            const int dex = exn - ex;
            ex = exn;
            sc = __builtin_amdgcn_ldexpf(1.f, ex);
            thr = __builtin_amdgcn_ldexpf(1.f, 14 - ex);
            mx_float4 fa, fb;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int dq = __shfl(dex, 4 * g4 + q);
                eC[q] -= dq;
                fa[q] = __builtin_amdgcn_ldexpf(1.f, dq / 2);
                fb[q] = __builtin_amdgcn_ldexpf(1.f, dq - dq / 2);
            }
#pragma unroll
            for (int sl = 0; sl < NR; ++sl) acc[sl] = acc[sl] * fa * fb;
        }
        mx_half8 ah, al;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float xs = v[e] * sc;
            const _Float16 h = (_Float16)xs;
            ah[e] = h;
            al[e] = (_Float16)(xs - (float)h);
        }
        acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[0], acc[s], 0, 0, 0);
        acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[0], acc[s], 0, 0, 0);
        acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[0], acc[s], 0, 0, 0);
#pragma unroll
        for (int b = 1; b <= H; ++b) {
            mx_float4 z = mx_float4{0.f, 0.f, 0.f, 0.f};
            z = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[b], z, 0, 0, 0);
            z = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[b], z, 0, 0, 0);
            z = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[b], z, 0, 0, 0);
            acc[(s + b) % NR] += z;
            acc[(s + NR - b) % NR] += z;
        }
        const int vr = j0 - H + rr;   // the loaded fine row, before mirroring
        const bool vin = FAST || (vr >= 0 && vr < m);
        if constexpr (RES) {
            if (vin) {
                mx_float4 cr = mx_float4{0.f, 0.f, 0.f, 0.f};
                cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, rx, cr, 0, 0, 0);
                cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, rx, cr, 0, 0, 0);
                if (ln < 8) {
                    mx_float4 cu;
#pragma unroll
                    for (int q = 0; q < 4; ++q) cu[q] = __builtin_amdgcn_ldexpf(cr[q], eC[q] + p.wexp);
                    *reinterpret_cast<mx_float4*>(&lds_r[buf][ln * kMxRPitch + 16 * w + 4 * g4]) = cu;
                }
            }
        }
        constexpr int os = (s + NR - H) % NR;   // the ring row that completes
        if (orow) {
            mx_float4 o;
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = __builtin_amdgcn_ldexpf(acc[os][q], eC[q]);
            o *= p.wconst;
            if constexpr (PRO) {
                // the window moves past the first output row where its base
                // does: at odd rows for the 4-point stencil, at even ones
                // for the quadratic
                if (FAST ? (j & 1) != (PS::OFF[0] == PS::OFF[1] ? 1 : 0) : cmove_at(rr)) {
#pragma unroll
                    for (int u = 0; u + 1 < NP; ++u) XV[u] = XV[u + 1];
                    mx_float4 x = mx_float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int t = 0; t < NP; ++t)
                        x += cxw[t] * *reinterpret_cast<const mx_float4*>(&lds_c[w][slot][(cci + t) * 16 + 4 * g4]);
                    XV[NP - 1] = x;
                }
                const bool oj = (j & 1) != 0;
#pragma unroll
                for (int u = 0; u < NP; ++u) o += (oj ? PS::C[1][u] : PS::C[0][u]) * XV[u];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) lds_o[buf][(VHB + 16 * w + 4 * g4 + q) * kMxPitch + ln] = o[q];
        }
        acc[os] = mx_float4{0.f, 0.f, 0.f, 0.f};
        // this wave's LDS writes done, then the block's barrier (no fence:
        // it would also wait for the input rows in flight); nothing is
        // scheduled across it
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        // ================= phase B =================
        if constexpr (RES) {
            if (vin) {
                const float xr[2] = {lds_r[buf][(2 * w) * kMxRPitch + lane],
                                     lds_r[buf][(2 * w + 1) * kMxRPitch + lane]};
                using RS = Stencil<QP>;
                auto store_coarse = [&](int J, const float* val) {
                    if (J >= Jlo && J < Jhi) {
                        const __amdgpu_buffer_rsrc_t ro =
                            buf_rsrc(p.rout + (size_t)K * ((size_t)(i0 / 2 + 2 * w) + (size_t)nc * J));
                        buf_st(val[0], ro, vkB, 0);
                        buf_st(val[1], ro, vkB, K * 4);
                    }
                };
                constexpr int NO = RS::NR / 2;
                constexpr int CP = (RS::ROFF + RS::NR - 1) & 1;
                if (FAST || (vr >= 8 && vr <= m - 9)) {
                    // interior rows: the row enters the NO open coarse rows
                    // with the stencil's odd or even taps, and with the odd
                    // ones it closes the oldest
                    if ((vr & 1) == CP) {
#pragma unroll
                        for (int q = 0; q < NO; ++q) {
                            racc[q][0] = __builtin_fmaf(RS::R[RS::NR - 1 - 2 * q], xr[0], racc[q][0]);
                            racc[q][1] = __builtin_fmaf(RS::R[RS::NR - 1 - 2 * q], xr[1], racc[q][1]);
                        }
                        store_coarse((vr - RS::ROFF - RS::NR + 1) >> 1, racc[0]);
#pragma unroll
                        for (int q = 0; q + 1 < NO; ++q) {
                            racc[q][0] = racc[q + 1][0];
                            racc[q][1] = racc[q + 1][1];
                        }
                        racc[NO - 1][0] = racc[NO - 1][1] = 0.f;
                    } else {
#pragma unroll
                        for (int q = 0; q < NO; ++q) {
                            racc[q][0] = __builtin_fmaf(RS::R[RS::NR - 2 - 2 * q], xr[0], racc[q][0]);
                            racc[q][1] = __builtin_fmaf(RS::R[RS::NR - 2 - 2 * q], xr[1], racc[q][1]);
                        }
                    }
                } else {
                    const float4 ya = p.rytab[2 * vr], yb = p.rytab[2 * vr + 1];
                    const float yw[NQ] = {ya.x, ya.y, ya.z, ya.w, yb.x, yb.y};
                    const int2 meta = p.rymeta[vr];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        racc[q][0] = __builtin_fmaf(yw[q], xr[0], racc[q][0]);
                        racc[q][1] = __builtin_fmaf(yw[q], xr[1], racc[q][1]);
                    }
#pragma unroll
                    for (int e = 0; e < NQ; ++e)
                        if (e < meta.y) store_coarse(meta.x + e, racc[e]);
#pragma unroll
                    for (int sft = 1; sft <= NQ; ++sft) {
                        if (meta.y == sft) {
#pragma unroll
                            for (int q = 0; q < NQ; ++q) {
                                racc[q][0] = q + sft < NQ ? racc[q + sft < NQ ? q + sft : 0][0] : 0.f;
                                racc[q][1] = q + sft < NQ ? racc[q + sft < NQ ? q + sft : 0][1] : 0.f;
                            }
                        }
                    }
                }
            }
        }
        if (orow) {
            float out[4];
            if constexpr (VHB > 0) {
#pragma unroll
                for (int t = 0; t < 4; ++t) out[t] = 0.f;
#pragma unroll
                for (int d = 0; d < NVB; ++d) {
                    const float4 r4 = *reinterpret_cast<const float4*>(&lds_o[buf][(lane + d) * kMxPitch + 4 * w]);
                    out[0] = __builtin_fmaf(vc[d], r4.x, out[0]);
                    out[1] = __builtin_fmaf(vc[d], r4.y, out[1]);
                    out[2] = __builtin_fmaf(vc[d], r4.z, out[2]);
                    out[3] = __builtin_fmaf(vc[d], r4.w, out[3]);
                }
            } else {
                const float4 r4 = *reinterpret_cast<const float4*>(&lds_o[buf][lane * kMxPitch + 4 * w]);
                out[0] = r4.x;
                out[1] = r4.y;
                out[2] = r4.z;
                out[3] = r4.w;
            }
            const __amdgpu_buffer_rsrc_t ro = buf_rsrc(p.out + (size_t)j * rowstride + (size_t)i0 * K);
#pragma unroll
            for (int t = 0; t < 4; ++t) buf_st(out[t], ro, vkB, (4 * w + t) * K * 4);
        }
    };
    // Interior blocks start at a multiple of NR in [fast_lo, fast_hi - NR].
    int fast_lo = olo + NB - 1, fast_hi = nsteps - NB + 1;
    if constexpr (RES) {
        fast_lo = fast_lo > 8 + H - j0 ? fast_lo : 8 + H - j0;
        fast_hi = fast_hi < m - 8 + H - j0 ? fast_hi : m - 8 + H - j0;
    }
    using Steps = std::make_integer_sequence<int, NR>;
    auto general_block = [&](int rb) __attribute__((always_inline)) {
        mx_static_for(
            [&](auto S) __attribute__((always_inline)) {
                if (rb + decltype(S)::value < nsteps) step(rb + decltype(S)::value, S, std::false_type{});
            },
            Steps{});
    };
    // the interior blocks form one run: general blocks before and after it
    int rb = 0;
    for (; rb < nsteps && !(rb >= fast_lo && rb + NR <= fast_hi); rb += NR) general_block(rb);
    for (; rb + NR <= fast_hi; rb += NR)
        mx_static_for([&](auto S) __attribute__((always_inline)) { step(rb + decltype(S)::value, S, std::true_type{}); },
                      Steps{});
    for (; rb < nsteps; rb += NR) general_block(rb);
}

template <int H, bool PRO, int VHB, bool RES, bool QP>
hipError_t launch_radial_mx_variant(const RadialArgs& a, hipStream_t st) {
    RadialMxParams p{};
    p.in = a.in;
    p.out = a.out;
    p.wconst = a.wconst;
    p.coarse = a.coarse;
    p.band = a.band;
    p.bfrag = a.bfrag;
    p.rfrag = a.rfrag;
    p.wexp = a.wexp;
    p.K = a.K;
    p.n = a.n;
    p.m = a.m;
    p.TJ = a.TJ;
    p.lm = a.lm;
    p.nfilt = a.nfilt;
    p.rout = a.rout;
    p.rytab = a.rytab;
    p.rymeta = a.rymeta;
    p.ntile_i = a.n / kMxCols;
    p.ntile_j = (a.m + a.TJ - 1) / a.TJ;
    const long long nblocks = (long long)p.ntile_i * p.ntile_j * (a.K / kWave);
    hipLaunchKernelGGL((radial_mx_kernel<H, PRO, VHB, RES, kMxSlots, QP>), dim3((unsigned)nblocks), dim3(kBlock), 0,
                       st, p);
    return hipGetLastError();
}

// The pass kinds with a transfer epilogue, for the 4-point (QP = false) or
// the quadratic transfer.
template <int H, bool QP>
hipError_t launch_radial_mx_epi(const RadialArgs& a, hipStream_t st) {
    const bool pro = a.coarse != nullptr;
    const int vhb = a.band ? a.vhb : 0;
    if (a.rout) {
        if constexpr (H >= 3) {
            if (pro || vhb || !a.rfrag || a.n < 2 * kMxCols || (a.TJ & 1)) return hipErrorInvalidValue;
            return launch_radial_mx_variant<H, false, 0, true, QP>(a, st);
        } else {
            return hipErrorInvalidValue;
        }
    }
    if (vhb) {
        // the band's coefficients and rows do not fit beside the ring and
        // the weight fragments of half-width 8 (they would spill)
        if constexpr (H >= kMxMaxBandH) {
            return hipErrorInvalidValue;
        } else {
            if (a.lm < 1 || kWave % a.lm) return hipErrorInvalidValue;
            switch (vhb) {
                case 3: return launch_radial_mx_variant<H, true, 3, false, QP>(a, st);
                case 4: return launch_radial_mx_variant<H, true, 4, false, QP>(a, st);
                case 6: return launch_radial_mx_variant<H, true, 6, false, QP>(a, st);
                case 8: return launch_radial_mx_variant<H, true, 8, false, QP>(a, st);
                default: return hipErrorInvalidValue;
            }
        }
    }
    return launch_radial_mx_variant<H, true, 0, false, QP>(a, st);
}

// The matrix-core engine for half-width H: the shapes it is built for are
// checked here (every k-chunk and column tile full, 32-bit row offsets).
template <int H>
hipError_t launch_radial_mx_h(const RadialArgs& a, hipStream_t st) {
    if constexpr (kMxCols + 2 * H > kMxWin) {
        return hipErrorInvalidValue;
    } else {
        if (a.K % kWave || a.n % kMxCols || a.TJ < 1 || a.klo != 0 || !a.bfrag || a.wfield)
            return hipErrorInvalidValue;
        if ((long long)a.n * a.K * 4 >= (1ll << 31)) return hipErrorInvalidValue;
        const bool pro = a.coarse != nullptr;
        if (a.band && !pro) return hipErrorInvalidValue;
        if (!pro && !a.rout) return launch_radial_mx_variant<H, false, 0, false, false>(a, st);
        return a.quad ? launch_radial_mx_epi<H, true>(a, st) : launch_radial_mx_epi<H, false>(a, st);
    }
}

}  // namespace mgbf
