// Radial 2-D beta filter kernel (template), instantiated per half-width H in
// mgbf_radial_inst.hip.
#pragma once

#include "mgbf_device.hpp"
#include "mgbf_launch.hpp"

namespace mgbf {

template <int H>
__host__ __device__ constexpr int disc_extent(int b) {
    int a = H;
    while (a >= 0 && a * a + b * b >= (H + 1) * (H + 1)) --a;
    return a;
}

// ---------------------------------------------------------------------------
// Radial 2-D beta filter F (gather form, half-sample mirror halo) with the
// epilogue of its generation:
//   s(k,i,j) = scale(k,i,j) * sum_{a,b} w(|a|,|b|) in(k, mx(i+a), my(j+b))
//              [+ (P coarse)(k,i,j)]                       (PRO)
//   out(k,i,j) = sum_d c(L,d) s(k+d,i,j) over the column    (VHB > 0)
// scale = wfield (per point) or wconst. One wave computes TI x TJ outputs for
// its 64 k: input rows stream through registers; for every input row the
// x-pair sums P_a = v(i+a)+v(i-a) give the per-row partial sums
// Z_b = sum_a w(a,b) P_a, which are added to the output rows j = r-b and r+b
// held in a ring of 2H+1 accumulator rows.
// PRO adds the prolongation of the next coarser generation (direct1 on the
// 2-point mirrored coarse field): a window of 4 x-interpolated coarse rows
// slides down with the output rows, one new coarse row every other row.
// VHB applies a vertical band of half-width VHB across the lanes of the wave
// (each wave then holds whole columns: lm divides 64 and K % 64 == 0).
// RES restricts the input field to the next generation on the way (R = P^T
// with the folded edges): every loaded fine row is restricted in x over the
// wave's 2 coarse columns (interior columns 2I-3 .. 2I+4 with kW8, or
// 2I-2 .. 2I+3 with the quadratic weights under QP; the 2 columns at each
// domain edge with the edge table) and added into a shift register of the
// coarse rows whose windows are open (4 or 3 in the interior, up to 6 next
// to an edge, with weights from a per-fine-row table there); a coarse row
// is stored when its window closes.
// QP selects the quadratic transfer (direct2 / adjoint2) for PRO or RES.
// The wave owns coarse rows j0/2 .. (j0+TJ)/2 - 1, whose windows lie inside
// the rows it loads (checked on the host).
// ---------------------------------------------------------------------------
template <int H, int TI, bool DISC, bool PRO, int VHB, bool RES, bool SEP = false, bool QP = false>
__global__ void __launch_bounds__(kBlock) radial_kernel(RadialParams p) {
    static_assert(!QP || PRO || (RES && VHB == 0), "the quadratic transfer rides in PRO or RES");
    static_assert(!RES || (TI == 4 && H >= 3 && !PRO && VHB == 0), "restriction epilogue shape");
    static_assert(!(SEP && DISC), "a separable filter is read as a box");
    constexpr int NR = 2 * H + 1;
    // columns loaded on each side of the tile: H for the filter, 4 for the
    // restriction's 8-point stencils at the domain edges
    constexpr int HL = (RES && H < 4) ? 4 : H;
    constexpr int NC = TI + 2 * HL;
    constexpr int NCC = TI / 2 + 4;   // coarse columns of the tile (TI even)
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = (int)xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wv;
    if (item >= p.nitems) return;
    const int K = p.K, n = p.n, m = p.m;
    const int ti = item % p.ntile_i;
    const int rest = item / p.ntile_i;
    const int tj = rest % p.ntile_j;
    const int kc = rest / p.ntile_j;
    const int k = p.klo + kc * kWave + lane;
    const bool kvalid = k < K;
    const int kk = kvalid ? k : K - 1;
    const int i0 = ti * TI, j0 = tj * p.TJ;
    const size_t rowstride = (size_t)n * K;

    // byte offsets: per lane k, per column wave-uniform (buffer addressing)
    const int vk = kk * 4;
    int soff[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) soff[c] = mirror_half(i0 - HL + c, n) * K * 4;
    // with the restriction state the column offsets move to VGPRs (per-lane
    // offsets vk + soff) where that ends the SGPR spills at the same
    // occupancy (from the compiler's register counts: disc H = 5, 6 and
    // separable H = 6; elsewhere the VGPRs would cost a wave per SIMD)
    constexpr bool VOFF = RES && ((DISC && (H == 5 || H == 6)) || (SEP && H == 6));
    int voff[VOFF ? NC : 1];
    if constexpr (VOFF) {
#pragma unroll
        for (int c = 0; c < NC; ++c) voff[c] = vk + soff[c];
    }

    float acc[NR][TI];
#pragma unroll
    for (int s = 0; s < NR; ++s)
#pragma unroll
        for (int t = 0; t < TI; ++t) acc[s][t] = 0.f;

    // prolongation window: 4 x-interpolated coarse rows, and the raw values
    // of the row that enters next, loaded one window step ahead
    float XV[4][TI];
    float cvn[NCC];
    int csoff[NCC];
    int cbase = 0;
    const int nc = n / 2, mc = m / 2;
    const size_t crow = (size_t)nc * K;
    auto load_craw = [&](int r, float* dst) {
        const __amdgpu_buffer_rsrc_t rc = buf_rsrc(p.coarse + (size_t)mirror_half(r, mc) * crow);
#pragma unroll
        for (int c = 0; c < NCC; ++c) dst[c] = buf_ld(rc, vk, csoff[c]);
    };
    // QP: the quadratic transfer (lquart): fine index i reads coarse
    // (i>>1)-1 .. (i>>1)+1 with Stencil<true>::C[i&1]; the window then starts
    // at (j>>1)-1 and its fourth row is not read
    constexpr bool quadp = QP;
    using SQ = Stencil<true>;
    auto wbase = [&](int j) { return quadp ? (j >> 1) - 1 : ((j + 1) >> 1) - 2; };
    auto xinterp = [&](const float* cv, float* dst) {
        if constexpr (quadp) {
#pragma unroll
            for (int t = 0; t < TI; ++t)
                dst[t] = SQ::C[t & 1][0] * cv[t / 2 + 1] + SQ::C[t & 1][1] * cv[t / 2 + 2] +
                         SQ::C[t & 1][2] * cv[t / 2 + 3];
            return;
        }
#pragma unroll
        for (int t = 0; t < TI; ++t) {
            if ((t & 1) == 0)
                dst[t] = kQc[0] * cv[t / 2] + kQc[1] * cv[t / 2 + 1] + kQc[2] * cv[t / 2 + 2] +
                         kQc[3] * cv[t / 2 + 3];
            else
                dst[t] = kPc[0] * cv[(t + 1) / 2] + kPc[1] * cv[(t + 1) / 2 + 1] +
                         kPc[2] * cv[(t + 1) / 2 + 2] + kPc[3] * cv[(t + 1) / 2 + 3];
        }
    };
    if constexpr (PRO) {
        const int v0 = i0 / 2;
#pragma unroll
        for (int c = 0; c < NCC; ++c) csoff[c] = mirror_half(v0 - 2 + c, nc) * K * 4;
        cbase = wbase(j0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            load_craw(cbase + u, cvn);
            xinterp(cvn, XV[u]);
        }
        // prefetch the next coarse row, except where the band epilogue
        // already presses on the registers
        if constexpr (VHB == 0) load_craw(cbase + 4, cvn);
    }

    // vertical band coefficients of this lane's level (identity past nfilt)
    constexpr int NVB = VHB > 0 ? 2 * VHB + 1 : 1;
    float vc[NVB];
    if constexpr (VHB > 0) {
        if (k < p.nfilt) {
            const float* row = p.band + (k % p.lm) * NVB;
#pragma unroll
            for (int d = 0; d < NVB; ++d) vc[d] = row[d];
        } else {
#pragma unroll
            for (int d = 0; d < NVB; ++d) vc[d] = d == VHB ? 1.f : 0.f;
        }
    }

    const int tj_eff = p.TJ < m - j0 ? p.TJ : m - j0;
    const int nsteps = tj_eff + 2 * H;

    // fused restriction state
    constexpr int NRC = RES ? TI / 2 : 1;   // coarse columns of the wave
    constexpr int NQ = 6;                    // open coarse rows at most
    float racc[NQ][NRC];
    int xcls = 0;                            // 0 interior, 1 left edge, 2 right edge
    const int Jlo = j0 / 2, Jhi = (j0 + tj_eff) / 2, nc2 = n / 2;
    if constexpr (RES) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int c = 0; c < NRC; ++c) racc[q][c] = 0.f;
        xcls = i0 == 0 ? 1 : (i0 + TI >= n ? 2 : 0);
    }

    for (int rb = 0; rb < nsteps; rb += NR) {
        // one step per ring row, unrolled at compile time: the ring index
        // must be a constant, and a loop the compiler leaves rolled (large
        // half-widths with the restriction) puts the ring in scratch memory
        static_for([&](auto S) __attribute__((always_inline)) {
            constexpr int s = decltype(S)::value;
            const int rr = rb + s;
            if (rr < nsteps) {
                const int r = mirror_half(j0 - H + rr, m);
                const __amdgpu_buffer_rsrc_t rs = buf_rsrc(p.in + (size_t)r * rowstride);
                float v[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) v[c] = VOFF ? buf_ld(rs, voff[c], 0) : buf_ld(rs, vk, soff[c]);
                if constexpr (RES) {
                    const int vr = j0 - H + rr;   // fine row, before mirroring
                    if (vr >= 0 && vr < m) {
                        using RS = Stencil<QP>;
                        float xr[NRC];
                        if (xcls == 0) {          // fine columns 2I+ROFF .. 2I+ROFF+NR-1
#pragma unroll
                            for (int c = 0; c < NRC; ++c) {
                                float a = 0.f;
#pragma unroll
                                for (int t = 0; t < RS::NR; ++t)
                                    a = __builtin_fmaf(RS::R[t], v[HL + RS::ROFF + 2 * c + t], a);
                                xr[c] = a;
                            }
                        } else if (xcls == 1) {   // fine columns 0 .. 7
#pragma unroll
                            for (int c = 0; c < NRC; ++c) {
                                float a = 0.f;
#pragma unroll
                                for (int t = 0; t < 8; ++t)
                                    a = __builtin_fmaf(QP ? kEdgeRWQ.w[c][t] : kEdgeRW.w[c][t], v[HL + t], a);
                                xr[c] = a;
                            }
                        } else {                  // fine columns n-8 .. n-1
#pragma unroll
                            for (int c = 0; c < NRC; ++c) {
                                float a = 0.f;
#pragma unroll
                                for (int t = 0; t < 8; ++t)
                                    a = __builtin_fmaf(QP ? kEdgeRWQ.w[NRC - 1 - c][7 - t] : kEdgeRW.w[NRC - 1 - c][7 - t],
                                                       v[HL - 4 + t], a);
                                xr[c] = a;
                            }
                        }
                        if (vr >= 8 && vr <= m - 9) {
                            // interior fine row: coarse rows jb .. jb+NO-1 are
                            // open (windows 2J+ROFF .. 2J+ROFF+NR-1). A row of
                            // parity CP is the last of row jb's window: it reads
                            // taps NR-1, NR-3, .. of the open windows and closes
                            // row jb; a row of the other parity reads taps
                            // NR-2, NR-4, .. (4-point: even rows close, quadratic:
                            // odd rows)
                            constexpr int NO = RS::NR / 2;
                            constexpr int CP = (RS::ROFF + RS::NR - 1) & 1;
                            const bool closes = (vr & 1) == CP;
                            float yw[NO];
#pragma unroll
                            for (int q = 0; q < NO; ++q)
                                yw[q] = closes ? RS::R[RS::NR - 1 - 2 * q] : RS::R[RS::NR - 2 - 2 * q];
#pragma unroll
                            for (int q = 0; q < NO; ++q)
#pragma unroll
                                for (int c = 0; c < NRC; ++c)
                                    racc[q][c] = __builtin_fmaf(yw[q], xr[c], racc[q][c]);
                            if (closes) {
                                const int J = (vr - RS::ROFF - RS::NR + 1) >> 1;
                                if (J >= Jlo && J < Jhi && kvalid) {
                                    const __amdgpu_buffer_rsrc_t ro =
                                        buf_rsrc(p.rout + (size_t)K * ((size_t)(i0 / 2) + (size_t)nc2 * J));
#pragma unroll
                                    for (int c = 0; c < NRC; ++c) {
                                        if (i0 / 2 + c < nc2) buf_st(racc[0][c], ro, vk, c * K * 4);
                                    }
                                }
                                // slots NO .. 5 stay empty between the edges
#pragma unroll
                                for (int q = 0; q + 1 < NO; ++q)
#pragma unroll
                                    for (int c = 0; c < NRC; ++c) racc[q][c] = racc[q + 1][c];
#pragma unroll
                                for (int c = 0; c < NRC; ++c) racc[NO - 1][c] = 0.f;
                            }
                        } else {
                            // rows next to an edge: weights and closing rows
                            // from the table
                            const float4 ya = p.rytab[2 * vr], yb = p.rytab[2 * vr + 1];
                            const float yw[NQ] = {ya.x, ya.y, ya.z, ya.w, yb.x, yb.y};
                            const int2 meta = p.rymeta[vr];
#pragma unroll
                            for (int q = 0; q < NQ; ++q)
#pragma unroll
                                for (int c = 0; c < NRC; ++c)
                                    racc[q][c] = __builtin_fmaf(yw[q], xr[c], racc[q][c]);
                            // coarse rows meta.x .. meta.x + meta.y - 1 close here
#pragma unroll
                            for (int e = 0; e < NQ; ++e) {
                                const int J = meta.x + e;
                                if (e < meta.y && J >= Jlo && J < Jhi && kvalid) {
                                    const __amdgpu_buffer_rsrc_t ro =
                                        buf_rsrc(p.rout + (size_t)K * ((size_t)(i0 / 2) + (size_t)nc2 * J));
#pragma unroll
                                    for (int c = 0; c < NRC; ++c) {
                                        if (i0 / 2 + c < nc2) buf_st(racc[e][c], ro, vk, c * K * 4);
                                    }
                                }
                            }
#pragma unroll
                            for (int sft = 1; sft <= NQ; ++sft) {
                                if (meta.y == sft) {
#pragma unroll
                                    for (int q = 0; q < NQ; ++q)
#pragma unroll
                                        for (int c = 0; c < NRC; ++c)
                                            racc[q][c] = q + sft < NQ ? racc[q + sft < NQ ? q + sft : 0][c] : 0.f;
                                }
                            }
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < TI; ++t) {
                    float P[H + 1];
                    P[0] = v[t + HL];
#pragma unroll
                    for (int a = 1; a <= H; ++a) P[a] = v[t + HL + a] + v[t + HL - a];
                    if constexpr (SEP) {
                        // w(a,b) = fx(a) fy(b): one x-sum per row, fy(b) into rows r+-b
                        float q = p.W.w[0][0] * P[0];
#pragma unroll
                        for (int a = 1; a <= H; ++a) q = __builtin_fmaf(p.W.w[0][a], P[a], q);
#pragma unroll
                        for (int b = 0; b <= H; ++b) {
                            acc[(s + b) % NR][t] = __builtin_fmaf(p.W.w[1][b], q, acc[(s + b) % NR][t]);
                            if (b > 0)
                                acc[(s + NR - b) % NR][t] =
                                    __builtin_fmaf(p.W.w[1][b], q, acc[(s + NR - b) % NR][t]);
                        }
                        continue;
                    }
#pragma unroll
                    for (int b = 0; b <= H; ++b) {
                        const int A = DISC ? disc_extent<H>(b) : H;
                        // on the disc w(a,b) = w(b,a): one kernel argument each
                        float z = p.W.w[DISC ? 0 : b][DISC ? b : 0] * P[0];
#pragma unroll
                        for (int a = 1; a <= H; ++a)
                            if (a <= A)
                                z = __builtin_fmaf(p.W.w[DISC && a < b ? a : b][DISC && a < b ? b : a],
                                                   P[a], z);
                        acc[(s + b) % NR][t] += z;
                        if (b > 0) acc[(s + NR - b) % NR][t] += z;
                    }
                }
                // output row j = j0 + rr - 2H is complete
                const int jrel = rr - 2 * H;
                constexpr int slot_off = NR - H;   // (s - H) mod NR
                if (jrel >= 0 && jrel < tj_eff) {
                    const int j = j0 + jrel;
                    const size_t rowoff = (size_t)j * rowstride + (size_t)i0 * K;
                    float o[TI];
                    if (p.wfield) {
                        const __amdgpu_buffer_rsrc_t rw = buf_rsrc(p.wfield + rowoff);
#pragma unroll
                        for (int t = 0; t < TI; ++t) {
                            const int tc = i0 + t < n ? t : n - 1 - i0;
                            o[t] = buf_ld(rw, vk, tc * K * 4) * acc[(s + slot_off) % NR][t];
                        }
                    } else {
#pragma unroll
                        for (int t = 0; t < TI; ++t) o[t] = p.wconst * acc[(s + slot_off) % NR][t];
                    }
                    if constexpr (PRO) {
                        const int nb = wbase(j);
                        if (nb != cbase) {   // the window moves down one coarse row
#pragma unroll
                            for (int u = 0; u < 3; ++u)
#pragma unroll
                                for (int t = 0; t < TI; ++t) XV[u][t] = XV[u + 1][t];
                            if constexpr (VHB == 0) {
                                xinterp(cvn, XV[3]);
                                load_craw(nb + 4, cvn);
                            } else {
                                load_craw(nb + 3, cvn);
                                xinterp(cvn, XV[3]);
                            }
                            cbase = nb;
                        }
                        const bool oj = (j & 1) != 0;
                        float y0, y1, y2, y3;
                        if constexpr (quadp) {
                            y0 = oj ? SQ::C[1][0] : SQ::C[0][0];
                            y1 = oj ? SQ::C[1][1] : SQ::C[0][1];
                            y2 = oj ? SQ::C[1][2] : SQ::C[0][2];
                            y3 = 0.f;
                        } else {
                            y0 = oj ? kPc[0] : kQc[0];
                            y1 = oj ? kPc[1] : kQc[1];
                            y2 = oj ? kPc[2] : kQc[2];
                            y3 = oj ? kPc[3] : kQc[3];
                        }
#pragma unroll
                        for (int t = 0; t < TI; ++t)
                            o[t] += y0 * XV[0][t] + y1 * XV[1][t] + y2 * XV[2][t] + y3 * XV[3][t];
                    }
                    if constexpr (VHB > 0) {
                        // the band over the lanes: level k+d sits in lane
                        // lane+d, reached by d single-lane wave shifts (DPP);
                        // lanes shifted in from outside the wave read 0 and
                        // meet a zero band coefficient, as do lanes of the
                        // neighbouring column
#pragma unroll
                        for (int t = 0; t < TI; ++t) {
                            float up = o[t], dn = o[t];
                            float rsum = vc[VHB] * o[t];
#pragma unroll
                            for (int d = 1; d <= VHB; ++d) {
                                up = lane_from_above(up);
                                dn = lane_from_below(dn);
                                rsum = __builtin_fmaf(vc[VHB + d], up, rsum);
                                rsum = __builtin_fmaf(vc[VHB - d], dn, rsum);
                            }
                            o[t] = rsum;
                        }
                    }
                    if (kvalid) {
                        const __amdgpu_buffer_rsrc_t ro = buf_rsrc(p.out + rowoff);
#pragma unroll
                        for (int t = 0; t < TI; ++t)
                            if (i0 + t < n) buf_st(o[t], ro, vk, t * K * 4);
                    }
                }
#pragma unroll
                for (int t = 0; t < TI; ++t) acc[(s + slot_off) % NR][t] = 0.f;
            }
        }, std::make_integer_sequence<int, NR>{});
    }
}

template <int H, int TI, bool DISC, bool PRO, int VHB, bool RES = false, bool SEP = false, bool QP = false>
hipError_t launch_radial_variant(const RadialArgs& a, hipStream_t st) {
    RadialParams p{};
    p.in = a.in;
    p.out = a.out;
    p.wfield = a.wfield;
    p.wconst = a.wconst;
    p.coarse = a.coarse;
    p.band = a.band;
    p.K = a.K;
    p.klo = a.klo;
    p.n = a.n;
    p.m = a.m;
    p.TJ = a.TJ;
    p.lm = a.lm;
    p.nfilt = a.nfilt;
    p.W = a.coef;
    p.rout = a.rout;
    p.rytab = a.rytab;
    p.rymeta = a.rymeta;
    p.ntile_i = (a.n + TI - 1) / TI;
    p.ntile_j = (a.m + a.TJ - 1) / a.TJ;
    const int nkc = (a.K - a.klo + kWave - 1) / kWave;
    const long long nitems = (long long)p.ntile_i * p.ntile_j * nkc;
    p.nitems = (int)nitems;
    const unsigned nblocks = (unsigned)((nitems + kWavesPerBlock - 1) / kWavesPerBlock);
    hipLaunchKernelGGL((radial_kernel<H, TI, DISC, PRO, VHB, RES, SEP, QP>), dim3(nblocks), dim3(kBlock), 0, st,
                       p);
    return hipGetLastError();
}

// Dispatch over the built variants for one H and tile width TI.
template <int H, int TI>
hipError_t launch_radial_ti(const RadialArgs& a, hipStream_t st) {
    const bool pro = a.coarse != nullptr;
    const int vhb = a.band ? a.vhb : 0;
    if (a.quad && pro) {   // the quadratic prolongation epilogue, with or without band
        if (a.separable) {
            switch (vhb) {
                case 0: return launch_radial_variant<H, TI, false, true, 0, false, true, true>(a, st);
                case 3: return launch_radial_variant<H, TI, false, true, 3, false, true, true>(a, st);
                case 4: return launch_radial_variant<H, TI, false, true, 4, false, true, true>(a, st);
                case 6: return launch_radial_variant<H, TI, false, true, 6, false, true, true>(a, st);
                case 8: return launch_radial_variant<H, TI, false, true, 8, false, true, true>(a, st);
                default: return hipErrorInvalidValue;
            }
        }
        if (!a.exact_disc) {
            if (vhb) return hipErrorInvalidValue;
            return launch_radial_variant<H, TI, false, true, 0, false, false, true>(a, st);
        }
        switch (vhb) {
            case 0: return launch_radial_variant<H, TI, true, true, 0, false, false, true>(a, st);
            case 3: return launch_radial_variant<H, TI, true, true, 3, false, false, true>(a, st);
            case 4: return launch_radial_variant<H, TI, true, true, 4, false, false, true>(a, st);
            case 6: return launch_radial_variant<H, TI, true, true, 6, false, false, true>(a, st);
            case 8: return launch_radial_variant<H, TI, true, true, 8, false, false, true>(a, st);
            default: return hipErrorInvalidValue;
        }
    }
    if (a.rout) {   // restriction epilogue: plain pass, 4 columns per lane, H >= 3
        if constexpr (TI == 4 && H >= 3) {
            if (pro || vhb) return hipErrorInvalidValue;
            if (a.quad) {
                if (a.separable) return launch_radial_variant<H, 4, false, false, 0, true, true, true>(a, st);
                return a.exact_disc ? launch_radial_variant<H, 4, true, false, 0, true, false, true>(a, st)
                                    : launch_radial_variant<H, 4, false, false, 0, true, false, true>(a, st);
            }
            if (a.separable) return launch_radial_variant<H, 4, false, false, 0, true, true>(a, st);
            return a.exact_disc ? launch_radial_variant<H, 4, true, false, 0, true>(a, st)
                                : launch_radial_variant<H, 4, false, false, 0, true>(a, st);
        } else {
            return hipErrorInvalidValue;
        }
    }
    if (a.separable) {
        if (vhb == 0) return pro ? launch_radial_variant<H, TI, false, true, 0, false, true>(a, st)
                                 : launch_radial_variant<H, TI, false, false, 0, false, true>(a, st);
        if (!pro) return hipErrorInvalidValue;
        switch (vhb) {
            case 3: return launch_radial_variant<H, TI, false, true, 3, false, true>(a, st);
            case 4: return launch_radial_variant<H, TI, false, true, 4, false, true>(a, st);
            case 6: return launch_radial_variant<H, TI, false, true, 6, false, true>(a, st);
            case 8: return launch_radial_variant<H, TI, false, true, 8, false, true>(a, st);
            default: return hipErrorInvalidValue;
        }
    }
    if (a.exact_disc) {
        if (vhb == 0) return pro ? launch_radial_variant<H, TI, true, true, 0>(a, st)
                                 : launch_radial_variant<H, TI, true, false, 0>(a, st);
        if (!pro) return hipErrorInvalidValue;
        switch (vhb) {
            case 3: return launch_radial_variant<H, TI, true, true, 3>(a, st);
            case 4: return launch_radial_variant<H, TI, true, true, 4>(a, st);
            case 6: return launch_radial_variant<H, TI, true, true, 6>(a, st);
            case 8: return launch_radial_variant<H, TI, true, true, 8>(a, st);
            default: return hipErrorInvalidValue;
        }
    }
    if (vhb != 0) return hipErrorInvalidValue;
    return pro ? launch_radial_variant<H, TI, false, true, 0>(a, st)
               : launch_radial_variant<H, TI, false, false, 0>(a, st);
}

template <int H>
hipError_t launch_radial_h(const RadialArgs& a, hipStream_t st) {
    switch (a.TI) {
        case 2: return launch_radial_ti<H, 2>(a, st);
        case 4: return launch_radial_ti<H, 4>(a, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mgbf
