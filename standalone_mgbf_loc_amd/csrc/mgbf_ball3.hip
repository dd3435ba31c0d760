// The 3-D filter of procedures 4 and 5 on the 3-D variables, in the
// streaming form of the radial kernel. One translation unit per horizontal
// half-width: compiled once for each MGBF_BALL_H in 1..4.
//
// The taps are sign-symmetric in every axis, w(a,b,c) = w(|a|,|b|,|c|), so
// with D_c the plane filter of the taps at level offset +-c and Z_c the
// level operator that adds the two mirrored shifts by +-c (Z_0 = I),
//   F = sum_c D_c (x) Z_c,    F^T = sum_c D_c (x) Z_c^T
// (the plane filters are symmetric under the half-sample mirror). Level and
// plane operators commute, so every loaded input row is first paired in the
// level direction, X_c = Z_c x (or Z_c^T x), across the lanes of the wave
// (a wave holds whole columns: lm divides 64), and then filtered in the
// plane like the radial kernel does: x-pair sums of X_c, per-row partial
// sums over the taps of every c, a ring of 2H+1 accumulator rows.
//
// Level pairs by DPP lane shifts: with ext the column extended by the
// whole-sample mirror (or by zeros, ZL), U_d(L) = ext(L+d) and
// D_d(L) = ext(L-d) follow from U_{d-1} and D_{d-1} by one shift each:
//   U_d = (L == lm-1) ? D_{d-1}(L-1) : U_{d-1}(L+1)   (0 under ZL)
//   D_d = (L == 0)    ? U_{d-1}(L+1) : D_{d-1}(L-1)   (0 under ZL)
// and Z_c x = U_c + D_c. The transpose of the mirrored pair differs from it
// in four lanes of a column only:
//   Z_c^T x = (1 + [L = lm-1-c] - [L = 0]) U_c + (1 + [L = c] - [L = lm-1]) D_c
// (lm >= 2 HZ + 2 keeps the four lanes apart). Under ZL, Z_c^T = Z_c.
#include "mgbf_device.hpp"
#include "mgbf_launch.hpp"

#ifndef MGBF_BALL_H
#error "MGBF_BALL_H must be defined"
#endif

namespace mgbf {

namespace {

// The ball's taps w(a,b,c) = wr[a^2 + b^2 + c^2] group by m = a^2 + c^2:
// ball_first: (a, c) is the first pair of its m in the loop order (c, then
// a); ball_m_used: some pair a <= H, c <= HZ has a^2 + c^2 = m.
template <int H, int HZ>
__device__ constexpr bool ball_first(int a, int c) {
    for (int c2 = 0; c2 <= c; ++c2)
        for (int a2 = 0; a2 <= H; ++a2) {
            if (c2 == c && a2 >= a) return true;
            if (a2 * a2 + c2 * c2 == a * a + c * c) return false;
        }
    return true;
}

template <int H, int HZ>
__device__ constexpr bool ball_m_used(int m) {
    for (int c = 0; c <= HZ; ++c)
        for (int a = 0; a <= H; ++a)
            if (a * a + c * c == m) return true;
    return false;
}

template <int H, int HZ, bool BALL, bool T, bool ZL, bool PRO, bool RES = false, bool QR = false>
__global__ void __launch_bounds__(kBlock) ball3_kernel(Ball3Params p) {
    static_assert(!RES || !PRO, "restriction epilogue shape");
    static_assert(!QR || RES, "QR selects the quadratic restriction (lquart)");
    constexpr int TI = 4;
    constexpr int NR = 2 * H + 1;
    // columns loaded on each side: H for the filter, 4 for the restriction's
    // 8-point stencils at the domain edges
    constexpr int HL = (RES && H < 4) ? 4 : H;
    constexpr int NC = TI + 2 * HL;
    // rows loaded on each side: H for the filter, 3 for the restriction's
    // 8-row windows; the D extra rows feed the restriction only
    constexpr int HR = (RES && H < 3) ? 3 : H;
    constexpr int D = HR - H;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = (int)xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wv;
    if (item >= p.nitems) return;
    const int K = p.K, n = p.n, m = p.m, lm = p.lm;
    const int ti = item % p.ntile_i;
    const int rest = item / p.ntile_i;
    const int tj = rest % p.ntile_j;
    const int kc = rest / p.ntile_j;
    const int k = kc * kWave + lane;
    const bool kvalid = k < p.K3;
    const int kk = k < K ? k : K - 1;
    const int i0 = ti * TI, j0 = tj * p.TJ;
    const size_t rowstride = (size_t)n * K;

    // this lane's level and the lanes of its column's ends
    const int L = k % lm;
    const bool top = L == lm - 1, bot = L == 0;
    float cu[HZ + 1], cd[HZ + 1];
#pragma unroll
    for (int c = 1; c <= HZ; ++c) {
        cu[c] = 1.f + (L == lm - 1 - c ? 1.f : 0.f) - (bot ? 1.f : 0.f);
        cd[c] = 1.f + (L == c ? 1.f : 0.f) - (top ? 1.f : 0.f);
    }

    const int vk = kk * 4;
    int soff[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) soff[c] = mirror_half(i0 - HL + c, n) * K * 4;

    float acc[NR][TI];
#pragma unroll
    for (int s = 0; s < NR; ++s)
#pragma unroll
        for (int t = 0; t < TI; ++t) acc[s][t] = 0.f;

    // PRO: the prolongation of the next coarser generation (direct1 on the
    // half-sample mirrored coarse field) added to the output, as the radial
    // kernel does: a window of 4 x-interpolated coarse rows slides down with
    // the output rows, the next raw coarse row loaded one step ahead
    constexpr int NCC = TI / 2 + 4;
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
    auto xinterp = [&](const float* cv, float* dst) {
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
        cbase = ((j0 + 1) >> 1) - 2;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            load_craw(cbase + u, cvn);
            xinterp(cvn, XV[u]);
        }
        load_craw(cbase + 4, cvn);
    }

    const int tj_eff = p.TJ < m - j0 ? p.TJ : m - j0;
    const int nsteps = tj_eff + 2 * HR;

    // RES: the restriction of the input to the next generation rides on the
    // loaded rows, as in the radial kernel: each fine row restricted in x
    // over the wave's 2 coarse columns, then added into a shift register of
    // the open coarse rows (4 inside, up to 6 next to an edge; per-row
    // weights from a table there); a coarse row is stored when it closes
    constexpr int NRC = RES ? TI / 2 : 1;
    constexpr int NQ = 6;
    float racc[NQ][NRC];
    int xcls = 0;   // 0 interior, 1 left edge, 2 right edge
    const int Jlo = j0 / 2, Jhi = (j0 + tj_eff) / 2, nc2 = n / 2;
    if constexpr (RES) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int c = 0; c < NRC; ++c) racc[q][c] = 0.f;
        xcls = i0 == 0 ? 1 : (i0 + TI >= n ? 2 : 0);
    }
    auto store_coarse = [&](int J, int q) {
        const __amdgpu_buffer_rsrc_t ro = buf_rsrc(p.rout + (size_t)K * ((size_t)(i0 / 2) + (size_t)nc2 * J));
#pragma unroll
        for (int c = 0; c < NRC; ++c)
            if (i0 / 2 + c < nc2) buf_st(racc[q][c], ro, vk, c * K * 4);
    };
    for (int rb = 0; rb < nsteps; rb += NR) {
#pragma unroll
        for (int s = 0; s < NR; ++s) {
            const int rr = rb + s;
            if (rr < nsteps) {
                const int r = mirror_half(j0 - HR + rr, m);
                const __amdgpu_buffer_rsrc_t rs = buf_rsrc(p.in + (size_t)r * rowstride);
                float vraw[NC];
#pragma unroll
                for (int col = 0; col < NC; ++col) vraw[col] = buf_ld(rs, vk, soff[col]);
                // filter row f of the ring (the first D loaded rows feed the restriction only)
                const int f = rr - D;
                const bool frow = f >= 0 && f < tj_eff + 2 * H;
                // level pairs X[c][col] of the loaded row
                float X[HZ + 1][NC];
#pragma unroll
                for (int col = 0; col < NC; ++col) {
                    const float v = vraw[col];
                    X[0][col] = v;
                    float u = v, d = v;
#pragma unroll
                    for (int c = 1; c <= HZ; ++c) {
                        const float fa = lane_from_above(u);
                        const float fb = lane_from_below(d);
                        if constexpr (ZL) {
                            u = top ? 0.f : fa;
                            d = bot ? 0.f : fb;
                        } else {
                            u = top ? fb : fa;
                            d = bot ? fa : fb;
                        }
                        if constexpr (T && !ZL)
                            X[c][col] = __builtin_fmaf(cu[c], u, cd[c] * d);
                        else
                            X[c][col] = u + d;
                    }
                }
                if constexpr (RES) {
                    const int vr = j0 - HR + rr;   // fine row, before mirroring
                    if (vr >= 0 && vr < m) {
                        using RS = Stencil<QR>;
                        float xr[NRC];
                        if (xcls == 0) {          // fine columns 2I+ROFF .. 2I+ROFF+NR-1
#pragma unroll
                            for (int c = 0; c < NRC; ++c) {
                                float a = 0.f;
#pragma unroll
                                for (int t = 0; t < RS::NR; ++t)
                                    a = __builtin_fmaf(RS::R[t], vraw[HL + RS::ROFF + 2 * c + t], a);
                                xr[c] = a;
                            }
                        } else if (xcls == 1) {   // fine columns 0 .. 7
#pragma unroll
                            for (int c = 0; c < NRC; ++c) {
                                float a = 0.f;
#pragma unroll
                                for (int t = 0; t < 8; ++t)
                                    a = __builtin_fmaf(QR ? kEdgeRWQ.w[c][t] : kEdgeRW.w[c][t], vraw[HL + t], a);
                                xr[c] = a;
                            }
                        } else {                  // fine columns n-8 .. n-1
#pragma unroll
                            for (int c = 0; c < NRC; ++c) {
                                float a = 0.f;
#pragma unroll
                                for (int t = 0; t < 8; ++t)
                                    a = __builtin_fmaf(QR ? kEdgeRWQ.w[NRC - 1 - c][7 - t] : kEdgeRW.w[NRC - 1 - c][7 - t],
                                                       vraw[HL - 4 + t], a);
                                xr[c] = a;
                            }
                        }
                        if (vr >= 8 && vr <= m - 9) {
                            // interior fine row: coarse rows jb .. jb+NO-1 are
                            // open; a row of parity CP closes row jb and reads
                            // taps NR-1, NR-3, .. of their windows, a row of
                            // the other parity taps NR-2, NR-4, ..
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
                                for (int c = 0; c < NRC; ++c) racc[q][c] = __builtin_fmaf(yw[q], xr[c], racc[q][c]);
                            if (closes) {
                                const int J = (vr - RS::ROFF - RS::NR + 1) >> 1;
                                if (J >= Jlo && J < Jhi && kvalid) store_coarse(J, 0);
#pragma unroll
                                for (int q = 0; q + 1 < NO; ++q)
#pragma unroll
                                    for (int c = 0; c < NRC; ++c) racc[q][c] = racc[q + 1][c];
#pragma unroll
                                for (int c = 0; c < NRC; ++c) racc[NO - 1][c] = 0.f;
                            }
                        } else {
                            // rows next to an edge: weights and closing rows from the table
                            const float4 ya = p.rytab[2 * vr], yb = p.rytab[2 * vr + 1];
                            const float yw[NQ] = {ya.x, ya.y, ya.z, ya.w, yb.x, yb.y};
                            const int2 meta = p.rymeta[vr];
#pragma unroll
                            for (int q = 0; q < NQ; ++q)
#pragma unroll
                                for (int c = 0; c < NRC; ++c) racc[q][c] = __builtin_fmaf(yw[q], xr[c], racc[q][c]);
#pragma unroll
                            for (int e = 0; e < NQ; ++e) {
                                const int J = meta.x + e;
                                if (e < meta.y && J >= Jlo && J < Jhi && kvalid) store_coarse(J, e);
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
                    if (!frow) break;
                    if constexpr (BALL) {
                        // w(a,b,c) = wr[a^2 + b^2 + c^2] on the ball: the
                        // pair sums are first added by m = a^2 + c^2, which
                        // shares the weight of every b, then Z_b = sum_m
                        // wr[m + b^2] S_m
                        constexpr int R2 = (H + 1) * (H + 1);   // a^2 + b^2 + c^2 < R2
                        constexpr int MM = H * H + HZ * HZ + 1;
                        float S[MM];
#pragma unroll
                        for (int c = 0; c <= HZ; ++c)
#pragma unroll
                            for (int a = 0; a <= H; ++a) {
                                const int mac = a * a + c * c;
                                if (mac >= R2) continue;
                                const float P = a == 0 ? X[c][t + HL] : X[c][t + HL + a] + X[c][t + HL - a];
                                if (ball_first<H, HZ>(a, c))
                                    S[mac] = P;
                                else
                                    S[mac] += P;
                            }
                        float Z[H + 1];
#pragma unroll
                        for (int b = 0; b <= H; ++b) {
                            Z[b] = 0.f;
#pragma unroll
                            for (int mm = 0; mm < MM; ++mm)
                                if (ball_m_used<H, HZ>(mm) && mm + b * b < R2)
                                    Z[b] = __builtin_fmaf(p.W.wr[mm + b * b], S[mm], Z[b]);
                        }
#pragma unroll
                        for (int b = 0; b <= H; ++b) {
                            acc[(s + NR - D + b) % NR][t] += Z[b];
                            if (b > 0) acc[(s + 2 * NR - D - b) % NR][t] += Z[b];
                        }
                    } else {
                        // the box of line products w(a,b,c) = fx(a) fy(b) fz(c):
                        // R = sum_c fz(c) sum_a fx(a) P_a(X_c), row r+-b gets fy(b) R
                        float R = 0.f;
#pragma unroll
                        for (int c = 0; c <= HZ; ++c) {
                            float q = p.W.fx[0] * X[c][t + HL];
#pragma unroll
                            for (int a = 1; a <= H; ++a)
                                q = __builtin_fmaf(p.W.fx[a], X[c][t + HL + a] + X[c][t + HL - a], q);
                            R = __builtin_fmaf(p.W.fz[c], q, R);
                        }
#pragma unroll
                        for (int b = 0; b <= H; ++b) {
                            acc[(s + NR - D + b) % NR][t] =
                                __builtin_fmaf(p.W.fy[b], R, acc[(s + NR - D + b) % NR][t]);
                            if (b > 0)
                                acc[(s + 2 * NR - D - b) % NR][t] =
                                    __builtin_fmaf(p.W.fy[b], R, acc[(s + 2 * NR - D - b) % NR][t]);
                        }
                    }
                }
                // output row j = j0 + f - 2H is complete
                const int jrel = f - 2 * H;
                constexpr int slot_off = 2 * NR - D - H;   // (f - H) mod NR
                if (frow && jrel >= 0 && jrel < tj_eff) {
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
                        const int nb = ((j + 1) >> 1) - 2;
                        if (nb != cbase) {   // the window moves down one coarse row
#pragma unroll
                            for (int u = 0; u < 3; ++u)
#pragma unroll
                                for (int t = 0; t < TI; ++t) XV[u][t] = XV[u + 1][t];
                            xinterp(cvn, XV[3]);
                            load_craw(nb + 4, cvn);
                            cbase = nb;
                        }
                        const bool oj = (j & 1) != 0;
                        const float y0 = oj ? kPc[0] : kQc[0], y1 = oj ? kPc[1] : kQc[1];
                        const float y2 = oj ? kPc[2] : kQc[2], y3 = oj ? kPc[3] : kQc[3];
#pragma unroll
                        for (int t = 0; t < TI; ++t)
                            o[t] += y0 * XV[0][t] + y1 * XV[1][t] + y2 * XV[2][t] + y3 * XV[3][t];
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
        }
    }
}

template <int H, int HZ, bool BALL, bool T, bool ZL, bool PRO, bool RES = false, bool QR = false>
hipError_t launch_variant(const Ball3Args& a, hipStream_t st) {
    Ball3Params p{};
    p.in = a.in;
    p.out = a.out;
    p.coarse = a.coarse;
    p.rout = a.rout;
    p.rytab = a.rytab;
    p.rymeta = a.rymeta;
    p.wfield = a.wfield;
    p.wconst = a.wconst;
    p.K = a.K;
    p.n = a.n;
    p.m = a.m;
    p.TJ = a.TJ;
    p.lm = a.lm;
    p.K3 = a.K3;
    p.W = a.coef;
    p.ntile_i = (a.n + 3) / 4;
    p.ntile_j = (a.m + a.TJ - 1) / a.TJ;
    const long long nkc = (a.K3 + kWave - 1) / kWave;
    const long long nitems = (long long)p.ntile_i * p.ntile_j * nkc;
    if (nitems >= (1ll << 31)) return hipErrorInvalidValue;
    p.nitems = (int)nitems;
    const unsigned nblocks = (unsigned)((nitems + kWavesPerBlock - 1) / kWavesPerBlock);
    hipLaunchKernelGGL((ball3_kernel<H, HZ, BALL, T, ZL, PRO, RES, QR>), dim3(nblocks), dim3(kBlock), 0, st, p);
    return hipGetLastError();
}

template <int H, int HZ>
hipError_t launch_hz(const Ball3Args& a, hipStream_t st) {
    const bool pro = a.coarse != nullptr;
    if (pro && (a.transpose || (a.n & 1) || (a.m & 1))) return hipErrorInvalidValue;
    if (a.rout && ((a.n & 1) || (a.m & 1) || (a.TJ & 1))) return hipErrorInvalidValue;
    if (a.zero_levels) {
        // zeros beyond the column ends: Z_c^T = Z_c, the box of the line products
        if (a.ball) return hipErrorInvalidValue;
        if (a.rout) {
            if (pro) return hipErrorInvalidValue;
            return a.quad ? launch_variant<H, HZ, false, false, true, false, true, true>(a, st)
                          : launch_variant<H, HZ, false, false, true, false, true>(a, st);
        }
        return pro ? launch_variant<H, HZ, false, false, true, true>(a, st)
                   : launch_variant<H, HZ, false, false, true, false>(a, st);
    }
    if (!a.ball || HZ != H) return hipErrorInvalidValue;
    if constexpr (HZ == H) {
        if (a.rout) {   // the transposed pass with the fused restriction
            if (!a.transpose || pro) return hipErrorInvalidValue;
            return a.quad ? launch_variant<H, HZ, true, true, false, false, true, true>(a, st)
                          : launch_variant<H, HZ, true, true, false, false, true>(a, st);
        }
        if (a.transpose) return launch_variant<H, HZ, true, true, false, false>(a, st);
        return pro ? launch_variant<H, HZ, true, false, false, true>(a, st)
                   : launch_variant<H, HZ, true, false, false, false>(a, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

#define MGBF_CAT2(x, y) x##y
#define MGBF_CAT(x, y) MGBF_CAT2(x, y)

hipError_t MGBF_CAT(launch_ball3_H, MGBF_BALL_H)(const Ball3Args& a, hipStream_t st) {
    constexpr int H = MGBF_BALL_H;
    switch (a.HZ) {
        case 0: return launch_hz<H, 0>(a, st);
        case 1: return launch_hz<H, 1>(a, st);
        case 2: return launch_hz<H, 2>(a, st);
        case 3: return launch_hz<H, 3>(a, st);
        case 4: return launch_hz<H, 4>(a, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mgbf
