// C ABI of the MGBF operator (declared in include/mgbf.h).
//
// The operator composition follows filtering_rad2_ens
// (src.mgbf/mg_filtering.f90:1000-1147) and filtering_rad2_bckg (:560-680):
//   u_1 = Vf^T x; u_{g+1} = R u_g            (sup_vrbeta1T_ens, upsending_ens)
//   z_g = F A_g F^T u_g                       (rbetaT, bocoT_2d, weighting_ens, boco_2d, rbeta)
//   s_gm = z_gm; s_g = z_g + P s_{g+1}        (downsending_ens)
//   y = Vf s_1                                (sup_vrbeta1_ens)
// On one device the reference's halo exchanges are the mirror boundary of
// the global domain and its up/down re-tiling is addressing, so every
// generation is one device array. F^T is applied as F: the radial filter is
// symmetric under the half-sample mirror when the aspect tensor is diagonal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "mgbf.h"
#include "mgbf_coef.hpp"
#include "mgbf_launch.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(MGBF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                          \
    do {                                             \
        hipError_t e_ = (expr);                      \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

// Device selection for the duration of one call.
struct DeviceGuard {
    int prev = -1;
    bool changed = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) {
            changed = hipSetDevice(dev) == hipSuccess;
        }
    }
    ~DeviceGuard() {
        if (changed) (void)hipSetDevice(prev);
    }
};

}  // namespace

// Window table of a transfer along the stacked k (launch_kwin).
struct VMap {
    const int* s = nullptr;
    const float* w = nullptr;
    int width = 0;
    // the same windows per 64-level chunk of the filter stack, for the tile
    // kernel that fuses the map (lo is null when they do not fit it)
    mgbf::XferVArgs fz{};
};

struct mgbf_ctx {
    mgbf::Setup s;
    int device = 0;
    int K = 0, gm = 0;
    int nfilt = 0;          // vertically filtered part of the stack (0: none)
    bool filt_g1 = true;
    // bumped by every call that changes what the launches read (weights,
    // tiles, engine choice): a group graph captured at an older version is
    // captured again
    unsigned long long version = 0;
    bool quad = false;      // lquart: quadratic inter-generation transfer
    int vf_last = 0;        // vertical band that ends W A W^T in the reference's order
    mgbf::RadialCoef coef{};
    int nx[mgbf::kMaxGen] = {}, ny[mgbf::kMaxGen] = {};
    long long gsize[mgbf::kMaxGen] = {};
    long long goff[mgbf::kMaxGen] = {};
    void* d_tables = nullptr;
    // vertical bands: 0 = Vf, 1 = Vf^T, 2 = Vf Vf^T
    const float* d_band[3] = {};
    // procedure 4 (filtering_rad3): tap lists of the 3-D filter and of the
    // 2-D filter of its 2-D variables
    bool rad3 = false;
    // the streaming 3-D filter on the 3-D variables (procedures 4 and 5),
    // when the taps and the column length allow it; else the gather
    bool ball3 = false;
    mgbf::Ball3Coef b3{};
    int b3H = 0, b3HZ = 0, b3ball = 0;
    const int4* d_t3 = nullptr;
    const float* d_w3 = nullptr;
    const int4* d_t2 = nullptr;
    const float* d_w2 = nullptr;
    mgbf::VBandW band_int[3] = {};
    int band_hb[3] = {}, band_ilo[3] = {}, band_ihi[3] = {};
    bool commute = true;     // weights constant along the stack: B = Vf Vf^T H
    // fusion into the radial passes: 0 none, 1 prolongation, 2 prolongation and
    // band (up-sweep epilogues), 3 also the restriction (down-sweep)
    int fuse = 3;
    // the generation-1 vertical band as a pass of its own after a kind-1
    // pass, where that is faster than the fused kind 2 (autotuner), or
    // pinned by a kind-2 tile with ti = 0
    bool band_split = false;
    // radial tiles per generation and pass kind (0 plain, 1 +P, 2 +P and band,
    // 3 plain with the restriction to the next generation)
    int tile_ti[mgbf::kMaxGen][4] = {};
    int tile_tj[mgbf::kMaxGen][4] = {};
    int tile_mfma[mgbf::kMaxGen][4] = {};   // 1: matrix-core engine for this pass
    // fused restriction: per-fine-row y tables, and whether generation g's
    // restriction has the structure the fused kernel assumes
    const float4* d_rytab[mgbf::kMaxGen] = {};
    const int2* d_rymeta[mgbf::kMaxGen] = {};
    bool res_ok[mgbf::kMaxGen] = {};
    // matrix-core engine: B fragments of the weight bands and, per
    // generation, of the x restriction (3 tile classes)
    const void* d_bfrag = nullptr;
    const void* d_rfrag[mgbf::kMaxGen] = {};
    bool mx_res_ok[mgbf::kMaxGen] = {};
    int wexp = 0;
    bool mfma_ok = false;                    // filter, half-width and stack fit the engine
    const int* d_rxs[mgbf::kMaxGen] = {};
    const float* d_rxw[mgbf::kMaxGen] = {};
    const int* d_rys[mgbf::kMaxGen] = {};
    const float* d_ryw[mgbf::kMaxGen] = {};
    const int* d_pxs[mgbf::kMaxGen] = {};
    const float* d_pxw[mgbf::kMaxGen] = {};
    const int* d_pys[mgbf::kMaxGen] = {};
    const float* d_pyw[mgbf::kMaxGen] = {};
    const float* wfield[mgbf::kMaxGen] = {};
    float wconst[mgbf::kMaxGen] = {};
    // per-point aspect tensors (mgbf_set_aspect): the weights of every
    // point's taps over the (2H+1)^2 box; F^T is then a filter of its own
    float* d_aw[mgbf::kMaxGen] = {};
    bool aniso[mgbf::kMaxGen] = {};
    // analysis <-> filter transfer: separable window tables (direct: filter ->
    // analysis, adjoint: its transpose) and the vertical k windows
    mgbf::XferArgs xf_dir{}, xf_adj{};
    // the same transfers by 8 x 8 output tiles, when their blocks fit
    bool xt_ok[2] = {false, false};   // [0] direct (filter -> analysis), [1] adjoint
    bool xfer_gather = false;         // MGBF_XFER_GATHER=1: the per-point gather instead
    bool xfer_split = false;          // MGBF_XFER_SPLIT=1: the vertical map as a pass of its own
    bool need_ws_a = false;           // a vertical map runs as a pass of its own
    mgbf::XferTileArgs xt[2]{};
    // vertical transfer of the _all2 (kv) and the _all (kw) routines:
    // direct (d) and adjoint (a) window tables; null when lm_a = lm or the
    // routine family cannot run
    VMap kv_d, kv_a, kw_d, kw_a;
    float* ws_a = nullptr;  // analysis grid x K, when the vertical transfer is not the identity
    float* ws_u = nullptr;  // multigrid-sized
    float* ws_t = nullptr;  // multigrid-sized
    double* d_dot = nullptr;
    // per-phase timers
    struct Rec {
        int phase;
        int gen;   // generation (0-based) of a per-generation launch, or -1
        hipEvent_t a, b;
        double bytes;
    };
    bool timing = false;
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double t_ms[MGBF_NTIMERS] = {};
    double t_bytes[MGBF_NTIMERS] = {};
    long long t_n[MGBF_NTIMERS] = {};
    // the same, split by generation
    double tg_ms[mgbf::kMaxGen][MGBF_NTIMERS] = {};
    double tg_bytes[mgbf::kMaxGen][MGBF_NTIMERS] = {};
    long long tg_n[mgbf::kMaxGen][MGBF_NTIMERS] = {};
};

namespace {

hipEvent_t take_event(mgbf_ctx* c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets the launches of one timer phase with events when timing is on.
struct Phase {
    mgbf_ctx* c;
    int ph;
    hipStream_t st;
    double bytes;
    int gen;
    hipEvent_t a = nullptr;
    Phase(mgbf_ctx* c_, int ph_, hipStream_t st_, double bytes_, int gen_ = -1)
        : c(c_), ph(ph_), st(st_), bytes(bytes_), gen(gen_) {
        if (c->timing) {
            a = take_event(c);
            if (a && hipEventRecord(a, st) != hipSuccess) {
                c->pool.push_back(a);
                a = nullptr;
            }
        }
    }
    ~Phase() {
        if (!a) return;
        hipEvent_t b = take_event(c);
        if (b && hipEventRecord(b, st) == hipSuccess) {
            c->recs.push_back({ph, gen, a, b, bytes});
        } else {
            c->pool.push_back(a);
            if (b) c->pool.push_back(b);
        }
    }
};

int drain_timers(mgbf_ctx* c) {
    int status = MGBF_OK;
    for (auto& r : c->recs) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess)
            status = fail(MGBF_ERR_HIP, "timer event failure");
        c->t_ms[r.phase] += ms;
        c->t_bytes[r.phase] += r.bytes;
        c->t_n[r.phase] += 1;
        if (r.gen >= 0 && r.gen < mgbf::kMaxGen) {
            c->tg_ms[r.gen][r.phase] += ms;
            c->tg_bytes[r.gen][r.phase] += r.bytes;
            c->tg_n[r.gen][r.phase] += 1;
        }
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->recs.clear();
    return status;
}

// Default rows per wave: the tallest tile that still gives 16384 waves.
int pick_tj(int K, int n, int m, int ti) {
    const long long nkc = (K + 63) / 64;
    const long long ntile_i = (n + ti - 1) / ti;
    int tj = m;
    while (tj > 16 && nkc * ntile_i * ((m + tj - 1) / tj) < 16384) tj = (tj + 1) / 2;
    return tj;
}

enum { BAND_VF = 0, BAND_VFT = 1, BAND_VVT = 2 };

// The inter-generation restriction along one axis, as the fused kernels
// apply it: interior coarse I reads fine 2I+roff .. 2I+roff+nr-1 with r[];
// the 2 coarse indices next to an edge read the first 8 fine indices with
// edge[I][] (R = P^T with the coarse indices of P mirrored half-sample into
// the domain; the far edge is the mirror image). The host twin of the
// kernels' compile-time stencils: 4-point (direct1) or quadratic (direct2).
struct ResStencil {
    int nr, roff;
    double r[8];
    double edge[2][8];
};

ResStencil res_stencil(bool quad) {
    ResStencil S{};
    const double kp[4] = {-9. / 128., 111. / 128., 29. / 128., -3. / 128.};
    const double kq[4] = {-3. / 128., 29. / 128., 111. / 128., -9. / 128.};
    const double ka[3] = {5. / 32., 30. / 32., -3. / 32.};
    const double kb[3] = {-3. / 32., 30. / 32., 5. / 32.};
    // fine i reads coarse (i >> 1) + off[i & 1] + t, t < np, with c[i & 1][t]
    const int np = quad ? 3 : 4;
    const int off[2] = {quad ? -1 : -2, -1};
    const double* c[2] = {quad ? ka : kq, quad ? kb : kp};
    S.nr = 2 * np;
    S.roff = quad ? -2 : -3;
    for (int i = 0; i < 2 * np + 8; ++i) {
        const int par = i & 1, u = i >> 1;
        for (int t = 0; t < np; ++t) {
            int I = u + off[par] + t;
            // interior weights from a coarse index far from the edges
            if (I == np && i - 2 * np - S.roff >= 0 && i - 2 * np - S.roff < S.nr)
                S.r[i - 2 * np - S.roff] += c[par][t];
            if (i >= 8) continue;
            if (I < 0) I = -1 - I;
            if (I < 2) S.edge[I][i] += c[par][t];
        }
    }
    return S;
}

// Does the restriction of a generation with nf x mf fine points have the
// shape the fused kernels assume? Along x (nf % 4 == 0) every coarse row of
// the table must equal, over the fine columns the kernel reads, the stencil
// of its tile class (edge, interior) and have no weight outside them. Along
// y the per-fine-row table (lowest open coarse row Jb, rows closing at v,
// weights into rows Jb .. Jb+5, padded to 8) is built here; a window closes
// at its last nonzero weight, every fine row must feed only those 6 rows
// (NR/2 in the interior, up to 6 next to an edge, where the edge rows share
// a window), and windows must open and close in row order.
bool build_res_tables(const mgbf::AxisTable& rx, const mgbf::AxisTable& ry, int nf, int mf, bool quad,
                      std::vector<float>& tab, std::vector<int>& meta) {
    if (nf % 4 || nf < 8 || mf < 8 || rx.width != 8 || ry.width != 8) return false;
    const ResStencil S = res_stencil(quad);
    const int nc = nf / 2, mc = mf / 2;
    auto close = [](double a, double b) { return std::fabs(a - b) <= 1e-12; };
    for (int I = 0; I < nc; ++I) {
        // the fine columns the kernel reads for I, and its weight on each
        int lo, len;
        if (I < 2) lo = 0, len = 8;
        else if (I >= nc - 2) lo = nf - 8, len = 8;
        else lo = 2 * I + S.roff, len = S.nr;
        auto kern = [&](int col) {
            if (I < 2) return S.edge[I][col];
            if (I >= nc - 2) return S.edge[nc - 1 - I][nf - 1 - col];
            return S.r[col - lo];
        };
        for (int t = 0; t < 8; ++t) {
            const int col = rx.start[I] + t;
            const double w = rx.w[(size_t)I * 8 + t];
            if (col < lo || col >= lo + len) {
                if (w != 0.0) return false;
            } else if (!close(w, kern(col))) {
                return false;
            }
        }
        for (int col = lo; col < lo + len; ++col) {
            const int t = col - rx.start[I];
            if ((t < 0 || t >= 8) && kern(col) != 0.0) return false;
        }
    }
    // window ends: the last fine row with a nonzero weight
    std::vector<int> end(mc);
    for (int J = 0; J < mc; ++J) {
        int e = ry.start[J];
        for (int t = 0; t < 8; ++t)
            if (ry.w[(size_t)J * 8 + t] != 0.0) e = ry.start[J] + t;
        end[J] = e;
    }
    for (int J = 1; J < mc; ++J)
        if (ry.start[J] < ry.start[J - 1] || end[J] < end[J - 1]) return false;
    tab.assign((size_t)mf * 8, 0.f);
    meta.assign((size_t)mf * 2, 0);
    for (int v = 0; v < mf; ++v) {
        int jb = 0, ne = 0;
        for (int J = 0; J < mc; ++J) {
            if (end[J] < v) ++jb;
            if (end[J] == v) ++ne;
        }
        for (int J = 0; J < mc; ++J) {
            if (v < ry.start[J] || v > end[J]) continue;
            if (J < jb || J > jb + 5) return false;
            tab[(size_t)v * 8 + (J - jb)] = (float)ry.w[(size_t)J * 8 + (v - ry.start[J])];
        }
        if (ne > 6) return false;
        meta[(size_t)v * 2] = jb;
        meta[(size_t)v * 2 + 1] = ne;
    }
    return true;
}

// Can the restriction of generation g ride in its first radial pass with
// output rows per wave TJ? The wave of output rows j0 .. j0+TJ-1 loads fine
// rows j0-H .. j0+TJ-1+H and owns coarse rows j0/2 .. (j0+TJ)/2-1, whose
// windows (8 rows, 6 for the quadratic transfer) must lie inside what it
// loads.
bool fused_ok(const mgbf_ctx* c, int g, int tj, int H) {
    if (g + 1 >= c->gm || !c->res_ok[g] || H < 3 || tj < 2 || (tj & 1)) return false;
    const int m = c->ny[g];
    // each coarse row's window: its first and last fine row of nonzero weight
    const mgbf::AxisTable& ry = c->s.ry[g];
    auto first = [&](int J) {
        for (int t = 0; t < ry.width; ++t)
            if (ry.w[(size_t)J * ry.width + t] != 0.0) return ry.start[J] + t;
        return ry.start[J];
    };
    auto last = [&](int J) {
        for (int t = ry.width - 1; t >= 0; --t)
            if (ry.w[(size_t)J * ry.width + t] != 0.0) return ry.start[J] + t;
        return ry.start[J];
    };
    for (int j0 = 0; j0 < m; j0 += tj) {
        const int te = tj < m - j0 ? tj : m - j0;
        for (int J = j0 / 2; J < (j0 + te) / 2; ++J)
            if (first(J) < j0 - H || last(J) > j0 + te - 1 + H) return false;
    }
    return true;
}

// Can the vertical band `which` be fused into the last radial pass of
// generation 1? Each wave must hold whole columns, and the variant built.
bool can_fuse_band(const mgbf_ctx* c, int which) {
    const int hb = c->band_hb[which];
    const int lm = c->s.cfg.lm;
    return c->nfilt > 0 && (c->s.exact_disc || c->s.separable) && c->gm > 1 && (64 % lm) == 0 &&
           (c->K % 64) == 0 &&
           (hb == 3 || hb == 4 || hb == 6 || hb == 8);
}

// Can pass `kind` of generation g run on the matrix-core engine? Full
// 16-column tiles and 64-k chunks, and for the restriction exact x
// fragments of 3 tile classes (either transfer stencil).
bool mx_ok(const mgbf_ctx* c, int g, int kind) {
    if (!c->mfma_ok || c->nx[g] % 16) return false;
    if (kind == 3) return c->mx_res_ok[g] && c->nx[g] >= 32 && c->s.H >= 3;
    if (kind == 2) return c->s.H < mgbf::kMxMaxBandH;
    return true;
}

// out = scale * F(in) [+ P coarse] [then vertical band `band`] on generation g
int radial(mgbf_ctx* c, int g, const float* in, float* out, const float* wfield, float wconst,
           hipStream_t st, int phase = MGBF_T_HFILT, const float* coarse = nullptr,
           int band = -1, float* rout = nullptr, bool transpose = false) {
    Phase ph(c, phase, st,
             4.0 * ((double)c->gsize[g] * (wfield ? 3 : 2) +
                    (coarse || rout ? (double)c->gsize[g + 1] : 0.0)),
             g);
    if (c->aniso[g]) {
        // per-point aspect tensors: F or F^T on its own, then the weighting;
        // no epilogue rides in it
        if (coarse || rout || band >= 0)
            return fail(MGBF_ERR_ARG, "internal: epilogue fused into a per-point aspect pass");
        mgbf::AspectArgs aa{in, out, c->d_aw[g], c->K, c->nx[g], c->ny[g], c->s.H};
        HIP_TRY(mgbf::launch_aspect(aa, transpose, st), "aspect filter launch");
        if (wfield || wconst != 1.0f)
            HIP_TRY(mgbf::launch_scale(out, out, wfield, wconst, 0, c->gsize[g], st), "scale launch");
        return MGBF_OK;
    }
    mgbf::RadialArgs a{};
    a.in = in;
    a.out = out;
    a.wfield = wfield;
    a.wconst = wconst;
    a.coarse = coarse;
    if (band >= 0) {
        a.band = c->d_band[band];
        a.vhb = c->band_hb[band];
    }
    a.lm = c->s.cfg.lm;
    a.nfilt = c->nfilt;
    a.K = c->K;
    a.n = c->nx[g];
    a.m = c->ny[g];
    a.H = c->s.H;
    const int kind = rout ? 3 : (band >= 0 ? 2 : (coarse ? 1 : 0));
    a.TI = c->tile_ti[g][kind];
    a.TJ = c->tile_tj[g][kind];
    if (rout) {
        a.rout = rout;
        a.rytab = c->d_rytab[g];
        a.rymeta = c->d_rymeta[g];
    }
    // the matrix-core engine takes constant weights: a per-point weight
    // field follows its pass as a scale pass of its own (the tuner times the
    // two together against the vector kernel, which reads the field in its
    // epilogue)
    const bool mx = c->tile_mfma[g][kind] && mx_ok(c, g, kind);
    if (mx) {
        a.mfma = true;
        a.bfrag = c->d_bfrag;
        a.rfrag = rout ? c->d_rfrag[g] : nullptr;
        a.wexp = c->wexp;
        a.wfield = nullptr;
        a.wconst = wfield ? 1.0f : wconst;
    }
    a.exact_disc = c->s.exact_disc;
    a.separable = c->s.separable;
    a.quad = c->quad;
    a.coef = c->coef;
    HIP_TRY(mgbf::launch_radial(a, st), "radial filter launch");
    if (mx && wfield)
        HIP_TRY(mgbf::launch_scale(out, out, wfield, wconst, 0, c->gsize[g], st), "scale launch");
    return MGBF_OK;
}

int restrict_g(mgbf_ctx* c, int g, const float* fine, float* coarse, hipStream_t st) {
    Phase ph(c, MGBF_T_UPSEND, st, 4.0 * (double)(c->gsize[g] + c->gsize[g + 1]), g);
    mgbf::TransferArgs a{fine, coarse, c->K, c->nx[g], c->ny[g],
                         c->d_rxs[g], c->d_rxw[g], c->d_rys[g], c->d_ryw[g], c->quad};
    HIP_TRY(mgbf::launch_restrict(a, st), "restriction launch");
    return MGBF_OK;
}

int prolong_g(mgbf_ctx* c, int g, const float* coarse, const float* addend, float* fine,
              hipStream_t st) {
    Phase ph(c, MGBF_T_DNSEND, st,
             4.0 * (double)(c->gsize[g + 1] + c->gsize[g] * (addend ? 2 : 1)), g);
    mgbf::TransferArgs a{coarse, fine, c->K, c->nx[g], c->ny[g],
                         c->d_pxs[g], c->d_pxw[g], c->d_pys[g], c->d_pyw[g], c->quad};
    HIP_TRY(mgbf::launch_prolong(a, addend, st), "prolongation launch");
    return MGBF_OK;
}

int vertical(mgbf_ctx* c, const float* in, float* out, int which, hipStream_t st) {
    Phase ph(c, which == BAND_VFT ? MGBF_T_VFILTT : MGBF_T_VFILT, st, 8.0 * (double)c->gsize[0], 0);
    if (c->nfilt == 0) {
        if (in != out) HIP_TRY(mgbf::launch_copy(in, out, c->gsize[0], st), "copy launch");
        return MGBF_OK;
    }
    mgbf::VBandArgs a{};
    a.in = in;
    a.out = out;
    a.band = c->d_band[which];
    a.wint = c->band_int[which];
    a.ilo = c->band_ilo[which];
    a.ihi = c->band_ihi[which];
    a.HB = c->band_hb[which];
    a.K = c->K;
    a.lm = c->s.cfg.lm;
    a.nfilt = c->nfilt;
    a.npts = (long long)c->nx[0] * c->ny[0];
    HIP_TRY(mgbf::launch_vband(a, st), "vertical filter launch");
    return MGBF_OK;
}

#define CALL(x)                      \
    do {                             \
        int st_ = (x);               \
        if (st_ != MGBF_OK) return st_; \
    } while (0)

// Restriction chain u_{g+1} = R u_g from u_1 (in U0 or a caller array).
int restrict_chain(mgbf_ctx* c, const float* u1, float* U, hipStream_t st) {
    for (int g = 0; g + 1 < c->gm; ++g)
        CALL(restrict_g(c, g, g == 0 ? u1 : U + c->goff[g], U + c->goff[g + 1], st));
    return MGBF_OK;
}

// Down-sweep: the first radial pass of every filtered generation,
// T_g = A_g F u_g (A_g = 1 unless `weighted`), written at T + goff[g], and
// the restriction u_{g+1} = R u_g into U + goff[g+1]. The restriction rides
// in the radial pass where the tiles allow (fusion level 3), else it runs on
// its own first. u_1 is read from u1.
int down_sweep(mgbf_ctx* c, const float* u1, float* U, float* T, bool weighted, hipStream_t st) {
    for (int g = 0; g < c->gm; ++g) {
        const float* ug = g == 0 ? u1 : U + c->goff[g];
        const bool last = g + 1 >= c->gm;
        const bool filt = g > 0 || c->filt_g1;
        float* next = last ? nullptr : U + c->goff[g + 1];
        const float* wf = weighted ? c->wfield[g] : nullptr;
        const float wc = weighted ? c->wconst[g] : 1.0f;
        if (filt && !last && c->fuse >= 3 && !c->aniso[g] && fused_ok(c, g, c->tile_tj[g][3], c->s.H)) {
            CALL(radial(c, g, ug, T + c->goff[g], wf, wc, st, MGBF_T_HFILTT, nullptr, -1, next));
            continue;
        }
        if (!last) CALL(restrict_g(c, g, ug, next, st));
        if (filt) CALL(radial(c, g, ug, T + c->goff[g], wf, wc, st, MGBF_T_HFILTT, nullptr, -1, nullptr, true));
    }
    return MGBF_OK;
}

// Up-sweep from the coarsest generation: for g = gm..1 (1-based)
//   s_g = F in_g + P s_{g+1},  in_1 = in1, in_g = inb + goff[g] (g >= 2)
// (in_g = T_g after a down-sweep for W A W^T; the control vector's v_g for W,
// scaled by A_g^{1/2} first when sqrt_w). s_g is written into U. The last
// pass writes y = band(s_1) directly when the vertical band fuses (or there
// is none), else s_1 goes to U0 and a separate vertical pass follows. An
// unfiltered generation 1 gives s_1 = u1 + P s_2.
int upsweep(mgbf_ctx* c, const float* in1, const float* inb, const float* u1, float* y, int band,
            bool sqrt_w, hipStream_t st) {
    float* U = c->ws_u;
    float* T = c->ws_t;
    // the radial kernels fuse the 4-point or the quadratic prolongation,
    // and the band after either
    const bool fuse = band >= 0 && c->fuse >= 2 && can_fuse_band(c, band) && !c->aniso[0] && !c->band_split;
    for (int g = c->gm - 1; g >= 0; --g) {
        const float* src = g == 0 ? in1 : inb + c->goff[g];
        const float* coarse = g + 1 < c->gm ? U + c->goff[g + 1] : nullptr;
        const float* pro = c->fuse >= 1 && !c->aniso[g] ? coarse : nullptr;
        const bool direct = g == 0 && (fuse || band < 0);
        float* dst = direct ? y : U + c->goff[g];
        const int fb = (g == 0 && fuse) ? band : -1;
        if (g == 0 && !c->filt_g1) {
            float* s1 = band >= 0 ? U : y;
            if (coarse) CALL(prolong_g(c, 0, coarse, u1, s1, st));
            else if (u1 != s1) HIP_TRY(mgbf::launch_copy(u1, s1, c->gsize[0], st), "copy launch");
            return band >= 0 ? vertical(c, U, y, band, st) : MGBF_OK;
        }
        if (sqrt_w) {
            HIP_TRY(mgbf::launch_scale(src, T, c->wfield[g], c->wconst[g], 1, c->gsize[g], st),
                    "scale launch");
            src = T;
        }
        CALL(radial(c, g, src, dst, nullptr, 1.0f, st, MGBF_T_HFILT, pro, fb));
        if (coarse && !pro) CALL(prolong_g(c, g, coarse, dst, dst, st));
    }
    if (band >= 0 && !fuse) return vertical(c, U, y, band, st);
    return MGBF_OK;
}

// W A W^T with weights that are the same along the stack: the horizontal
// multigrid operator H acts identically on every k and so commutes with the
// vertical filter, B = Vf H Vf^T = (Vf Vf^T) H; one banded vertical pass,
// fused into the last radial pass when possible.
int apply_wawt_commuted(mgbf_ctx* c, const float* x, float* y, hipStream_t st) {
    CALL(down_sweep(c, x, c->ws_u, c->ws_t, true, st));
    return upsweep(c, c->ws_t, c->ws_t, x, y, c->nfilt ? BAND_VVT : -1, false, st);
}

int apply_wawt(mgbf_ctx* c, const float* x, float* y, hipStream_t st) {
    bool fields = false;
    for (int g = 0; g < c->gm; ++g) fields = fields || c->wfield[g] != nullptr;
    if (!fields && c->commute) return apply_wawt_commuted(c, x, y, st);
    float* U = c->ws_u;
    CALL(vertical(c, x, U, BAND_VFT, st));
    CALL(down_sweep(c, U, U, c->ws_t, true, st));
    return upsweep(c, c->ws_t, c->ws_t, U, y, c->nfilt ? c->vf_last : -1, false, st);
}

int apply_wt(mgbf_ctx* c, const float* x, float* v, bool sqrt_weights, hipStream_t st) {
    float* U = c->ws_u;
    CALL(vertical(c, x, U, BAND_VFT, st));
    CALL(down_sweep(c, U, U, v, false, st));
    for (int g = 0; g < c->gm; ++g) {
        if (g == 0 && !c->filt_g1) {
            HIP_TRY(mgbf::launch_copy(U, v, c->gsize[0], st), "copy launch");
            continue;
        }
        if (sqrt_weights)
            HIP_TRY(mgbf::launch_scale(v + c->goff[g], v + c->goff[g], c->wfield[g], c->wconst[g], 1,
                                       c->gsize[g], st),
                    "scale launch");
    }
    return MGBF_OK;
}

// y = W v: the control vector's generations are filtered once and summed up
// the hierarchy, then Vf.
int apply_w(mgbf_ctx* c, const float* v, float* y, bool sqrt_weights, hipStream_t st) {
    return upsweep(c, v, v, v, y, c->nfilt ? BAND_VF : -1, sqrt_weights, st);
}

// ---------------------------------------------------------------------------
// Procedures 4 and 5, filtering_rad3 / filtering_lin3 (mg_filtering.f90:
// 123-557): every generation filtered by F (3-D variables over (i, j, L),
// 2-D variables in the plane), no separate vertical filter:
//   y = sum_g P^(g-1) F_g A_g F_g^T R^(g-1) x.
// ---------------------------------------------------------------------------
// Chunk tables of a vertical map fused into the tile kernel (XferVArgs)
// from its window table t over the stacked levels: the direct map
// (analysis levels from filter levels) or its adjoint. False when a window
// spans more than 4 levels, the direct windows leave a 64-level chunk of
// the filter stack or do not advance chunk by chunk, or a chunk's windows
// cover more than 128 analysis levels.
struct VFuseHost {
    std::vector<int> lo, cnt, r0;
    std::vector<double> w;
};

bool build_vfuse(const mgbf::AxisTable& t, int K, bool adjoint, VFuseHost& f) {
    constexpr int kChunk = 64;   // the lanes of a wave
    const int nkc = (K + kChunk - 1) / kChunk;
    if (t.n_out < 1 || (adjoint && t.n_out != K)) return false;
    f.lo.assign(nkc, 0);
    f.cnt.assign(nkc, 0);
    f.r0.assign(t.n_out, 0);
    f.w.assign((size_t)t.n_out * 4, 0.0);
    // first and last level of output o's nonzero taps (-1: none)
    auto span = [&](int o, int& first, int& last) {
        first = last = -1;
        for (int u = 0; u < t.width; ++u)
            if (t.w[(size_t)o * t.width + u] != 0.0) {
                if (first < 0) first = t.start[o] + u;
                last = t.start[o] + u;
            }
    };
    auto put = [&](int o, int first, int base) {
        f.r0[o] = first < 0 ? 0 : first - base;
        for (int u = 0; u < t.width; ++u) {
            const double wv = t.w[(size_t)o * t.width + u];
            if (wv != 0.0) f.w[(size_t)o * 4 + (t.start[o] + u - first)] = wv;
        }
    };
    if (!adjoint) {
        int cur = 0;
        for (int o = 0; o < t.n_out; ++o) {
            int first, last;
            span(o, first, last);
            if (first >= 0) {
                const int ch = first / kChunk;
                if (ch < cur || last / kChunk != ch || last - first >= 4 || last >= K) return false;
                cur = ch;
            }
            if (f.cnt[cur] == 0) f.lo[cur] = o;
            if (++f.cnt[cur] > 2 * kChunk) return false;
            put(o, first, kChunk * cur);
        }
        return true;
    }
    for (int ch = 0; ch < nkc; ++ch) {
        const int o0 = ch * kChunk, o1 = std::min(o0 + kChunk, K);
        int lo = -1, hi = -1;
        for (int o = o0; o < o1; ++o) {
            int first, last;
            span(o, first, last);
            if (first < 0) continue;
            if (last - first >= 4) return false;
            lo = lo < 0 ? first : std::min(lo, first);
            hi = std::max(hi, last);
        }
        if (hi < 0) continue;
        if (hi - lo + 1 > 2 * kChunk) return false;
        f.lo[ch] = lo;
        f.cnt[ch] = hi - lo + 1;
        for (int o = o0; o < o1; ++o) {
            int first, last;
            span(o, first, last);
            put(o, first, lo);
        }
    }
    return true;
}

// Dense per-tile blocks of a 1-D window table for the tiled transfer: tile
// q covers outputs q*TB .. q*TB+TB-1 and reads inputs s0[q] .. s0[q]+S-1
// with weights W[q][t][c]. S is the widest span over the tiles, rounded up
// to a built width for the x axis (round_x); false when S exceeds the input,
// the largest built width, or 64 rows.
bool tile_blocks(const mgbf::AxisTable& t, int nin, int TB, bool round_x, int& S, std::vector<int>& s0,
                 std::vector<double>& W) {
    const int nt = (t.n_out + TB - 1) / TB;
    if (nt < 1) return false;
    std::vector<int> lo(nt, nin), hi(nt, -1);
    for (int r = 0; r < t.n_out; ++r)
        for (int u = 0; u < t.width; ++u)
            if (t.w[(size_t)r * t.width + u] != 0.0) {
                const int col = t.start[r] + u;
                lo[r / TB] = std::min(lo[r / TB], col);
                hi[r / TB] = std::max(hi[r / TB], col);
            }
    int span = 1;
    for (int q = 0; q < nt; ++q)
        if (hi[q] >= lo[q]) span = std::max(span, hi[q] - lo[q] + 1);
    if (round_x) {
        const int sizes[6] = {8, 12, 16, 20, 24, 32};
        int r = 0;
        for (int z : sizes)
            if (!r && z >= span) r = z;
        if (!r) return false;
        span = r;
    } else if (span > 64) {
        return false;
    }
    if (span > nin) return false;
    S = span;
    s0.assign(nt, 0);
    W.assign((size_t)nt * TB * S, 0.0);
    for (int q = 0; q < nt; ++q) {
        int st = hi[q] >= lo[q] ? lo[q] : 0;
        if (st + S > nin) st = nin - S;
        s0[q] = st;
    }
    for (int r = 0; r < t.n_out; ++r) {
        const int q = r / TB, tt = r % TB;
        for (int u = 0; u < t.width; ++u) {
            const double w = t.w[(size_t)r * t.width + u];
            if (w == 0.0) continue;
            const int col = t.start[r] + u - s0[q];
            if (col < 0 || col >= S) return false;
            W[((size_t)q * TB + tt) * S + col] += w;
        }
    }
    return true;
}

// Plans the streaming 3-D filter: procedure 4's ball weights by |g|^2, or
// procedure 5's line factors; false (the gather then runs) when the taps,
// the half-widths or the column length are outside what it is built for.
bool plan_ball3(const mgbf::Setup& s, const mgbf_config& cfg, mgbf::Ball3Coef& co, int& H, int& HZ,
                int& ball) {
    co = mgbf::Ball3Coef{};
    const int lm = cfg.lm;
    if (s.dims.km3_all < 1 || lm < 1 || 64 % lm != 0) return false;   // whole columns per wavefront
    if (cfg.mgbf_proc == 4) {
        H = HZ = 0;
        for (const auto& t : s.taps3) {
            H = std::max(H, std::max(std::abs(t.gx), std::abs(t.gy)));
            HZ = std::max(HZ, std::abs(t.gz));
        }
        if (H < 1 || H != HZ || H > mgbf::kBall3MaxH) return false;
        std::vector<double> wr((H + 1) * (H + 1), 0.0);
        std::vector<char> seen(wr.size(), 0);
        for (const auto& t : s.taps3) {
            const int r2 = t.gx * t.gx + t.gy * t.gy + t.gz * t.gz;
            if (r2 >= (H + 1) * (H + 1)) return false;
            if (seen[r2] && std::fabs(wr[r2] - t.w) > 1e-12 * std::fabs(t.w)) return false;
            wr[r2] = t.w;
            seen[r2] = 1;
        }
        for (size_t r = 0; r < wr.size(); ++r) co.wr[r] = (float)wr[r];
        ball = 1;
    } else if (cfg.mgbf_proc == 5) {
        const std::vector<double>* f = s.hex_f;
        if (f[0].empty() || f[1].empty() || f[2].empty()) return false;
        H = std::max(1, (int)std::max(f[0].size(), f[1].size()) - 1);
        HZ = (int)f[2].size() - 1;
        if (H > mgbf::kBall3MaxH || HZ > mgbf::kBall3MaxHZ) return false;
        for (size_t a = 0; a < f[0].size(); ++a) co.fx[a] = (float)f[0][a];
        for (size_t b = 0; b < f[1].size(); ++b) co.fy[b] = (float)f[1][b];
        for (size_t z = 0; z < f[2].size(); ++z) co.fz[z] = (float)f[2][z];
        if (!s.zero_levels) return false;
        ball = 0;
    } else {
        return false;
    }
    return lm >= 2 * HZ + 2;
}

// Rows per wave of the radial pass over the 2-D variables (k >= K3) of
// procedures 4 and 5: even, so that a restriction can ride in it.
int rad2d_tj(const mgbf_ctx* c, int g) {
    const int K3 = c->s.dims.km3_all * c->s.cfg.lm;
    int tj = pick_tj(c->K - K3, c->nx[g], c->ny[g], 4);
    if (tj & 1) tj = tj + 1 <= c->ny[g] ? tj + 1 : tj - 1;
    return tj;
}

// The 3-D filter of generation g (F, or F^T with transpose). With `scaled`
// given, the output is also multiplied by the generation's weights when the
// streaming pass runs, and *scaled says whether it was.
// coarse (forward only, when ball3_pro(c)): the prolongation of generation
// g+1 is added to the output, s_g = F in + P s_{g+1}.
// rout (transposed only): the restriction of `in` to generation g+1.
// The 2-D variables after the 3-D ones (k >= K3) take the same epilogues in
// a radial pass of their own.
int ball(mgbf_ctx* c, int g, const float* in, float* out, bool transpose, hipStream_t st,
         bool* scaled = nullptr, const float* coarse = nullptr, float* rout = nullptr) {
    Phase ph(c, transpose ? MGBF_T_HFILTT : MGBF_T_HFILT, st,
             8.0 * (double)c->gsize[g] + (coarse || rout ? 4.0 * (double)c->gsize[g + 1] : 0.0), g);
    const int K3 = c->s.dims.km3_all * c->s.cfg.lm;
    if (scaled) *scaled = false;
    if (coarse && (transpose || !c->ball3))
        return fail(MGBF_ERR_ARG, "internal: prolongation fused into a 3-D pass that cannot take it");
    if (rout && (!transpose || !c->ball3))
        return fail(MGBF_ERR_ARG, "internal: restriction fused into a 3-D pass that cannot take it");
    if (c->ball3) {
        const bool fuse = scaled != nullptr;
        mgbf::Ball3Args b{};
        b.in = in;
        b.out = out;
        b.coarse = coarse;
        b.rout = rout;
        b.rytab = rout ? c->d_rytab[g] : nullptr;
        b.rymeta = rout ? c->d_rymeta[g] : nullptr;
        b.wfield = fuse ? c->wfield[g] : nullptr;
        b.wconst = fuse ? c->wconst[g] : 1.f;
        if (fuse) *scaled = true;
        b.K = c->K;
        b.n = c->nx[g];
        b.m = c->ny[g];
        b.lm = c->s.cfg.lm;
        b.K3 = K3;
        b.H = c->b3H;
        b.HZ = c->b3HZ;
        b.TJ = pick_tj(K3, b.n, b.m, 4);
        b.ball = c->b3ball;
        b.transpose = transpose ? 1 : 0;
        b.zero_levels = c->s.zero_levels ? 1 : 0;
        b.quad = c->quad ? 1 : 0;
        b.coef = c->b3;
        HIP_TRY(mgbf::launch_ball3(b, st), "3-D filter launch");
        if (K3 >= c->K) return MGBF_OK;
        // the 2-D variables (k >= K3): the procedure's 2-D filter (radial disc
        // or line product, even in x and y, so F^T = F) streams through the
        // radial kernel from k = K3 on, with the same weighting, restriction
        // or prolongation as the 3-D pass
        mgbf::RadialArgs r{};
        r.in = in;
        r.out = out;
        r.wfield = fuse ? c->wfield[g] : nullptr;
        r.wconst = fuse ? c->wconst[g] : 1.f;
        r.coarse = coarse;
        r.lm = c->s.cfg.lm;
        r.nfilt = c->nfilt;
        r.K = c->K;
        r.klo = K3;
        r.n = c->nx[g];
        r.m = c->ny[g];
        // the restriction needs 3 halo rows: a narrower filter runs at
        // half-width 3 with zero taps beyond its own (coef is zero there)
        r.H = rout ? std::max(c->s.H, 3) : c->s.H;
        r.TI = 4;
        r.TJ = rad2d_tj(c, g);
        if (rout) {
            r.rout = rout;
            r.rytab = c->d_rytab[g];
            r.rymeta = c->d_rymeta[g];
        }
        r.exact_disc = c->s.exact_disc;
        r.separable = c->s.separable;
        r.quad = c->quad;
        r.coef = c->coef;
        HIP_TRY(mgbf::launch_radial(r, st), "2-D filter launch");
        return MGBF_OK;
    }
    mgbf::BallArgs a{};
    a.in = in;
    a.out = out;
    a.t3 = c->d_t3;
    a.w3 = c->d_w3;
    a.n3 = (int)c->s.taps3.size();
    a.t2 = c->d_t2;
    a.w2 = c->d_w2;
    a.n2 = (int)c->s.taps2.size();
    a.K = c->K;
    a.n = c->nx[g];
    a.m = c->ny[g];
    a.lm = c->s.cfg.lm;
    a.K3 = c->s.dims.km3_all * c->s.cfg.lm;
    a.transpose = transpose ? 1 : 0;
    a.zero_levels = c->s.zero_levels ? 1 : 0;
    a.klo = c->ball3 ? K3 : 0;   // the 2-D variables only, after the streaming pass
    HIP_TRY(mgbf::launch_ball(a, st), "3-D filter launch");
    return MGBF_OK;
}

// s_g = in_g + P s_{g+1} from the coarsest generation up, in place in U
// (generation 1 into y)
int rad3_upsweep(mgbf_ctx* c, float* U, float* y, hipStream_t st) {
    for (int g = c->gm - 2; g >= 0; --g)
        CALL(prolong_g(c, g, U + c->goff[g + 1], U + c->goff[g], g == 0 ? y : U + c->goff[g], st));
    if (c->gm == 1) HIP_TRY(mgbf::launch_copy(U, y, c->gsize[0], st), "copy launch");
    return MGBF_OK;
}

// The streaming 3-D pass takes the prolongation of the up-sweep in its
// epilogue when the transfer is the 4-point one (the 2-D variables after the
// 3-D ones take it in their radial pass).
bool ball3_pro(const mgbf_ctx* c) {
    return c->ball3 && !c->quad;
}

// Forward passes from the coarsest generation up, each adding the
// prolongation of the one below: s_g = F T_g + P s_{g+1} (into U, y at g = 0).
int rad3_forward_upsweep(mgbf_ctx* c, const float* T, float* U, float* y, hipStream_t st) {
    for (int g = c->gm - 1; g >= 0; --g)
        CALL(ball(c, g, T + c->goff[g], g == 0 ? y : U + c->goff[g], false, st, nullptr,
                  g + 1 < c->gm ? U + c->goff[g + 1] : nullptr));
    return MGBF_OK;
}

// The transposed 3-D passes of the down-sweep, T_g = F^T u_g (scaled by A_g
// when `scaled` is given and the pass can take it), each carrying the
// restriction u_{g+1} = R u_g where the streaming kernel and its tiles allow,
// else preceded by a restriction pass of its own.
int rad3_down(mgbf_ctx* c, const float* x, float* U, float* T, int g, bool* scaled, hipStream_t st) {
    const float* ug = g == 0 ? x : U + c->goff[g];
    float* next = g + 1 < c->gm ? U + c->goff[g + 1] : nullptr;
    const int K3 = c->s.dims.km3_all * c->s.cfg.lm;
    // the streaming pass loads at least 3 halo rows when it restricts
    // (the 2-D variables after K3 restrict in their radial pass, at H >= 3)
    const bool res = next && c->ball3 && c->fuse >= 3 &&
                     fused_ok(c, g, pick_tj(K3, c->nx[g], c->ny[g], 4), std::max(c->b3H, 3)) &&
                     (K3 >= c->K || fused_ok(c, g, rad2d_tj(c, g), std::max(c->s.H, 3)));
    if (next && !res) CALL(restrict_g(c, g, ug, next, st));
    return ball(c, g, ug, T + c->goff[g], true, st, scaled, nullptr, res ? next : nullptr);
}

int apply_wawt_rad3(mgbf_ctx* c, const float* x, float* y, hipStream_t st) {
    float* U = c->ws_u;
    float* T = c->ws_t;
    const bool pro = ball3_pro(c);
    for (int g = 0; g < c->gm; ++g) {
        bool scaled = false;
        CALL(rad3_down(c, x, U, T, g, &scaled, st));
        if (!scaled)
            HIP_TRY(mgbf::launch_scale(T + c->goff[g], T + c->goff[g], c->wfield[g], c->wconst[g], 0,
                                       c->gsize[g], st),
                    "scale launch");
        if (!pro) CALL(ball(c, g, T + c->goff[g], U + c->goff[g], false, st));
    }
    if (pro) return rad3_forward_upsweep(c, T, U, y, st);
    return rad3_upsweep(c, U, y, st);
}

int apply_wt_rad3(mgbf_ctx* c, const float* x, float* v, bool sqrt_weights, hipStream_t st) {
    for (int g = 0; g < c->gm; ++g) {
        // the transposed pass writes v_g; v + goff stands in for T
        CALL(rad3_down(c, x, c->ws_u, v, g, nullptr, st));
        if (sqrt_weights)
            HIP_TRY(mgbf::launch_scale(v + c->goff[g], v + c->goff[g], c->wfield[g], c->wconst[g], 1,
                                       c->gsize[g], st),
                    "scale launch");
    }
    return MGBF_OK;
}

int apply_w_rad3(mgbf_ctx* c, const float* v, float* y, bool sqrt_weights, hipStream_t st) {
    float* U = c->ws_u;
    float* T = c->ws_t;
    if (ball3_pro(c)) {
        if (sqrt_weights) {
            for (int g = 0; g < c->gm; ++g)
                HIP_TRY(mgbf::launch_scale(v + c->goff[g], T + c->goff[g], c->wfield[g], c->wconst[g], 1,
                                           c->gsize[g], st),
                        "scale launch");
            return rad3_forward_upsweep(c, T, U, y, st);
        }
        return rad3_forward_upsweep(c, v, U, y, st);
    }
    for (int g = 0; g < c->gm; ++g) {
        const float* src = v + c->goff[g];
        if (sqrt_weights) {
            HIP_TRY(mgbf::launch_scale(src, T + c->goff[g], c->wfield[g], c->wconst[g], 1, c->gsize[g],
                                       st),
                    "scale launch");
            src = T + c->goff[g];
        }
        CALL(ball(c, g, src, U + c->goff[g], false, st));
    }
    return rad3_upsweep(c, U, y, st);
}

// The vertical map v rides in the tile kernel of direction e (0 direct,
// 1 adjoint).
bool fused_vmap(const mgbf_ctx* c, const VMap& v, int e) {
    return v.fz.lo && c->xt_ok[e] && !c->xfer_gather && !c->xfer_split;
}

// filt_to_anal_all2 / filt_to_anal_all: analysis field = T_v T_h
// (generation-1 field), the horizontal interpolation first, then lm -> lm_a
// levels with the routine family's vertical map (vd: its direct table).
int apply_f2a(mgbf_ctx* c, const VMap& vd, const char* refusal, const float* in, float* out,
              hipStream_t st) {
    const mgbf_dims& d = c->s.dims;
    if (refusal) return fail(MGBF_ERR_CONFIG, refusal);
    if (c->s.anal_identity) {
        if (in != out) HIP_TRY(mgbf::launch_copy(in, out, c->gsize[0], st), "copy launch");
        return MGBF_OK;
    }
    if (in == out) return fail(MGBF_ERR_ARG, "the analysis transfer needs distinct in and out");
    if (vd.s && fused_vmap(c, vd, 0)) {
        mgbf::XferTileArgs a = c->xt[0];
        a.in = in;
        a.out = out;
        HIP_TRY(mgbf::launch_xfer_tile_v(a, vd.fz, false, st), "transfer launch");
        return MGBF_OK;
    }
    float* mid = vd.s ? c->ws_a : out;
    if (c->xt_ok[0] && !c->xfer_gather) {
        mgbf::XferTileArgs a = c->xt[0];
        a.in = in;
        a.out = mid;
        HIP_TRY(mgbf::launch_xfer_tile(a, st), "transfer launch");
    } else {
        mgbf::XferArgs a = c->xf_dir;
        a.in = in;
        a.out = mid;
        HIP_TRY(mgbf::launch_xfer2d(a, st), "transfer launch");
    }
    if (vd.s)
        HIP_TRY(mgbf::launch_kwin(mid, out, c->K, d.km_a_all, (long long)d.na * d.ma, vd.s, vd.w, vd.width, st),
                "vertical transfer launch");
    return MGBF_OK;
}

// anal_to_filt_all2 / anal_to_filt_all: the transpose of the matching
// filt_to_anal (va: the adjoint vertical table).
int apply_a2f(mgbf_ctx* c, const VMap& va, const char* refusal, const float* in, float* out,
              hipStream_t st) {
    const mgbf_dims& d = c->s.dims;
    if (refusal) return fail(MGBF_ERR_CONFIG, refusal);
    if (c->s.anal_identity) {
        if (in != out) HIP_TRY(mgbf::launch_copy(in, out, c->gsize[0], st), "copy launch");
        return MGBF_OK;
    }
    if (in == out) return fail(MGBF_ERR_ARG, "the analysis transfer needs distinct in and out");
    if (va.s && fused_vmap(c, va, 1)) {
        mgbf::XferTileArgs a = c->xt[1];
        a.in = in;
        a.out = out;
        HIP_TRY(mgbf::launch_xfer_tile_v(a, va.fz, true, st), "transfer launch");
        return MGBF_OK;
    }
    const float* src = in;
    if (va.s) {
        HIP_TRY(mgbf::launch_kwin(in, c->ws_a, d.km_a_all, c->K, (long long)d.na * d.ma, va.s, va.w, va.width,
                                  st),
                "vertical transfer launch");
        src = c->ws_a;
    }
    if (c->xt_ok[1] && !c->xfer_gather) {
        mgbf::XferTileArgs a = c->xt[1];
        a.in = src;
        a.out = out;
        HIP_TRY(mgbf::launch_xfer_tile(a, st), "transfer launch");
    } else {
        mgbf::XferArgs a = c->xf_adj;
        a.in = src;
        a.out = out;
        HIP_TRY(mgbf::launch_xfer2d(a, st), "transfer launch");
    }
    return MGBF_OK;
}

int apply_weights(mgbf_ctx* c, const float* v, float* out, bool use_sqrt, hipStream_t st) {
    for (int g = 0; g < c->gm; ++g) {
        if (g == 0 && !c->filt_g1) {
            if (v != out)
                HIP_TRY(mgbf::launch_copy(v, out, c->gsize[0], st), "copy launch");
            continue;
        }
        HIP_TRY(mgbf::launch_scale(v + c->goff[g], out + c->goff[g], c->wfield[g], c->wconst[g],
                                   use_sqrt ? 1 : 0, c->gsize[g], st),
                "scale launch");
    }
    return MGBF_OK;
}

// Floats of each workspace (multigrid U and T, the analysis-grid
// intermediate), each rounded up to 256 bytes.
struct WsSizes {
    size_t u = 0, t = 0, a = 0;
    size_t total() const { return u + t + a; }
};

WsSizes ws_sizes(const mgbf_ctx* c) {
    auto r = [](size_t n) { return (n + 63) & ~size_t(63); };
    WsSizes w;
    w.u = r((size_t)c->s.dims.mg_size);
    w.t = w.u;
    w.a = c->need_ws_a ? r((size_t)c->K * c->s.dims.na * c->s.dims.ma) : 0;
    return w;
}

// Allocates the multigrid workspaces on first use.
int ensure_workspace(mgbf_ctx* c) {
    const WsSizes w = ws_sizes(c);
    if (!c->ws_u) HIP_TRY(hipMalloc(&c->ws_u, sizeof(float) * w.u), "device allocation (workspace)");
    if (!c->ws_t) HIP_TRY(hipMalloc(&c->ws_t, sizeof(float) * w.t), "device allocation (workspace)");
    if (w.a && !c->ws_a) HIP_TRY(hipMalloc(&c->ws_a, sizeof(float) * w.a), "device allocation (workspace)");
    return MGBF_OK;
}

// mgbf_apply on workspaces already in place.
int apply_op(mgbf_ctx* c, int op, const float* in, float* out, hipStream_t st) {
    if (c->rad3) {
        switch (op) {
            case MGBF_OP_WAWT: return apply_wawt_rad3(c, in, out, st);
            case MGBF_OP_WT: return apply_wt_rad3(c, in, out, false, st);
            case MGBF_OP_W: return apply_w_rad3(c, in, out, false, st);
            case MGBF_OP_BSQRT: return apply_w_rad3(c, in, out, true, st);
            case MGBF_OP_BSQRTT: return apply_wt_rad3(c, in, out, true, st);
            default: break;
        }
    }
    switch (op) {
        case MGBF_OP_WAWT: return apply_wawt(c, in, out, st);
        case MGBF_OP_WT: return apply_wt(c, in, out, false, st);
        case MGBF_OP_W: return apply_w(c, in, out, false, st);
        case MGBF_OP_A: return apply_weights(c, in, out, false, st);
        case MGBF_OP_ASQRT: return apply_weights(c, in, out, true, st);
        case MGBF_OP_BSQRT: return apply_w(c, in, out, true, st);
        case MGBF_OP_BSQRTT: return apply_wt(c, in, out, true, st);
        case MGBF_OP_A2F: return apply_a2f(c, c->kv_a, c->s.all2_refusal, in, out, st);
        case MGBF_OP_F2A: return apply_f2a(c, c->kv_d, c->s.all2_refusal, in, out, st);
        case MGBF_OP_A2F_ALL: return apply_a2f(c, c->kw_a, c->s.all_refusal, in, out, st);
        case MGBF_OP_F2A_ALL: return apply_f2a(c, c->kw_d, c->s.all_refusal, in, out, st);
        default: return fail(MGBF_ERR_ARG, "unknown operation");
    }
}

template <typename T>
T* carve(char*& p, size_t count) {
    T* r = reinterpret_cast<T*>(p);
    size_t bytes = (sizeof(T) * count + 255) & ~size_t(255);
    p += bytes;
    return r;
}

size_t carve_size(size_t bytes) { return (bytes + 255) & ~size_t(255); }

}  // namespace

extern "C" {

const char* mgbf_version(void) { return "mgbf-mi355x 0.1.0"; }

const char* mgbf_last_error(void) { return g_last_error.c_str(); }

int mgbf_config_default(mgbf_config* cfg) {
    if (!cfg) return fail(MGBF_ERR_ARG, "cfg is NULL");
    std::memset(cfg, 0, sizeof *cfg);
    cfg->hx = cfg->hy = 6;
    cfg->hz = 3;
    cfg->p = 2;
    cfg->mg_ampl02 = 49.0 / 8.0;   // reach (h+1) = 7 = halo + 1
    cfg->mg_ampl01 = 16.0 / 7.0;   // reach 4 = hz + 1
    cfg->mg_ampl03 = 1.0;
    cfg->mg_weig1 = 1.0;
    cfg->mg_weig2 = 0.5;
    cfg->mg_weig3 = 0.25;
    cfg->mg_weig4 = 0.125;
    cfg->mg_weig5 = 0.0625;
    cfg->mgbf_proc = 6;
    cfg->n_ens = 1;
    cfg->lm_a = cfg->lm = 32;
    cfg->km2 = 0;
    cfg->km3 = 1;
    cfg->l_filt_g1 = 1;
    cfg->l_new_map = 1;
    cfg->l_vertical_filter = 1;
    cfg->gm_max = 4;
    cfg->nxPE = cfg->nyPE = 8;
    cfg->im_filt = cfg->jm_filt = 24;
    cfg->nm0 = cfg->mm0 = 192;
    return MGBF_OK;
}

int mgbf_derive(const mgbf_config* cfg, mgbf_dims* dims) {
    if (!cfg || !dims) return fail(MGBF_ERR_ARG, "NULL argument");
    std::string err;
    mgbf::Setup s;
    int st = mgbf::build_setup(*cfg, s, &err);
    if (st != MGBF_OK) return fail(st, err);
    *dims = s.dims;
    return MGBF_OK;
}

int mgbf_init(const mgbf_config* cfg, int device, mgbf_ctx** out) {
    if (!cfg || !out) return fail(MGBF_ERR_ARG, "NULL argument");
    *out = nullptr;
    auto* c = new mgbf_ctx();
    std::string err;
    int st = mgbf::build_setup(*cfg, c->s, &err);
    if (st != MGBF_OK) {
        delete c;
        return fail(st, err);
    }
    const mgbf::Setup& s = c->s;
    c->device = device;
    c->K = s.dims.km_all;
    c->gm = s.dims.gm;
    // l_filt_g1 is read by the ensemble procedures only: the background
    // procedures (1-3) and the 3-D ones (4, 5: filtering_rad3 and
    // filtering_lin3 weight through weighting_all -> weighting, which scales
    // generation 1 unconditionally, mg_filtering.f90:203, 427,
    // mg_generations.f90:133-155, 973-1009) filter and weight generation 1
    // whatever it says
    c->filt_g1 = (cfg->mgbf_proc >= 1 && cfg->mgbf_proc <= 5) ? true : cfg->l_filt_g1 != 0;
    // filtering_lin2_ens ends with sup_vrbeta1T_ens (mg_filtering.f90:1303-1305)
    c->vf_last = cfg->mgbf_proc == 7 ? BAND_VFT : BAND_VF;
    c->nfilt = cfg->l_vertical_filter ? s.dims.km3_all * cfg->lm : 0;
    c->quad = cfg->lquart != 0;
    c->rad3 = cfg->mgbf_proc == 4 || cfg->mgbf_proc == 5;
    {
        const char* xg = std::getenv("MGBF_XFER_GATHER");
        c->xfer_gather = xg && xg[0] == '1';
        const char* xs = std::getenv("MGBF_XFER_SPLIT");
        c->xfer_split = xs && xs[0] == '1';
    }
    if (c->rad3) {
        // MGBF_BALL_GATHER=1 keeps the gather for the 3-D variables too
        const char* bg = std::getenv("MGBF_BALL_GATHER");
        c->ball3 = !(bg && bg[0] == '1') && plan_ball3(s, *cfg, c->b3, c->b3H, c->b3HZ, c->b3ball);
    }
    {
        // MGBF_EXACT_ORDER=1 keeps the reference's order Vf^T ... Vf also for
        // weights that are constant along the stack
        const char* ev = std::getenv("MGBF_EXACT_ORDER");
        c->commute = !(ev && ev[0] == '1');
        // MGBF_FUSE=0|1|2 selects how much of the up-sweep rides in the
        // epilogue of the last radial pass (all levels compute the same)
        const char* fv = std::getenv("MGBF_FUSE");
        if (fv && fv[0] >= '0' && fv[0] <= '3') c->fuse = fv[0] - '0';
        // MGBF_MFMA=1 puts every eligible radial pass on the matrix cores
        const char* mv = std::getenv("MGBF_MFMA");
        if (mv && mv[0] == '1')
            for (int g = 0; g < mgbf::kMaxGen; ++g)
                for (int kind = 0; kind < 4; ++kind) c->tile_mfma[g][kind] = 1;
    }
    long long off = 0;
    for (int g = 0; g < c->gm; ++g) {
        c->nx[g] = s.nx[g];
        c->ny[g] = s.ny[g];
        c->gsize[g] = (long long)c->K * s.nx[g] * s.ny[g];
        c->goff[g] = off;
        off += c->gsize[g];
        c->wconst[g] = (float)s.wconst[g];
        for (int kind = 0; kind < 4; ++kind) {
            c->tile_ti[g][kind] = 4;
            c->tile_tj[g][kind] = pick_tj(c->K, s.nx[g], s.ny[g], 4);
            if (c->tile_mfma[g][kind]) c->tile_tj[g][kind] = pick_tj(c->K, s.nx[g], s.ny[g], 16);
        }
        // the fused restriction needs an even number of rows per wave
        int& tj3 = c->tile_tj[g][3];
        if (tj3 & 1) tj3 = tj3 + 1 <= s.ny[g] ? tj3 + 1 : tj3 - 1;
    }
    for (int b = 0; b <= mgbf::kMaxHaloDev; ++b)
        for (int a = 0; a <= mgbf::kMaxHaloDev; ++a)
            c->coef.w[b][a] = (b <= mgbf::kMaxHalo && a <= mgbf::kMaxHalo) ? (float)s.wsym[b][a] : 0.f;
    if (s.separable) {
        // the separable kernels read the two factors from rows 0 and 1
        c->coef = mgbf::RadialCoef{};
        for (int a = 0; a <= s.H; ++a) c->coef.w[0][a] = (float)s.sepx[a];
        for (int b = 0; b <= s.H; ++b) c->coef.w[1][b] = (float)s.sepy[b];
    }

    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || device < 0 || device >= dev_count) {
        delete c;
        return fail(MGBF_ERR_HIP, "no HIP device with the requested index");
    }
    DeviceGuard guard(device);

    // vertical band descriptors and their interior rows
    const std::vector<double>* bands[3] = {&s.vf, &s.vft, &s.vvt};
    const int lmv = cfg->lm;
    for (int b = 0; b < 3; ++b) {
        const int hb = b == BAND_VVT ? 2 * cfg->hz : cfg->hz;
        const int nb = 2 * hb + 1;
        if (hb > mgbf::kMaxVBand || hb < 1) {
            if (c->nfilt > 0) {
                delete c;
                return fail(MGBF_ERR_UNSUPPORTED, "vertical half-width outside 1..10");
            }
            continue;
        }
        c->band_hb[b] = hb;
        // rows equal to the middle row form the interior range
        const int mid = lmv / 2;
        auto same = [&](int L) {
            for (int t = 0; t < nb; ++t)
                if ((float)(*bands[b])[(size_t)L * nb + t] != (float)(*bands[b])[(size_t)mid * nb + t])
                    return false;
            return true;
        };
        int lo = mid, hi = mid + 1;
        while (lo > 0 && same(lo - 1)) --lo;
        while (hi < lmv && same(hi)) ++hi;
        c->band_ilo[b] = lo;
        c->band_ihi[b] = hi;
        for (int t = 0; t < nb; ++t) c->band_int[b].w[t] = (float)(*bands[b])[(size_t)mid * nb + t];
    }

    // Fragments of the matrix-core engine (B operand of
    // v_mfma_f32_16x16x32_f16: lane l holds window columns 8(l/16) .. +7 of
    // output column l%16):
    //  - the weight bands T_b(c, i) = w(|c - H - i|, b) (b = 0..H), scaled
    //    by 2^wexp and split into FP16 hi + lo, [b][hi/lo][lane][8];
    //  - per generation g < gm-1 the x restriction of a 16-column tile onto
    //    its 8 coarse columns (R = P^T with the folded edges, multiples of
    //    1/128, or of 1/32 for the quadratic transfer, exact in FP16) for
    //    the tile classes interior, left edge and
    //    right edge, [class][lane][8]; every tile of the generation must
    //    equal its class, else the restriction stays on the vector kernel.
    std::vector<_Float16> bfrag;
    std::vector<_Float16> rfrag[mgbf::kMaxGen];
    if (s.H <= 8 && (s.exact_disc || s.separable) && c->K % 64 == 0) {
        double maxw = 0;
        for (int b = 0; b <= s.H; ++b)
            for (int a = 0; a <= s.H; ++a) maxw = std::max(maxw, s.wsym[b][a]);
        c->wexp = maxw > 0 ? 10 - std::ilogb(maxw) : 0;
        bfrag.assign((size_t)(s.H + 1) * 2 * 64 * 8, (_Float16)0.f);
        for (int b = 0; b <= s.H; ++b)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) {
                    const int a = (l / 16) * 8 + e - s.H - (l % 16);
                    const int aa = a < 0 ? -a : a;
                    const float wv = aa <= s.H ? (float)std::ldexp(s.wsym[b][aa], c->wexp) : 0.f;
                    const _Float16 hi = (_Float16)wv;
                    const _Float16 lo = (_Float16)(wv - (float)hi);
                    bfrag[((size_t)(b * 2 + 0) * 64 + l) * 8 + e] = hi;
                    bfrag[((size_t)(b * 2 + 1) * 64 + l) * 8 + e] = lo;
                }
        c->mfma_ok = true;
        for (int g = 0; g + 1 < c->gm; ++g) {
            const mgbf::AxisTable& rt = s.rx[g];
            const int nf = s.nx[g];
            if (nf % 16 || nf < 32 || s.H < 3) continue;
            auto tile_frag = [&](int i0, std::vector<_Float16>& f) -> bool {
                std::vector<double> B(32 * 16, 0.0);
                for (int I = 0; I < 8; ++I) {
                    const int Ia = i0 / 2 + I;
                    for (int t = 0; t < rt.width; ++t) {
                        const double wv = rt.w[(size_t)Ia * rt.width + t];
                        if (wv == 0.0) continue;
                        const int cc = rt.start[Ia] + t - i0 + s.H;
                        if (cc < 0 || cc >= 32) return false;
                        B[cc * 16 + I] += wv;
                    }
                }
                f.assign(64 * 8, (_Float16)0.f);
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) {
                        const double wv = B[((l / 16) * 8 + e) * 16 + (l % 16)];
                        const _Float16 h = (_Float16)wv;
                        if ((double)(float)h != wv) return false;   // not exact in FP16
                        f[l * 8 + e] = h;
                    }
                return true;
            };
            const int nt = nf / 16;
            std::vector<_Float16> cls[3], ft;
            bool ok = tile_frag(0, cls[1]) && tile_frag(nf - 16, cls[2]);
            if (ok && nt >= 3) ok = tile_frag(16, cls[0]);
            else cls[0].assign(64 * 8, (_Float16)0.f);
            for (int t = 1; ok && t + 1 < nt; ++t) {
                ok = tile_frag(t * 16, ft);
                for (size_t q = 0; ok && q < ft.size(); ++q) ok = ft[q] == cls[0][q];
            }
            if (!ok) continue;
            rfrag[g].clear();
            for (int k = 0; k < 3; ++k) rfrag[g].insert(rfrag[g].end(), cls[k].begin(), cls[k].end());
            c->mx_res_ok[g] = true;
        }
    }

    // one device allocation for all coefficient tables
    size_t bytes = carve_size(sizeof(double) * (1 + mgbf::kDotBlocks)) + carve_size(sizeof(_Float16) * bfrag.size());
    for (int g = 0; g < mgbf::kMaxGen; ++g) bytes += carve_size(sizeof(_Float16) * rfrag[g].size());
    for (int b = 0; b < 3; ++b) bytes += carve_size(sizeof(float) * bands[b]->size());
    std::vector<float> rytab[mgbf::kMaxGen];
    std::vector<int> rymeta[mgbf::kMaxGen];
    for (int g = 0; g + 1 < c->gm; ++g) {
        const mgbf::AxisTable* tabs[4] = {&s.rx[g], &s.ry[g], &s.px[g], &s.py[g]};
        for (const auto* t : tabs)
            bytes += carve_size(sizeof(int) * t->n_out) + carve_size(sizeof(float) * t->w.size());
        c->res_ok[g] = build_res_tables(s.rx[g], s.ry[g], s.nx[g], s.ny[g], c->quad, rytab[g], rymeta[g]);
        bytes += carve_size(sizeof(float) * rytab[g].size()) + carve_size(sizeof(int) * rymeta[g].size());
    }
    {
        const mgbf::AxisTable* tabs[8] = {&s.ax_d, &s.ay_d, &s.ax_a, &s.ay_a,
                                          &s.kv_d, &s.kv_a, &s.kw_d, &s.kw_a};
        for (const auto* t : tabs)
            bytes += carve_size(sizeof(int) * t->start.size()) + carve_size(sizeof(float) * t->w.size());
    }
    if (c->rad3) {
        bytes += carve_size(sizeof(int4) * s.taps3.size()) + carve_size(sizeof(float) * s.taps3.size());
        bytes += carve_size(sizeof(int4) * s.taps2.size()) + carve_size(sizeof(float) * s.taps2.size());
    }
    // tiled transfer blocks: [0] direct (filter -> analysis), [1] adjoint
    struct Blocks {
        std::vector<int> x0, y0;
        std::vector<double> wx, wy;
        int SX = 0, SY = 0;
        bool ok = false;
    } tb[2];
    if (!s.anal_identity) {
        const mgbf_dims& d = s.dims;
        const mgbf::AxisTable* ax[2] = {&s.ax_d, &s.ax_a};
        const mgbf::AxisTable* ay[2] = {&s.ay_d, &s.ay_a};
        const int ninx[2] = {d.nx, d.na}, niny[2] = {d.ny, d.ma};
        for (int e = 0; e < 2; ++e) {
            Blocks& b = tb[e];
            // 32-bit row offsets inside an input row
            b.ok = (long long)ninx[e] * c->K * 4 < (1ll << 31) &&
                   tile_blocks(*ax[e], ninx[e], 8, true, b.SX, b.x0, b.wx) &&
                   tile_blocks(*ay[e], niny[e], 8, false, b.SY, b.y0, b.wy);
            if (b.ok)
                bytes += carve_size(sizeof(int) * b.x0.size()) + carve_size(sizeof(float) * b.wx.size()) +
                         carve_size(sizeof(int) * b.y0.size()) + carve_size(sizeof(float) * b.wy.size());
        }
    }
    // vertical maps fused into the tile kernels: kv_d, kv_a, kw_d, kw_a
    VFuseHost vfz[4];
    bool vfz_ok[4] = {};
    if (!s.anal_identity) {
        const mgbf::AxisTable* vt[4] = {&s.kv_d, &s.kv_a, &s.kw_d, &s.kw_a};
        for (int e = 0; e < 4; ++e) {
            const bool adj = e & 1;
            if (vt[e]->n_out == 0 || !tb[adj ? 1 : 0].ok) continue;
            vfz_ok[e] = build_vfuse(*vt[e], c->K, adj, vfz[e]);
            if (vfz_ok[e])
                bytes += carve_size(sizeof(int) * vfz[e].lo.size()) + carve_size(sizeof(int) * vfz[e].cnt.size()) +
                         carve_size(sizeof(int) * vfz[e].r0.size()) + carve_size(sizeof(float) * vfz[e].w.size());
        }
    }
    std::vector<char> host(bytes, 0);
    hipError_t e = hipMalloc(&c->d_tables, bytes);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipMalloc(tables)");
    }
    // Host staging and device addresses are carved side by side. A table
    // the size count above missed lands in a spill buffer instead of past
    // the staging array, and init fails before anything reaches the device.
    char* hp = host.data();
    char* const hend = host.data() + host.size();
    char* dp = static_cast<char*>(c->d_tables);
    bool overflow = false;
    std::vector<std::vector<char>> spill;
    auto take = [&](auto* tag, size_t count, auto** dev) {
        using T = std::remove_pointer_t<decltype(tag)>;
        const size_t nb = carve_size(sizeof(T) * count);
        *dev = carve<T>(dp, count);
        if (hp + nb > hend) {
            overflow = true;
            spill.emplace_back(nb + sizeof(T));
            return reinterpret_cast<T*>(spill.back().data());
        }
        return carve<T>(hp, count);
    };
    auto put_f = [&](const std::vector<double>& v) -> const float* {
        float* d = nullptr;
        float* h = take((float*)nullptr, v.size(), &d);
        for (size_t i = 0; i < v.size(); ++i) h[i] = (float)v[i];
        return d;
    };
    auto put_i = [&](const std::vector<int>& v) -> const int* {
        int* d = nullptr;
        int* h = take((int*)nullptr, v.size(), &d);
        for (size_t i = 0; i < v.size(); ++i) h[i] = v[i];
        return d;
    };
    for (int b = 0; b < 3; ++b) c->d_band[b] = put_f(*bands[b]);
    auto put_h = [&](const std::vector<_Float16>& v) -> const void* {
        if (v.empty()) return nullptr;
        _Float16* d = nullptr;
        _Float16* h = take((_Float16*)nullptr, v.size(), &d);
        for (size_t i = 0; i < v.size(); ++i) h[i] = v[i];
        return d;
    };
    c->d_bfrag = put_h(bfrag);
    for (int g = 0; g < mgbf::kMaxGen; ++g) c->d_rfrag[g] = put_h(rfrag[g]);
    take((double*)nullptr, 1 + mgbf::kDotBlocks, &c->d_dot);
    if (c->rad3) {
        auto put_taps = [&](size_t nt, auto tap, const int4** dt, const float** dw) {
            int4* d4 = nullptr;
            float* dwp = nullptr;
            int4* ht = take((int4*)nullptr, nt, &d4);
            float* hw = take((float*)nullptr, nt, &dwp);
            *dt = d4;
            *dw = dwp;
            for (size_t t = 0; t < nt; ++t) tap(t, ht[t], hw[t]);
        };
        put_taps(s.taps3.size(), [&](size_t t, int4& g, float& w) {
            g = make_int4(s.taps3[t].gx, s.taps3[t].gy, s.taps3[t].gz, 0);
            w = (float)s.taps3[t].w;
        }, &c->d_t3, &c->d_w3);
        put_taps(s.taps2.size(), [&](size_t t, int4& g, float& w) {
            g = make_int4(s.taps2[t].gx, s.taps2[t].gy, 0, 0);
            w = (float)s.taps2[t].w;
        }, &c->d_t2, &c->d_w2);
    }
    for (int g = 0; g + 1 < c->gm; ++g) {
        c->d_rxs[g] = put_i(s.rx[g].start);
        c->d_rxw[g] = put_f(s.rx[g].w);
        c->d_rys[g] = put_i(s.ry[g].start);
        c->d_ryw[g] = put_f(s.ry[g].w);
        c->d_pxs[g] = put_i(s.px[g].start);
        c->d_pxw[g] = put_f(s.px[g].w);
        c->d_pys[g] = put_i(s.py[g].start);
        c->d_pyw[g] = put_f(s.py[g].w);
        if (c->res_ok[g]) {
            float* df = nullptr;
            int* di = nullptr;
            float* hf = take((float*)nullptr, rytab[g].size(), &df);
            for (size_t i = 0; i < rytab[g].size(); ++i) hf[i] = rytab[g][i];
            int* hi = take((int*)nullptr, rymeta[g].size(), &di);
            for (size_t i = 0; i < rymeta[g].size(); ++i) hi[i] = rymeta[g][i];
            c->d_rytab[g] = reinterpret_cast<const float4*>(df);
            c->d_rymeta[g] = reinterpret_cast<const int2*>(di);
        }
    }
    if (!s.anal_identity) {
        const mgbf_dims& d = s.dims;
        c->xf_dir = mgbf::XferArgs{nullptr, nullptr, c->K, d.nx, d.ny, d.na, d.ma,
                                   put_i(s.ax_d.start), put_f(s.ax_d.w), s.ax_d.width,
                                   put_i(s.ay_d.start), put_f(s.ay_d.w), s.ay_d.width};
        c->xf_adj = mgbf::XferArgs{nullptr, nullptr, c->K, d.na, d.ma, d.nx, d.ny,
                                   put_i(s.ax_a.start), put_f(s.ax_a.w), s.ax_a.width,
                                   put_i(s.ay_a.start), put_f(s.ay_a.w), s.ay_a.width};
        for (int e = 0; e < 2; ++e) {
            if (!tb[e].ok) continue;
            const bool dir = e == 0;
            c->xt[e] = mgbf::XferTileArgs{nullptr, nullptr, c->K,
                                          dir ? d.nx : d.na, dir ? d.ny : d.ma, dir ? d.na : d.nx, dir ? d.ma : d.ny,
                                          put_i(tb[e].x0), put_f(tb[e].wx), tb[e].SX,
                                          put_i(tb[e].y0), put_f(tb[e].wy), tb[e].SY};
            c->xt_ok[e] = true;
        }
        auto put_v = [&](const mgbf::AxisTable& t, VMap& v, int e) {
            if (t.n_out == 0) return;
            v.s = put_i(t.start);
            v.w = put_f(t.w);
            v.width = t.width;
            if (vfz_ok[e])
                v.fz = mgbf::XferVArgs{put_i(vfz[e].lo), put_i(vfz[e].cnt), put_i(vfz[e].r0), put_f(vfz[e].w),
                                       d.km_a_all};
        };
        put_v(s.kv_d, c->kv_d, 0);
        put_v(s.kv_a, c->kv_a, 1);
        put_v(s.kw_d, c->kw_d, 2);
        put_v(s.kw_a, c->kw_a, 3);
        // the analysis-grid intermediate, for a vertical map on its own
        const VMap* vm[4] = {&c->kv_d, &c->kv_a, &c->kw_d, &c->kw_a};
        for (int e = 0; e < 4; ++e)
            if (vm[e]->s && !fused_vmap(c, *vm[e], e & 1)) c->need_ws_a = true;
    }
    if (overflow) {
        // a table was not counted above: its device address lies past the
        // allocation, so nothing may use it
        mgbf_finalize(c);
        return fail(MGBF_ERR_HIP, "internal: coefficient tables exceed their allocation");
    }
    e = hipMemcpy(c->d_tables, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        int r = hip_fail(e, "device allocation");
        mgbf_finalize(c);
        return r;
    }
    // the static scratch of mgbf_apply_noalloc on this device (its address
    // is looked up once per device, with the device current)
    st = mgbf::noalloc_setup(device, &err);
    if (st != MGBF_OK) {
        mgbf_finalize(c);
        return fail(st, err);
    }
    // the multigrid workspaces are allocated by the first call that needs
    // them (mgbf_apply, mgbf_autotune, mgbf_group_apply), so that a context
    // used only through mgbf_apply_noalloc holds nothing beyond its tables
    *out = c;
    return MGBF_OK;
}

int mgbf_finalize(mgbf_ctx* c) {
    if (!c) return MGBF_OK;
    DeviceGuard guard(c->device);
    (void)hipDeviceSynchronize();
    (void)drain_timers(c);
    for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
    if (c->ws_u) (void)hipFree(c->ws_u);
    if (c->ws_t) (void)hipFree(c->ws_t);
    if (c->d_tables) (void)hipFree(c->d_tables);
    if (c->ws_a) (void)hipFree(c->ws_a);
    for (int g = 0; g < mgbf::kMaxGen; ++g)
        if (c->d_aw[g]) (void)hipFree(c->d_aw[g]);
    delete c;
    return MGBF_OK;
}

int mgbf_get_dims(const mgbf_ctx* c, mgbf_dims* dims) {
    if (!c || !dims) return fail(MGBF_ERR_ARG, "NULL argument");
    *dims = c->s.dims;
    return MGBF_OK;
}

int mgbf_set_weights(mgbf_ctx* c, int gen, const float* field, double value) {
    if (!c) return fail(MGBF_ERR_ARG, "ctx is NULL");
    if (gen < 1 || gen > c->gm) return fail(MGBF_ERR_ARG, "generation out of range");
    c->wfield[gen - 1] = field;
    if (!field) c->wconst[gen - 1] = (float)value;
    ++c->version;
    return MGBF_OK;
}

int mgbf_set_aspect(mgbf_ctx* c, int gen, const double* pasp2) {
    if (!c) return fail(MGBF_ERR_ARG, "ctx is NULL");
    if (gen < 1 || gen > c->gm) return fail(MGBF_ERR_ARG, "generation out of range");
    const int g = gen - 1;
    DeviceGuard guard(c->device);
    if (!pasp2) {
        if (c->d_aw[g]) {
            HIP_TRY(hipDeviceSynchronize(), "synchronize before freeing the aspect weights");
            HIP_TRY(hipFree(c->d_aw[g]), "hipFree(aspect weights)");
        }
        c->d_aw[g] = nullptr;
        c->aniso[g] = false;
        ++c->version;
        return MGBF_OK;
    }
    const int proc = c->s.cfg.mgbf_proc;
    if ((proc != 1 && proc != 6) || c->s.separable || !c->s.exact_disc)
        return fail(MGBF_ERR_UNSUPPORTED,
                    "per-point aspect tensors are built for the radial procedures 1 and 6");
    std::vector<float> w;
    std::string err;
    const int st = mgbf::aspect_weights(c->s.cfg, c->s.H, c->nx[g], c->ny[g], pasp2, w, &err);
    if (st != MGBF_OK) return fail(st, err);
    if (!c->d_aw[g]) HIP_TRY(hipMalloc(&c->d_aw[g], sizeof(float) * w.size()), "hipMalloc(aspect weights)");
    HIP_TRY(hipMemcpy(c->d_aw[g], w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice),
            "aspect weights upload");
    c->aniso[g] = true;
    ++c->version;
    return MGBF_OK;
}

int mgbf_apply(mgbf_ctx* c, int op, const float* in, float* out, void* stream) {
    if (!c || !in || !out) return fail(MGBF_ERR_ARG, "NULL argument");
    DeviceGuard guard(c->device);
    CALL(ensure_workspace(c));
    return apply_op(c, op, in, out, static_cast<hipStream_t>(stream));
}

}  // extern "C"

// The workspaces of one call in caller-given device memory: U, T and the
// analysis-grid intermediate as slices of `base`
// (every one 256-byte aligned). Nothing is allocated and no timer event is
// recorded; the context's own workspaces are untouched.
int mgbf::apply_in_scratch(mgbf_ctx* c, int op, const float* in, float* out, void* stream,
                           float* (*scratch_of)(int device), size_t nfloats) {
    if (!c || !in || !out) return fail(MGBF_ERR_ARG, "NULL argument");
    float* const base = scratch_of(c->device);
    if (!base) return fail(MGBF_ERR_HIP, "the static scratch has no device address on this device");
    const WsSizes w = ws_sizes(c);
    if (w.total() > nfloats)
        return fail(MGBF_ERR_ARG,
                    "the workspace of this context (" + std::to_string(w.total()) +
                        " floats) exceeds the static scratch of mgbf_apply_noalloc (" +
                        std::to_string(nfloats) + " floats)");
    DeviceGuard guard(c->device);
    float* const keep_u = c->ws_u;
    float* const keep_t = c->ws_t;
    float* const keep_a = c->ws_a;
    const bool keep_timing = c->timing;
    c->ws_u = base;
    c->ws_t = base + w.u;
    c->ws_a = w.a ? base + w.u + w.t : nullptr;
    c->timing = false;
    const int status = apply_op(c, op, in, out, static_cast<hipStream_t>(stream));
    c->ws_u = keep_u;
    c->ws_t = keep_t;
    c->ws_a = keep_a;
    c->timing = keep_timing;
    return status;
}

extern "C" {

int mgbf_apply_stage(mgbf_ctx* c, int stage, int gen, const float* in, float* out,
                     void* stream) {
    if (!c || !in || !out) return fail(MGBF_ERR_ARG, "NULL argument");
    if (in == out) return fail(MGBF_ERR_ARG, "stages need distinct in and out");
    DeviceGuard guard(c->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int g = gen - 1;
    if (g < 0 || g >= c->gm) return fail(MGBF_ERR_ARG, "generation out of range");
    switch (stage) {
        case MGBF_STAGE_VF:
        case MGBF_STAGE_VFT:
            if (g != 0) return fail(MGBF_ERR_ARG, "the vertical filter acts on generation 1");
            return vertical(c, in, out, stage == MGBF_STAGE_VFT ? BAND_VFT : BAND_VF, st);
        case MGBF_STAGE_F:
        case MGBF_STAGE_FT:
            if (c->rad3) return ball(c, g, in, out, stage == MGBF_STAGE_FT, st);
            return radial(c, g, in, out, nullptr, 1.0f, st, MGBF_T_HFILT, nullptr, -1, nullptr,
                          stage == MGBF_STAGE_FT);
        case MGBF_STAGE_R:
            if (g + 1 >= c->gm) return fail(MGBF_ERR_ARG, "no coarser generation");
            return restrict_g(c, g, in, out, st);
        case MGBF_STAGE_P:
            if (g + 1 >= c->gm) return fail(MGBF_ERR_ARG, "no coarser generation");
            return prolong_g(c, g, in, nullptr, out, st);
        case MGBF_STAGE_A:
            HIP_TRY(mgbf::launch_scale(in, out, c->wfield[g], c->wconst[g], 0, c->gsize[g], st),
                    "scale launch");
            return MGBF_OK;
        default:
            return fail(MGBF_ERR_ARG, "unknown stage");
    }
}

int mgbf_dot(mgbf_ctx* c, const float* a, const float* b, long long n, double* result,
             void* stream) {
    if (!c || !a || !b || !result || n < 0) return fail(MGBF_ERR_ARG, "bad argument");
    DeviceGuard guard(c->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(mgbf::launch_dot(a, b, n, c->d_dot, st), "dot launch");
    HIP_TRY(hipMemcpyAsync(result, c->d_dot, sizeof(double), hipMemcpyDeviceToHost, st),
            "dot copy");
    HIP_TRY(hipStreamSynchronize(st), "dot sync");
    return MGBF_OK;
}

int mgbf_autotune(mgbf_ctx* c, void* stream) {
    if (!c) return fail(MGBF_ERR_ARG, "ctx is NULL");
    if (c->rad3) return MGBF_OK;   // procedure 4 runs no radial kernel: nothing to tune
    DeviceGuard guard(c->device);
    CALL(ensure_workspace(c));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t mgb = sizeof(float) * (size_t)c->s.dims.mg_size;
    HIP_TRY(hipMemsetAsync(c->ws_u, 0, mgb, st), "autotune memset");
    HIP_TRY(hipMemsetAsync(c->ws_t, 0, mgb, st), "autotune memset");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "event");
    HIP_TRY(hipEventCreate(&e1), "event");
    const bool timing = c->timing;
    c->timing = false;
    int status = MGBF_OK;
    const int vband = c->commute ? BAND_VVT : c->vf_last;
    // MGBF_AUTOTUNE_LOG=1 prints every candidate's time (3 passes) to stderr
    const char* logv = std::getenv("MGBF_AUTOTUNE_LOG");
    const bool log = logv && logv[0] == '1';
    // the restriction chain fills the coarse generations the candidates read
    if (status == MGBF_OK) status = restrict_chain(c, c->ws_u, c->ws_u, st);
    // kind 3 writes its restriction over the (zero) next generation of U,
    // which the restriction of zeros leaves unchanged
    // the pass with the restriction is the down-sweep's weighted one: timed
    // with the generation's weight field, if it has one
    auto time_pass = [&](int g, const float* coarse, int band, float* rout, float* ms) -> int {
        const float* wf = rout ? c->wfield[g] : nullptr;
        const float wc = wf ? c->wconst[g] : 1.0f;
        int r = radial(c, g, c->ws_u + c->goff[g], c->ws_t, wf, wc, st, MGBF_T_HFILT, coarse, band, rout);
        if (r != MGBF_OK) return r;
        if (hipEventRecord(e0, st) != hipSuccess) return fail(MGBF_ERR_HIP, "autotune timing");
        for (int rep = 0; rep < 3 && r == MGBF_OK; ++rep)
            r = radial(c, g, c->ws_u + c->goff[g], c->ws_t, wf, wc, st, MGBF_T_HFILT, coarse, band, rout);
        if (r != MGBF_OK) return r;
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(ms, e0, e1) != hipSuccess)
            return fail(MGBF_ERR_HIP, "autotune timing");
        return MGBF_OK;
    };
    for (int g = 0; g < c->gm && status == MGBF_OK; ++g) {
        const int n = c->nx[g], m = c->ny[g];
        if (c->aniso[g]) continue;   // the per-point aspect filter has no tiles
        float kind_ms[4] = {-1.f, -1.f, -1.f, -1.f};   // the fastest candidate per kind
        for (int kind = 0; kind < 4 && status == MGBF_OK; ++kind) {
            if (kind >= 1 && g + 1 >= c->gm) continue;
            if (kind == 2 && !(g == 0 && c->nfilt > 0 && can_fuse_band(c, vband))) continue;
            if (kind == 3 && !(c->fuse >= 3 && c->res_ok[g] && c->s.H >= 3)) continue;
            const float* coarse = kind == 1 || kind == 2 ? c->ws_u + c->goff[g + 1] : nullptr;
            float* rout = kind == 3 ? c->ws_u + c->goff[g + 1] : nullptr;
            const int band = kind == 2 ? vband : -1;
            float best = 1e30f;
            int bti = c->tile_ti[g][kind], btj = c->tile_tj[g][kind], bmf = c->tile_mfma[g][kind];
            const bool try_mfma = mx_ok(c, g, kind);
            // engines: TI = 2, TI = 4 (vector), 16 (matrix cores)
            const int engines[3] = {2, 4, 16};
            for (int en : engines) {
                if (en == 16 && !try_mfma) continue;
                if (kind == 3 && en == 2) continue;
                // waves of one row of tiles (every column tile and k chunk)
                const long long strip = (long long)((c->K + 63) / 64) * ((n + en - 1) / en) * (en == 16 ? 4 : 1);
                // tile heights: a fixed ladder, and for q = 1 .. 16 the
                // lowest height whose launch stays within q x 1024 waves
                // (q / occupancy rounds of the chip's 1024 SIMDs). On a
                // small field the ladder's heights leave the last round of
                // blocks part empty, or take many rounds of short tiles
                // that each repeat the 2H rows of the ring's start.
                int tjs[32];
                int ntj = 0;
                auto add_tj = [&](int tj) {
                    if (tj < 1 || tj > m) return;
                    for (int q2 = 0; q2 < ntj; ++q2)
                        if (tjs[q2] == tj) return;
                    tjs[ntj++] = tj;
                };
                for (int tj : {16, 24, 32, 48, 64, 96, 128, 192, 256, 384}) add_tj(tj);
                add_tj(m);
                for (long long q = 1; q <= 16; ++q) {
                    const long long segs = q * 1024 / strip;
                    if (segs < 1) continue;
                    int tj = (int)((m + segs - 1) / segs);
                    if (kind == 3) tj += tj & 1;
                    add_tj(tj);
                }
                for (int q = 0; q < ntj && status == MGBF_OK; ++q) {
                    const int tj = tjs[q];
                    if (kind == 3 && !fused_ok(c, g, tj, c->s.H)) continue;
                    const long long waves = strip * ((m + tj - 1) / tj);
                    if (waves < 512 && tj != 16) continue;
                    c->tile_ti[g][kind] = en == 16 ? 4 : en;
                    c->tile_tj[g][kind] = tj;
                    c->tile_mfma[g][kind] = en == 16 ? 1 : 0;
                    float ms = 0.f;
                    status = time_pass(g, coarse, band, rout, &ms);
                    if (log && status == MGBF_OK)
                        std::fprintf(stderr, "[autotune] gen %d kind %d engine %d TJ %d: %.4f ms\n",
                                     g + 1, kind, en, tj, ms / 3);
                    if (status == MGBF_OK && ms < best) {
                        best = ms;
                        bti = c->tile_ti[g][kind];
                        btj = tj;
                        bmf = c->tile_mfma[g][kind];
                    }
                }
            }
            c->tile_ti[g][kind] = bti;
            c->tile_tj[g][kind] = btj;
            c->tile_mfma[g][kind] = bmf;
            if (best < 1e30f) kind_ms[kind] = best;
        }
        // the band fused into the last pass (kind 2), or a kind-1 pass and the
        // band as a pass of its own: where the band's half-width keeps the
        // fused pass off the matrix cores, the split form can be faster
        if (g == 0 && status == MGBF_OK && kind_ms[1] >= 0.f && kind_ms[2] >= 0.f) {
            float vms = 0.f;
            status = vertical(c, c->ws_u, c->ws_t, vband, st);
            if (status == MGBF_OK && hipEventRecord(e0, st) != hipSuccess)
                status = fail(MGBF_ERR_HIP, "autotune timing");
            for (int rep = 0; rep < 3 && status == MGBF_OK; ++rep) status = vertical(c, c->ws_u, c->ws_t, vband, st);
            if (status == MGBF_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                      hipEventElapsedTime(&vms, e0, e1) != hipSuccess))
                status = fail(MGBF_ERR_HIP, "autotune timing");
            if (status == MGBF_OK) {
                c->band_split = kind_ms[1] + vms < kind_ms[2];
                if (log)
                    std::fprintf(stderr, "[autotune] gen 1 band: fused %.4f ms, split %.4f + %.4f ms\n",
                                 kind_ms[2] / 3, kind_ms[1] / 3, vms / 3);
            }
        }
    }
    c->timing = timing;
    ++c->version;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return status;
}

int mgbf_get_tile(const mgbf_ctx* c, int gen, int kind, int* ti, int* tj) {
    if (!c || !ti || !tj || gen < 1 || gen > c->gm || kind < 0 || kind > 3)
        return fail(MGBF_ERR_ARG, "bad argument");
    *ti = c->tile_mfma[gen - 1][kind] && mx_ok(c, gen - 1, kind) ? 16 : c->tile_ti[gen - 1][kind];
    if (gen == 1 && kind == 2 && c->band_split) *ti = 0;
    *tj = c->tile_tj[gen - 1][kind];
    return MGBF_OK;
}

int mgbf_set_tile(mgbf_ctx* c, int gen, int kind, int ti, int tj) {
    if (!c || gen < 1 || gen > c->gm || kind < 0 || kind > 3)
        return fail(MGBF_ERR_ARG, "bad argument");
    if (ti == 0 && gen == 1 && kind == 2 && tj >= 1) {
        // the vertical band as a pass of its own, after the kind-1 pass
        c->band_split = true;
        ++c->version;
        return MGBF_OK;
    }
    if (!(ti == 2 || ti == 4 || ti == 16) || tj < 1)
        return fail(MGBF_ERR_ARG, "tile must be ti in {2, 4, 16} (or 0 for generation 1's kind 2) and tj >= 1");
    if (kind == 3 && (ti == 2 || (tj & 1)))
        return fail(MGBF_ERR_ARG, "the pass with the restriction takes ti = 4 or 16 and an even tj");
    if (ti == 16 && !mx_ok(c, gen - 1, kind))
        return fail(MGBF_ERR_UNSUPPORTED,
                    "the matrix-core engine needs half-width <= 8, a disc or product filter, a stack of "
                    "whole 64-k chunks and 16-column tiles (with the vertical band: half-width < 8; with "
                    "the restriction: half-width >= 3, tiles of exact FP16 x weights, at least 32 columns)");
    c->tile_ti[gen - 1][kind] = ti == 16 ? 4 : ti;
    c->tile_tj[gen - 1][kind] = tj;
    c->tile_mfma[gen - 1][kind] = ti == 16 ? 1 : 0;
    if (gen == 1 && kind == 2) c->band_split = false;
    ++c->version;
    return MGBF_OK;
}

}  // extern "C"

// A group applies one operation over several contexts (regional domains of
// uneven size) as one hipGraph: captured once per set of buffers, with one
// branch stream per domain so independent domains overlap, then replayed.
struct mgbf_group {
    std::vector<mgbf_ctx*> ctx;
    int device = 0;
    std::vector<hipStream_t> branch;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int key_op = -1;
    std::vector<const float*> key_in;
    std::vector<float*> key_out;
    std::vector<unsigned long long> key_version;   // each context's version at capture
};

namespace {

void group_drop_graph(mgbf_group* g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->exec = nullptr;
    g->graph = nullptr;
    g->key_op = -1;
}

int group_capture(mgbf_group* g, int op, const float* const* in, float* const* out) {
    group_drop_graph(g);
    const int n = (int)g->ctx.size();
    // workspaces are allocated before the capture begins
    for (int i = 0; i < n; ++i) {
        DeviceGuard guard(g->ctx[i]->device);
        CALL(ensure_workspace(g->ctx[i]));
    }
    std::vector<hipEvent_t> ev(n + 1, nullptr);
    for (auto& e : ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "group event");
    int status = MGBF_OK;
    hipError_t e = hipStreamBeginCapture(g->cap, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) status = hip_fail(e, "graph capture begin");
    if (status == MGBF_OK) {
        (void)hipEventRecord(ev[n], g->cap);   // fork
        for (int i = 0; i < n && status == MGBF_OK; ++i) {
            hipStream_t bs = g->branch[i];
            (void)hipStreamWaitEvent(bs, ev[n], 0);
            mgbf_ctx* c = g->ctx[i];
            const bool timing = c->timing;
            c->timing = false;
            status = apply_op(c, op, in[i], out[i], bs);
            c->timing = timing;
            (void)hipEventRecord(ev[i], bs);
            (void)hipStreamWaitEvent(g->cap, ev[i], 0);   // join
        }
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(g->cap, &graph);
        if (status == MGBF_OK && e != hipSuccess) status = hip_fail(e, "graph capture end");
        if (status == MGBF_OK) {
            g->graph = graph;
            e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) status = hip_fail(e, "graph instantiate");
        } else if (graph) {
            (void)hipGraphDestroy(graph);
        }
    }
    for (auto& ee : ev) (void)hipEventDestroy(ee);
    if (status != MGBF_OK) {
        group_drop_graph(g);
        return status;
    }
    g->key_op = op;
    g->key_in.assign(in, in + n);
    g->key_out.assign(out, out + n);
    g->key_version.resize(n);
    for (int i = 0; i < n; ++i) g->key_version[i] = g->ctx[i]->version;
    return MGBF_OK;
}

}  // namespace

extern "C" {

int mgbf_group_create(mgbf_ctx* const* ctxs, int n, mgbf_group** out) {
    if (!ctxs || !out || n < 1) return fail(MGBF_ERR_ARG, "bad argument");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return fail(MGBF_ERR_ARG, "NULL context in group");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return fail(MGBF_ERR_ARG, "a context appears twice in the group");
        if (ctxs[i]->device != ctxs[0]->device) return fail(MGBF_ERR_ARG, "group spans devices");
    }
    auto* g = new mgbf_group();
    g->ctx.assign(ctxs, ctxs + n);
    g->device = ctxs[0]->device;
    DeviceGuard guard(g->device);
    hipError_t e = hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking);
    g->branch.assign(n, nullptr);
    for (int i = 0; i < n && e == hipSuccess; ++i)
        e = hipStreamCreateWithFlags(&g->branch[i], hipStreamNonBlocking);
    if (e != hipSuccess) {
        mgbf_group_destroy(g);
        return hip_fail(e, "group streams");
    }
    *out = g;
    return MGBF_OK;
}

int mgbf_group_apply(mgbf_group* g, int op, const float* const* in, float* const* out,
                     void* stream) {
    if (!g || !in || !out) return fail(MGBF_ERR_ARG, "NULL argument");
    const int n = (int)g->ctx.size();
    for (int i = 0; i < n; ++i)
        if (!in[i] || !out[i]) return fail(MGBF_ERR_ARG, "NULL buffer in group apply");
    DeviceGuard guard(g->device);
    bool same = g->exec && g->key_op == op &&
                std::equal(g->key_in.begin(), g->key_in.end(), in) &&
                std::equal(g->key_out.begin(), g->key_out.end(), out);
    for (int i = 0; i < n && same; ++i) same = g->key_version[i] == g->ctx[i]->version;
    if (!same) {
        int st = group_capture(g, op, in, out);
        if (st != MGBF_OK) return st;
    }
    HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)), "graph launch");
    return MGBF_OK;
}

int mgbf_group_destroy(mgbf_group* g) {
    if (!g) return MGBF_OK;
    DeviceGuard guard(g->device);
    group_drop_graph(g);
    for (hipStream_t s : g->branch)
        if (s) (void)hipStreamDestroy(s);
    if (g->cap) (void)hipStreamDestroy(g->cap);
    delete g;
    return MGBF_OK;
}

int mgbf_timers_enable(mgbf_ctx* c, int enable) {
    if (!c) return fail(MGBF_ERR_ARG, "ctx is NULL");
    DeviceGuard guard(c->device);
    int st = drain_timers(c);
    c->timing = enable != 0;
    for (int i = 0; i < MGBF_NTIMERS; ++i) {
        c->t_ms[i] = 0;
        c->t_bytes[i] = 0;
        c->t_n[i] = 0;
        for (int g = 0; g < mgbf::kMaxGen; ++g) {
            c->tg_ms[g][i] = 0;
            c->tg_bytes[g][i] = 0;
            c->tg_n[g][i] = 0;
        }
    }
    return st;
}

int mgbf_timers_read(mgbf_ctx* c, double* ms, long long* launches, double* bytes) {
    if (!c) return fail(MGBF_ERR_ARG, "ctx is NULL");
    DeviceGuard guard(c->device);
    int st = drain_timers(c);
    for (int i = 0; i < MGBF_NTIMERS; ++i) {
        if (ms) ms[i] = c->t_ms[i];
        if (launches) launches[i] = c->t_n[i];
        if (bytes) bytes[i] = c->t_bytes[i];
    }
    return st;
}

int mgbf_timers_read_gen(mgbf_ctx* c, int gen, double* ms, long long* launches, double* bytes) {
    if (!c) return fail(MGBF_ERR_ARG, "ctx is NULL");
    if (gen < 1 || gen > c->gm) return fail(MGBF_ERR_ARG, "generation out of range");
    DeviceGuard guard(c->device);
    int st = drain_timers(c);
    for (int i = 0; i < MGBF_NTIMERS; ++i) {
        if (ms) ms[i] = c->tg_ms[gen - 1][i];
        if (launches) launches[i] = c->tg_n[gen - 1][i];
        if (bytes) bytes[i] = c->tg_bytes[gen - 1][i];
    }
    return st;
}

}  // extern "C"
