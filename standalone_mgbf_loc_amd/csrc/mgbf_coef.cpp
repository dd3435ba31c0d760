// Host-side FP64 setup of the MGBF operator.
//
// Derived sizes follow init_mg_parameter / def_maxgen / def_ngens
// (src.mgbf/mg_parameter.f90:179-635). Filter coefficients follow
// def_mg_weights (src.mgbf/mg_intstate.f90:296-433): the aspect tensor is
// turned into the Cholesky factor of its inverse (cholaspect2,
// src.mgbf/jp_pbfil.f90:90-107), the taps are the lattice points the
// reference's vrbeta2 loop visits (jp_pbfil.f90:909-926) and the line-sum
// normalizer is getlinesum2 (jp_pbfil.f90:182-218). The transfer constants
// are those of def_offset_coef (src.mgbf/mg_interpolate.f90:76-84) applied as
// direct1 / adjoint do (src.mgbf/mg_generations.f90:1117-1250) with the
// 2-point half-sample mirror of boco_2d / bocoT_2d, or with lquart as
// direct2 / adjoint2 do (mg_generations.f90:1253-1374) with a 1-point one.
#include "mgbf_coef.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace mgbf {
namespace {

constexpr double kEps = 1.e-12;  // tap-boundary tolerance of jp_pbfil.f90

const double kPCoef[4] = {-9. / 128., 111. / 128., 29. / 128., -3. / 128.};
const double kQCoef[4] = {-3. / 128., 29. / 128., 111. / 128., -9. / 128.};
const double kACoef[3] = {5. / 32., 30. / 32., -3. / 32.};   // a_coef
const double kBCoef[3] = {-3. / 32., 30. / 32., 5. / 32.};   // b_coef

double ipow(double x, int p) {
    double r = 1.0;
    for (int i = 0; i < p; ++i) r *= x;
    return r;
}

std::string fmt(const char* f, int a = 0, int b = 0, int c = 0) {
    char buf[256];
    std::snprintf(buf, sizeof buf, f, a, b, c);
    return buf;
}

// def_maxgen (mg_parameter.f90:578-611)
int def_maxgen(int nxm, int nym) {
    int npx = nxm, gx = 1;
    for (;;) {
        npx = (npx + 1) / 2;
        gx += 1;
        if (npx == 1) break;
    }
    int npy = nym, gy = 1;
    for (;;) {
        npy = (npy + 1) / 2;
        gy += 1;
        if (npy == 1) break;
    }
    return gx < gy ? gx : gy;
}

int mirror_half(int q, int n) {  // 0-based half-sample mirror
    if (q < 0) return -1 - q;
    if (q >= n) return 2 * n - 1 - q;
    return q;
}

int mirror_whole(int q, int n) {  // 0-based whole-sample mirror
    if (q < 0) return -q;
    if (q >= n) return 2 * (n - 1) - q;
    return q;
}

// Prolongation along one axis, fine nf <- coarse nf/2 (dense, FP64), with a
// flag matrix of structural nonzeros. 4-point (direct1): fine i1 (1-based)
// reads coarse iL-1 .. iL+2, iL = i1/2, padded by boco_2d(.,2,2). Quadratic
// (direct2): odd i1 reads coarse iL-1 .. iL+1 with a_coef, iL = (i1+1)/2,
// even i1 the same with b_coef, iL = i1/2, padded by boco_2d(.,1,1).
void interp_dense(int nf, bool quad, std::vector<double>& P, std::vector<char>& nz) {
    const int nc = nf / 2;
    P.assign((size_t)nf * nc, 0.0);
    nz.assign((size_t)nf * nc, 0);
    for (int i1 = 1; i1 <= nf; ++i1) {
        const bool odd = (i1 & 1) != 0;
        const int np = quad ? 3 : 4;
        const int lo = quad ? (odd ? (i1 + 1) / 2 : i1 / 2) - 1 : i1 / 2 - 1;
        const double* c = quad ? (odd ? kACoef : kBCoef) : (odd ? kQCoef : kPCoef);
        for (int t = 0; t < np; ++t) {
            const int I1 = lo + t;                     // 1-based coarse
            const int I0 = mirror_half(I1 - 1, nc);    // boco_2d pad
            P[(size_t)(i1 - 1) * nc + I0] += c[t];
            nz[(size_t)(i1 - 1) * nc + I0] = 1;
        }
    }
}

// Packs the rows of a (rows x cols) banded matrix into a window table.
bool pack_rows(int rows, int cols, const std::vector<double>& M, const std::vector<char>& nz,
               bool transpose, int width, AxisTable& t) {
    t.n_out = rows;
    t.width = width;
    t.start.assign(rows, 0);
    t.w.assign((size_t)rows * width, 0.0);
    for (int r = 0; r < rows; ++r) {
        int lo = -1, hi = -1;
        for (int c = 0; c < cols; ++c) {
            const size_t e = transpose ? (size_t)c * rows + r : (size_t)r * cols + c;
            if (nz[e]) {
                if (lo < 0) lo = c;
                hi = c;
            }
        }
        if (lo < 0) {
            lo = 0;
            hi = -1;
        }
        if (hi - lo + 1 > width) return false;
        // keep the window inside [0, cols)
        int s = lo;
        if (s + width > cols) s = cols - width;
        if (s < 0) s = 0;
        t.start[r] = s;
        for (int c = lo; c <= hi; ++c) {
            const size_t e = transpose ? (size_t)c * rows + r : (size_t)r * cols + c;
            t.w[(size_t)r * width + (c - s)] = M[e];
        }
    }
    return true;
}

// Widest structural row (or column, transposed) span of a banded matrix.
int max_span(int rows, int cols, const std::vector<char>& nz, bool transpose) {
    int wmax = 0;
    for (int r = 0; r < rows; ++r) {
        int lo = -1, hi = -1;
        for (int c = 0; c < cols; ++c) {
            const size_t e = transpose ? (size_t)c * rows + r : (size_t)r * cols + c;
            if (nz[e]) {
                if (lo < 0) lo = c;
                hi = c;
            }
        }
        if (lo >= 0 && hi - lo + 1 > wmax) wmax = hi - lo + 1;
    }
    return wmax;
}

// One column of the vertical transfer, lm filter levels -> la analysis
// levels, as a dense la x lm matrix (row n, column k, both 0-based).
// vmap 0, the linear map (l_vertical_direct_spec2 / l_vertical_direct_spec,
// mg_interpolate.f90:595-631, 1006-1044): odd analysis level n (1-based)
// copies filter level (n+1)/2, even n averages levels n/2 and n/2+1. vmap 1,
// lwq_vertical_direct_spec (:511-553) with the weights of lwq_vertical_coef
// (:301-370): analysis height y = (n-1)(lm-1)/(la-1) over filter heights
// 0..lm-1, kref = floor(y)+1, the two quadratic interpolants through
// kref-1..kref+1 and kref..kref+2 (1-based levels) blended linearly between
// kref and kref+1; at kref = 1 (and lm-1) the weight of the ghost level
// below (above) the column folds onto its whole-sample mirror image; the end
// levels copy.
std::vector<double> vertical_column(int lm, int la, int vmap) {
    std::vector<double> V((size_t)la * lm, 0.0);
    auto at = [&](int n, int k) -> double& { return V[(size_t)(n - 1) * lm + (k - 1)]; };
    if (vmap == 0) {
        for (int n = 1; n <= la; ++n) {
            if (n & 1) {
                at(n, (n + 1) / 2) = 1.0;
            } else {
                at(n, n / 2) += 0.5;
                at(n, n / 2 + 1) += 0.5;
            }
        }
        return V;
    }
    const double dy = 1.0 * (lm - 1) / (la - 1);
    for (int n = 2; n <= la - 1; ++n) {
        const double y = (n - 1) * dy;
        const int kr = (int)(y + 1.0);
        // nodes at heights kr-2 .. kr+1 (levels kr-1 .. kr+2)
        double z[4], ql[3], qr[3];
        for (int t = 0; t < 4; ++t) z[t] = kr - 2 + t;
        auto lag3 = [&](const double* zz, double* l) {
            l[0] = (y - zz[1]) * (y - zz[2]) / ((zz[0] - zz[1]) * (zz[0] - zz[2]));
            l[1] = (y - zz[0]) * (y - zz[2]) / ((zz[1] - zz[0]) * (zz[1] - zz[2]));
            l[2] = (y - zz[0]) * (y - zz[1]) / ((zz[2] - zz[0]) * (zz[2] - zz[1]));
        };
        lag3(z, ql);
        lag3(z + 1, qr);
        const double bl = z[2] - y, br = y - z[1];
        double c[4] = {bl * ql[0], bl * ql[1] + br * qr[0], bl * ql[2] + br * qr[1], br * qr[2]};
        if (kr == 1) {
            c[2] += c[0];
            c[0] = 0.0;
        }
        if (kr == lm - 1) {
            c[1] += c[3];
            c[3] = 0.0;
        }
        for (int t = 0; t < 4; ++t)
            if (c[t] != 0.0) at(n, kr - 1 + t) += c[t];
    }
    at(1, 1) = 1.0;
    at(la, lm) = 1.0;
    return V;
}

// Direct (rows: analysis k) and transpose (rows: filter k) window tables
// of the vertical transfer tiled over ncol columns, windows inside each
// column.
bool vertical_transfer_tables(int ncol, int lm, int la, int vmap, AxisTable& dir, AxisTable& adj) {
    const std::vector<double> V = vertical_column(lm, la, vmap);
    std::vector<char> nz(V.size());
    for (size_t e = 0; e < V.size(); ++e) nz[e] = V[e] != 0.0;
    AxisTable d1, a1;
    if (!pack_rows(la, lm, V, nz, false, max_span(la, lm, nz, false), d1) ||
        !pack_rows(lm, la, V, nz, true, max_span(lm, la, nz, true), a1))
        return false;
    auto tile = [&](const AxisTable& one, int nrow, int ncolk, AxisTable& t) {
        t.n_out = ncol * nrow;
        t.width = one.width;
        t.start.resize(t.n_out);
        t.w.resize((size_t)t.n_out * t.width);
        for (int c = 0; c < ncol; ++c)
            for (int r = 0; r < nrow; ++r) {
                t.start[(size_t)c * nrow + r] = c * ncolk + one.start[r];
                for (int u = 0; u < t.width; ++u)
                    t.w[((size_t)c * nrow + r) * t.width + u] = one.w[(size_t)r * t.width + u];
            }
    };
    tile(d1, la, lm, dir);
    tile(a1, lm, la, adj);
    return true;
}

}  // namespace

bool anal_axis_matrix(int npe, int nloc_a, int nloc_f, int scheme, std::vector<double>& T) {
    // Per PE tile the reference places the analysis points at cell centres
    // xa(n) = dxa/2 + dxa (n-1) and the filter points at xf(i) = dxf/2 +
    // dxf (i-1), with the tile length equal to nloc_a (mg_parameter.f90:
    // 518-541), so dxa = 1 and dxf = nloc_a/nloc_f. With i0 the first i in
    // 1-ib .. im+ib-1 (ib = 6) with xa(n) < xf(i) (lsqr_mg_coef,
    // mg_interpolate.f90:129-149), analysis point n takes
    //   lsqr: filter points i0-2 .. i0+1, the quadratic Lagrange interpolants
    //         through the first and the last three blended linearly between
    //         the middle two (:156-191);
    //   lin:  points i0-1, i0, linear weights (:269-296);
    //   quad: points i0-1 .. i0+1, quadratic Lagrange weights (:231-267).
    // Filter indices outside the global domain take the half-sample mirror
    // (boco_2d); inside, tiles join up seamlessly.
    const int na = npe * nloc_a, nf = npe * nloc_f;
    T.assign((size_t)na * nf, 0.0);
    const double len = 1.0 * nloc_a;
    const double dxa = len / nloc_a, dxf = len / nloc_f;
    const double xa0 = 0.5 * dxa, xf0 = 0.5 * dxf;
    const int ib = 6;
    // Lagrange basis of the nodes z[0..2] at x
    auto lag3 = [](double x, const double* z, double* l) {
        l[0] = (x - z[1]) * (x - z[2]) / ((z[0] - z[1]) * (z[0] - z[2]));
        l[1] = (x - z[0]) * (x - z[2]) / ((z[1] - z[0]) * (z[1] - z[2]));
        l[2] = (x - z[0]) * (x - z[1]) / ((z[2] - z[0]) * (z[2] - z[1]));
    };
    for (int n = 1; n <= nloc_a; ++n) {
        const double x = xa0 + dxa * (n - 1);
        int i0 = 0;
        for (int i = 1 - ib; i <= nloc_f + ib - 1 && !i0; ++i)
            if (x < xf0 + dxf * (i - 1)) i0 = i;
        if (!i0) return false;
        int first = 0, width = 0;
        double c[4] = {};
        if (scheme == 1) {
            first = i0 - 1;
            width = 2;
            const double z0 = xf0 + dxf * (first - 1), z1 = xf0 + dxf * first;
            c[0] = (z1 - x) / (z1 - z0);
            c[1] = (x - z0) / (z1 - z0);
        } else if (scheme == 2) {
            first = i0 - 1;
            width = 3;
            double z[3];
            for (int t = 0; t < 3; ++t) z[t] = xf0 + dxf * (first + t - 1);
            lag3(x, z, c);
        } else {
            first = i0 - 2;
            width = 4;
            double z[4];
            for (int t = 0; t < 4; ++t) z[t] = xf0 + dxf * (first + t - 1);
            double ql[3], qr[3];
            lag3(x, z, ql);
            lag3(x, z + 1, qr);
            const double bl = (z[2] - x) / (z[2] - z[1]);   // weight of the left interpolant
            const double br = (x - z[1]) / (z[2] - z[1]);
            c[0] = bl * ql[0];
            c[1] = bl * ql[1] + br * qr[0];
            c[2] = bl * ql[2] + br * qr[1];
            c[3] = br * qr[2];
        }
        for (int p = 0; p < npe; ++p) {
            const int row = p * nloc_a + n - 1;
            for (int t = 0; t < width; ++t) {
                int I = p * nloc_f + first + t;   // 1-based global filter index
                if (I < 1) I = 1 - I;
                if (I > nf) I = 2 * nf + 1 - I;
                if (I < 1 || I > nf) return false;
                T[(size_t)row * nf + (I - 1)] += c[t];
            }
        }
    }
    return true;
}

// The taps of one aspect tensor A = [[a00, a01], [a01, a11]]: cholaspect2
// (the lower Cholesky factor of A^-1, jp_pbfil.f90:90-107), the lattice
// offsets vrbeta2 visits (:909-926) and getlinesum2's normalizer
// (:182-218). 0 on success, 1 if the reach exceeds hy, 2 if it exceeds hx
// (getlinesum2 stops there).
int radial_taps_tensor(const mgbf_config& cfg, double a00, double a11, double a01, std::vector<Tap2>& taps) {
    taps.clear();
    const int p = cfg.p;
    const double rmom2_2 = 1.0 / std::sqrt(2.0 * p + 4.0);
    // inverse, then its lower Cholesky factor
    const double det = a00 * a11 - a01 * a01;
    const double i00 = a11 / det, i11 = a00 / det, i10 = -a01 / det;
    const double b00 = std::sqrt(i00);
    const double b10 = i10 / b00;
    const double b11 = std::sqrt(i11 - b10 * b10);
    const double exx = b00 * rmom2_2, eyy = b11 * rmom2_2, eyx = b10 * rmom2_2;
    const double y = 1.0 / eyy;
    const int gyl = (int)std::ceil(-y + kEps), gym = (int)std::floor(y - kEps);
    if (gyl < -cfg.hy || gym > cfg.hy) return 1;
    std::vector<double> rrc;
    for (int gy = gyl; gy <= gym; ++gy) {
        const double yy = gy;
        const double xc = -yy * eyx;
        const double rrx = (yy * eyy) * (yy * eyy);
        const double x = std::sqrt(1.0 - rrx);
        const int gxl = (int)std::ceil((xc - x) / exx + kEps);
        const int gxm = (int)std::floor((xc + x) / exx - kEps);
        if (gxl < -cfg.hx || gxm > cfg.hx) return 2;
        for (int gx = gxl; gx <= gxm; ++gx) {
            const double t = gx * exx - xc;
            taps.push_back({gx, gy, 0.0});
            rrc.push_back(1.0 - (rrx + t * t));
        }
    }
    double s = 0;
    for (double r : rrc) s += ipow(r, p);
    const double ss = 1.0 / s;
    for (size_t t = 0; t < taps.size(); ++t) taps[t].w = ss * ipow(rrc[t], p);
    return 0;
}

bool radial_taps(const mgbf_config& cfg, std::vector<Tap2>& taps) {
    // aspect tensor pasp2 with p_del = p_eps = 0 (mg_intstate.f90:316-319, 383-390)
    return radial_taps_tensor(cfg, cfg.mg_ampl02, cfg.mg_ampl02, 0.0, taps) == 0;
}

int aspect_weights(const mgbf_config& cfg, int H, int n, int m, const double* pasp2, std::vector<float>& w,
                   std::string* err) {
    const int NB = 2 * H + 1, nt = NB * NB;
    w.assign((size_t)n * m * nt, 0.f);
    std::vector<Tap2> taps;
    for (size_t q = 0; q < (size_t)n * m; ++q) {
        const double a00 = pasp2[3 * q], a11 = pasp2[3 * q + 1], a01 = pasp2[3 * q + 2];
        if (!(a00 > 0.0) || !(a11 > 0.0) || !(a00 * a11 - a01 * a01 > 0.0)) {
            *err = fmt("aspect tensor at point %d is not positive definite", (int)q);
            return MGBF_ERR_ARG;
        }
        const int r = radial_taps_tensor(cfg, a00, a11, a01, taps);
        if (r != 0) {
            *err = r == 1 ? "In getlinesum2; filter reach becomes too large for hy"
                          : "In getlinesum2; filter reach becomes too large for hx";
            return MGBF_ERR_CONFIG;
        }
        for (const Tap2& t : taps) {
            if (std::abs(t.gx) > H || std::abs(t.gy) > H) {
                *err = "aspect taps beyond the radial half-width";
                return MGBF_ERR_CONFIG;
            }
            w[q * nt + (size_t)(t.gy + H) * NB + (t.gx + H)] = (float)t.w;
        }
    }
    return MGBF_OK;
}

// 1-D beta taps of getlinesum1 / vrbeta1 (jp_pbfil.f90:149-177, 856-885)
// for the factor el of the aspect: offsets gx with |gx el'| < 1, el' = el /
// sqrt(2p+3), weights (1 - (gx el')^2)^p normalized to sum 1. False if the
// reach exceeds the halo.
bool beta1_taps(double el, int halo, int p, std::vector<int>& gx, std::vector<double>& w) {
    gx.clear();
    w.clear();
    const double rmom2_1 = 1.0 / std::sqrt(2.0 * p + 3.0);
    const double exx = el * rmom2_1;
    const double x = 1.0 / exx;
    const int gxl = (int)std::ceil(-x + kEps), gxm = (int)std::floor(x - kEps);
    if (gxl < -halo || gxm > halo) return false;
    double s = 0;
    for (int g = gxl; g <= gxm; ++g) {
        const double rr = (g * exx) * (g * exx);
        gx.push_back(g);
        w.push_back(ipow(1.0 - rr, p));
        s += w.back();
    }
    for (double& v : w) v /= s;
    return true;
}

bool vertical_taps(const mgbf_config& cfg, std::vector<int>& gx, std::vector<double>& w) {
    return beta1_taps(1.0 / std::sqrt(cfg.mg_ampl01), cfg.hz, cfg.p, gx, w);   // cholaspect1
}

// Triad line filters of filtering_lin2_* (jp_pbfil3.f90). init_mg_line
// (mg_intstate.f90:436-480) runs after def_mg_weights has replaced pasp2 by
// the Cholesky factor of its inverse, so the 3-vector that reaches tritform
// is (L11, L22, L12) = (1/sqrt(a), 1/sqrt(a), 0), a = mg_ampl02 (p_del =
// p_eps = 0). tritform resolves it onto three lattice lines (triad) and
// turns each line's spread**2 into a real half-span (hstform); each line is
// then filtered by the mixture of the normalized beta filters of the two
// neighbouring integer half-spans (blinfil) along its generator.
namespace {

constexpr int kLineNH = 6;     // nh (mg_parameter.f90:562)
constexpr int kLineNFil = 8;   // nfil = nh + 2

// Normalizer and spread**2 of the beta filter (1-(i/h)^2)^p, |i| < h, for
// h = 1..nh: the discrete sums that inimomtab evaluates in closed form.
void line_moments(int p, double* bnorm, double* bsprds) {
    for (int h = 1; h <= kLineNH; ++h) {
        double s0 = 0.0, s2 = 0.0;
        for (int i = -h + 1; i <= h - 1; ++i) {
            const double z = (double)i / h;
            const double w = ipow(1.0 - z * z, p);
            s0 += w;
            s2 += (double)i * i * w;
        }
        bnorm[h] = 1.0 / s0;
        bsprds[h] = s2 / s0;
    }
}

// hstform (jp_pbfil3.f90:1399-1430): spread**2 -> real half-span by linear
// interpolation in the table of spreads**2; false when negative or beyond it.
bool line_hstform(const double* bsprds, double& hs) {
    const bool neg = hs < 0.0;
    for (int h = 2; h <= kLineNH; ++h)
        if (bsprds[h] >= hs) {
            hs = h - (bsprds[h] - hs) / (bsprds[h] - bsprds[h - 1]);
            return !neg;
        }
    return false;
}

// blinfil (jp_pbfil3.f90:1463-1501): the half dibeta filter fil[0..h] of
// half-span hs, the mixture of the normalized betas of half-spans h+1 and h,
// h = int(hs); empty when h < 1 or h+1 leaves the table.
std::vector<double> line_blinfil(const double* bnorm, int p, double hs) {
    const int h = (int)hs, hp = h + 1;
    if (h < 1 || hp > kLineNH || hp > kLineNFil) return {};
    std::vector<double> fil(h + 1, 0.0);
    const double whp = (hs - h) * bnorm[hp], wh = (hp - hs) * bnorm[h];
    for (int i = 0; i <= h; ++i) {
        const double z = (double)i / hp;
        fil[i] = whp * ipow(1.0 - z * z, p);
    }
    for (int i = 0; i < h; ++i) {
        const double z = (double)i / h;
        fil[i] += wh * ipow(1.0 - z * z, p);
    }
    return fil;
}

bool is_identity(const std::vector<double>& fil) {
    bool id = !fil.empty() && fil[0] == 1.0;
    for (size_t i = 1; i < fil.size(); ++i) id = id && fil[i] == 0.0;
    return id;
}

}  // namespace

// Line decompositions of the aspect tensor that init_mg_line hands to
// tritform / hextform (mg_intstate.f90:436-480). A namelist reaches diagonal
// tensors only (def_mg_weights fixes p_del = p_eps = 0, mg_intstate.f90:
// 316-319), so both are built in closed form for A = diag(a_1, .., a_d),
// every a_e > 0, and anything else is refused (false):
//   A = sum_c w_c l_c l_c^T, w_c >= 0, over the lines of a superbase.
// The axes with w = a_e already give A; every other line of the superbase
// carries w = 0 and its filter is the identity. The colors (the order in
// which the reference applies the lines) are those of the reference's
// starting superbase, which its flip iterations (jp_pbfil3.f90 triad and
// hexad) leave alone when no starting weight is negative, as here.
//
// 2-D (tritform): colors 1, 2 = the x and y axes with a_1, a_2; color 3 =
// the idle diagonal (-1, -1). a = (A11, A22, A12).
bool line_triad(const double a[3], int l[3][2], double w[3]) {
    if (a[2] != 0.0 || !(a[0] > 0.0) || !(a[1] > 0.0)) return false;
    const int lines[3][2] = {{1, 0}, {0, 1}, {-1, -1}};
    const double wt[3] = {a[0], a[1], 0.0};
    for (int c = 0; c < 3; ++c) {
        l[c][0] = lines[c][0];
        l[c][1] = lines[c][1];
        w[c] = wt[c];
    }
    return true;
}

// 3-D (hextform): colors 1, 2, 3 = the x, y and level axes with a_1, a_2,
// a_3; colors 4, 5, 6 = the idle face diagonals (1,-1,0), (-1,0,1),
// (0,1,-1); color 7 = the null generator. a = (A11, A22, A33, A23, A31, A12).
bool line_hexad(const double a[6], int l7[7][3], double w7[7]) {
    if (a[3] != 0.0 || a[4] != 0.0 || a[5] != 0.0) return false;
    if (!(a[0] > 0.0) || !(a[1] > 0.0) || !(a[2] > 0.0)) return false;
    const int diag[3][3] = {{1, -1, 0}, {-1, 0, 1}, {0, 1, -1}};
    for (int c = 0; c < 7; ++c) {
        w7[c] = c < 3 ? a[c] : 0.0;
        for (int e = 0; e < 3; ++e) l7[c][e] = c < 3 ? (e == c ? 1 : 0) : (c < 6 ? diag[c - 3][e] : 0);
    }
    return true;
}

int hex_taps(const mgbf_config& cfg, std::vector<Tap3>& taps, bool* zero_levels, std::string* err,
             std::vector<double>* factors) {
    taps.clear();
    *zero_levels = true;
    if (factors)
        for (int e = 0; e < 3; ++e) factors[e].assign(1, 1.0);
    if (!cfg.mgbf_line) {
        // without init_mg_line no color has a line: every call copies
        taps.push_back({0, 0, 0, 1.0});
        return MGBF_OK;
    }
    if (cfg.p < 1 || cfg.p > 6) {
        *err = " In inimomtab; prescribed exponent p out of bounds";
        return MGBF_ERR_CONFIG;
    }
    double bnorm[kLineNH + 1], bsprds[kLineNH + 1];
    line_moments(cfg.p, bnorm, bsprds);
    // the Cholesky factor of pasp3 = mg_ampl03 I (p_del = p_eps = 0) as
    // t33_to_6 hands it to hextform: (s, s, s, 0, 0, 0)
    const double el = 1.0 / std::sqrt(cfg.mg_ampl03);
    const double asp[6] = {el, el, el, 0.0, 0.0, 0.0};
    int l7[7][3];
    double w7[7];
    if (!line_hexad(asp, l7, w7)) {
        *err = " In hextform; hexad, failed; check aspect tensor";
        return MGBF_ERR_CONFIG;
    }
    // hextform: the six non-null lines in color order, each as a dibeta filter
    std::vector<double> f[3] = {{1.0}, {1.0}, {1.0}};
    int axis_color[3] = {0, 0, 0};
    int nline = 0;
    for (int c = 1; c <= 7; ++c) {
        const int* g = l7[c - 1];
        if (std::abs(g[0]) + std::abs(g[1]) + std::abs(g[2]) == 0) continue;
        ++nline;
        double hs = w7[c - 1];
        if (!line_hstform(bsprds, hs)) {
            *err = fmt(" In hextform; hstform failed at i=%d", nline);
            return MGBF_ERR_CONFIG;
        }
        std::vector<double> fil = line_blinfil(bnorm, cfg.p, hs);
        if (fil.empty()) {
            *err = "hexad line half-span outside the moment table";
            return MGBF_ERR_CONFIG;
        }
        if (is_identity(fil)) continue;
        int ax = -1;
        for (int e = 0; e < 3; ++e)
            if (std::abs(g[e]) == 1 && g[(e + 1) % 3] == 0 && g[(e + 2) % 3] == 0) ax = e;
        if (ax < 0 || axis_color[ax]) {
            *err = "line filters off the grid axes are not built";
            return MGBF_ERR_UNSUPPORTED;
        }
        const int halo[3] = {cfg.hx, cfg.hy, cfg.hz};
        if ((int)fil.size() - 1 > halo[ax]) {
            *err = "line filter reach exceeds the halo";
            return MGBF_ERR_CONFIG;
        }
        f[ax] = fil;
        axis_color[ax] = c;
    }
    if (nline != 6) {
        *err = " In hextform; inconsistent hexad generator set found";
        return MGBF_ERR_CONFIG;
    }
    // filtering_lin3 mirrors the level halo once, before the color loop;
    // every call ends with a = b, which zeroes the halos: only a line of
    // color 1 reads the mirror, the others read zeros beyond the column ends
    // (and the adjoint folds no level halo)
    if (axis_color[2] == 1) {
        *err = "a level line of the first color is not built";
        return MGBF_ERR_UNSUPPORTED;
    }
    if (factors)
        for (int e = 0; e < 3; ++e) factors[e] = f[e];
    const int hx = (int)f[0].size() - 1, hy = (int)f[1].size() - 1, hz = (int)f[2].size() - 1;
    for (int c = -hz; c <= hz; ++c)
        for (int b = -hy; b <= hy; ++b)
            for (int a = -hx; a <= hx; ++a)
                taps.push_back({a, b, c, f[0][std::abs(a)] * f[1][std::abs(b)] * f[2][std::abs(c)]});
    return MGBF_OK;
}

int line_taps(const mgbf_config& cfg, std::vector<Tap2>& taps, std::string* err) {
    taps.clear();
    if (!cfg.mgbf_line) {
        // without init_mg_line every generator is 0: each line is the identity
        taps.push_back({0, 0, 1.0});
        return MGBF_OK;
    }
    if (cfg.p < 1 || cfg.p > 6) {
        *err = " In inimomtab; prescribed exponent p out of bounds";
        return MGBF_ERR_CONFIG;
    }
    double bnorm[kLineNH + 1], bsprds[kLineNH + 1];
    line_moments(cfg.p, bnorm, bsprds);
    const double el = 1.0 / std::sqrt(cfg.mg_ampl02);
    const double asp[3] = {el, el, 0.0};
    int l[3][2];
    double hs[3];
    if (!line_triad(asp, l, hs)) {
        *err = " In tritform; triad failed; check aspect tensor";
        return MGBF_ERR_CONFIG;
    }
    // 1-D half filters along x and y (identity unless a line says otherwise)
    std::vector<double> fx{1.0}, fy{1.0};
    bool have_x = false, have_y = false;
    for (int c = 0; c < 3; ++c) {
        if (!line_hstform(bsprds, hs[c])) {
            *err = fmt(" In tritform; hstform failed at i=%d", c + 1);
            return MGBF_ERR_CONFIG;
        }
        if (l[c][0] == 0 && l[c][1] == 0) continue;
        // a failing blinfil makes vdibeta2 leave the field as it is
        const std::vector<double> fil = line_blinfil(bnorm, cfg.p, hs[c]);
        if (fil.empty() || is_identity(fil)) continue;
        const int h = (int)fil.size() - 1;
        const bool xline = l[c][1] == 0 && std::abs(l[c][0]) == 1;
        const bool yline = l[c][0] == 0 && std::abs(l[c][1]) == 1;
        if ((xline && have_x) || (yline && have_y) || (!xline && !yline)) {
            *err = "line filters off the grid axes are not built";
            return MGBF_ERR_UNSUPPORTED;
        }
        if ((xline && h > cfg.hx) || (yline && h > cfg.hy)) {
            *err = "line filter reach exceeds the halo";
            return MGBF_ERR_CONFIG;
        }
        (xline ? fx : fy) = fil;
        (xline ? have_x : have_y) = true;
    }
    // the lines apply one after the other with the half-sample mirror in
    // between: on the grid axes that is the product filter fx(a) fy(b)
    const int hxl = (int)fx.size() - 1, hyl = (int)fy.size() - 1;
    for (int b = -hyl; b <= hyl; ++b)
        for (int a = -hxl; a <= hxl; ++a) taps.push_back({a, b, fx[std::abs(a)] * fy[std::abs(b)]});
    return MGBF_OK;
}

// getlinesum3's reach test (jp_pbfil.f90:220-262) for pasp3 = mg_ampl03 I
// (p_del = p_sig = p_rho = 0): the lattice ball |g| < sqrt(mg_ampl03 (2p+5))
// visited z, then y, then x, with the message of the first halo exceeded.
bool ball_reach_ok(const mgbf_config& cfg, std::string* err) {
    const double e = std::sqrt(1.0 / cfg.mg_ampl03) / std::sqrt(2.0 * cfg.p + 5.0);
    const double z = 1.0 / e;
    const int gzl = (int)std::ceil(-z + kEps), gzm = (int)std::floor(z - kEps);
    if (gzl < -cfg.hz || gzm > cfg.hz) {
        *err = "In getlinesum3; filter reach becomes too large for hz";
        return false;
    }
    for (int gz = gzl; gz <= gzm; ++gz) {
        const double rry = (gz * e) * (gz * e);
        const double y = std::sqrt(1.0 - rry);
        const int gyl = (int)std::ceil(-y / e + kEps), gym = (int)std::floor(y / e - kEps);
        if (gyl < -cfg.hy || gym > cfg.hy) {
            *err = "In getlinesum3; filter reach becomes too large for hy";
            return false;
        }
        for (int gy = gyl; gy <= gym; ++gy) {
            const double rrx = rry + (gy * e) * (gy * e);
            const double x = std::sqrt(std::max(0.0, 1.0 - rrx));
            const int gxl = (int)std::ceil(-x / e + kEps), gxm = (int)std::floor(x / e - kEps);
            if (gxl < -cfg.hx || gxm > cfg.hx) {
                *err = "In getlinesum3; filter reach becomes too large for hx";
                return false;
            }
        }
    }
    return true;
}

// Taps of the 3-D filter of procedure 4 (getlinesum3 + the vrbeta3 loop,
// jp_pbfil.f90:220-262, 930-977) for pasp3 = mg_ampl03 I: the lattice ball
// |g| < sqrt(mg_ampl03 (2p+5)) visited z, y, x, weights ss3 (1-|g e|^2)^p.
void ball_taps(const mgbf_config& cfg, std::vector<Tap3>& taps) {
    taps.clear();
    const double e = std::sqrt(1.0 / cfg.mg_ampl03) / std::sqrt(2.0 * cfg.p + 5.0);
    const double z = 1.0 / e;
    const int gzl = (int)std::ceil(-z + kEps), gzm = (int)std::floor(z - kEps);
    std::vector<double> rrc;
    for (int gz = gzl; gz <= gzm; ++gz) {
        const double rry = (gz * e) * (gz * e);
        const double y = std::sqrt(1.0 - rry);
        const int gyl = (int)std::ceil(-y / e + kEps), gym = (int)std::floor(y / e - kEps);
        for (int gy = gyl; gy <= gym; ++gy) {
            const double rrx = rry + (gy * e) * (gy * e);
            const double x = std::sqrt(std::max(0.0, 1.0 - rrx));
            const int gxl = (int)std::ceil(-x / e + kEps), gxm = (int)std::floor(x / e - kEps);
            for (int gx = gxl; gx <= gxm; ++gx) {
                taps.push_back({gx, gy, gz, 0.0});
                rrc.push_back(1.0 - (rrx + (gx * e) * (gx * e)));
            }
        }
    }
    double s = 0;
    for (double r : rrc) s += ipow(r, cfg.p);
    for (size_t t = 0; t < taps.size(); ++t) taps[t].w = ipow(rrc[t], cfg.p) / s;
}

bool fast_taps(const mgbf_config& cfg, std::vector<Tap2>& taps) {
    // the fast filter's x and y factors keep the aspect amplitude as the
    // factor itself: paspx = paspy = mg_ampl02 reach getlinesum1 and vrbeta1
    // without cholaspect (mg_intstate.f90:370-375, 422-431)
    std::vector<int> gx, gy;
    std::vector<double> wx, wy;
    if (!beta1_taps(cfg.mg_ampl02, cfg.hx, cfg.p, gx, wx) || !beta1_taps(cfg.mg_ampl02, cfg.hy, cfg.p, gy, wy))
        return false;
    taps.clear();
    for (size_t b = 0; b < gy.size(); ++b)
        for (size_t a = 0; a < gx.size(); ++a) taps.push_back({gx[a], gy[b], wx[a] * wy[b]});
    return true;
}

const char* transfer_all2_refusal(const mgbf_config& c) {
    if (c.lm_a == c.lm) return nullptr;
    // l_vertical_*_spec2 (mg_interpolate.f90:963-1044) run over km3*n_ens
    // columns of lm_a levels: only the 3-D variables, with every analysis
    // level written and every filter level inside the column
    if (c.km2 > 0)
        return "anal_to_filt_all2 / filt_to_anal_all2 with lm_a > lm need km2 = 0 (the vertical map "
               "covers the 3-D variables only)";
    if (c.lm_a % 2 == 0 || c.lm_a > 2 * c.lm - 1)
        return "anal_to_filt_all2 / filt_to_anal_all2 with lm_a > lm need an odd lm_a <= 2 lm - 1 (the "
               "linear vertical map writes odd levels from level (n+1)/2)";
    return nullptr;
}

const char* transfer_all_refusal(const mgbf_config& c) {
    // S2C_ens / C2S_ens (mg_transfer.f90:269-336) move the 3-D variables
    // only: the 2-D ones would keep whatever the target array held
    if (c.km2 > 0)
        return "anal_to_filt_all / filt_to_anal_all need km2 = 0 (S2C_ens and C2S_ens carry the 3-D "
               "variables only)";
    if (c.lm_a == c.lm) return nullptr;
    if (c.l_lin_vertical && (c.lm_a % 2 == 0 || c.lm_a > 2 * c.lm - 1))
        return "anal_to_filt_all / filt_to_anal_all with l_lin_vertical need an odd lm_a <= 2 lm - 1";
    if (!c.l_lin_vertical && c.lm < 3)
        return "anal_to_filt_all / filt_to_anal_all with the lwq vertical map need lm >= 3 (its edge "
               "branches touch three levels)";
    return nullptr;
}

int derive(const mgbf_config& c, mgbf_dims& d, int* nx, int* ny, std::string* err) {
    d = mgbf_dims{};
    if (c.nxPE < 1 || c.nyPE < 1 || c.im_filt < 1 || c.jm_filt < 1 || c.lm < 1 || c.n_ens < 1 ||
        c.km3 < 0 || c.km2 < 0 || c.gm_max < 1 || c.p < 0 || c.hx < 0 || c.hy < 0 || c.hz < 0) {
        *err = "invalid namelist value (sizes must be positive)";
        return MGBF_ERR_CONFIG;
    }
    if (c.mgbf_proc < 1 || c.mgbf_proc > 8) {
        *err = fmt("mgbf_proc=%d is not built (built: 1 filtering_rad2_bckg, 2 filtering_lin2_bckg, "
                   "3 filtering_fast_bckg, 4 filtering_rad3, 5 filtering_lin3, 6 filtering_rad2_ens, "
                   "7 filtering_lin2_ens, 8 filtering_fast_ens)",
                   c.mgbf_proc);
        return MGBF_ERR_UNSUPPORTED;
    }
    // l_loc and l_loc_vertical change no result: l_loc only builds the MPI
    // neighbour tables of init_domain_loc (mg_entrymod.f90:94-96) and the
    // w*_loc weight arrays (mg_intstate.f90:145-152, 325-330), and the only
    // readers of those, filtering_rad2_z_loc_g3/g4 (mg_filtering.f90:1510-1830),
    // are private and never called by filtering_procedure (:100-117).
    // km, km_all (mg_parameter.f90:293-297)
    d.km = c.km2 + c.lm * c.km3;
    const bool bckg = c.mgbf_proc >= 1 && c.mgbf_proc <= 5;
    if (bckg) {
        if (c.n_ens != 1) {
            *err = fmt("mgbf_proc=%d (background) needs n_ens=1", c.mgbf_proc);
            return MGBF_ERR_CONFIG;
        }
        d.km_all = d.km;
        d.km3_all = c.km3;
    } else {
        d.km_all = d.km * c.n_ens;
        d.km3_all = c.km3 * c.n_ens;
        if (c.km2 != 0 && c.l_vertical_filter) {
            // sup_vrbeta1_ens strides the stack by lm only (mg_filtering.f90:2066-2069)
            *err = "ensemble procedures (6, 7, 8) with km2>0 and l_vertical_filter: the reference's "
                   "ensemble vertical filter assumes km2=0";
            return MGBF_ERR_CONFIG;
        }
    }
    if (d.km_all < 1) {
        *err = "km_all must be positive";
        return MGBF_ERR_CONFIG;
    }
    // analysis grid divisibility (mg_parameter.f90:447-455)
    const int nm = c.nm0 / c.nxPE, mm = c.mm0 / c.nyPE;
    if (nm * c.nxPE != c.nm0) {
        *err = "nm0 is not divisible by nxm";
        return MGBF_ERR_CONFIG;
    }
    if (mm * c.nyPE != c.mm0) {
        *err = "mm0 is not divisible by nym";
        return MGBF_ERR_CONFIG;
    }
    // analysis <-> filter transfer (mg_transfer.f90:58-203, 339-464,
    // lsqr_mg_coef mg_interpolate.f90:89-298): every horizontal scheme fits
    // its halo when the filter grid is no finer than the analysis grid
    if (nm < c.im_filt || mm < c.jm_filt) {
        *err = "filter grid finer than the analysis grid (im_filt > nm0/nxPE or jm_filt > mm0/nyPE): "
               "the transfer assumes im <= nm and jm <= mm";
        return MGBF_ERR_UNSUPPORTED;
    }
    if (c.lm_a < c.lm) {
        *err = fmt("lm_a=%d < lm=%d: the analysis columns must hold at least the filter levels", c.lm_a, c.lm);
        return MGBF_ERR_CONFIG;
    }
    // the driver calls the _all2 routines with l_new_map and the _all ones
    // without (RBETA_TEST.f90:48-52, 92-96)
    if (const char* why = c.l_new_map ? transfer_all2_refusal(c) : transfer_all_refusal(c)) {
        *err = why;
        return MGBF_ERR_CONFIG;
    }
    // generations (mg_parameter.f90:322-328)
    int gm = def_maxgen(c.nxPE, c.nyPE);
    if (gm > c.gm_max) gm = c.gm_max;
    if (gm > kMaxGen) {
        *err = fmt("gm=%d exceeds the built maximum %d", gm, kMaxGen);
        return MGBF_ERR_CONFIG;
    }
    d.gm = gm;
    d.nx = c.nxPE * c.im_filt;
    d.ny = c.nyPE * c.jm_filt;
    long long fs = 0, ms = 0;
    const int hmax = c.hx > c.hy ? c.hx : c.hy;
    for (int g = 0; g < gm; ++g) {
        const int n = d.nx >> g, m = d.ny >> g;
        if (g + 1 < gm && ((n & 1) || (m & 1))) {
            *err = fmt("generation %d grid %dx%d is not even: cannot restrict", g + 1, n, m);
            return MGBF_ERR_CONFIG;
        }
        if (n < hmax || m < hmax || n < 2 || m < 2) {
            *err = fmt("generation %d grid %dx%d is smaller than the halo", g + 1, n, m);
            return MGBF_ERR_CONFIG;
        }
        if ((long long)d.km_all * (n + 16) * 4 >= (1ll << 31)) {
            // the kernels address within one grid row with 32-bit byte offsets
            *err = fmt("generation %d row of %d x %d values exceeds 2 GiB", g + 1, n, d.km_all);
            return MGBF_ERR_CONFIG;
        }
        nx[g] = n;
        ny[g] = m;
        const long long sz = (long long)d.km_all * n * m;
        if (g == 0) fs = sz;
        ms += sz;
    }
    d.field_size = fs;
    d.mg_size = ms;
    d.na = c.nm0;
    d.ma = c.mm0;
    d.km_a_all = (bckg ? 1 : c.n_ens) * (c.km2 + c.km3 * c.lm_a);
    d.anal_size = (long long)d.km_a_all * d.na * d.ma;
    if ((c.l_vertical_filter || ((c.mgbf_proc == 4 || c.mgbf_proc == 5) && c.km3 > 0)) && c.lm <= c.hz) {
        *err = "lm must exceed hz for the vertical mirror";
        return MGBF_ERR_CONFIG;
    }
    return MGBF_OK;
}

int build_setup(const mgbf_config& cfg, Setup& s, std::string* err) {
    s = Setup{};
    s.cfg = cfg;
    int st = derive(cfg, s.dims, s.nx, s.ny, err);
    if (st != MGBF_OK) return st;
    // def_mg_weights normalizes every filter of the namelist whatever the
    // procedure (mg_intstate.f90:422-431), and getlinesum stops on the first
    // reach beyond its halo (jp_pbfil.f90:149-262): paspx and paspy (the
    // fast factors, mg_ampl02 as given) along x and y, pasp1 along z, pasp2
    // in the plane, pasp3 in the volume
    {
        std::vector<int> g;
        std::vector<double> w;
        std::vector<Tap2> t2;
        if (!beta1_taps(cfg.mg_ampl02, cfg.hx, cfg.p, g, w) ||
            !beta1_taps(cfg.mg_ampl02, cfg.hy, cfg.p, g, w) || !vertical_taps(cfg, g, w)) {
            *err = "In getlinesum1; filter reach fx becomes too large for hx";
            return MGBF_ERR_CONFIG;
        }
        if (!radial_taps(cfg, t2)) {
            *err = "In getlinesum2; filter reach becomes too large for hx/hy";
            return MGBF_ERR_CONFIG;
        }
        if (!ball_reach_ok(cfg, err)) return MGBF_ERR_CONFIG;
    }
    const bool fast = cfg.mgbf_proc == 3 || cfg.mgbf_proc == 8;
    const bool line = cfg.mgbf_proc == 2 || cfg.mgbf_proc == 5 || cfg.mgbf_proc == 7;
    if (line) {
        // filtering_lin2_*: the triad of dibeta line filters
        st = line_taps(cfg, s.taps2, err);
        if (st != MGBF_OK) return st;
    } else if (fast) {
        // filtering_fast_*: the horizontal filter is the product of 1-D beta
        // filters along x and y (rbeta over each row and column)
        if (!fast_taps(cfg, s.taps2)) {
            *err = "In getlinesum1; filter reach fx becomes too large for hx";
            return MGBF_ERR_CONFIG;
        }
    } else if (!radial_taps(cfg, s.taps2)) {
        *err = "In getlinesum2; filter reach becomes too large for hx/hy";
        return MGBF_ERR_CONFIG;
    }
    if (!vertical_taps(cfg, s.taps1_gx, s.taps1_w)) {
        *err = "In getlinesum1; filter reach fx becomes too large for hz";
        return MGBF_ERR_CONFIG;
    }
    if (cfg.mgbf_proc == 4) ball_taps(cfg, s.taps3);   // filtering_rad3 (reach checked above)
    if (cfg.mgbf_proc == 5) {                           // filtering_lin3
        st = hex_taps(cfg, s.taps3, &s.zero_levels, err, s.hex_f);
        if (st != MGBF_OK) return st;
    }
    s.dims.ntaps2 = (int)s.taps2.size();
    s.dims.ntaps3 = (int)s.taps3.size();
    s.dims.ntaps1 = (int)s.taps1_gx.size();
    s.H = cfg.hx > cfg.hy ? cfg.hx : cfg.hy;
    if (fast || line) {
        // the kernels need only the taps' extent, which a raw factor can keep
        // far inside the halo
        int ext = 1;
        for (const Tap2& t : s.taps2) ext = std::max(ext, std::max(std::abs(t.gx), std::abs(t.gy)));
        s.H = ext;
    }
    if (s.H > kMaxHalo || s.H < 1) {
        *err = fmt("radial half-width %d outside the built range 1..%d", s.H, kMaxHalo);
        return MGBF_ERR_UNSUPPORTED;
    }
    // fold the taps into w(|a|,|b|); the filter must be even in a and in b
    bool seen[kMaxHalo + 1][kMaxHalo + 1] = {};
    for (const Tap2& t : s.taps2) {
        const int a = t.gx < 0 ? -t.gx : t.gx, b = t.gy < 0 ? -t.gy : t.gy;
        if (seen[b][a]) {
            if (std::fabs(s.wsym[b][a] - t.w) > 1e-15 * std::fabs(t.w)) {
                *err = "radial taps are not mirror symmetric";
                return MGBF_ERR_UNSUPPORTED;
            }
        }
        seen[b][a] = true;
        s.wsym[b][a] = t.w;
    }
    // exact lattice disc a^2 + b^2 < (H+1)^2 with weights even under a <-> b
    // (an isotropic aspect tensor)? The disc kernels then skip the lattice
    // points outside the disc and read w(a,b) and w(b,a) from one register.
    {
        const int R2 = (s.H + 1) * (s.H + 1);
        bool exact = cfg.hx == cfg.hy;
        for (int b = 0; b <= s.H && exact; ++b)
            for (int a = 0; a <= s.H; ++a) {
                if (seen[b][a] != (a * a + b * b < R2)) exact = false;
                if (std::fabs(s.wsym[b][a] - s.wsym[a][b]) > 1e-15 * std::fabs(s.wsym[b][a]))
                    exact = false;
            }
        s.exact_disc = exact;
    }
    // a product w(a,b) = fx(a) fy(b) (the fast and line procedures' taps)?
    // The separable kernels then sum the x-pairs once per row and scale by
    // fy(b) per output row.
    if (!s.exact_disc && s.wsym[0][0] != 0.0) {
        bool sep = true;
        for (int a = 0; a <= s.H; ++a) s.sepx[a] = s.wsym[0][a];
        for (int b = 0; b <= s.H; ++b) s.sepy[b] = s.wsym[b][0] / s.wsym[0][0];
        for (int b = 0; b <= s.H && sep; ++b)
            for (int a = 0; a <= s.H; ++a) {
                const double prod = s.sepx[a] * s.sepy[b];
                if (std::fabs(s.wsym[b][a] - prod) > 1e-12 * std::fabs(s.wsym[0][0])) sep = false;
            }
        s.separable = sep;
    }
    // vertical band matrices (sup_vrbeta1_ens / sup_vrbeta1T_ens)
    {
        const int lm = cfg.lm, hz = cfg.hz, W = 2 * hz + 1;
        std::vector<double> V((size_t)lm * lm, 0.0);
        for (int L = 0; L < lm; ++L)
            for (size_t t = 0; t < s.taps1_gx.size(); ++t) {
                const int src = mirror_whole(L + s.taps1_gx[t], lm);
                V[(size_t)L * lm + src] += s.taps1_w[t];
            }
        s.vf.assign((size_t)lm * W, 0.0);
        s.vft.assign((size_t)lm * W, 0.0);
        // Vf Vf^T (band 2hz): the vertical part of W A W^T when the weights
        // are the same along the stack, where it commutes with the
        // horizontal multigrid operator. filtering_lin2_ens ends with
        // sup_vrbeta1T_ens where the others end with Vf
        // (mg_filtering.f90:1303-1305): its product is Vf^T Vf^T.
        {
            const bool tt = cfg.mgbf_proc == 7;
            const int W2 = 4 * hz + 1;
            s.vvt.assign((size_t)lm * W2, 0.0);
            for (int L = 0; L < lm; ++L)
                for (int c = L - 2 * hz; c <= L + 2 * hz; ++c) {
                    if (c < 0 || c >= lm) continue;
                    double acc = 0;
                    for (int q = 0; q < lm; ++q)
                        acc += (tt ? V[(size_t)q * lm + L] : V[(size_t)L * lm + q]) * V[(size_t)c * lm + q];
                    s.vvt[(size_t)L * W2 + (c - L + 2 * hz)] = acc;
                }
        }
        for (int L = 0; L < lm; ++L)
            for (int t = -hz; t <= hz; ++t) {
                const int c = L + t;
                if (c < 0 || c >= lm) continue;
                s.vf[(size_t)L * W + (t + hz)] = V[(size_t)L * lm + c];
                s.vft[(size_t)L * W + (t + hz)] = V[(size_t)c * lm + L];
            }
    }
    // transfer tables for each generation pair g -> g+1
    for (int g = 0; g + 1 < s.dims.gm; ++g) {
        std::vector<double> P;
        std::vector<char> nz;
        for (int axis = 0; axis < 2; ++axis) {
            const int nf = axis == 0 ? s.nx[g] : s.ny[g];
            interp_dense(nf, cfg.lquart != 0, P, nz);
            AxisTable& pt = axis == 0 ? s.px[g] : s.py[g];
            AxisTable& rt = axis == 0 ? s.rx[g] : s.ry[g];
            if (!pack_rows(nf, nf / 2, P, nz, false, kPWidth, pt) ||
                !pack_rows(nf / 2, nf, P, nz, true, kRWidth, rt)) {
                *err = "transfer stencil wider than its table";
                return MGBF_ERR_CONFIG;
            }
        }
    }
    // generation weights (mg_intstate.f90:333-351): g >= 5 share mg_weig5
    const double wv[5] = {cfg.mg_weig1, cfg.mg_weig2, cfg.mg_weig3, cfg.mg_weig4, cfg.mg_weig5};
    for (int g = 0; g < s.dims.gm; ++g) s.wconst[g] = wv[g < 4 ? g : 4];
    // analysis <-> filter transfer tables
    const int nm = cfg.nm0 / cfg.nxPE, mm = cfg.mm0 / cfg.nyPE;
    s.anal_identity = nm == cfg.im_filt && mm == cfg.jm_filt && cfg.lm_a == cfg.lm;
    // l_lin_horizontal wins over l_quad_horizontal (mg_transfer.f90:368-382)
    const int scheme = cfg.l_lin_horizontal ? 1 : cfg.l_quad_horizontal ? 2 : 0;
    if (!s.anal_identity) {
        const int axes[2][3] = {{cfg.nxPE, nm, cfg.im_filt}, {cfg.nyPE, mm, cfg.jm_filt}};
        AxisTable* dirs[2] = {&s.ax_d, &s.ay_d};
        AxisTable* adjs[2] = {&s.ax_a, &s.ay_a};
        for (int a = 0; a < 2; ++a) {
            const int na = axes[a][0] * axes[a][1], nf = axes[a][0] * axes[a][2];
            std::vector<double> T;
            if (!anal_axis_matrix(axes[a][0], axes[a][1], axes[a][2], scheme, T)) {
                *err = "analysis-filter interpolation: stencil outside the mirrored domain";
                return MGBF_ERR_CONFIG;
            }
            std::vector<char> nz(T.size());
            for (size_t e = 0; e < T.size(); ++e) nz[e] = T[e] != 0.0;
            const int wd = max_span(na, nf, nz, false), wa = max_span(nf, na, nz, true);
            if (wd > kMaxXferWidth || wa > kMaxXferWidth || !pack_rows(na, nf, T, nz, false, wd, *dirs[a]) ||
                !pack_rows(nf, na, T, nz, true, wa, *adjs[a])) {
                *err = fmt("analysis-filter interpolation wider than %d points", kMaxXferWidth);
                return MGBF_ERR_UNSUPPORTED;
            }
        }
    }
    s.all2_refusal = transfer_all2_refusal(cfg);
    s.all_refusal = transfer_all_refusal(cfg);
    if (cfg.lm_a != cfg.lm) {
        bool ok = true;
        if (!s.all2_refusal) ok = vertical_transfer_tables(s.dims.km3_all, cfg.lm, cfg.lm_a, 0, s.kv_d, s.kv_a);
        if (ok && !s.all_refusal)
            ok = vertical_transfer_tables(s.dims.km3_all, cfg.lm, cfg.lm_a, cfg.l_lin_vertical ? 0 : 1, s.kw_d,
                                          s.kw_a);
        if (!ok) {
            *err = "vertical transfer tables";
            return MGBF_ERR_CONFIG;
        }
    }
    return MGBF_OK;
}

}  // namespace mgbf
