/*
 * FP64 CPU restatement of the MGBF localization operator: the oracle.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests, the smoke check and the CPU-baseline
 * leg of the benchmark load this code, as the checker; the product library
 * never links or calls it.
 *
 * Every routine restates one reference routine (paths relative to the
 * reference root, all in src.mgbf/), in the reference's own formulation on a
 * single global domain:
 *   - cholaspect1/2, getlinesum1/2 ............ jp_pbfil.f90:74-107, 149-218
 *   - dinvmtf / dl1lmf (2x2 inverse, Cholesky) . jp_pmat.f90:147-183, 867-897
 *   - vrbeta1 / vrbeta1T (1-D gather/scatter) .. jp_pbfil.f90:856-885, 981-1009
 *   - vrbeta2 / vrbeta2T (2-D gather/scatter) .. jp_pbfil.f90:888-928, 1011-1050
 *     with a per-point aspect field (el, ss per point; pasp2 of
 *     def_mg_weights, mg_intstate.f90:383-390) ... mgbfo_radial_aspect
 *   - boco_2d (halo fill, half-sample mirror) .. mg_bocos.f90:225-262, 340-377
 *   - bocoT_2d (adjoint fold) .................. mg_bocos.f90:929-957, 1033-1061
 *   - adjoint / direct1 (R^T-side, P) .......... mg_generations.f90:1117-1187, 1190-1250
 *   - def_offset_coef constants ................ mg_interpolate.f90:61-86
 *   - adjoint2 / direct2 (lquart) .............. mg_generations.f90:1253-1374
 *   - inimomtab, hstform, blinfil .............. jp_pbfil3.f90:378-457, 1399-1430, 1463-1501
 *   - triad / tritform ......................... jp_pbfil3.f90:496-532, 573-624
 *   - vdibeta2 / vdibeta2t ..................... jp_pbfil3.f90:1801-1837, 2345-2382
 *   - init_mg_line ............................. mg_intstate.f90:436-480
 *   - filtering_lin2_ens / _bckg composition ... mg_filtering.f90:1150-1313, 683-822
 *   - cholaspect3, getlinesum3 ................. jp_pbfil.f90:108-128, 220-262
 *   - vrbeta3 / vrbeta3T ....................... jp_pbfil.f90:930-977, 1053-1100
 *   - sup_vrbeta3 / sup_vrbeta3T ............... mg_filtering.f90:1923-2039
 *   - filtering_rad3 composition ............... mg_filtering.f90:123-257
 *   - hexad / hextform ......................... jp_pbfil3.f90:702-750, 798-904
 *   - vdibetax3 / vdibetax3t ................... jp_pbfil3.f90:1928-1977, 2475-2525
 *   - filtering_lin3 composition ............... mg_filtering.f90:260-557
 *   - upsending2_ens / downsending2_ens ........ mg_generations.f90:497-609
 *   - upsending_ens / downsending_ens .......... mg_generations.f90:382-494
 *   - weighting_ens ............................ mg_generations.f90:1012-1052
 *   - sup_vrbeta1_ens / sup_vrbeta1T_ens ....... mg_filtering.f90:2042-2151
 *   - filtering_rad2_ens composition ........... mg_filtering.f90:1000-1147
 * On one domain the reference's MPI halo exchanges and up/down re-tiling
 * (mg_bocos.f90 upsend_all/downsend_all) reduce to the global mirror
 * boundary at the domain edges and to plain addressing.
 */
#include "mgbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS 1.e-12

/* interpolation constants, mg_interpolate.f90:76-79 */
static const double PCOEF[4] = {-9. / 128., 111. / 128., 29. / 128., -3. / 128.};
static const double QCOEF[4] = {-3. / 128., 29. / 128., 111. / 128., -9. / 128.};
/* quadratic constants a_coef, b_coef, mg_interpolate.f90:81-84 */
static const double ACOEF[3] = {5. / 32., 30. / 32., -3. / 32.};
static const double BCOEF[3] = {-3. / 32., 30. / 32., 5. / 32.};

int mgbfo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double ipow(double x, int p) {
    double r = 1.0;
    for (int i = 0; i < p; ++i) r *= x;
    return r;
}

/* ------------------------------------------------------------------------
 * Padded field: i in [1-hx, n+hx], j in [1-hy, m+hy] (1-based like the
 * reference), k fastest.
 * ---------------------------------------------------------------------- */
typedef struct {
    double* d;
    int K, n, m, hx, hy;
} pf_t;

static inline size_t pf_off(const pf_t* f, int k, int i, int j) {
    return (size_t)k + (size_t)f->K * ((size_t)(i + f->hx - 1) +
                                       (size_t)(f->n + 2 * f->hx) * (size_t)(j + f->hy - 1));
}
#define PF(f, k, i, j) ((f)->d[pf_off((f), (k), (i), (j))])
#define PCOL(f, i, j) (&(f)->d[pf_off((f), 0, (i), (j))])

static pf_t pf_alloc(int K, int n, int m, int hx, int hy) {
    pf_t f;
    f.K = K; f.n = n; f.m = m; f.hx = hx; f.hy = hy;
    f.d = (double*)calloc((size_t)K * (n + 2 * hx) * (size_t)(m + 2 * hy), sizeof(double));
    return f;
}
static void pf_free(pf_t* f) { free(f->d); f->d = NULL; }
static void pf_zero(pf_t* f) {
    memset(f->d, 0, sizeof(double) * (size_t)f->K * (f->n + 2 * f->hx) * (size_t)(f->m + 2 * f->hy));
}
static void pf_from_interior(pf_t* f, const double* in) {
    const int K = f->K, n = f->n, m = f->m;
    for (int j = 1; j <= m; ++j)
        for (int i = 1; i <= n; ++i)
            memcpy(PCOL(f, i, j), in + (size_t)K * ((size_t)(i - 1) + (size_t)n * (j - 1)),
                   sizeof(double) * K);
}
static void pf_to_interior(const pf_t* f, double* out) {
    const int K = f->K, n = f->n, m = f->m;
    for (int j = 1; j <= m; ++j)
        for (int i = 1; i <= n; ++i)
            memcpy(out + (size_t)K * ((size_t)(i - 1) + (size_t)n * (j - 1)), PCOL(f, i, j),
                   sizeof(double) * K);
}

/* ------------------------------------------------------------------------
 * Coefficients
 * ---------------------------------------------------------------------- */

/* 2x2 symmetric aspect tensor -> Cholesky factor of its inverse (cholaspect2:
   inv then l1lm, jp_pbfil.f90:90-107). el is column-major el[r + 2c]. */
static void cholaspect2(const double a[4], double el[4]) {
    /* inverse of a 2x2 matrix (dinvmtf reduces to the adjugate formula) */
    double det = a[0] * a[3] - a[2] * a[1];
    double inv[4];
    inv[0] = a[3] / det;  inv[3] = a[0] / det;
    inv[2] = -a[2] / det; inv[1] = -a[1] / det;
    /* lower Cholesky factor, column by column (dl1lmf, jp_pmat.f90:867-897) */
    double b00 = sqrt(inv[0]);
    double b10 = inv[1] / b00;
    double b11 = sqrt(inv[3] - b10 * b10);
    el[0] = b00; el[1] = b10; el[2] = 0.0; el[3] = b11;
}

static void aspect2(const mgbfo_params* P, double a[4]) {
    /* mg_intstate.f90:383-390 */
    a[0] = P->ampl02 * (1. + P->p_del);
    a[3] = P->ampl02 * (1. - P->p_del);
    a[2] = P->ampl02 * P->p_eps;
    a[1] = P->ampl02 * P->p_eps;
}

/* Visit the taps of the 2-D filter in the reference's loop order
   (vrbeta2, jp_pbfil.f90:909-926). Returns count, or -1 if reach > halo. */
static int taps2_raw(const mgbfo_params* P, int maxt, int* gxo, int* gyo, double* rrco) {
    const double rmom2_2 = 1.0 / sqrt((double)(2 * P->p) + 4.0);
    double a[4], el[4];
    aspect2(P, a);
    cholaspect2(a, el);
    const double exx = el[0] * rmom2_2, eyy = el[3] * rmom2_2, eyx = el[1] * rmom2_2;
    double y = 1.0 / eyy;
    int gyl = (int)ceil(-y + EPS), gym = (int)floor(y - EPS);
    if (gyl < -P->hy || gym > P->hy) return -1;
    int nt = 0;
    for (int gy = gyl; gy <= gym; ++gy) {
        double yy = gy;
        double xc = -yy * eyx;
        double rrx = (yy * eyy) * (yy * eyy);
        double x = sqrt(1.0 - rrx);
        int gxl = (int)ceil((xc - x) / exx + EPS), gxm = (int)floor((xc + x) / exx - EPS);
        if (gxl < -P->hx || gxm > P->hx) return -1;
        for (int gx = gxl; gx <= gxm; ++gx) {
            double xx = gx;
            double t = xx * exx - xc;
            double rr = rrx + t * t;
            if (nt < maxt) {
                gxo[nt] = gx;
                gyo[nt] = gy;
                rrco[nt] = 1.0 - rr;
            }
            ++nt;
        }
    }
    return nt;
}

double mgbfo_ss2(const mgbfo_params* P) {
    /* getlinesum2, jp_pbfil.f90:182-218 */
    int cap = (2 * P->hx + 1) * (2 * P->hy + 1);
    int* gx = malloc(sizeof(int) * cap);
    int* gy = malloc(sizeof(int) * cap);
    double* rrc = malloc(sizeof(double) * cap);
    int nt = taps2_raw(P, cap, gx, gy, rrc);
    double s = 0;
    for (int t = 0; t < nt; ++t) s += ipow(rrc[t], P->p);
    free(gx); free(gy); free(rrc);
    return nt > 0 ? 1.0 / s : 0.0;
}

int mgbfo_taps2(const mgbfo_params* P, int maxt, int* gx, int* gy, double* w) {
    int cap = (2 * P->hx + 1) * (2 * P->hy + 1);
    int* tgx = malloc(sizeof(int) * cap);
    int* tgy = malloc(sizeof(int) * cap);
    double* rrc = malloc(sizeof(double) * cap);
    int nt = taps2_raw(P, cap, tgx, tgy, rrc);
    if (nt > 0) {
        double ss = mgbfo_ss2(P);
        for (int t = 0; t < nt && t < maxt; ++t) {
            gx[t] = tgx[t];
            gy[t] = tgy[t];
            w[t] = ss * ipow(rrc[t], P->p);
        }
    }
    free(tgx); free(tgy); free(rrc);
    return nt;
}

static int taps1_raw(const mgbfo_params* P, int maxt, int* gxo, double* rrco) {
    /* cholaspect1 + getlinesum1 / vrbeta1 loop, jp_pbfil.f90:74-88, 149-180, 872-884 */
    const double rmom2_1 = 1.0 / sqrt((double)(2 * P->p) + 3.0);
    double el = 1.0 / sqrt(P->ampl01);
    double exx = el * rmom2_1;
    double x = 1.0 / exx;
    int gxl = (int)ceil(-x + EPS), gxm = (int)floor(x - EPS);
    if (gxl < -P->hz || gxm > P->hz) return -1;
    int nt = 0;
    for (int gx = gxl; gx <= gxm; ++gx) {
        double xx = gx;
        double rr = (xx * exx) * (xx * exx);
        if (nt < maxt) {
            gxo[nt] = gx;
            rrco[nt] = 1.0 - rr;
        }
        ++nt;
    }
    return nt;
}

double mgbfo_ss1(const mgbfo_params* P) {
    int cap = 2 * P->hz + 1;
    int* gx = malloc(sizeof(int) * cap);
    double* rrc = malloc(sizeof(double) * cap);
    int nt = taps1_raw(P, cap, gx, rrc);
    double s = 0;
    for (int t = 0; t < nt; ++t) s += ipow(rrc[t], P->p);
    free(gx); free(rrc);
    return nt > 0 ? 1.0 / s : 0.0;
}

int mgbfo_taps1(const mgbfo_params* P, int maxt, int* gx, double* w) {
    int cap = 2 * P->hz + 1;
    int* tgx = malloc(sizeof(int) * cap);
    double* rrc = malloc(sizeof(double) * cap);
    int nt = taps1_raw(P, cap, tgx, rrc);
    if (nt > 0) {
        double ss = mgbfo_ss1(P);
        for (int t = 0; t < nt && t < maxt; ++t) {
            gx[t] = tgx[t];
            w[t] = ss * ipow(rrc[t], P->p);
        }
    }
    free(tgx); free(rrc);
    return nt;
}

/* 1-D taps of the fast filter: getlinesum1 / vrbeta1 with el = ampl02 as
   given (paspx, paspy are not passed through cholaspect,
   mg_intstate.f90:370-375, 422-431). Returns count, or -1 past the halo. */
static int taps_fast_raw(const mgbfo_params* P, int halo, int maxt, int* gxo, double* wo) {
    const double rmom2_1 = 1.0 / sqrt((double)(2 * P->p) + 3.0);
    const double exx = P->ampl02 * rmom2_1;
    const double x = 1.0 / exx;
    int gxl = (int)ceil(-x + EPS), gxm = (int)floor(x - EPS);
    if (gxl < -halo || gxm > halo) return -1;
    int nt = 0;
    double s = 0.0;
    for (int gx = gxl; gx <= gxm; ++gx, ++nt) {
        double rr = ((double)gx * exx) * ((double)gx * exx);
        double v = ipow(1.0 - rr, P->p);
        s += v;
        if (nt < maxt) { gxo[nt] = gx; wo[nt] = v; }
    }
    for (int t = 0; t < nt && t < maxt; ++t) wo[t] /= s;
    return nt;
}

int mgbfo_taps_fast(const mgbfo_params* P, int maxt, int* gx, double* w) {
    return taps_fast_raw(P, P->hx, maxt, gx, w);
}

/* ------------------------------------------------------------------------
 * Halo fill and fold (boco_2d / bocoT_2d on the whole domain)
 * ---------------------------------------------------------------------- */

/* boco_2d: rows (N/S) over interior i first, then columns (E/W) over the
   extended rows (mg_bocos.f90:225-262, 340-377). */
static void boco_2d(pf_t* W, int nbx, int nby) {
    const int K = W->K, n = W->n, m = W->m;
    for (int j = 1; j <= nby; ++j)
        for (int i = 1; i <= n; ++i)
            memcpy(PCOL(W, i, -nby + j), PCOL(W, i, nby + 1 - j), sizeof(double) * K);
    for (int j = 1; j <= nby; ++j)
        for (int i = 1; i <= n; ++i)
            memcpy(PCOL(W, i, m + j), PCOL(W, i, m + 1 - j), sizeof(double) * K);
    for (int j = 1 - nby; j <= m + nby; ++j)
        for (int i = 1; i <= nbx; ++i)
            memcpy(PCOL(W, -nbx + i, j), PCOL(W, nbx + 1 - i, j), sizeof(double) * K);
    for (int j = 1 - nby; j <= m + nby; ++j)
        for (int i = 1; i <= nbx; ++i)
            memcpy(PCOL(W, n + i, j), PCOL(W, n + 1 - i, j), sizeof(double) * K);
}

static inline void vadd(double* dst, const double* src, int K) {
    for (int k = 0; k < K; ++k) dst[k] += src[k];
}

/* bocoT_2d: E/W fold over the extended rows first, then N/S over interior i
   (mg_bocos.f90:929-957, 1033-1061). */
static void bocoT_2d(pf_t* W, int nbx, int nby) {
    const int K = W->K, n = W->n, m = W->m;
    for (int j = 1 - nby; j <= m + nby; ++j)
        for (int i = 1; i <= nbx; ++i) vadd(PCOL(W, i, j), PCOL(W, 1 - i, j), K);
    for (int j = 1 - nby; j <= m + nby; ++j)
        for (int i = 1; i <= nbx; ++i) vadd(PCOL(W, n - nbx + i, j), PCOL(W, n + 1 + nbx - i, j), K);
    for (int j = 1; j <= nby; ++j)
        for (int i = 1; i <= n; ++i) vadd(PCOL(W, i, j), PCOL(W, i, 1 - j), K);
    for (int j = 1; j <= nby; ++j)
        for (int i = 1; i <= n; ++i) vadd(PCOL(W, i, m - nby + j), PCOL(W, i, m + 1 + nby - j), K);
}

/* ------------------------------------------------------------------------
 * Radial beta filters
 * ---------------------------------------------------------------------- */

typedef struct {
    int nt;
    int* gx;
    int* gy;
    double* w;
} taps_t;

static taps_t taps2_make(const mgbfo_params* P) {
    taps_t T;
    int cap = (2 * P->hx + 1) * (2 * P->hy + 1);
    T.gx = malloc(sizeof(int) * cap);
    T.gy = malloc(sizeof(int) * cap);
    T.w = malloc(sizeof(double) * cap);
    T.nt = mgbfo_taps2(P, cap, T.gx, T.gy, T.w);
    if (T.nt < 0) abort(); /* the reference stops: "filter reach becomes too large" */
    return T;
}
static taps_t taps1_make(const mgbfo_params* P) {
    taps_t T;
    int cap = 2 * P->hz + 1;
    T.gx = malloc(sizeof(int) * cap);
    T.gy = NULL;
    T.w = malloc(sizeof(double) * cap);
    T.nt = mgbfo_taps1(P, cap, T.gx, T.w);
    if (T.nt < 0) abort();
    return T;
}
static void taps_free(taps_t* T) { free(T->gx); free(T->gy); free(T->w); }

/* vrbeta2: gather over the disc from the padded input; interior output, halo
   of the result is zero (b=0; a=b). jp_pbfil.f90:888-928 */
static void vrbeta2(const taps_t* T, pf_t* a) {
    const int K = a->K, n = a->n, m = a->m;
    pf_t b = pf_alloc(K, n, m, a->hx, a->hy);
#pragma omp parallel for schedule(static)
    for (int iy = 1; iy <= m; ++iy) {
        double* tb = malloc(sizeof(double) * K);
        for (int ix = 1; ix <= n; ++ix) {
            memset(tb, 0, sizeof(double) * K);
            for (int t = 0; t < T->nt; ++t) {
                const double frow = T->w[t];
                const double* src = PCOL(a, ix + T->gx[t], iy + T->gy[t]);
                for (int k = 0; k < K; ++k) tb[k] += frow * src[k];
            }
            memcpy(PCOL(&b, ix, iy), tb, sizeof(double) * K);
        }
        free(tb);
    }
    free(a->d);
    a->d = b.d;
}

/* vrbeta2T: scatter from each interior point into the padded result (halo
   included). jp_pbfil.f90:1011-1050. Parallel over k: each k is independent. */
static void vrbeta2T(const taps_t* T, pf_t* a) {
    const int K = a->K, n = a->n, m = a->m;
    pf_t b = pf_alloc(K, n, m, a->hx, a->hy);
#pragma omp parallel
    {
        int nth = 1, tid = 0;
#ifdef _OPENMP
        nth = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        int k0 = (int)((long)K * tid / nth), k1 = (int)((long)K * (tid + 1) / nth);
        for (int iy = 1; iy <= m; ++iy)
            for (int ix = 1; ix <= n; ++ix) {
                const double* ta = PCOL(a, ix, iy);
                for (int t = 0; t < T->nt; ++t) {
                    const double frow = T->w[t];
                    double* dst = PCOL(&b, ix + T->gx[t], iy + T->gy[t]);
                    for (int k = k0; k < k1; ++k) dst[k] += frow * ta[k];
                }
            }
    }
    free(a->d);
    a->d = b.d;
}

/* ------------------------------------------------------------------------
 * Inter-generation transfer: adjoint (P^T) and direct1 (P)
 * Coarse arrays have a 2-point halo: indices -1 .. n/2+2.
 * ---------------------------------------------------------------------- */

/* adjoint: F(K,1:n,1:m) -> W(K,-1:nL+2,-1:mL+2), mg_generations.f90:1117-1187.
   Scatter; parallel over k ranges. */
static void adjoint_op(const pf_t* F, pf_t* W) {
    const int K = F->K, n = F->n, m = F->m;
    const int nL = n / 2, mL = m / 2;
    /* W_AUX(K, 1:n, -1:mL+2) */
    pf_t A = pf_alloc(K, n, mL, 0, 2);
    pf_zero(W);
#pragma omp parallel
    {
        int nth = 1, tid = 0;
#ifdef _OPENMP
        nth = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        int k0 = (int)((long)K * tid / nth), k1 = (int)((long)K * (tid + 1) / nth);
        for (int j = m; j >= 2; j -= 2) {
            int jL = j / 2;
            for (int i = n; i >= 1; --i) {
                const double* f = PCOL(F, i, j);
                for (int t = 3; t >= 0; --t) {
                    double* dst = PCOL(&A, i, jL - 1 + t);
                    for (int k = k0; k < k1; ++k) dst[k] += PCOEF[t] * f[k];
                }
            }
        }
        for (int j = m - 1; j >= 1; j -= 2) {
            int jL = j / 2;
            for (int i = n; i >= 1; --i) {
                const double* f = PCOL(F, i, j);
                for (int t = 3; t >= 0; --t) {
                    double* dst = PCOL(&A, i, jL - 1 + t);
                    for (int k = k0; k < k1; ++k) dst[k] += QCOEF[t] * f[k];
                }
            }
        }
        for (int jL = mL + 2; jL >= -1; --jL) {
            for (int i = n - 1; i >= 1; i -= 2) {
                int iL = i / 2;
                const double* a = PCOL(&A, i, jL);
                for (int t = 3; t >= 0; --t) {
                    double* dst = PCOL(W, iL - 1 + t, jL);
                    for (int k = k0; k < k1; ++k) dst[k] += QCOEF[t] * a[k];
                }
            }
            for (int i = n; i >= 2; i -= 2) {
                int iL = i / 2;
                const double* a = PCOL(&A, i, jL);
                for (int t = 3; t >= 0; --t) {
                    double* dst = PCOL(W, iL - 1 + t, jL);
                    for (int k = k0; k < k1; ++k) dst[k] += PCOEF[t] * a[k];
                }
            }
        }
    }
    (void)nL;
    pf_free(&A);
}

/* direct1: W(K,-1:nL+2,-1:mL+2) -> F(K,1:n,1:m), mg_generations.f90:1190-1250.
   Gather; F here is an interior-only output written into a padded field. */
static void direct1_op(const pf_t* W, pf_t* F) {
    const int K = F->K, n = F->n, m = F->m;
    const int mL = m / 2;
    pf_t A = pf_alloc(K, n, mL, 0, 2); /* W_AUX(K,1:n,-1:mL+2) */
#pragma omp parallel for schedule(static)
    for (int jL = -1; jL <= mL + 2; ++jL) {
        for (int i = 1; i <= n; ++i) {
            int iL = i / 2;
            const double* c = (i & 1) ? QCOEF : PCOEF;
            double* dst = PCOL(&A, i, jL);
            const double* w0 = PCOL(W, iL - 1, jL);
            const double* w1 = PCOL(W, iL, jL);
            const double* w2 = PCOL(W, iL + 1, jL);
            const double* w3 = PCOL(W, iL + 2, jL);
            for (int k = 0; k < K; ++k) dst[k] = c[0] * w0[k] + c[1] * w1[k] + c[2] * w2[k] + c[3] * w3[k];
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= m; ++j) {
        int jL = j / 2;
        const double* c = (j & 1) ? QCOEF : PCOEF;
        for (int i = 1; i <= n; ++i) {
            double* dst = PCOL(F, i, j);
            const double* a0 = PCOL(&A, i, jL - 1);
            const double* a1 = PCOL(&A, i, jL);
            const double* a2 = PCOL(&A, i, jL + 1);
            const double* a3 = PCOL(&A, i, jL + 2);
            for (int k = 0; k < K; ++k) dst[k] = c[0] * a0[k] + c[1] * a1[k] + c[2] * a2[k] + c[3] * a3[k];
        }
    }
    pf_free(&A);
}

/* ------------------------------------------------------------------------
 * Quadratic inter-generation transfer (lquart): adjoint2 and direct2.
 * Coarse arrays have a 1-point halo: indices 0 .. n/2+1. Fine index i
 * (1-based) reads coarse iL-1 .. iL+1 with iL = (i+1)/2 and a_coef for odd
 * i, iL = i/2 and b_coef for even i; y as x.
 * ---------------------------------------------------------------------- */

/* coarse base index and constants of fine index i (1-based) */
static inline int quad_base(int i, const double** c) {
    *c = (i & 1) ? ACOEF : BCOEF;
    return ((i & 1) ? (i + 1) / 2 : i / 2) - 1;
}

/* adjoint2: F(K,1:n,1:m) -> W(K,0:nL+1,0:mL+1), mg_generations.f90:1253-1315.
   Scatter: first along y into W_AUX(K,1:n,0:mL+1), then along x. */
static void adjoint2_op(const pf_t* F, pf_t* W) {
    const int K = F->K, n = F->n, m = F->m;
    const int mL = m / 2;
    pf_t A = pf_alloc(K, n, mL, 0, 1);
    pf_zero(W);
#pragma omp parallel
    {
        int nth = 1, tid = 0;
#ifdef _OPENMP
        nth = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        int k0 = (int)((long)K * tid / nth), k1 = (int)((long)K * (tid + 1) / nth);
        for (int j = m; j >= 1; --j) {
            const double* c;
            const int b = quad_base(j, &c);
            for (int i = n; i >= 1; --i) {
                const double* f = PCOL(F, i, j);
                for (int t = 2; t >= 0; --t) {
                    double* dst = PCOL(&A, i, b + t);
                    for (int k = k0; k < k1; ++k) dst[k] += c[t] * f[k];
                }
            }
        }
        for (int jL = mL + 1; jL >= 0; --jL)
            for (int i = n; i >= 1; --i) {
                const double* c;
                const int b = quad_base(i, &c);
                const double* a = PCOL(&A, i, jL);
                for (int t = 2; t >= 0; --t) {
                    double* dst = PCOL(W, b + t, jL);
                    for (int k = k0; k < k1; ++k) dst[k] += c[t] * a[k];
                }
            }
    }
    pf_free(&A);
}

/* direct2: W(K,0:nL+1,0:mL+1) -> F(K,1:n,1:m), mg_generations.f90:1318-1374.
   Gather: along x into W_AUX(K,1:n,0:mL+1), then along y. */
static void direct2_op(const pf_t* W, pf_t* F) {
    const int K = F->K, n = F->n, m = F->m;
    const int mL = m / 2;
    pf_t A = pf_alloc(K, n, mL, 0, 1);
#pragma omp parallel for schedule(static)
    for (int jL = 0; jL <= mL + 1; ++jL)
        for (int i = 1; i <= n; ++i) {
            const double* c;
            const int b = quad_base(i, &c);
            double* dst = PCOL(&A, i, jL);
            const double* w0 = PCOL(W, b, jL);
            const double* w1 = PCOL(W, b + 1, jL);
            const double* w2 = PCOL(W, b + 2, jL);
            for (int k = 0; k < K; ++k) dst[k] = c[0] * w0[k] + c[1] * w1[k] + c[2] * w2[k];
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= m; ++j) {
        const double* c;
        const int b = quad_base(j, &c);
        for (int i = 1; i <= n; ++i) {
            double* dst = PCOL(F, i, j);
            const double* a0 = PCOL(&A, i, b);
            const double* a1 = PCOL(&A, i, b + 1);
            const double* a2 = PCOL(&A, i, b + 2);
            for (int k = 0; k < K; ++k) dst[k] = c[0] * a0[k] + c[1] * a1[k] + c[2] * a2[k];
        }
    }
    pf_free(&A);
}

/* ------------------------------------------------------------------------
 * Vertical filter (sup_vrbeta1_ens / sup_vrbeta1T_ens)
 * Column (i,j): W(kc, L) = V((kc-1)*lm + L, i, j) for kc=1..nvcol, L=1..lm.
 * ---------------------------------------------------------------------- */
static void vertical_op(const mgbfo_params* P, pf_t* V, int adjoint) {
    const int lm = P->lm, hz = P->hz, nv = P->nvcol;
    const int n = V->n, m = V->m;
    taps_t T = taps1_make(P);
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= m; ++j) {
        double* Wc = malloc(sizeof(double) * (lm + 2 * hz));
        double* Bc = malloc(sizeof(double) * (lm + 2 * hz));
#define WZ(L) Wc[(L) + hz - 1]
#define BZ(L) Bc[(L) + hz - 1]
        for (int i = 1; i <= n; ++i) {
            double* col = PCOL(V, i, j);
            for (int kc = 0; kc < nv; ++kc) {
                double* c = col + (size_t)kc * lm;
                for (int L = 1; L <= lm; ++L) WZ(L) = c[L - 1];
                for (int L = 1; L <= hz; ++L) {
                    WZ(1 - L) = WZ(1 + L);   /* whole-sample mirror, mg_filtering.f90:2072-2075 */
                    WZ(lm + L) = WZ(lm - L);
                }
                for (int L = 1 - hz; L <= lm + hz; ++L) BZ(L) = 0.0;
                if (!adjoint) {
                    /* vrbeta1: gather into the interior */
                    for (int ix = 1; ix <= lm; ++ix) {
                        double tb = 0;
                        for (int t = 0; t < T.nt; ++t) tb += T.w[t] * WZ(ix + T.gx[t]);
                        BZ(ix) = tb;
                    }
                } else {
                    /* vrbeta1T: scatter from the interior, then fold the halo
                       (mg_filtering.f90:2134-2137) */
                    for (int ix = 1; ix <= lm; ++ix) {
                        double ta = WZ(ix);
                        for (int t = 0; t < T.nt; ++t) BZ(ix + T.gx[t]) += T.w[t] * ta;
                    }
                    for (int L = 1; L <= hz; ++L) {
                        BZ(1 + L) += BZ(1 - L);
                        BZ(lm - L) += BZ(lm + L);
                    }
                }
                for (int L = 1; L <= lm; ++L) c[L - 1] = BZ(L);
            }
        }
#undef WZ
#undef BZ
        free(Wc);
        free(Bc);
    }
    taps_free(&T);
}

/* ------------------------------------------------------------------------
 * Single-stage wrappers on interior arrays
 * ---------------------------------------------------------------------- */

/* ------------------------------------------------------------------------
 * Fast (separable) horizontal filter: filtering_fast_ens / _bckg,
 * mg_filtering.f90:1316-1506, 825-997
 * ---------------------------------------------------------------------- */

typedef struct {
    int nx, ny;                  /* taps along x and y */
    int gx[64], gy[64];
    double wx[64], wy[64];
} fast_t;

static fast_t fast_make(const mgbfo_params* P) {
    fast_t F;
    F.nx = taps_fast_raw(P, P->hx, 64, F.gx, F.wx);
    F.ny = taps_fast_raw(P, P->hy, 64, F.gy, F.wy);
    if (F.nx < 0 || F.ny < 0 || F.nx > 64 || F.ny > 64) abort();   /* the reference stops */
    return F;
}

/* F: bocoy, rbeta along y over the interior columns, bocox, rbeta along x
   over the interior rows (gathers; the halo of each result is zero).
   F^T: rbetaT along x (scatter into the x halo), bocoTx, rbetaT along y,
   bocoTy. bocox / bocoy are the x and y halves of boco_2d, the same
   half-sample mirror. */
static void fast_filter(const fast_t* F, pf_t* a, int adjoint) {
    const int K = a->K, n = a->n, m = a->m, hx = a->hx, hy = a->hy;
    if (!adjoint) {
        for (int j = 1; j <= hy; ++j)
            for (int i = 1; i <= n; ++i) {
                memcpy(PCOL(a, i, 1 - j), PCOL(a, i, j), sizeof(double) * K);
                memcpy(PCOL(a, i, m + j), PCOL(a, i, m + 1 - j), sizeof(double) * K);
            }
        pf_t b = pf_alloc(K, n, m, hx, hy);
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= n; ++i) {
                double* d = PCOL(&b, i, j);
                for (int t = 0; t < F->ny; ++t) {
                    const double* src = PCOL(a, i, j + F->gy[t]);
                    for (int k = 0; k < K; ++k) d[k] += F->wy[t] * src[k];
                }
            }
        free(a->d);
        a->d = b.d;
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= hx; ++i) {
                memcpy(PCOL(a, 1 - i, j), PCOL(a, i, j), sizeof(double) * K);
                memcpy(PCOL(a, n + i, j), PCOL(a, n + 1 - i, j), sizeof(double) * K);
            }
        pf_t c = pf_alloc(K, n, m, hx, hy);
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= n; ++i) {
                double* d = PCOL(&c, i, j);
                for (int t = 0; t < F->nx; ++t) {
                    const double* src = PCOL(a, i + F->gx[t], j);
                    for (int k = 0; k < K; ++k) d[k] += F->wx[t] * src[k];
                }
            }
        free(a->d);
        a->d = c.d;
    } else {
        pf_t b = pf_alloc(K, n, m, hx, hy);
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= n; ++i) {
                const double* src = PCOL(a, i, j);
                for (int t = 0; t < F->nx; ++t) {
                    double* d = PCOL(&b, i + F->gx[t], j);
                    for (int k = 0; k < K; ++k) d[k] += F->wx[t] * src[k];
                }
            }
        free(a->d);
        a->d = b.d;
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= hx; ++i) {
                vadd(PCOL(a, i, j), PCOL(a, 1 - i, j), K);
                vadd(PCOL(a, n + 1 - i, j), PCOL(a, n + i, j), K);
            }
        pf_t c = pf_alloc(K, n, m, hx, hy);
#pragma omp parallel for schedule(static)
        for (int i = 1; i <= n; ++i)
            for (int j = 1; j <= m; ++j) {
                const double* src = PCOL(a, i, j);
                for (int t = 0; t < F->ny; ++t) {
                    double* d = PCOL(&c, i, j + F->gy[t]);
                    for (int k = 0; k < K; ++k) d[k] += F->wy[t] * src[k];
                }
            }
        free(a->d);
        a->d = c.d;
        for (int j = 1; j <= hy; ++j)
            for (int i = 1; i <= n; ++i) {
                vadd(PCOL(a, i, j), PCOL(a, i, 1 - j), K);
                vadd(PCOL(a, i, m + 1 - j), PCOL(a, i, m + j), K);
            }
    }
}

/* ------------------------------------------------------------------------
 * 3-D radial filter (procedure 4, filtering_rad3): the 3-D variables of the
 * stack are filtered over (i, j, L) by the beta filter of the 3x3 aspect
 * pasp3 (mg_intstate.f90:395-410; p_sig = p_rho = 0), the 2-D variables by
 * the 2-D filter of pasp2. L has the whole-sample mirror of sup_vrbeta3.
 * ---------------------------------------------------------------------- */

/* cholaspect3 (jp_pbfil.f90:108-128): inverse, then the lower Cholesky
   factor (dl1lmf), of a symmetric 3x3 matrix a (row-major). */
static void cholaspect3(const double a[9], double el[9]) {
    double inv[9];
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8],
                 c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    inv[0] = c00 / det;
    inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    inv[3] = c01 / det;
    inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    inv[6] = c02 / det;
    inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
    for (int e = 0; e < 9; ++e) el[e] = 0.0;
    for (int j = 0; j < 3; ++j) {
        double sj = inv[j * 3 + j];
        for (int q = 0; q < j; ++q) sj -= el[j * 3 + q] * el[j * 3 + q];
        el[j * 3 + j] = sqrt(sj);
        for (int i = j + 1; i < 3; ++i) {
            double si = inv[i * 3 + j];
            for (int q = 0; q < j; ++q) si -= el[i * 3 + q] * el[j * 3 + q];
            el[i * 3 + j] = si / el[j * 3 + j];
        }
    }
}

/* Taps of the 3-D filter in vrbeta3's loop order (gz, gy, gx;
   jp_pbfil.f90:955-971). Returns the count, or -1 beyond a halo. */
static int taps3_raw(const mgbfo_params* P, int maxt, int* gxo, int* gyo, int* gzo, double* rrco) {
    const double rmom2_3 = 1.0 / sqrt((double)(2 * P->p) + 5.0);
    const double a3 = P->ampl03;
    const double a[9] = {a3 * (1. + P->p_del), a3 * P->p_eps, 0.0,
                         a3 * P->p_eps, a3, 0.0,
                         0.0, 0.0, a3 * (1. - P->p_del)};
    double el[9];
    cholaspect3(a, el);
    const double exx = el[0] * rmom2_3, eyy = el[4] * rmom2_3, ezz = el[8] * rmom2_3;
    const double eyx = el[3] * rmom2_3, ezx = el[6] * rmom2_3, ezy = el[7] * rmom2_3;
    const double z0 = 1.0 / ezz;
    const int gzl = (int)ceil(-z0 + EPS), gzm = (int)floor(z0 - EPS);
    if (gzl < -P->hz || gzm > P->hz) return -1;
    int nt = 0;
    for (int gz = gzl; gz <= gzm; ++gz) {
        const double z = gz, yc = -z * ezy;
        const double rry = (z * ezz) * (z * ezz), y0 = sqrt(1.0 - rry);
        const int gyl = (int)ceil((yc - y0) / eyy + EPS), gym = (int)floor((yc + y0) / eyy - EPS);
        if (gyl < -P->hy || gym > P->hy) return -1;
        for (int gy = gyl; gy <= gym; ++gy) {
            const double y = gy, xc = -y * eyx - z * ezx;
            const double rrx = rry + (y * eyy - yc) * (y * eyy - yc), x0 = sqrt(1.0 - rrx);
            const int gxl = (int)ceil((xc - x0) / exx + EPS), gxm = (int)floor((xc + x0) / exx - EPS);
            if (gxl < -P->hx || gxm > P->hx) return -1;
            for (int gx = gxl; gx <= gxm; ++gx) {
                const double x = gx;
                const double rr = rrx + (x * exx - xc) * (x * exx - xc);
                if (nt < maxt) {
                    gxo[nt] = gx;
                    gyo[nt] = gy;
                    gzo[nt] = gz;
                    rrco[nt] = 1.0 - rr;
                }
                ++nt;
            }
        }
    }
    return nt;
}

typedef struct {
    int nt;
    int *gx, *gy, *gz;
    double* w;
} taps3_t;

/* taps with weights ss3 rrc^p (getlinesum3 then vrbeta3) */
static taps3_t taps3_make(const mgbfo_params* P) {
    const int cap = (2 * P->hx + 1) * (2 * P->hy + 1) * (2 * P->hz + 1);
    taps3_t T;
    T.gx = malloc(sizeof(int) * cap);
    T.gy = malloc(sizeof(int) * cap);
    T.gz = malloc(sizeof(int) * cap);
    T.w = malloc(sizeof(double) * cap);
    T.nt = taps3_raw(P, cap, T.gx, T.gy, T.gz, T.w);
    if (T.nt < 0) abort();   /* "In getlinesum3; filter reach becomes too large" */
    double s = 0;
    for (int t = 0; t < T.nt; ++t) s += ipow(T.w[t], P->p);
    const double ss = 1.0 / s;
    for (int t = 0; t < T.nt; ++t) T.w[t] = ss * ipow(T.w[t], P->p);
    return T;
}

static void taps3_free(taps3_t* T) {
    free(T->gx); free(T->gy); free(T->gz); free(T->w);
}

int mgbfo_taps3(const mgbfo_params* P, int maxt, int* gx, int* gy, int* gz, double* w) {
    taps3_t T = taps3_make(P);
    for (int t = 0; t < T.nt && t < maxt; ++t) {
        gx[t] = T.gx[t];
        gy[t] = T.gy[t];
        gz[t] = T.gz[t];
        w[t] = T.w[t];
    }
    const int nt = T.nt;
    taps3_free(&T);
    return nt;
}

/* The filter of filtering_rad3 on one generation of the stack (padded
   field a with halos hx, hy). Forward: boco_2d, then the 3-D variables
   through sup_vrbeta3 (levels padded by the whole-sample mirror, vrbeta3
   gathers into the interior) and the 2-D variables through vrbeta2.
   Adjoint: vrbeta3T scatters from the interior into the padded volume, the
   level halo folds back (sup_vrbeta3T), the 2-D variables scatter
   (vrbeta2T), then bocoT_2d. */
static void rad3_filter(const mgbfo_params* P, const taps_t* T2, const taps3_t* T3, pf_t* a,
                        int adjoint) {
    const int K = a->K, n = a->n, m = a->m, hx = a->hx, hy = a->hy, hz = P->hz;
    const int lm = P->lm, nv = P->nvcol, K3 = nv * lm, K2 = K - K3;
    const int nxp = n + 2 * hx, nyp = m + 2 * hy, nlp = lm + 2 * hz;
    /* one 3-D variable as W(i, j, L), padded in all three */
#define W3(w, i, j, L) (w)[((size_t)((L) + hz - 1) * nyp + (size_t)((j) + hy - 1)) * nxp + (size_t)((i) + hx - 1)]
    if (!adjoint) boco_2d(a, hx, hy);
    pf_t b = pf_alloc(K, n, m, hx, hy);
#pragma omp parallel for schedule(dynamic)
    for (int v = 0; v < nv; ++v) {
        double* W = calloc((size_t)nxp * nyp * nlp, sizeof(double));
        double* B = calloc((size_t)nxp * nyp * nlp, sizeof(double));
        for (int L = 1; L <= lm; ++L)
            for (int j = 1 - hy; j <= m + hy; ++j)
                for (int i = 1 - hx; i <= n + hx; ++i) W3(W, i, j, L) = PF(a, v * lm + L - 1, i, j);
        if (!adjoint) {
            for (int L = 1; L <= hz; ++L)
                for (int j = 1 - hy; j <= m + hy; ++j)
                    for (int i = 1 - hx; i <= n + hx; ++i) {
                        W3(W, i, j, 1 - L) = W3(W, i, j, 1 + L);
                        W3(W, i, j, lm + L) = W3(W, i, j, lm - L);
                    }
            for (int L = 1; L <= lm; ++L)
                for (int j = 1; j <= m; ++j)
                    for (int i = 1; i <= n; ++i) {
                        double tb = 0;
                        for (int t = 0; t < T3->nt; ++t)
                            tb += T3->w[t] * W3(W, i + T3->gx[t], j + T3->gy[t], L + T3->gz[t]);
                        PF(&b, v * lm + L - 1, i, j) = tb;
                    }
        } else {
            for (int L = 1; L <= lm; ++L)
                for (int j = 1; j <= m; ++j)
                    for (int i = 1; i <= n; ++i) {
                        const double ta = W3(W, i, j, L);
                        for (int t = 0; t < T3->nt; ++t)
                            W3(B, i + T3->gx[t], j + T3->gy[t], L + T3->gz[t]) += T3->w[t] * ta;
                    }
            for (int L = 1; L <= hz; ++L)
                for (int j = 1 - hy; j <= m + hy; ++j)
                    for (int i = 1 - hx; i <= n + hx; ++i) {
                        W3(B, i, j, 1 + L) += W3(B, i, j, 1 - L);
                        W3(B, i, j, lm - L) += W3(B, i, j, lm + L);
                    }
            for (int L = 1; L <= lm; ++L)
                for (int j = 1 - hy; j <= m + hy; ++j)
                    for (int i = 1 - hx; i <= n + hx; ++i) PF(&b, v * lm + L - 1, i, j) = W3(B, i, j, L);
        }
        free(W);
        free(B);
    }
#undef W3
    if (K2 > 0) {
        /* the 2-D variables: the stack's tail as its own padded field */
        pf_t a2 = pf_alloc(K2, n, m, hx, hy);
        for (int j = 1 - hy; j <= m + hy; ++j)
            for (int i = 1 - hx; i <= n + hx; ++i)
                memcpy(PCOL(&a2, i, j), PCOL(a, i, j) + K3, sizeof(double) * K2);
        if (!adjoint) vrbeta2(T2, &a2);
        else vrbeta2T(T2, &a2);
        for (int j = 1 - hy; j <= m + hy; ++j)
            for (int i = 1 - hx; i <= n + hx; ++i)
                memcpy(PCOL(&b, i, j) + K3, PCOL(&a2, i, j), sizeof(double) * K2);
        pf_free(&a2);
    }
    free(a->d);
    a->d = b.d;
    if (adjoint) bocoT_2d(a, hx, hy);
}

/* ------------------------------------------------------------------------
 * Triad line filters (procedures 2 and 7, filtering_lin2_bckg / _ens).
 * init_mg_line resolves the aspect 3-vector of each point onto three
 * lattice lines (triad), turns each line's spread**2 into a real half-span
 * (hstform), and each line is filtered by the "dibeta" mixture of the beta
 * filters of the two neighbouring integer half-spans (blinfil), gathered
 * along the line (vdibeta2) or scattered (vdibeta2t).
 * ---------------------------------------------------------------------- */
#define LINE_NH 6     /* nh, mg_parameter.f90:562 */
#define LINE_NFIL 8   /* nfil = nh + 2, mg_parameter.f90:563 */

/* inimomtab (jp_pbfil3.f90:378-457): for the beta exponent p and integer
   half-spans h = 1..nh, bnorm[h] = 1 / (0th moment) and bsprds[h] =
   (2nd moment) / (0th moment), from the continuum moments of (1-(x/h)^2)^p
   plus their Euler-Maclaurin corrections (polynomials n0pk, n2pk in h^2).
   Arrays are indexed 1..nh. Returns 1 for p or nh out of range. */
int mgbfo_momtab(int p, int nh, double* bnorm, double* bsprds) {
    static const double n0pk[6][3] = {
        {-1, 0, 0}, {-1, 0, 0}, {-5, 14, 0}, {-63, 240, 0},
        {-1575, 6930, -2640}, {-68409, 327600, -216216}};
    static const double n2pk[6][4] = {
        {1, -5, 0, 0}, {5, -21, 0, 0}, {63, -285, 126, 0}, {1575, -7623, 5280, 0},
        {68409, -348075, 306306, -34320}, {4729725, -24969285, 25552800, -5405400}};
    if (p < 1 || p > 6 || nh < 2 || nh > 1000) return 1;
    double ffac[17];            /* ffac[i + 1] = i!! for i = -1 .. 15 */
    ffac[0] = 1.0;
    ffac[1] = 1.0;
    for (int i = 1; i <= 15; ++i) ffac[i + 1] = i * ffac[i - 1];
#define FFAC(i) ffac[(i) + 1]
    const int p2 = 2 * p, mk0 = (p - 1) / 2, mk2 = mk0 + 1;
    for (int h = 1; h <= nh; ++h) {
        const double x = h, xx = x * x;
        const double m0 = 2.0 * FFAC(p2) * x / FFAC(p2 + 1);
        const double m2 = 2.0 * FFAC(p2) * (x * x * x) / FFAC(p2 + 3);
        double hm0 = 0.0, hm2 = 0.0;
        for (int k = 0; k <= mk0; ++k) hm0 += n0pk[p - 1][k] * ipow(xx, k);
        for (int k = 0; k <= mk2; ++k) hm2 += n2pk[p - 1][k] * ipow(xx, k);
        const double cm0 = m0 + hm0 / (FFAC(p2 + 1) * ipow(x, p2 - 1));
        const double cm2 = m2 + hm2 / (FFAC(p2 + 3) * ipow(x, p2 - 1));
        bnorm[h] = 1.0 / cm0;
        bsprds[h] = cm2 / cm0;
    }
#undef FFAC
    return 0;
}

/* hstform (jp_pbfil3.f90:1399-1430): spread**2 -> real half-span hs by
   linear interpolation in the table of spreads**2. Returns 1 when hs is
   negative or beyond the table. */
int mgbfo_hstform(int p, double* hs) {
    double bnorm[LINE_NH + 1], bsprds[LINE_NH + 1];
    if (mgbfo_momtab(p, LINE_NH, bnorm, bsprds)) return 1;
    const int ff = *hs < 0.0;
    for (int h = 2; h <= LINE_NH; ++h)
        if (bsprds[h] >= *hs) {
            *hs = h - (bsprds[h] - *hs) / (bsprds[h] - bsprds[h - 1]);
            return ff;
        }
    return 1;
}

/* blinfil (jp_pbfil3.f90:1463-1501): half of the symmetric dibeta filter of
   real half-span hspan: fil[0..h], h = int(hspan), the mixture of the
   normalized beta filters of half-spans h+1 and h with weights hspan-h and
   h+1-hspan. Returns 1 when h < 1 or h+1 exceeds nh or nfil. */
int mgbfo_blinfil(int p, double hspan, int* h, double* fil) {
    double bnorm[LINE_NH + 1], bsprds[LINE_NH + 1];
    if (mgbfo_momtab(p, LINE_NH, bnorm, bsprds)) return 1;
    const int hh = (int)hspan, hp = hh + 1;
    *h = hh;
    if (hh < 1 || hp > LINE_NH || hp > LINE_NFIL) return 1;
    const double whp = (hspan - hh) * bnorm[hp];
    const double wh = (hp - hspan) * bnorm[hh];
    for (int i = 0; i <= hh; ++i) {
        double z = (double)i / hp;
        z = z * z;
        fil[i] = whp * ipow(1.0 - z, p);
    }
    for (int i = 0; i < hh; ++i) {
        double z = (double)i / hh;
        z = z * z;
        fil[i] += wh * ipow(1.0 - z, p);
    }
    return 0;
}

/* triad (jp_pbfil3.f90:573-624): resolve the aspect 3-vector
   (A11, A22, A12) onto three integer line generators (ltri[2c], ltri[2c+1])
   with weights wtri[c] >= 0, starting from the lines (1,0), (0,1), (-1,-1)
   and flipping the most negative one until none is. lui (columns c) maps
   the aspect to the weights. Returns 1 when 200 iterations do not end it. */
int mgbfo_triad(const double aspect[3], int ltri[6], double wtri[3]) {
    int L[3][2] = {{1, 0}, {0, 1}, {-1, -1}};
    int lui[3][3] = {{1, 0, -1}, {0, 1, -1}, {0, 0, 1}};   /* lui[c][r] */
    for (int c = 0; c < 3; ++c)
        wtri[c] = aspect[0] * lui[c][0] + aspect[1] * lui[c][1] + aspect[2] * lui[c][2];
    int it;
    for (it = 1; it <= 200; ++it) {
        int k = 0;
        for (int c = 1; c < 3; ++c)
            if (wtri[c] < wtri[k]) k = c;
        const double dw = 2.0 * wtri[k];
        if (dw >= -1.e-14) break;
        const int l = (k + 1) % 3, m = (l + 1) % 3;
        int dl[3];
        for (int r = 0; r < 3; ++r) dl[r] = 2 * lui[k][r];
        L[l][0] = -L[l][0];
        L[l][1] = -L[l][1];
        L[k][0] = -L[l][0] - L[m][0];
        L[k][1] = -L[l][1] - L[m][1];
        for (int c = 0; c < 3; ++c) {
            const int sg = c == k ? -1 : 1;
            for (int r = 0; r < 3; ++r) lui[c][r] += dl[r] * sg;
            wtri[c] += sg * dw;
        }
    }
    for (int c = 0; c < 3; ++c) {
        ltri[2 * c] = L[c][0];
        ltri[2 * c + 1] = L[c][1];
    }
    return it > 200;
}

/* The three lines of the domain: init_mg_line on the aspect pasp2 after
   def_mg_weights has replaced it by the Cholesky factor of its inverse
   (cholaspect), which is what reaches t22_to_3 and tritform. Without the
   tables (mgbf_line off) every generator is 0 and each line is the
   identity. The tables do not vary over the domain (p_del = p_eps fixed). */
typedef struct {
    int dx[3], dy[3];
    int h[3];                 /* discrete half-span, or -1: the line is the identity */
    double fil[3][LINE_NFIL + 1];
} line_t;

static line_t line_make(const mgbfo_params* P) {
    line_t Lf;
    memset(&Lf, 0, sizeof Lf);
    for (int c = 0; c < 3; ++c) Lf.h[c] = -1;
    if (!P->line_tables) return Lf;
    double a[4], el[4];
    aspect2(P, a);
    cholaspect2(a, el);
    const double v[3] = {el[0], el[3], el[2]};   /* t22_to_3: (1,1), (2,2), (1,2) */
    int ltri[6];
    double hs[3];
    if (mgbfo_triad(v, ltri, hs)) abort();       /* tritform failure: tables unset */
    for (int c = 0; c < 3; ++c)
        if (mgbfo_hstform(P->p, &hs[c])) abort();
    for (int c = 0; c < 3; ++c) {
        Lf.dx[c] = ltri[2 * c];
        Lf.dy[c] = ltri[2 * c + 1];
        if (Lf.dx[c] == 0 && Lf.dy[c] == 0) continue;
        int h;
        /* blinfil failing returns from vdibeta2 with the field untouched */
        if (mgbfo_blinfil(P->p, hs[c], &h, Lf.fil[c])) continue;
        const int ax = abs(Lf.dx[c]) * h, ay = abs(Lf.dy[c]) * h;
        if (ax > P->hx || ay > P->hy) abort();   /* reads beyond the halo */
        Lf.h[c] = h;
    }
    return Lf;
}

/* vdibeta2 / vdibeta2t on one line c of a padded field: the result replaces
   the field, halo included (the reference's a = b). */
static void dibeta_line(const line_t* Lf, int c, pf_t* a, int adjoint) {
    const int K = a->K, n = a->n, m = a->m;
    const int dx = Lf->dx[c], dy = Lf->dy[c], h = Lf->h[c];
    const int ident = dx == 0 && dy == 0;
    if (!ident && h < 0) return;
    pf_t b = pf_alloc(K, n, m, a->hx, a->hy);
    if (!adjoint) {
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= n; ++i) {
                double* d = PCOL(&b, i, j);
                const double* s0 = PCOL(a, i, j);
                if (ident) {
                    memcpy(d, s0, sizeof(double) * K);
                    continue;
                }
                for (int k = 0; k < K; ++k) d[k] = Lf->fil[c][0] * s0[k];
                for (int t = 1; t <= h; ++t) {
                    const double f = Lf->fil[c][t];
                    const double* sp = PCOL(a, i + dx * t, j + dy * t);
                    const double* sm = PCOL(a, i - dx * t, j - dy * t);
                    for (int k = 0; k < K; ++k) d[k] += f * (sp[k] + sm[k]);
                }
            }
    } else {
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= n; ++i) {
                const double* s0 = PCOL(a, i, j);
                if (ident) {
                    vadd(PCOL(&b, i, j), s0, K);
                    continue;
                }
                double* d0 = PCOL(&b, i, j);
                for (int k = 0; k < K; ++k) d0[k] += Lf->fil[c][0] * s0[k];
                for (int t = 1; t <= h; ++t) {
                    const double f = Lf->fil[c][t];
                    double* dp = PCOL(&b, i + dx * t, j + dy * t);
                    double* dm = PCOL(&b, i - dx * t, j - dy * t);
                    for (int k = 0; k < K; ++k) {
                        const double v = f * s0[k];
                        dp[k] += v;
                        dm[k] += v;
                    }
                }
            }
    }
    free(a->d);
    a->d = b.d;
}

/* The line filter of filtering_lin2_*: forward boco_2d + dibeta for lines
   1, 2, 3; adjoint dibetat + bocoT_2d for lines 3, 2, 1. */
static void line_filter(const mgbfo_params* P, const line_t* Lf, pf_t* a, int adjoint) {
    if (!adjoint) {
        for (int c = 0; c < 3; ++c) {
            boco_2d(a, P->hx, P->hy);
            dibeta_line(Lf, c, a, 0);
        }
    } else {
        for (int c = 2; c >= 0; --c) {
            dibeta_line(Lf, c, a, 1);
            bocoT_2d(a, P->hx, P->hy);
        }
    }
}

/* ------------------------------------------------------------------------
 * Hexad line filters of the 3-D variables (procedure 5, filtering_lin3).
 * init_mg_line resolves the 6-vector of pasp3 (after cholaspect3) onto six
 * lattice lines with "colors" 1..7 (hexad, hextform); filtering_lin3 runs
 * the color of each call, 1..7 forward and 7..1 adjoint, each line a dibeta
 * filter (vdibetax3 / vdibetax3t). The level halo is filled once by the
 * whole-sample mirror before the forward loop; boco_3d / bocoT_3d refresh
 * and fold only the x/y halos of the interior levels.
 * ---------------------------------------------------------------------- */

/* hexad (jp_pbfil3.f90:798-904): aspect (A11, A22, A33, A23, A31, A12) ->
   seven generators lhex7[3c..3c+2] (colors c+1 = 1..7, one of them null)
   and weights whex7[c]. Returns 1 when 200 iterations do not end it. */
int mgbfo_hexad(const double aspect[6], int lhex7[21], double whex7[7]) {
    static const int deflhex[7][3] = {{0, 0, 0}, {1, -1, 0}, {0, 1, -1}, {0, 0, 1},
                                      {-1, 0, 1}, {0, 1, 0}, {1, 0, 0}};
    static const int deflui[7][6] = {{0, 0, 0, 0, 0, 0},  {0, 0, 0, 0, 0, -1}, {0, 0, 0, -1, 0, 0},
                                     {0, 0, 1, 1, 1, 0},  {0, 0, 0, 0, -1, 0}, {0, 1, 0, 1, 0, 1},
                                     {1, 0, 0, 0, 1, 1}};
    /* Mmats(:,:,d), d = 1..6, column-major: mm[d-1][col][row] */
    static const int mm[6][3][3] = {{{1, 1, -1}, {1, 0, 0}, {1, 0, -1}},
                                    {{-1, 1, 0}, {-1, 1, 1}, {0, 1, 0}},
                                    {{0, -1, 1}, {1, -1, 0}, {1, 0, 0}},
                                    {{0, 0, 1}, {0, -1, 1}, {1, -1, 1}},
                                    {{-1, 0, 1}, {0, 0, 1}, {-1, 1, 0}},
                                    {{0, 1, 0}, {1, 0, -1}, {0, 1, -1}}};
    static const int ttriad[6] = {5, 3, 3, 6, 5, 6};
    static const int sstriad[13] = {-1, -1, 1, -1, 1, 1, 1, -1, -1, 1, -1, 1, 1};  /* -6..6 */
    static const int line3[3] = {1, 2, 4};
    static const int jcol[7] = {7, 4, 6, 3, 5, 2, 1};
    int lhex[7][3], lui[7][6], kset[3][3];
    double whex[7];
    memcpy(lhex, deflhex, sizeof lhex);
    memcpy(lui, deflui, sizeof lui);
    int hcol = 0;
    for (int c = 0; c < 7; ++c) {
        whex[c] = 0.0;
        for (int r = 0; r < 6; ++r) whex[c] += aspect[r] * lui[c][r];
    }
    for (int i = 0; i < 3; ++i)
        for (int e = 0; e < 3; ++e) kset[i][e] = lhex[((hcol - line3[i]) % 7 + 7) % 7][e];
    int it, kcol = 0;
    for (it = 1; it <= 200; ++it) {
        kcol = 0;
        for (int c = 1; c < 7; ++c)
            if (whex[c] < whex[kcol]) kcol = c;
        const double dw = whex[kcol];
        if (dw >= -1.e-14) break;
        const int dcol = ((kcol - hcol) % 7 + 7) % 7;
        hcol = kcol;
        const int L = (hcol + ttriad[dcol - 1]) % 7;
        int knew[3][3];
        for (int j = 0; j < 3; ++j)
            for (int e = 0; e < 3; ++e) {
                int acc = 0;
                for (int i = 0; i < 3; ++i) acc += kset[i][e] * mm[dcol - 1][j][i];
                knew[j][e] = acc;
            }
        memcpy(kset, knew, sizeof kset);
        int dl[6];
        for (int r = 0; r < 6; ++r) dl[r] = lui[hcol][r];
        for (int c = 0; c < 7; ++c) {
            const int sg = sstriad[c - L + 6];
            for (int r = 0; r < 6; ++r) lui[c][r] += dl[r] * sg;
            whex[c] += sg * dw;
        }
    }
    if (it > 200) return 1;
    for (int i = 0; i < 3; ++i) {
        const int ip = (i + 1) % 3;
        for (int e = 0; e < 3; ++e) {
            lhex[((hcol - line3[i]) % 7 + 7) % 7][e] = kset[i][e];
            lhex[(hcol + line3[i]) % 7][e] = kset[i][e] - kset[ip][e];
        }
    }
    for (int e = 0; e < 3; ++e) lhex[kcol][e] = 0;
    for (int i = 0; i < 7; ++i) {
        const int j = jcol[i] - 1;
        for (int e = 0; e < 3; ++e) lhex7[3 * j + e] = lhex[i][e];
        whex7[j] = whex[i];
    }
    return 0;
}

/* The six lines of the domain's 3-D variables (hextform on the Cholesky
   factor of pasp3's inverse, as init_mg_line sees it): qcol[j] is the color
   of line j, generator g[j], half filter fil[j][0..h[j]]. */
typedef struct {
    int n;                 /* lines (6) */
    int qcol[8];           /* qcol[1..n]; qcol[0] = qcol[7] = 0 */
    int g[7][3];
    int h[7];
    double fil[7][LINE_NFIL + 1];
} hex_t;

static hex_t hex_make(const mgbfo_params* P) {
    hex_t X;
    memset(&X, 0, sizeof X);
    if (!P->line_tables) return X;   /* no lines: every color copies */
    const double a3 = P->ampl03;
    const double a[9] = {a3 * (1. + P->p_del), a3 * P->p_eps, 0.0,
                         a3 * P->p_eps, a3, 0.0,
                         0.0, 0.0, a3 * (1. - P->p_del)};
    double el[9];
    cholaspect3(a, el);
    /* t33_to_6 of the lower factor: (1,1), (2,2), (3,3), (2,3), (3,1), (1,2);
       the upper entries (2,3) and (1,2) are 0 */
    const double asp[6] = {el[0], el[4], el[8], 0.0, el[6], 0.0};
    int lh7[21];
    double wh7[7];
    if (mgbfo_hexad(asp, lh7, wh7)) abort();
    int j = 1;
    double hs[7];
    for (int c = 1; c <= 7; ++c) {
        const int* g = lh7 + 3 * (c - 1);
        if (abs(g[0]) + abs(g[1]) + abs(g[2]) == 0) continue;
        if (j > 6) abort();    /* inconsistent hexad generator set */
        X.qcol[j] = c;
        X.g[j][0] = g[0];
        X.g[j][1] = g[1];
        X.g[j][2] = g[2];
        hs[j] = wh7[c - 1];
        ++j;
    }
    if (j != 7) abort();
    X.n = 6;
    for (int l = 1; l <= 6; ++l) {
        if (mgbfo_hstform(P->p, &hs[l])) abort();
        int h;
        /* a failing blinfil leaves the call after moving one point's line
           counter: not a state to restate */
        if (mgbfo_blinfil(P->p, hs[l], &h, X.fil[l])) abort();
        X.h[l] = h;
        if (abs(X.g[l][0]) * h > P->hx || abs(X.g[l][1]) * h > P->hy || abs(X.g[l][2]) * h > P->hz)
            abort();           /* reads beyond the halos */
    }
    return X;
}

int mgbfo_hexlines(const mgbfo_params* P, int* qcol, int* g, int* h, double* fil) {
    hex_t X = hex_make(P);
    for (int l = 1; l <= 6; ++l) {
        qcol[l - 1] = X.qcol[l];
        for (int e = 0; e < 3; ++e) g[3 * (l - 1) + e] = X.g[l][e];
        h[l - 1] = X.h[l];
        for (int t = 0; t <= LINE_NFIL; ++t) fil[(l - 1) * (LINE_NFIL + 1) + t] = X.fil[l][t];
    }
    return X.n;
}

/* The filter of filtering_lin3 on one generation of the stack: the 2-D
   variables through the triad lines of pasp2 (dibeta / dibetat with
   boco_2d / bocoT_2d), the 3-D variables through the hexad lines over
   (i, j, L). */
static void lin3_filter(const mgbfo_params* P, const line_t* L2, const hex_t* X, pf_t* a,
                        int adjoint) {
    const int K = a->K, n = a->n, m = a->m, hx = a->hx, hy = a->hy, hz = P->hz;
    const int lm = P->lm, nv = P->nvcol, K3 = nv * lm, K2 = K - K3;
    const int nxp = n + 2 * hx, nyp = m + 2 * hy, nlp = lm + 2 * hz;
    const size_t vol = (size_t)nxp * nyp * nlp;
#define W3(w, i, j, L) (w)[((size_t)((L) + hz - 1) * nyp + (size_t)((j) + hy - 1)) * nxp + (size_t)((i) + hx - 1)]
    pf_t b = pf_alloc(K, n, m, hx, hy);
    memcpy(b.d, a->d, sizeof(double) * (size_t)K * nxp * nyp);
#pragma omp parallel for schedule(dynamic)
    for (int v = 0; v < nv; ++v) {
        double* W = calloc(vol, sizeof(double));
        double* B = malloc(vol * sizeof(double));
        for (int L = 1; L <= lm; ++L)
            for (int j = 1 - hy; j <= m + hy; ++j)
                for (int i = 1 - hx; i <= n + hx; ++i) W3(W, i, j, L) = PF(a, v * lm + L - 1, i, j);
        /* boco_3d / bocoT_3d: the x/y halos of the interior levels */
        pf_t lev = {NULL, 1, n, m, hx, hy};
        double* levbuf = malloc(sizeof(double) * (size_t)nxp * nyp);
        lev.d = levbuf;
        int jc = 0;
        if (!adjoint) {
            for (int L = 1; L <= hz; ++L)
                for (int j = 1 - hy; j <= m + hy; ++j)
                    for (int i = 1 - hx; i <= n + hx; ++i) {
                        W3(W, i, j, 1 - L) = W3(W, i, j, 1 + L);
                        W3(W, i, j, lm + L) = W3(W, i, j, lm - L);
                    }
            for (int icol = 1; icol <= 7; ++icol) {
                for (int L = 1; L <= lm; ++L) {
                    memcpy(levbuf, &W3(W, 1 - hx, 1 - hy, L), sizeof(double) * (size_t)nxp * nyp);
                    boco_2d(&lev, hx, hy);
                    memcpy(&W3(W, 1 - hx, 1 - hy, L), levbuf, sizeof(double) * (size_t)nxp * nyp);
                }
                if (icol == 1) jc = 1;
                const int hit = jc <= 6 && X->qcol[jc] == icol;
                memset(B, 0, vol * sizeof(double));
                if (!hit) {
                    for (int L = 1; L <= lm; ++L)
                        for (int j = 1; j <= m; ++j)
                            for (int i = 1; i <= n; ++i) W3(B, i, j, L) = W3(W, i, j, L);
                } else {
                    const int l = jc++;
                    const int* g = X->g[l];
                    for (int L = 1; L <= lm; ++L)
                        for (int j = 1; j <= m; ++j)
                            for (int i = 1; i <= n; ++i) {
                                double t = X->fil[l][0] * W3(W, i, j, L);
                                for (int q = 1; q <= X->h[l]; ++q)
                                    t += X->fil[l][q] * (W3(W, i + g[0] * q, j + g[1] * q, L + g[2] * q) +
                                                         W3(W, i - g[0] * q, j - g[1] * q, L - g[2] * q));
                                W3(B, i, j, L) = t;
                            }
                }
                memcpy(W, B, vol * sizeof(double));
            }
        } else {
            for (int icol = 7; icol >= 1; --icol) {
                /* the level halo is filled here too; the scatter from the
                   interior below replaces the whole array */
                if (icol == 7) jc = 6;
                const int hit = jc >= 1 && X->qcol[jc] == icol;
                memset(B, 0, vol * sizeof(double));
                if (!hit) {
                    for (int L = 1; L <= lm; ++L)
                        for (int j = 1; j <= m; ++j)
                            for (int i = 1; i <= n; ++i) W3(B, i, j, L) += W3(W, i, j, L);
                } else {
                    const int l = jc--;
                    const int* g = X->g[l];
                    for (int L = 1; L <= lm; ++L)
                        for (int j = 1; j <= m; ++j)
                            for (int i = 1; i <= n; ++i) {
                                const double at = W3(W, i, j, L);
                                W3(B, i, j, L) += X->fil[l][0] * at;
                                for (int q = 1; q <= X->h[l]; ++q) {
                                    const double f = X->fil[l][q] * at;
                                    W3(B, i + g[0] * q, j + g[1] * q, L + g[2] * q) += f;
                                    W3(B, i - g[0] * q, j - g[1] * q, L - g[2] * q) += f;
                                }
                            }
                }
                memcpy(W, B, vol * sizeof(double));
                for (int L = 1; L <= lm; ++L) {
                    memcpy(levbuf, &W3(W, 1 - hx, 1 - hy, L), sizeof(double) * (size_t)nxp * nyp);
                    bocoT_2d(&lev, hx, hy);
                    memcpy(&W3(W, 1 - hx, 1 - hy, L), levbuf, sizeof(double) * (size_t)nxp * nyp);
                }
            }
        }
        for (int L = 1; L <= lm; ++L)
            for (int j = 1 - hy; j <= m + hy; ++j)
                for (int i = 1 - hx; i <= n + hx; ++i) PF(&b, v * lm + L - 1, i, j) = W3(W, i, j, L);
        free(levbuf);
        free(W);
        free(B);
    }
#undef W3
    if (K2 > 0) {
        pf_t a2 = pf_alloc(K2, n, m, hx, hy);
        for (int j = 1 - hy; j <= m + hy; ++j)
            for (int i = 1 - hx; i <= n + hx; ++i)
                memcpy(PCOL(&a2, i, j), PCOL(a, i, j) + K3, sizeof(double) * K2);
        line_filter(P, L2, &a2, adjoint);
        for (int j = 1 - hy; j <= m + hy; ++j)
            for (int i = 1 - hx; i <= n + hx; ++i)
                memcpy(PCOL(&b, i, j) + K3, PCOL(&a2, i, j), sizeof(double) * K2);
        pf_free(&a2);
    }
    free(a->d);
    a->d = b.d;
}

/* One line of the domain's line filter, for the pins: generator and the
   half filter fil[0..h]; h = -1 for an identity line. */
int mgbfo_line(const mgbfo_params* P, int c, int* dx, int* dy, int* h, double* fil) {
    if (c < 0 || c > 2) return 1;
    line_t Lf = line_make(P);
    *dx = Lf.dx[c];
    *dy = Lf.dy[c];
    *h = Lf.h[c];
    for (int t = 0; t <= (Lf.h[c] > 0 ? Lf.h[c] : 0); ++t) fil[t] = Lf.fil[c][t];
    return 0;
}

/* The horizontal filter of the procedure: the radial 2-D beta filter with
   boco_2d / bocoT_2d, the fast separable one, or the triad line filters. */
typedef struct {
    taps_t T;
    fast_t F;
    line_t L;
    taps3_t T3;
    hex_t X;
} hfilt_t;

static hfilt_t hfilt_make(const mgbfo_params* P) {
    hfilt_t H;
    memset(&H, 0, sizeof H);
    if (P->line || P->lin3) H.L = line_make(P);
    else if (P->fast) H.F = fast_make(P);
    else H.T = taps2_make(P);
    if (P->rad3) H.T3 = taps3_make(P);
    if (P->lin3) H.X = hex_make(P);
    return H;
}

static void hfilt_free(hfilt_t* H) {
    if (H->T.nt > 0) taps_free(&H->T);
    if (H->T3.gx) taps3_free(&H->T3);
}

static void hfilter(const mgbfo_params* P, const hfilt_t* H, pf_t* a, int adjoint) {
    if (P->rad3) {
        rad3_filter(P, &H->T, &H->T3, a, adjoint);
    } else if (P->lin3) {
        lin3_filter(P, &H->L, &H->X, a, adjoint);
    } else if (P->line) {
        line_filter(P, &H->L, a, adjoint);
    } else if (P->fast) {
        fast_filter(&H->F, a, adjoint);
    } else if (!adjoint) {
        boco_2d(a, P->hx, P->hy);
        vrbeta2(&H->T, a);
    } else {
        vrbeta2T(&H->T, a);
        bocoT_2d(a, P->hx, P->hy);
    }
}

void mgbfo_radial(const mgbfo_params* P, int K, int n, int m, const double* in, double* out,
                  int adjoint) {
    hfilt_t H = hfilt_make(P);
    pf_t a = pf_alloc(K, n, m, P->hx, P->hy);
    pf_from_interior(&a, in);
    hfilter(P, &H, &a, adjoint);
    pf_to_interior(&a, out);
    pf_free(&a);
    hfilt_free(&H);
}

/* ------------------------------------------------------------------------
 * Per-point aspect tensors: pasp2(:,:,ix,iy) of def_mg_weights
 * (mg_intstate.f90:383-390) with p_del and p_eps varying over the domain.
 * cholaspect2 per point (jp_pbfil.f90:90-107), getlinesum2 per point
 * (:182-218), and vrbeta2 / vrbeta2T with el(:,:,ix,iy) and ss(ix,iy)
 * (:888-928, 1011-1050), between boco_2d and bocoT_2d.
 * ---------------------------------------------------------------------- */
typedef struct {
    double exx, eyy, eyx, ss;
} pel_t;

/* el of one point, rescaled by rmom2_2, and its line sum; -1 if a reach
   exceeds the halo (getlinesum2 stops there). a3 = (A11, A22, A12). */
static int point_el(const mgbfo_params* P, const double a3[3], pel_t* e) {
    const double a[4] = {a3[0], a3[2], a3[2], a3[1]};
    double el[4];
    cholaspect2(a, el);
    const double rmom2_2 = 1.0 / sqrt((double)(2 * P->p) + 4.0);
    e->exx = el[0] * rmom2_2;
    e->eyy = el[3] * rmom2_2;
    e->eyx = el[1] * rmom2_2;
    const double y = 1.0 / e->eyy;
    const int gyl = (int)ceil(-y + EPS), gym = (int)floor(y - EPS);
    if (gyl < -P->hy || gym > P->hy) return -1;
    double s = 0.0;
    for (int gy = gyl; gy <= gym; ++gy) {
        const double xc = -gy * e->eyx, rrx = (gy * e->eyy) * (gy * e->eyy);
        const double x = sqrt(1.0 - rrx);
        const int gxl = (int)ceil((xc - x) / e->exx + EPS), gxm = (int)floor((xc + x) / e->exx - EPS);
        if (gxl < -P->hx || gxm > P->hx) return -1;
        for (int gx = gxl; gx <= gxm; ++gx) {
            const double t = gx * e->exx - xc;
            s += ipow(1.0 - (rrx + t * t), P->p);
        }
    }
    e->ss = 1.0 / s;
    return 0;
}

/* The taps of one point in vrbeta2's loop order: calls f(gx, gy, frow). */
#define PEL_TAPS(E, BODY)                                                              \
    do {                                                                               \
        const double y_ = 1.0 / (E)->eyy;                                              \
        for (int gy = (int)ceil(-y_ + EPS); gy <= (int)floor(y_ - EPS); ++gy) {        \
            const double xc_ = -gy * (E)->eyx, rrx_ = (gy * (E)->eyy) * (gy * (E)->eyy); \
            const double x_ = sqrt(1.0 - rrx_);                                        \
            const int gxl_ = (int)ceil((xc_ - x_) / (E)->exx + EPS);                   \
            const int gxm_ = (int)floor((xc_ + x_) / (E)->exx - EPS);                  \
            for (int gx = gxl_; gx <= gxm_; ++gx) {                                    \
                const double t_ = gx * (E)->exx - xc_;                                 \
                const double frow = (E)->ss * ipow(1.0 - (rrx_ + t_ * t_), P->p);      \
                BODY;                                                                  \
            }                                                                          \
        }                                                                              \
    } while (0)

int mgbfo_radial_aspect(const mgbfo_params* P, int K, int n, int m, const double* asp,
                        const double* in, double* out, int adjoint) {
    pel_t* E = malloc(sizeof(pel_t) * (size_t)n * m);
    for (int q = 0; q < n * m; ++q)
        if (point_el(P, asp + 3 * (size_t)q, &E[q]) != 0) {
            free(E);
            return -1;
        }
    pf_t a = pf_alloc(K, n, m, P->hx, P->hy);
    pf_t b = pf_alloc(K, n, m, P->hx, P->hy);
    pf_from_interior(&a, in);
    if (!adjoint) {
        boco_2d(&a, P->hx, P->hy);
#pragma omp parallel for schedule(static)
        for (int iy = 1; iy <= m; ++iy)
            for (int ix = 1; ix <= n; ++ix) {
                const pel_t* e = &E[(ix - 1) + (size_t)n * (iy - 1)];
                double* tb = PCOL(&b, ix, iy);
                PEL_TAPS(e, {
                    const double* src = PCOL(&a, ix + gx, iy + gy);
                    for (int k = 0; k < K; ++k) tb[k] += frow * src[k];
                });
            }
    } else {
        /* scatter, parallel over k: each k is independent */
#pragma omp parallel
        {
            int nth = 1, tid = 0;
#ifdef _OPENMP
            nth = omp_get_num_threads();
            tid = omp_get_thread_num();
#endif
            const int k0 = (int)((long)K * tid / nth), k1 = (int)((long)K * (tid + 1) / nth);
            for (int iy = 1; iy <= m; ++iy)
                for (int ix = 1; ix <= n; ++ix) {
                    const pel_t* e = &E[(ix - 1) + (size_t)n * (iy - 1)];
                    const double* ta = PCOL(&a, ix, iy);
                    PEL_TAPS(e, {
                        double* dst = PCOL(&b, ix + gx, iy + gy);
                        for (int k = k0; k < k1; ++k) dst[k] += frow * ta[k];
                    });
                }
        }
        bocoT_2d(&b, P->hx, P->hy);
    }
    pf_to_interior(&b, out);
    pf_free(&a);
    pf_free(&b);
    free(E);
    return 0;
}

/* The line sums ss(ix,iy) of a per-point aspect field (getlinesum2); -1 if
   a reach exceeds the halo. */
int mgbfo_ss2_aspect(const mgbfo_params* P, int n, int m, const double* asp, double* ss) {
    for (int q = 0; q < n * m; ++q) {
        pel_t e;
        if (point_el(P, asp + 3 * (size_t)q, &e) != 0) return -1;
        ss[q] = e.ss;
    }
    return 0;
}

void mgbfo_vertical(const mgbfo_params* P, int n, int m, const double* in, double* out,
                    int adjoint) {
    pf_t a = pf_alloc(P->K, n, m, 0, 0);
    pf_from_interior(&a, in);
    vertical_op(P, &a, adjoint);
    pf_to_interior(&a, out);
    pf_free(&a);
}

/* R = bocoT_2d(.,2,2) o adjoint (upsending_ens, mg_generations.f90:404-426),
   or with quad R2 = bocoT_2d(.,1,1) o adjoint2 (upsending2_ens,
   mg_generations.f90:497-547). */
static void restrict_pf(const pf_t* fine, pf_t* coarse_int /* (K,n/2,m/2,2,2) */, int quad) {
    if (quad) {
        adjoint2_op(fine, coarse_int);
        bocoT_2d(coarse_int, 1, 1);
    } else {
        adjoint_op(fine, coarse_int);
        bocoT_2d(coarse_int, 2, 2);
    }
}

/* P = direct1 o boco_2d(.,2,2) (downsending_ens, mg_generations.f90:461-491),
   or with quad P2 = direct2 o boco_2d(.,1,1) (downsending2_ens,
   mg_generations.f90:550-609). coarse has a halo of 2, fine none. */
static void prolong_pf(pf_t* coarse, pf_t* fine, int quad) {
    if (quad) {
        boco_2d(coarse, 1, 1);
        direct2_op(coarse, fine);
    } else {
        boco_2d(coarse, 2, 2);
        direct1_op(coarse, fine);
    }
}

static void restrict_any(int K, int n, int m, const double* fine, double* coarse, int quad) {
    pf_t F = pf_alloc(K, n, m, 0, 0);
    pf_from_interior(&F, fine);
    pf_t W = pf_alloc(K, n / 2, m / 2, 2, 2);
    restrict_pf(&F, &W, quad);
    pf_to_interior(&W, coarse);
    pf_free(&F);
    pf_free(&W);
}

static void prolong_any(int K, int n, int m, const double* coarse, double* fine, int quad) {
    pf_t W = pf_alloc(K, n / 2, m / 2, 2, 2);
    pf_from_interior(&W, coarse);
    pf_t F = pf_alloc(K, n, m, 0, 0);
    prolong_pf(&W, &F, quad);
    pf_to_interior(&F, fine);
    pf_free(&W);
    pf_free(&F);
}

void mgbfo_restrict(int K, int n, int m, const double* fine, double* coarse) {
    restrict_any(K, n, m, fine, coarse, 0);
}
void mgbfo_prolong(int K, int n, int m, const double* coarse, double* fine) {
    prolong_any(K, n, m, coarse, fine, 0);
}
void mgbfo_restrict2(int K, int n, int m, const double* fine, double* coarse) {
    restrict_any(K, n, m, fine, coarse, 1);
}
void mgbfo_prolong2(int K, int n, int m, const double* coarse, double* fine) {
    prolong_any(K, n, m, coarse, fine, 1);
}

void mgbfo_pad(int K, int n, int m, int h, const double* in, double* padded) {
    pf_t a = pf_alloc(K, n, m, h, h);
    pf_from_interior(&a, in);
    boco_2d(&a, h, h);
    memcpy(padded, a.d, sizeof(double) * (size_t)K * (n + 2 * h) * (size_t)(m + 2 * h));
    pf_free(&a);
}

void mgbfo_fold(int K, int n, int m, int h, const double* padded, double* out) {
    pf_t a = pf_alloc(K, n, m, h, h);
    memcpy(a.d, padded, sizeof(double) * (size_t)K * (n + 2 * h) * (size_t)(m + 2 * h));
    bocoT_2d(&a, h, h);
    pf_to_interior(&a, out);
    pf_free(&a);
}

/* ------------------------------------------------------------------------
 * The operator
 * ---------------------------------------------------------------------- */

size_t mgbfo_mg_size(const mgbfo_params* P) {
    size_t s = 0;
    for (int g = 0; g < P->gm; ++g) s += (size_t)P->K * (size_t)(P->nx >> g) * (size_t)(P->ny >> g);
    return s;
}

/* weighting_ens (mg_generations.f90:1012-1052): interior points only */
static void weight_pf(const mgbfo_params* P, int gi, pf_t* f, const double* wconst,
                      const double* const* wfield) {
    const int K = f->K, n = f->n, m = f->m;
    const double* wf = (wfield && wfield[gi]) ? wfield[gi] : NULL;
    const double wc = wconst ? wconst[gi] : 1.0;
    for (int j = 1; j <= m; ++j)
        for (int i = 1; i <= n; ++i) {
            double* c = PCOL(f, i, j);
            if (wf) {
                const double* w = wf + (size_t)K * ((size_t)(i - 1) + (size_t)n * (j - 1));
                for (int k = 0; k < K; ++k) c[k] = w[k] * c[k];
            } else {
                for (int k = 0; k < K; ++k) c[k] = wc * c[k];
            }
        }
}

/* Down-sweep shared by W A W^T and W^T: u[0] = Vf^T x, u[g] = R u[g-1]. */
static void upsweep_inputs(const mgbfo_params* P, const double* x, pf_t* G) {
    const int gm = P->gm;
    pf_from_interior(&G[0], x);
    if (P->l_vertical_filter && !P->rad3 && !P->lin3) vertical_op(P, &G[0], 1);
    for (int g = 1; g < gm; ++g) {
        /* adjoint reads the interior of the finer generation only */
        pf_t Fi = pf_alloc(P->K, G[g - 1].n, G[g - 1].m, 0, 0);
        pf_t Wi = pf_alloc(P->K, G[g].n, G[g].m, 2, 2);
        for (int j = 1; j <= Fi.m; ++j)
            for (int i = 1; i <= Fi.n; ++i)
                memcpy(PCOL(&Fi, i, j), PCOL(&G[g - 1], i, j), sizeof(double) * P->K);
        restrict_pf(&Fi, &Wi, P->lquart);
        for (int j = 1; j <= Wi.m; ++j)
            for (int i = 1; i <= Wi.n; ++i)
                memcpy(PCOL(&G[g], i, j), PCOL(&Wi, i, j), sizeof(double) * P->K);
        pf_free(&Fi);
        pf_free(&Wi);
    }
}

/* Up-sweep: G[g-1] += P G[g] for g = gm-1 .. 1; then Vf on G[0] (Vf^T
   with final_vft, as filtering_lin2_ens ends, mg_filtering.f90:1303-1305). */
static void downsweep_outputs(const mgbfo_params* P, pf_t* G, double* y, int final_vft) {
    const int gm = P->gm;
    for (int g = gm - 1; g >= 1; --g) {
        pf_t Wi = pf_alloc(P->K, G[g].n, G[g].m, 2, 2);
        for (int j = 1; j <= Wi.m; ++j)
            for (int i = 1; i <= Wi.n; ++i)
                memcpy(PCOL(&Wi, i, j), PCOL(&G[g], i, j), sizeof(double) * P->K);
        pf_t Fp = pf_alloc(P->K, G[g - 1].n, G[g - 1].m, 0, 0);
        prolong_pf(&Wi, &Fp, P->lquart);
        for (int j = 1; j <= Fp.m; ++j)
            for (int i = 1; i <= Fp.n; ++i) vadd(PCOL(&G[g - 1], i, j), PCOL(&Fp, i, j), P->K);
        pf_free(&Wi);
        pf_free(&Fp);
    }
    if (P->l_vertical_filter && !P->rad3 && !P->lin3) vertical_op(P, &G[0], final_vft);
    pf_to_interior(&G[0], y);
}

static pf_t* gens_alloc(const mgbfo_params* P) {
    pf_t* G = malloc(sizeof(pf_t) * P->gm);
    for (int g = 0; g < P->gm; ++g) G[g] = pf_alloc(P->K, P->nx >> g, P->ny >> g, P->hx, P->hy);
    return G;
}
static void gens_free(const mgbfo_params* P, pf_t* G) {
    for (int g = 0; g < P->gm; ++g) pf_free(&G[g]);
    free(G);
}

void mgbfo_wawt(const mgbfo_params* P, const double* wconst, const double* const* wfield,
                const double* x, double* y) {
    /* filtering_rad2_ens, mg_filtering.f90:1000-1147 (and the _bckg, fast
       and lin2 procedures, which share the skeleton) */
    const int gm = P->gm;
    hfilt_t H = hfilt_make(P);
    pf_t* G = gens_alloc(P);
    upsweep_inputs(P, x, G);
    for (int g = 0; g < gm; ++g) {
        if (g == 0 && !P->l_filt_g1) continue;
        hfilter(P, &H, &G[g], 1);
        weight_pf(P, g, &G[g], wconst, wfield);
        hfilter(P, &H, &G[g], 0);
    }
    downsweep_outputs(P, G, y, P->final_vft);
    gens_free(P, G);
    hfilt_free(&H);
}

void mgbfo_wt(const mgbfo_params* P, const double* x, double* v) {
    const int gm = P->gm;
    hfilt_t H = hfilt_make(P);
    pf_t* G = gens_alloc(P);
    upsweep_inputs(P, x, G);
    size_t off = 0;
    for (int g = 0; g < gm; ++g) {
        if (!(g == 0 && !P->l_filt_g1)) hfilter(P, &H, &G[g], 1);
        pf_to_interior(&G[g], v + off);
        off += (size_t)P->K * G[g].n * G[g].m;
    }
    gens_free(P, G);
    hfilt_free(&H);
}

void mgbfo_w(const mgbfo_params* P, const double* v, double* y) {
    const int gm = P->gm;
    hfilt_t H = hfilt_make(P);
    pf_t* G = gens_alloc(P);
    size_t off = 0;
    for (int g = 0; g < gm; ++g) {
        pf_from_interior(&G[g], v + off);
        off += (size_t)P->K * G[g].n * G[g].m;
        if (!(g == 0 && !P->l_filt_g1)) hfilter(P, &H, &G[g], 0);
    }
    downsweep_outputs(P, G, y, 0);
    gens_free(P, G);
    hfilt_free(&H);
}

void mgbfo_weight(const mgbfo_params* P, const double* wconst, const double* const* wfield,
                  double* v) {
    size_t off = 0;
    for (int g = 0; g < P->gm; ++g) {
        const int n = P->nx >> g, m = P->ny >> g;
        const size_t sz = (size_t)P->K * n * m;
        if (!(g == 0 && !P->l_filt_g1)) {
            const double* wf = (wfield && wfield[g]) ? wfield[g] : NULL;
            const double wc = wconst ? wconst[g] : 1.0;
            for (size_t e = 0; e < sz; ++e) v[off + e] *= wf ? wf[e] : wc;
        }
        off += sz;
    }
}

/* ------------------------------------------------------------------------
 * analysis <-> filter transfer
 * ---------------------------------------------------------------------- */

/* Stencil width and halo of the horizontal schemes of anal_to_filt /
   filt_to_anal (mg_transfer.f90:339-464): 0 lsqr (4 points, ibm = 3),
   1 lin (2 points, ibm = 1), 2 quad (3 points, ibm = 2). */
static int anal_width(int scheme) { return scheme == 1 ? 2 : scheme == 2 ? 3 : 4; }
static int anal_halo(int scheme) { return scheme == 1 ? 1 : scheme == 2 ? 2 : 3; }

/* lsqr_mg_coef (mg_interpolate.f90:89-298) on one PE tile: analysis points
   at xa(n) = dxa/2 + dxa (n-1), filter points at xf(i) = dxf/2 + dxf (i-1)
   over a tile of length nloc_a (mg_parameter.f90:518-541). With i0 the first
   i in 1-ib .. im+ib-1 (ib = 6) with xa(n) < xf(i):
     lsqr: iref = i0-2; the 4 weights blend the quadratic interpolants through
           points iref..iref+2 and iref+1..iref+3 linearly between points
           iref+1 and iref+2 (:156-191);
     lin:  irefL = i0-1; the linear weights Lx0, Lx1 (:269-296);
     quad: irefq = i0-1; the 3-point Lagrange weights qx0..qx2 (:231-267).
   Returns -1 if no i0 is found. */
static int anal_coef(int nloc_a, int nloc_f, int scheme, int* iref, double* cx)
{
    const double len = (double)nloc_a;
    const double dxa = len / nloc_a, dxf = len / nloc_f;
    const int w = anal_width(scheme);
    for (int n = 1; n <= nloc_a; ++n) {
        const double x = 0.5 * dxa + dxa * (n - 1);
        int i0 = 0, found = 0;
        for (int i = 1 - 6; i <= nloc_f + 6 - 1; ++i)
            if (x < 0.5 * dxf + dxf * (i - 1)) { i0 = i; found = 1; break; }
        if (!found) return -1;
        double* c = cx + (size_t)w * (n - 1);
        if (scheme == 1) {
            const double z0 = 0.5 * dxf + dxf * (i0 - 2), z1 = 0.5 * dxf + dxf * (i0 - 1);
            iref[n - 1] = i0 - 1;
            c[0] = (z1 - x) / (z1 - z0);
            c[1] = (x - z0) / (z1 - z0);
            continue;
        }
        if (scheme == 2) {
            double z[3];
            for (int t = 0; t < 3; ++t) z[t] = 0.5 * dxf + dxf * (i0 - 2 + t);
            iref[n - 1] = i0 - 1;
            c[0] = (x - z[1]) * (x - z[2]) / ((z[0] - z[1]) * (z[0] - z[2]));
            c[1] = (x - z[0]) * (x - z[2]) / ((z[1] - z[0]) * (z[1] - z[2]));
            c[2] = (x - z[0]) * (x - z[1]) / ((z[2] - z[0]) * (z[2] - z[1]));
            continue;
        }
        double z[4];
        for (int t = 0; t < 4; ++t) z[t] = 0.5 * dxf + dxf * (i0 - 3 + t);
        /* left quadratic through z0,z1,z2 and right one through z1,z2,z3 */
        const double l0 = (x - z[1]) * (x - z[2]) / ((z[0] - z[1]) * (z[0] - z[2]));
        const double l1 = (x - z[0]) * (x - z[2]) / ((z[1] - z[0]) * (z[1] - z[2]));
        const double l2 = (x - z[0]) * (x - z[1]) / ((z[2] - z[0]) * (z[2] - z[1]));
        const double r0 = (x - z[2]) * (x - z[3]) / ((z[1] - z[2]) * (z[1] - z[3]));
        const double r1 = (x - z[1]) * (x - z[3]) / ((z[2] - z[1]) * (z[2] - z[3]));
        const double r2 = (x - z[1]) * (x - z[2]) / ((z[3] - z[1]) * (z[3] - z[2]));
        const double wl = (z[2] - x) / (z[2] - z[1]), wr = (x - z[1]) / (z[2] - z[1]);
        iref[n - 1] = i0 - 2;
        c[0] = wl * l0;
        c[1] = wl * l1 + wr * r0;
        c[2] = wl * l2 + wr * r1;
        c[3] = wr * r2;
    }
    return 0;
}

/* lwq_vertical_coef (mg_interpolate.f90:301-370) for nm analysis levels
   over im filter levels: filter level i at height x(i) = i-1, analysis level
   n at y(n) = (n-1)(im-1)/(nm-1) (the last one pinned to x(im)); interior n
   takes kref = floor(y)+1 and the 4-point lsqr-type weights on x(kref-1..
   kref+2), with the weight of x(0) folded onto x(2) when kref = 1 and that
   of x(im+1) onto x(im-1) when kref = im-1 (whole-sample mirror). n = 1 and
   n = nm copy levels 1 and im. */
void mgbfo_lwq_coef(int nm, int im, double* c, int* kref)
{
    const double dy = 1.0 * (im - 1) / (nm - 1);
    for (int n = 2; n <= nm - 1; ++n) {
        const double y = (n - 1) * dy;
        const int i = (int)(y + 1.0);
        const double x1 = i - 2, x2 = i - 1, x3 = i, x4 = i + 1;
        const double d1 = y - x1, d2 = y - x2, d3 = y - x3, d4 = y - x4;
        const double d13 = d1 * d3, d23 = 0.5 * d2 * d3, d24 = d2 * d4;
        double* cn = c + 4 * (n - 1);
        kref[n - 1] = i;
        cn[0] = -d23 * d3;
        cn[1] = (d13 + 0.5 * d24) * d3;
        cn[2] = -(0.5 * d13 + d24) * d2;
        cn[3] = d23 * d2;
        if (i == 1) { cn[2] += cn[0]; cn[0] = 0.0; }
        if (i == im - 1) { cn[1] += cn[3]; cn[3] = 0.0; }
    }
    kref[0] = 1;
    c[0] = 0.0; c[1] = 1.0; c[2] = 0.0; c[3] = 0.0;
    kref[nm - 1] = im;
    c[4 * (nm - 1) + 0] = 0.0; c[4 * (nm - 1) + 1] = 1.0;
    c[4 * (nm - 1) + 2] = 0.0; c[4 * (nm - 1) + 3] = 0.0;
}

/* One column: lm filter levels f -> la analysis levels o. vmap 0:
   l_vertical_direct_spec(2) (:595-631, :1006-1044), odd levels copy and even
   ones average; vmap 1: lwq_vertical_direct_spec (:511-553) with kref / c
   from mgbfo_lwq_coef. */
static void vert_direct(int vmap, int lm, int la, const int* kref, const double* c, const double* f,
                        double* o)
{
    if (vmap == 0) {
        int k = 1;
        for (int n = 1; n <= la; n += 2, ++k) o[n - 1] = f[k - 1];
        k = 1;
        for (int n = 2; n <= la - 1; n += 2, ++k) o[n - 1] = 0.5 * (f[k - 1] + f[k]);
        return;
    }
    for (int n = 2; n <= la - 1; ++n) {
        const int k = kref[n - 1];
        const double* cn = c + 4 * (n - 1);
        if (k == 1)
            o[n - 1] = cn[1] * f[k - 1] + cn[2] * f[k] + cn[3] * f[k + 1];
        else if (k == lm - 1)
            o[n - 1] = cn[0] * f[k - 2] + cn[1] * f[k - 1] + cn[2] * f[k];
        else if (k == lm)
            o[n - 1] = cn[1] * f[k - 1];
        else
            o[n - 1] = cn[0] * f[k - 2] + cn[1] * f[k - 1] + cn[2] * f[k] + cn[3] * f[k + 1];
    }
    o[0] = f[0];
    o[la - 1] = f[lm - 1];
}

/* The transpose of vert_direct: l_vertical_adjoint_spec(2) (:556-592,
   :963-1003) or lwq_vertical_adjoint_spec (:461-508). f is cleared first. */
static void vert_adjoint(int vmap, int lm, int la, const int* kref, const double* c, const double* o,
                         double* f)
{
    for (int k = 0; k < lm; ++k) f[k] = 0.0;
    if (vmap == 0) {
        int k = 1;
        for (int n = 2; n <= la - 1; n += 2, ++k) {
            f[k - 1] += 0.5 * o[n - 1];
            f[k] += 0.5 * o[n - 1];
        }
        k = 1;
        for (int n = 1; n <= la; n += 2, ++k) f[k - 1] += o[n - 1];
        return;
    }
    for (int n = 2; n <= la - 1; ++n) {
        const int k = kref[n - 1];
        const double* cn = c + 4 * (n - 1);
        const double w = o[n - 1];
        if (k == 1) {
            f[0] += cn[1] * w; f[1] += cn[2] * w; f[2] += cn[3] * w;
        } else if (k == lm - 1) {
            f[lm - 3] += cn[0] * w; f[lm - 2] += cn[1] * w; f[lm - 1] += cn[2] * w;
        } else if (k == lm) {
            f[k - 1] += cn[1] * w;
        } else {
            f[k - 2] += cn[0] * w; f[k - 1] += cn[1] * w; f[k] += cn[2] * w; f[k + 1] += cn[3] * w;
        }
    }
    f[0] += o[0];
    f[lm - 1] += o[la - 1];
}

/* The vertical map must stay inside each column and write every analysis
   level: the linear one needs an odd la with (la+1)/2 <= lm (an even la
   leaves level la unwritten), the lwq one lm >= 3 (its edge branches touch
   3 levels). */
static void anal_check(const mgbfo_params* P, const mgbfo_anal* A, int K2)
{
    const int lm = P->lm, la = A->lm_a;
    if (la < lm) abort();
    if (la != lm && K2 != 0) abort();   /* the maps run over the 3-D columns only */
    if (la != lm && A->vmap == 0 && ((la + 1) / 2 > lm || la % 2 == 0)) abort();
    if (la != lm && A->vmap == 1 && lm < 3) abort();
    if (A->scheme < 0 || A->scheme > 2 || A->vmap < 0 || A->vmap > 1) abort();
}

void mgbfo_filt_to_anal(const mgbfo_params* P, const mgbfo_anal* A, const double* vf, double* wa)
{
    const int K = P->K, nx = P->nx, ny = P->ny;
    const int im = nx / A->npx, jm = ny / A->npy, nm = A->na / A->npx, mm = A->ma / A->npy;
    const int lm = P->lm, la = A->lm_a, nv = P->nvcol;
    const int K2 = K - nv * lm, Ka = nv * la + K2;
    anal_check(P, A, K2);
    const int w = anal_width(A->scheme), XH = anal_halo(A->scheme);
    int* iref = (int*)malloc(sizeof(int) * nm);
    int* jref = (int*)malloc(sizeof(int) * mm);
    double* cx = (double*)malloc(sizeof(double) * w * nm);
    double* cy = (double*)malloc(sizeof(double) * w * mm);
    if (anal_coef(nm, im, A->scheme, iref, cx) || anal_coef(mm, jm, A->scheme, jref, cy)) abort();
    int* kref = (int*)malloc(sizeof(int) * la);
    double* cv = (double*)malloc(sizeof(double) * 4 * la);
    if (la != lm && A->vmap == 1) mgbfo_lwq_coef(la, lm, cv, kref);
    /* boco_2d: the whole-domain field with a half-sample mirror halo of XH */
    const int px = nx + 2 * XH, py = ny + 2 * XH;
    double* V = (double*)malloc(sizeof(double) * (size_t)K * px * py);
    mgbfo_pad(K, nx, ny, XH, vf, V);
    /* *_direct_offset on every PE tile, at K levels of the filter */
    double* W = (double*)malloc(sizeof(double) * (size_t)K * A->na * A->ma);
    double* VX = (double*)malloc(sizeof(double) * (size_t)K * nm * (jm + 2 * XH));
    for (int tj = 0; tj < A->npy; ++tj)
        for (int ti = 0; ti < A->npx; ++ti) {
            for (int j = 1 - XH; j <= jm + XH; ++j)
                for (int n = 1; n <= nm; ++n) {
                    double* vx = VX + (size_t)K * ((size_t)(j - 1 + XH) * nm + (n - 1));
                    for (int k = 0; k < K; ++k) vx[k] = 0.0;
                    for (int t = 0; t < w; ++t) {
                        const int gi = ti * im + iref[n - 1] + t;   /* 1-based */
                        const int gj = tj * jm + j;
                        const double* v = V + (size_t)K * ((size_t)(gj - 1 + XH) * px + (gi - 1 + XH));
                        for (int k = 0; k < K; ++k) vx[k] += cx[w * (n - 1) + t] * v[k];
                    }
                }
            for (int m = 1; m <= mm; ++m)
                for (int n = 1; n <= nm; ++n) {
                    const int gn = ti * nm + n, gm = tj * mm + m;
                    double* o = W + (size_t)K * ((size_t)(gm - 1) * A->na + (gn - 1));
                    for (int k = 0; k < K; ++k) o[k] = 0.0;
                    for (int t = 0; t < w; ++t) {
                        const int j = jref[m - 1] + t;
                        const double* vx = VX + (size_t)K * ((size_t)(j - 1 + XH) * nm + (n - 1));
                        for (int k = 0; k < K; ++k) o[k] += cy[w * (m - 1) + t] * vx[k];
                    }
                }
        }
    /* the vertical map over the nv columns (a copy when lm_a = lm) */
    const long long npts = (long long)A->na * A->ma;
    for (long long q = 0; q < npts; ++q) {
        const double* f = W + (size_t)K * q;
        double* o = wa + (size_t)Ka * q;
        if (la == lm) {
            for (int k = 0; k < K; ++k) o[k] = f[k];
            continue;
        }
        for (int e = 0; e < nv; ++e) vert_direct(A->vmap, lm, la, kref, cv, f + e * lm, o + e * la);
        for (int k = 0; k < K2; ++k) o[nv * la + k] = f[nv * lm + k];
    }
    free(VX); free(W); free(V); free(cv); free(kref); free(cy); free(cx); free(jref); free(iref);
}

void mgbfo_anal_to_filt(const mgbfo_params* P, const mgbfo_anal* A, const double* wa, double* vf)
{
    const int K = P->K, nx = P->nx, ny = P->ny;
    const int im = nx / A->npx, jm = ny / A->npy, nm = A->na / A->npx, mm = A->ma / A->npy;
    const int lm = P->lm, la = A->lm_a, nv = P->nvcol;
    const int K2 = K - nv * lm, Ka = nv * la + K2;
    anal_check(P, A, K2);
    const int w = anal_width(A->scheme), XH = anal_halo(A->scheme);
    int* iref = (int*)malloc(sizeof(int) * nm);
    int* jref = (int*)malloc(sizeof(int) * mm);
    double* cx = (double*)malloc(sizeof(double) * w * nm);
    double* cy = (double*)malloc(sizeof(double) * w * mm);
    if (anal_coef(nm, im, A->scheme, iref, cx) || anal_coef(mm, jm, A->scheme, jref, cy)) abort();
    int* kref = (int*)malloc(sizeof(int) * la);
    double* cv = (double*)malloc(sizeof(double) * 4 * la);
    if (la != lm && A->vmap == 1) mgbfo_lwq_coef(la, lm, cv, kref);
    /* the adjoint vertical map (a copy when lm_a = lm) */
    const long long npts = (long long)A->na * A->ma;
    double* W = (double*)calloc((size_t)K * npts, sizeof(double));
    for (long long q = 0; q < npts; ++q) {
        const double* o = wa + (size_t)Ka * q;
        double* f = W + (size_t)K * q;
        if (la == lm) {
            for (int k = 0; k < K; ++k) f[k] = o[k];
            continue;
        }
        for (int e = 0; e < nv; ++e) vert_adjoint(A->vmap, lm, la, kref, cv, o + e * la, f + e * lm);
        for (int k = 0; k < K2; ++k) f[nv * lm + k] = o[nv * la + k];
    }
    /* *_adjoint_offset on every PE tile into the padded field */
    const int px = nx + 2 * XH, py = ny + 2 * XH;
    double* V = (double*)calloc((size_t)K * px * py, sizeof(double));
    double* VX = (double*)malloc(sizeof(double) * (size_t)K * nm * (jm + 2 * XH));
    for (int tj = 0; tj < A->npy; ++tj)
        for (int ti = 0; ti < A->npx; ++ti) {
            for (size_t e = 0; e < (size_t)K * nm * (jm + 2 * XH); ++e) VX[e] = 0.0;
            for (int m = 1; m <= mm; ++m)
                for (int n = 1; n <= nm; ++n) {
                    const int gn = ti * nm + n, gm = tj * mm + m;
                    const double* o = W + (size_t)K * ((size_t)(gm - 1) * A->na + (gn - 1));
                    for (int t = 0; t < w; ++t) {
                        const int j = jref[m - 1] + t;
                        double* vx = VX + (size_t)K * ((size_t)(j - 1 + XH) * nm + (n - 1));
                        for (int k = 0; k < K; ++k) vx[k] += o[k] * cy[w * (m - 1) + t];
                    }
                }
            for (int n = 1; n <= nm; ++n)
                for (int j = 1 - XH; j <= jm + XH; ++j) {
                    const double* vx = VX + (size_t)K * ((size_t)(j - 1 + XH) * nm + (n - 1));
                    for (int t = 0; t < w; ++t) {
                        const int gi = ti * im + iref[n - 1] + t;
                        const int gj = tj * jm + j;
                        double* v = V + (size_t)K * ((size_t)(gj - 1 + XH) * px + (gi - 1 + XH));
                        for (int k = 0; k < K; ++k) v[k] += vx[k] * cx[w * (n - 1) + t];
                    }
                }
        }
    /* bocoT_2d: fold the halo back with the half-sample mirror */
    mgbfo_fold(K, nx, ny, XH, V, vf);
    free(VX); free(V); free(W); free(cv); free(kref); free(cy); free(cx); free(jref); free(iref);
}
