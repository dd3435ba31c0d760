/*
 * FP64 CPU restatement of the MGBF localization operator (test infrastructure).
 *
 * This is the oracle: the checker that the HIP product path is compared with.
 * It is never linked into the product library. It follows the reference's own
 * formulation step by step (padded halo arrays, scatter-form adjoints, the
 * explicit halo fill and fold order of the reference's boundary routines) on a
 * single global domain. Field layout: k fastest, then x (i), then y (j), as
 * the reference's VALL(k,i,j).
 */
#ifndef MGBF_ORACLE_H
#define MGBF_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mgbfo_params {
    int nx, ny;            /* global generation-1 grid */
    int K;                 /* stacked k extent (km_all) */
    int lm;                /* levels per vertical column */
    int nvcol;             /* km3_all: vertical columns in the stack; the first nvcol*lm
                              entries are vertically filtered, the rest pass through */
    int hx, hy, hz;        /* halo widths = maximum filter reach */
    int p;                 /* beta exponent (integer) */
    int gm;                /* number of generations */
    double ampl01, ampl02; /* aspect amplitudes: vertical (pasp01), horizontal (pasp02) */
    double p_del, p_eps;   /* horizontal aspect perturbations (the reference fixes them at 0) */
    int l_filt_g1;         /* filter and weight generation 1 */
    int l_vertical_filter; /* apply the vertical 1-D filter */
    int fast;              /* filtering_fast_* (procedures 3, 8): the horizontal filter is the
                              product of 1-D beta filters along x and y */
    int lquart;            /* quadratic inter-generation transfer (adjoint2 / direct2 with a
                              1-point halo) instead of the 4-point one */
    int line;              /* filtering_lin2_* (procedures 2, 7): the horizontal filter is
                              the triad of dibeta line filters */
    int line_tables;       /* mgbf_line: init_mg_line has set the line tables (without them
                              every line generator is 0 and each line is the identity) */
    int final_vft;         /* W A W^T ends with Vf^T instead of Vf (filtering_lin2_ens) */
    int rad3;              /* filtering_rad3 (procedure 4): the first nvcol*lm entries of the
                              stack are nvcol 3-D variables filtered over (i, j, L), the rest
                              2-D variables; no separate vertical filter */
    double ampl03;         /* 3-D aspect amplitude (pasp03 = mg_ampl03) */
    int lin3;              /* filtering_lin3 (procedure 5): the 3-D variables through the hexad
                              line filters over (i, j, L), the 2-D ones through the triad lines */
} mgbfo_params;

/* ---- coefficient generation (jp_pbfil.f90 cholaspect/getlinesum) ---- */
/* Returns the number of taps; fills up to maxt entries (gx, gy, w) of the
   normalized 2-D tap table. Returns -1 if the reach exceeds (hx,hy). */
int mgbfo_taps2(const mgbfo_params* P, int maxt, int* gx, int* gy, double* w);
/* Same for the 1-D vertical filter (gx only). */
int mgbfo_taps1(const mgbfo_params* P, int maxt, int* gx, double* w);
/* The 1-D factor of the fast filter along x (halo hx): paspx = ampl02 enters
   getlinesum1 / vrbeta1 without cholaspect (mg_intstate.f90:370-375, 427). */
int mgbfo_taps_fast(const mgbfo_params* P, int maxt, int* gx, double* w);
/* Line-sum normalizers ss = 1/sum(rrc^p) (getlinesum1/2). */
double mgbfo_ss1(const mgbfo_params* P);
double mgbfo_ss2(const mgbfo_params* P);

/* ---- single stages on interior-only arrays (K, n, m), k fastest ---- */
/* Radial 2-D filter F (adjoint=0: halo fill + gather) or its adjoint
   (adjoint=1: scatter + halo fold), on a generation of size n x m. */
void mgbfo_radial(const mgbfo_params* P, int K, int n, int m,
                  const double* in, double* out, int adjoint);
/* F (adjoint=0) or F^T (adjoint=1) on a generation of size n x m whose
   aspect tensor varies per point: asp holds (A11, A22, A12) of point (i, j)
   (0-based) at 3*(i + n*j). Returns 0, or -1 if a reach exceeds the halo. */
int mgbfo_radial_aspect(const mgbfo_params* P, int K, int n, int m, const double* asp,
                        const double* in, double* out, int adjoint);
/* The per-point line sums ss = 1/sum(rrc^p) of such a field (getlinesum2). */
int mgbfo_ss2_aspect(const mgbfo_params* P, int n, int m, const double* asp, double* ss);
/* Vertical filter Vf / Vf^T on an (K, n, m) field. */
void mgbfo_vertical(const mgbfo_params* P, int n, int m,
                    const double* in, double* out, int adjoint);
/* Restriction R = Fold_2 o P^T: fine (K, n, m) -> coarse (K, n/2, m/2). */
void mgbfo_restrict(int K, int n, int m, const double* fine, double* coarse);
/* Prolongation P = interp o Pad_2: coarse (K, n/2, m/2) -> fine (K, n, m). */
void mgbfo_prolong(int K, int n, int m, const double* coarse, double* fine);
/* The quadratic pair (lquart): R2 = Fold_1 o adjoint2, P2 = direct2 o Pad_1. */
void mgbfo_restrict2(int K, int n, int m, const double* fine, double* coarse);
void mgbfo_prolong2(int K, int n, int m, const double* coarse, double* fine);
/* Halo fill (boco_2d, half-sample mirror) of an interior array into a padded
   array (K, n+2h, m+2h); and its adjoint fold back to interior. */
void mgbfo_pad(int K, int n, int m, int h, const double* in, double* padded);
void mgbfo_fold(int K, int n, int m, int h, const double* padded, double* out);

/* Taps of the 3-D filter (gx, gy, gz, w) in vrbeta3's order; returns the count
   (fills up to maxt). Aborts beyond a halo. */
int mgbfo_taps3(const mgbfo_params* P, int maxt, int* gx, int* gy, int* gz, double* w);

/* hexad: aspect (A11, A22, A33, A23, A31, A12) -> generators lhex7[3c..3c+2] and
   weights whex7[c] of colors c+1 = 1..7. Nonzero when it does not converge. */
int mgbfo_hexad(const double aspect[6], int lhex7[21], double whex7[7]);
/* The six hexad lines of the domain's 3-D variables: colors qcol[6], generators
   g[18], half-spans h[6], half filters fil[6][9]. Returns the number of lines. */
int mgbfo_hexlines(const mgbfo_params* P, int* qcol, int* g, int* h, double* fil);

/* ---- triad line filters (jp_pbfil3.f90) ---- */
/* inimomtab: bnorm[1..nh], bsprds[1..nh] for exponent p. Nonzero on bad p/nh. */
int mgbfo_momtab(int p, int nh, double* bnorm, double* bsprds);
/* hstform with the nh = 6 table of exponent p: spread**2 -> half-span. */
int mgbfo_hstform(int p, double* hs);
/* blinfil: half dibeta filter fil[0..*h] of half-span hspan (nfil = 8). */
int mgbfo_blinfil(int p, double hspan, int* h, double* fil);
/* triad: aspect (A11, A22, A12) -> line generators ltri[2c..2c+1], weights wtri[c]. */
int mgbfo_triad(const double aspect[3], int ltri[6], double wtri[3]);
/* Line c (0..2) of the domain's line filter: generator, half-span (-1: identity), fil[0..h]. */
int mgbfo_line(const mgbfo_params* P, int c, int* dx, int* dy, int* h, double* fil);

/* ---- the operator ----
   Generation g (0-based gi = g-1) has size (nx>>gi, ny>>gi). The multigrid
   control vector v is the concatenation of the gm generation fields.
   wconst[gi] is the constant weight of generation g; wfield[gi], if wfield
   and wfield[gi] are non-NULL, is a per-point weight field that replaces it. */
size_t mgbfo_mg_size(const mgbfo_params* P);
/* y = W A W^T x (filtering_rad2_ens). */
void mgbfo_wawt(const mgbfo_params* P, const double* wconst, const double* const* wfield,
                const double* x, double* y);
/* v = W^T x: u_1 = Vf^T x, u_{g+1} = R u_g, v_g = F^T u_g (v_1 = u_1 if !l_filt_g1). */
void mgbfo_wt(const mgbfo_params* P, const double* x, double* v);
/* y = W v: s_gm = F v_gm, s_g = F v_g + P s_{g+1}, y = Vf s_1. */
void mgbfo_w(const mgbfo_params* P, const double* v, double* y);
/* Scales generation g of a multigrid vector by its weight (A). */
void mgbfo_weight(const mgbfo_params* P, const double* wconst, const double* const* wfield,
                  double* v);

/* Number of OpenMP threads the oracle uses (1 without OpenMP). */
int mgbfo_num_threads(void);

/* ---- analysis <-> filter transfer (mg_transfer.f90:58-203, 339-464) ---- */
/* The analysis grid: na x ma global points (nm0 x mm0) split over an
   npx x npy PE grid like the filter grid (nx = npx*im, ny = npy*jm), and
   lm_a >= lm levels per column. scheme: the horizontal interpolation of
   anal_to_filt / filt_to_anal, 0 lsqr (4 points, halo 3), 1 lin
   (l_lin_horizontal, 2 points, halo 1), 2 quad (l_quad_horizontal, 3 points,
   halo 2). vmap: the vertical map for lm_a > lm, 0 the linear one
   (l_vertical_*_spec2 of the _all2 routines; l_vertical_*_spec of the _all
   routines with l_lin_vertical), 1 lwq_vertical_*_spec (the _all routines
   without l_lin_vertical). */
typedef struct mgbfo_anal {
    int na, ma;
    int npx, npy;
    int lm_a;
    int scheme;
    int vmap;
} mgbfo_anal;
/* filt_to_anal: boco_2d halo, the *_direct_offset interpolation per PE tile
   (mg_interpolate.f90:634-681, 750-795, 858-903, coefficients of
   lsqr_mg_coef :89-298), then the vertical map (:511-553, 595-631,
   1006-1044). vf: K x nx x ny; wa: Ka x na x ma with Ka = nvcol*lm_a +
   (K - nvcol*lm). With lm_a > lm the columns must cover the stack. */
void mgbfo_filt_to_anal(const mgbfo_params* P, const mgbfo_anal* A, const double* vf, double* wa);
/* anal_to_filt: the adjoint vertical map (:461-508, 556-592, 963-1003), the
   *_adjoint_offset interpolation per PE tile (:684-747, 798-855, 906-960)
   into the halo-padded field, then bocoT_2d. */
void mgbfo_anal_to_filt(const mgbfo_params* P, const mgbfo_anal* A, const double* wa, double* vf);
/* lwq_vertical_coef (mg_interpolate.f90:301-370): nm analysis levels over im
   filter levels; c holds 4 weights per analysis level, kref the filter
   level (1-based) of the second weight. */
void mgbfo_lwq_coef(int nm, int im, double* c, int* kref);

#ifdef __cplusplus
}
#endif
#endif
