/*
 * mgbf.h -- C ABI of the MI355X-native MGBF localization operator.
 *
 * The operator is B = W A W^T with W = Vf . D . F (vertical 1-D beta filter,
 * multigrid prolongation chain, radial 2-D beta filter on every generation)
 * and A the per-generation weights. B^{1/2} = W A^{1/2}.
 *
 * The reference exposes this path as a Fortran module call surface over
 * module-global state (src.mgbf/RBETA_TEST.f90:34-126):
 *   mg_initialize()                 src.mgbf/mg_entrymod.f90:36-219
 *   anal_to_filt_all2()             src.mgbf/mg_transfer.f90:154-183
 *   filtering_procedure(mgbf_proc)  src.mgbf/mg_filtering.f90:67-120
 *   filt_to_anal_all2()             src.mgbf/mg_transfer.f90:186-203
 *   mg_finalize()                   src.mgbf/mg_entrymod.f90:222-273
 * The entry points below replace that lifecycle with an explicit context; the
 * Fortran module mg_entrymod of this package binds them with ISO_C_BINDING.
 *
 * Conventions:
 *   - every entry point returns MGBF_OK (0) or a negative status; the message
 *     of the last failure on the calling thread is mgbf_last_error();
 *   - fields are caller-owned device buffers of float (FP32), layout k fastest,
 *     then x (i), then y (j): element (k,i,j) at k + K*(i + nx*j), K = km_all,
 *     which is the memory order of the reference's WORKA/VALL(k,i,j);
 *   - a multigrid (control) vector is the concatenation of the gm generation
 *     fields, generation g of size K*(nx>>(g-1))*(ny>>(g-1));
 *   - the context owns its device workspace; one context per host thread or
 *     stream; calls on one context must not overlap;
 *   - `stream` is a hipStream_t (NULL = the null stream).
 */
#ifndef MGBF_H
#define MGBF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MGBF_OK 0
#define MGBF_ERR_ARG (-1)      /* bad pointer or argument */
#define MGBF_ERR_CONFIG (-2)   /* configuration rejected (the reference would `stop`) */
#define MGBF_ERR_HIP (-3)      /* HIP runtime failure */
#define MGBF_ERR_UNSUPPORTED (-4) /* option of the reference not built here */

/*
 * Every key of the namelist &PARAMETERS_MGBETA, with the reference's meaning
 * (src.mgbf/mg_parameter.f90:190-208). Logicals are ints (0/1).
 */
typedef struct mgbf_config {
    double mg_ampl01, mg_ampl02, mg_ampl03; /* aspect amplitudes: vertical, horizontal, 3-D */
    double mg_weig1, mg_weig2, mg_weig3, mg_weig4, mg_weig5; /* generation weights (g>=5 -> weig5) */
    int hx, hy, hz;  /* halo widths = maximum filter reach */
    int p;           /* beta exponent (integer in the reference) */
    int mgbf_line;   /* init_mg_line: the triad (procedures 2, 7) and hexad (procedure 5)
                        line filters; without it every line is the identity. Lines off
                        the grid axes (a non-diagonal aspect tensor, which def_mg_weights
                        never builds) are refused with MGBF_ERR_UNSUPPORTED */
    int mgbf_proc;   /* filtering procedure (filtering_procedure, mg_filtering.f90:67-120),
                        all built: 1 rad2_bckg, 2 lin2_bckg, 3 fast_bckg, 4 rad3, 5 lin3,
                        6 rad2_ens, 7 lin2_ens, 8 fast_ens. l_filt_g1 is read by 6-8 only.
                        Procedure 7 ends with Vf^T as the reference's does, so its W A W^T
                        is not symmetric; km2 > 0 is refused with 6 (the reference's
                        ensemble vertical filter strides the stack wrongly there,
                        mg_filtering.f90:2066-2069) */
    int n_ens;       /* ensemble members, stacked in k */
    int lm_a, lm;    /* analysis / filter vertical levels */
    int km2, km3;    /* 2-D and 3-D variables */
    int l_loc, l_loc_vertical, l_filt_g1, l_lin_vertical, l_lin_horizontal,
        l_quad_horizontal, l_new_map, l_vertical_filter, ldelta, lquart, lhelm;
    int gm_max;      /* requested number of generations */
    int nm0, mm0;    /* global analysis grid */
    int nxPE, nyPE;  /* reference PE grid: fixes the global grid and caps gm */
    int im_filt, jm_filt; /* filter tile size per PE */
} mgbf_config;

/* Derived sizes (init_mg_parameter, src.mgbf/mg_parameter.f90:179-575). */
typedef struct mgbf_dims {
    int gm;          /* generations: min(def_maxgen(nxPE,nyPE), gm_max) */
    int nx, ny;      /* global generation-1 grid: nxPE*im_filt, nyPE*jm_filt */
    int km;          /* km2 + lm*km3 */
    int km_all;      /* K: stacked k extent = km*n_ens */
    int km3_all;     /* vertical columns in the stack */
    int ntaps2;      /* taps of the radial 2-D filter */
    int ntaps1;      /* taps of the vertical 1-D filter */
    long long field_size; /* K*nx*ny */
    long long mg_size;    /* elements of a multigrid vector */
    int na, ma;           /* global analysis grid: nm0 x mm0 */
    int km_a_all;         /* stacked k extent on the analysis grid: n_ens*(km2 + km3*lm_a) */
    long long anal_size;  /* km_a_all*na*ma (WORKA) */
    int ntaps3;           /* taps of the 3-D filter (procedure 4), else 0 */
} mgbf_dims;

typedef struct mgbf_ctx mgbf_ctx;

/* operations of mgbf_apply */
enum {
    MGBF_OP_WAWT = 0,  /* y = W A W^T x   (filtering_procedure(mgbf_proc), mg_filtering.f90:67-120:
                          procedures 1-8): field -> field */
    MGBF_OP_WT = 1,    /* v = W^T x       : field -> multigrid vector */
    MGBF_OP_W = 2,     /* y = W v         : multigrid vector -> field */
    MGBF_OP_A = 3,     /* v = A v         : multigrid -> multigrid */
    MGBF_OP_ASQRT = 4, /* v = A^{1/2} v   : multigrid -> multigrid */
    MGBF_OP_BSQRT = 5, /* y = W A^{1/2} v : multigrid vector -> field */
    MGBF_OP_BSQRTT = 6, /* v = A^{1/2} W^T x : field -> multigrid vector */
    MGBF_OP_A2F = 7,   /* anal_to_filt_all2 (mg_transfer.f90:154-183): analysis field
                          (km_a_all x na x ma) -> generation-1 field; the transpose of F2A */
    MGBF_OP_F2A = 8,   /* filt_to_anal_all2 (mg_transfer.f90:186-203): generation-1 field ->
                          analysis field: the horizontal interpolation of filt_to_anal
                          (:406-464; lsqr_direct_offset mg_interpolate.f90:634-681, or
                          lin_direct_offset :858-903 with l_lin_horizontal, or
                          quad_direct_offset :750-795 with l_quad_horizontal), then lm -> lm_a
                          levels (l_vertical_direct_spec2, :1006-1044; odd lm_a <= 2 lm - 1,
                          km2 = 0); a copy when the grids agree */
    MGBF_OP_A2F_ALL = 9,  /* anal_to_filt_all (mg_transfer.f90:58-110): the transpose of
                             F2A_ALL */
    MGBF_OP_F2A_ALL = 10  /* filt_to_anal_all (mg_transfer.f90:113-151): the horizontal
                             interpolation of F2A, then lm -> lm_a levels of the 3-D variables
                             (S2C_ens / C2S_ens, km2 = 0) with l_vertical_direct_spec
                             (:595-631) under l_lin_vertical, else lwq_vertical_direct_spec
                             (:511-553, weights of lwq_vertical_coef :301-370) */
};

/* single stages of mgbf_apply_stage (gen is 1-based) */
enum {
    MGBF_STAGE_VF = 0,  /* Vf   on a generation-1 field (sup_vrbeta1_ens) */
    MGBF_STAGE_VFT = 1, /* Vf^T on a generation-1 field (sup_vrbeta1T_ens) */
    MGBF_STAGE_F = 2,   /* F  on generation gen: boco_2d + rbeta (vrbeta2; procedure 4:
                           sup_vrbeta3, procedure 5: the hexad lines, for the 3-D variables) */
    MGBF_STAGE_FT = 3,  /* F^T on generation gen: rbetaT (vrbeta2T) + bocoT_2d (procedures 4
                           and 5: their transposes) */
    MGBF_STAGE_R = 4,   /* restriction gen -> gen+1: adjoint + bocoT_2d(.,2,2)
                           (lquart: adjoint2 + bocoT_2d(.,1,1)) */
    MGBF_STAGE_P = 5,   /* prolongation gen+1 -> gen: boco_2d(.,2,2) + direct1
                           (lquart: boco_2d(.,1,1) + direct2) */
    MGBF_STAGE_A = 6    /* weighting of generation gen (weighting_ens) */
};

/* Fills *cfg with the settings of a 4-generation ensemble run (mgbf_proc=6,
   radial half-width 6, vertical half-width 3, p=2, reach equal to the halo). */
int mgbf_config_default(mgbf_config* cfg);

/* Validates cfg and derives the sizes, as init_mg_parameter does
   (src.mgbf/mg_parameter.f90:179-575, def_maxgen :578-611); no device use. */
int mgbf_derive(const mgbf_config* cfg, mgbf_dims* dims);

/* mg_initialize (src.mgbf/mg_entrymod.f90:36-219) without MPI: validates,
   builds the FP64 coefficients (def_offset_coef, def_mg_weights ->
   cholaspect/getlinesum, src.mgbf/mg_intstate.f90:296-433), uploads the
   FP32 tables on `device`. */
int mgbf_init(const mgbf_config* cfg, int device, mgbf_ctx** ctx);

/* mg_finalize (src.mgbf/mg_entrymod.f90:222-273): frees the context. */
int mgbf_finalize(mgbf_ctx* ctx);

int mgbf_get_dims(const mgbf_ctx* ctx, mgbf_dims* dims);

/* Replaces the weight of generation gen (1-based): a_diff_f (gen 1) or
   a_diff_h (gen >= 2) of src.mgbf/mg_intstate.f90:333-351. `field` is a
   caller-owned device array of K*n_g*m_g floats that must stay valid while
   the context uses it, or NULL to use the constant `value`. */
int mgbf_set_weights(mgbf_ctx* ctx, int gen, const float* field, double value);

/* Per-point aspect tensors of generation gen (1-based): the pasp2(:,:,i,j)
   that def_mg_weights builds from the p_del and p_eps fields
   (src.mgbf/mg_intstate.f90:383-390), which the reference's filters read
   per point (vrbeta2 / vrbeta2T with el(:,:,ix,iy) and ss(ix,iy),
   src.mgbf/jp_pbfil.f90:888-928, 1011-1050; getlinesum2 :182-218).
   pasp2 is a host array of n_g*m_g*3 doubles, (A11, A22, A12) of point
   (i, j) at 3*(i + n_g*j), or NULL to return to the namelist's
   mg_ampl02 * I. The generation's F then gathers with every point's own
   ellipse and normalizer, and F^T (no longer F) is the transposed gather of
   the scatter and fold. Built for the radial procedures 1 and 6
   (MGBF_ERR_UNSUPPORTED otherwise); a reach beyond the halo is refused with
   the reference's message. The weights of each point's taps are kept on the
   device: (2H+1)^2 floats per point of the generation. */
int mgbf_set_aspect(mgbf_ctx* ctx, int gen, const double* pasp2);

/* Applies an operator (MGBF_OP_*). For WAWT, `in` and `out` may alias. The
   context's workspaces are allocated by the first call that needs them. */
int mgbf_apply(mgbf_ctx* ctx, int op, const float* in, float* out, void* stream);

/* mgbf_apply for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc). No
   counterpart in the reference. The call allocates nothing: instead of the
   context's workspaces (which mgbf_apply allocates on first use) it uses a
   static device array of MGBF_NOALLOC_SCRATCH_FLOATS floats (128 MiB), one
   per device, held from the library's load until the process ends. A
   context whose workspace (2 multigrid vectors and, for analysis transfers
   whose vertical map runs as a pass of its own, one analysis-grid field of
   K levels) exceeds it is refused with MGBF_ERR_ARG. Every call on a device
   shares that device's array and leaves its contents behind: calls that
   overlap on one device (any contexts, any streams) race. The timers record
   nothing during the call. Results equal mgbf_apply's bit for bit. */
#define MGBF_NOALLOC_SCRATCH_FLOATS (32ll << 20)
int mgbf_apply_noalloc(mgbf_ctx* ctx, int op, const float* in, float* out, void* stream);

/* Applies one stage (MGBF_STAGE_*) on generation gen; in and out must not alias. */
int mgbf_apply_stage(mgbf_ctx* ctx, int stage, int gen, const float* in, float* out,
                     void* stream);

/* FP64-accumulated inner product of two device float arrays of n elements;
   *result is written on the host after the stream has finished the sum. */
int mgbf_dot(mgbf_ctx* ctx, const float* a, const float* b, long long n, double* result,
             void* stream);

/* A group of contexts (regional domains of uneven size, each with its own
   grid, generations and filter) applied as one hipGraph. The reference has
   no counterpart: it runs one domain per launch of the program. */
typedef struct mgbf_group mgbf_group;
/* Creates a group over n distinct contexts of one device (the contexts must
   outlive the group and not be used elsewhere while it runs). */
int mgbf_group_create(mgbf_ctx* const* ctxs, int n, mgbf_group** grp);
/* Applies `op` (MGBF_OP_*) to every context i, in[i] -> out[i], as one graph
   launch on `stream`. The graph is captured on the first call with a given
   op and set of buffers and replayed while they stay the same. */
int mgbf_group_apply(mgbf_group* grp, int op, const float* const* in, float* const* out,
                     void* stream);
int mgbf_group_destroy(mgbf_group* grp);

/* Per-phase device timers, the phases of the reference's mg_timers
   (src.mgbf/mg_timers.f90:17-60, filtering_rad2_ens btim/etim brackets):
   hipEvents recorded on the caller's stream around every launch of a phase,
   with the algorithmic bytes (reads + writes) of each launch. The weighting
   (weight_tim) is fused into the adjoint radial filter (hfiltT). */
enum {
    MGBF_T_VFILTT = 0, /* Vf^T */
    MGBF_T_UPSEND = 1, /* restriction chain */
    MGBF_T_HFILTT = 2, /* F^T with the weighting A */
    MGBF_T_HFILT = 3,  /* F */
    MGBF_T_DNSEND = 4, /* prolongation and accumulation */
    MGBF_T_VFILT = 5,  /* Vf */
    MGBF_NTIMERS = 6
};
/* Turns the timers on (1) or off (0) and clears the accumulators. */
int mgbf_timers_enable(mgbf_ctx* ctx, int enable);
/* Waits for the recorded events and returns, per phase, the accumulated
   milliseconds, number of launches and algorithmic bytes (arrays of
   MGBF_NTIMERS entries; any may be NULL). */
int mgbf_timers_read(mgbf_ctx* ctx, double* ms, long long* launches, double* bytes);
/* The same accumulators restricted to the launches that act on generation
   gen (1..gm): a phase's own time, launches and bytes per generation. No
   counterpart in the reference (mg_timers sums a phase over generations). */
int mgbf_timers_read_gen(mgbf_ctx* ctx, int gen, double* ms, long long* launches, double* bytes);

/* Times the radial filter on this context's shapes for every candidate tile
   (output columns per lane x output rows per wavefront), for each generation
   and each pass kind (0 plain, 1 with the prolongation, 2 with the
   prolongation and the vertical band, 3 plain with the restriction to the
   next generation), and keeps the fastest. The candidate heights are a fixed
   ladder and the heights that fit the launch into whole rounds of the
   chip's waves. For generation 1 it also times the vertical band as a pass
   of its own after a kind-1 pass, and keeps that form when it beats kind 2.
   Uses the context's workspace; synchronizes `stream`. No counterpart in
   the reference, whose loops have no tiles. */
int mgbf_autotune(mgbf_ctx* ctx, void* stream);
/* The tile in use for generation gen (1-based) and pass kind: *ti is the
   output columns per lane of the vector kernel (2 or 4), or 16 when the pass
   runs on the matrix cores (16 output columns x 64 k per block of 4
   wavefronts), or 0 for generation 1's kind 2 when the vertical band runs as
   a pass of its own; *tj is the output rows per wavefront or block. */
int mgbf_get_tile(const mgbf_ctx* ctx, int gen, int kind, int* ti, int* tj);
/* Pins the tile of generation gen (1-based) and pass kind, in the terms of
   mgbf_get_tile (ti = 2 or 4, or 16 for the matrix cores, or 0 for
   generation 1's kind 2: the vertical band as a pass of its own; tj >= 1,
   even for kind 3), so that tiles found by mgbf_autotune in one process can
   be reused in another without timing again. The matrix cores take every
   pass kind of a generation whose width is a multiple of 16, with K a
   multiple of 64 and a disc or product filter of half-width <= 8, with
   either transfer stencil (kind 2 also: half-width < 8; kind 3 also: at
   least 32 columns, half-width >= 3, exact FP16 restriction weights); else
   MGBF_ERR_UNSUPPORTED. No counterpart in the reference. */
int mgbf_set_tile(mgbf_ctx* ctx, int gen, int kind, int ti, int tj);

/* Message of the last failure on this thread ("" if none). */
const char* mgbf_last_error(void);

/* Library version string. */
const char* mgbf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MGBF_H */
