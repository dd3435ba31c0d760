// Host-side FP64 setup of the MGBF operator: derived sizes, filter taps,
// vertical band matrices and inter-generation transfer tables.
#pragma once

#include <string>
#include <vector>

#include "mgbf.h"

namespace mgbf {

constexpr int kMaxGen = 12;
constexpr int kMaxHalo = 10;      // largest radial half-width with a built kernel
constexpr int kRWidth = 8;        // restriction window per coarse index
constexpr int kPWidth = 4;        // prolongation window per fine index
constexpr int kMaxXferWidth = 16; // analysis-filter interpolation window

// One tap of a beta filter (offsets from the output point, FP64 weight).
struct Tap3 {
    int gx, gy, gz;
    double w;
};

struct Tap2 {
    int gx, gy;
    double w;
};

// 1-D transfer table along one axis of one generation pair: for every output
// index a window start and kRWidth (restriction) or kPWidth (prolongation)
// weights; entries beyond the structural nonzeros are 0.
struct AxisTable {
    int n_out = 0, width = 0;
    std::vector<int> start;
    std::vector<double> w;   // n_out * width
};

struct Setup {
    mgbf_config cfg{};
    mgbf_dims dims{};
    int nx[kMaxGen] = {}, ny[kMaxGen] = {};
    std::vector<Tap2> taps2;           // reference loop order
    std::vector<Tap3> taps3;           // 3-D filter of procedures 4 (vrbeta3 loop order) and 5
    bool zero_levels = false;          // the 3-D filter reads zeros beyond the column ends (5)
    std::vector<double> hex_f[3];      // procedure 5: the x, y, level line factors of taps3
    std::vector<int> taps1_gx;         // vertical taps
    std::vector<double> taps1_w;
    int H = 0;                          // max(hx, hy) used by the radial kernels
    double wsym[kMaxHalo + 1][kMaxHalo + 1] = {};  // w(|a|,|b|) as [b][a]
    bool exact_disc = false;            // taps = lattice disc a^2+b^2 < (H+1)^2
    bool separable = false;             // w(a,b) = sepx[a] sepy[b] (not a disc)
    double sepx[kMaxHalo + 1] = {}, sepy[kMaxHalo + 1] = {};
    std::vector<double> vf, vft;       // lm x (2hz+1) band of Vf and Vf^T
    std::vector<double> vvt;           // lm x (4hz+1) band of Vf Vf^T
    AxisTable rx[kMaxGen], ry[kMaxGen];   // restriction g -> g+1 (coarse rows)
    AxisTable px[kMaxGen], py[kMaxGen];   // prolongation g+1 -> g (fine rows)
    double wconst[kMaxGen] = {};
    // analysis <-> filter transfer: direct = filter -> analysis, adjoint =
    // its transpose; the horizontal tables serve both routine families
    bool anal_identity = true;         // nm = im, mm = jm and lm_a = lm
    AxisTable ax_d, ax_a, ay_d, ay_a;  // horizontal, per axis
    AxisTable kv_d, kv_a;              // vertical of the _all2 routines (lm_a > lm)
    AxisTable kw_d, kw_a;              // vertical of the _all routines (lm_a > lm)
    const char* all2_refusal = nullptr;  // why the _all2 routines cannot run, or null
    const char* all_refusal = nullptr;   // why the _all routines cannot run, or null
};

// Dense 1-D analysis <- filter interpolation matrix of one axis over the
// whole domain (na = npe*nloc_a rows, nf = npe*nloc_f columns, row-major).
// scheme: 0 lsqr (4 points), 1 lin (l_lin_horizontal), 2 quad
// (l_quad_horizontal).
bool anal_axis_matrix(int npe, int nloc_a, int nloc_f, int scheme, std::vector<double>& T);
// Why anal_to_filt_all2 / filt_to_anal_all2 (resp. the _all routines) cannot
// run with this configuration, or null when they can.
const char* transfer_all2_refusal(const mgbf_config& cfg);
const char* transfer_all_refusal(const mgbf_config& cfg);

// init_mg_parameter-style validation and derived sizes. Returns MGBF_OK or
// an error status with a message in *err.
int derive(const mgbf_config& cfg, mgbf_dims& dims, int* nx, int* ny, std::string* err);

// Full FP64 setup (def_offset_coef + def_mg_weights). Returns MGBF_OK or an
// error status with a message in *err.
int build_setup(const mgbf_config& cfg, Setup& s, std::string* err);

// Radial taps of the isotropic / diagonal-aspect beta filter (cholaspect2 +
// getlinesum2 + the vrbeta2 tap loop). Returns false if the reach exceeds
// the halo (the reference stops there).
bool radial_taps(const mgbf_config& cfg, std::vector<Tap2>& taps);
// The per-point weights of a generation of n x m points whose aspect tensor
// varies (pasp2 of def_mg_weights, mg_intstate.f90:383-390: (A11, A22, A12)
// of point (i, j) at 3*(i + n*j)): per point the (2H+1)^2 offsets of the box
// |gx|, |gy| <= H at (gy+H)(2H+1) + gx+H, ss (rrc)^p on the taps vrbeta2
// visits, 0 elsewhere (cholaspect2 + getlinesum2 per point). MGBF_OK, or an
// error status with the reference's message when a reach exceeds hx or hy.
int aspect_weights(const mgbf_config& cfg, int H, int n, int m, const double* pasp2, std::vector<float>& w,
                   std::string* err);
bool vertical_taps(const mgbf_config& cfg, std::vector<int>& gx, std::vector<double>& w);
// 2-D taps of the triad line filters of procedures 2 and 7 (mgbf_line), or
// the identity without line tables. Returns MGBF_OK or an error status.
int line_taps(const mgbf_config& cfg, std::vector<Tap2>& taps, std::string* err);
// Line decompositions A = sum_c w_c l_c l_c^T (w_c >= 0) of a diagonal aspect
// tensor, by color: the triad of tritform (a = A11, A22, A12; 3 lines) and
// the hexad of hextform (a = A11, A22, A33, A23, A31, A12; 7 colors, one
// null). False for a tensor with an off-diagonal entry or a non-positive
// diagonal one, which no namelist reaches.
bool line_triad(const double a[3], int l[3][2], double w[3]);
bool line_hexad(const double a[6], int l7[7][3], double w7[7]);
// 3-D taps of the hexad line filters of procedure 5 (mgbf_line): the x, y and
// level lines as a product; *zero_levels: zeros beyond the column ends.
// factors (3 vectors, optional): the x, y and level line filters whose
// product the taps are, indexed by |offset|.
int hex_taps(const mgbf_config& cfg, std::vector<Tap3>& taps, bool* zero_levels, std::string* err,
             std::vector<double>* factors = nullptr);

}  // namespace mgbf
