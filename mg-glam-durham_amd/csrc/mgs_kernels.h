// mgs_kernels.h -- host-side launchers of the libmgsolve device kernels.
// Every launcher enqueues on the given stream and returns immediately.
#pragma once

#include "mgs_common.h"

namespace mgs {

// ---- layout conversion -------------------------------------------------
// lexicographic (x fastest) <-> colour-split, with element type conversion
void launch_pack_f64_from_f64(const double* lex, double* split, int n, hipStream_t s);
void launch_pack_f64_from_f32(const float* lex, double* split, int n, hipStream_t s);
void launch_unpack_f64_to_f64(const double* split, double* lex, int n, double scale, hipStream_t s);
void launch_unpack_f64_to_f32(const double* split, float* lex, int n, double scale, hipStream_t s);
void launch_fill_f64(double* p, size_t count, double v, hipStream_t s);
void launch_fill_f32(float* p, size_t count, float v, hipStream_t s);
void launch_copy_f64(const double* src, double* dst, size_t count, hipStream_t s);
void launch_copy_f32(const float* src, float* dst, size_t count, hipStream_t s);

// ---- plain FP64 f(R) path (the reference's operators in FP64) ----------
void launch_fr_smooth64(int frn, int level, double* u, const double* rho, const double* f, int n, int colour,
                        const FrCoef& c, hipStream_t s);
// residual at level (0: with delta, >0: minus f); res may be null (norm only);
// partials receives one double per block; returns the number of blocks
int launch_fr_residual64(int frn, int level, const double* u, const double* rho, const double* f, double* res,
                         double* partials, int n, const FrCoef& c, hipStream_t s);
// ru1/u1 non-null: the same pass also stores R u (restrict_scalar_field) there
void launch_fr_restrict_residual1_64(int frn, const double* u0, const double* delta, double* rres1, double* ru1,
                                     double* u1, int N, const FrCoef& c0, hipStream_t s);
void launch_restrict8_64(const double* fine, double* out1, double* out2, int nc, hipStream_t s);
void launch_fr_rhs64(int frn, const double* ru, const double* rho, const double* rres, double* f, int n,
                     const FrCoef& c, hipStream_t s);
// only1: colour-1 fine cells only, for a correction that a colour-0 f(R)
// relaxation pass follows (the colour-0 values would be overwritten unread)
void launch_correct64(const double* uH, const double* ru, double* ufine, int nc, int guard, bool only1,
                      hipStream_t s);

// deterministic sum of `count` partials into out[0] (single block)
void launch_sum_partials(const double* partials, int count, double* out, hipStream_t s);

// ---- mixed precision f(R) path: FP64 outer loop, FP32 delta-form FAS ----
// FP64 residual of u (partials for its norm), g = -r and xb = u/u_bg in FP32,
// and the inner increment d after its first colour-0 pass from zero;
// norm_only: the partials alone (the same values)
int launch_outer_residual(int frn, const double* u, const double* delta, float* g, float* xb, float* d,
                          double* partials, int n, const FrCoef& c, const DfCoef& dc, bool norm_only, hipStream_t s);
// the last colour-1 inner pass fused with u += u_bg (x(d) - xb) on all cells
void launch_df_smooth_update(int frn, const float* d, const float* xb, const float* rhs, double* u, int n,
                             const DfCoef& c, double ubg, hipStream_t s);
void launch_restrict_base(int frn, const float* xb_f, float* xb_c, int nc, hipStream_t s);
void launch_df_smooth(int frn, float* d, const float* xb, const float* rhs, int n, int colour, const DfCoef& c,
                      hipStream_t s);
void launch_df_residual(int frn, const float* d, const float* xb, const float* rhs, float* res, int n,
                        const DfCoef& c, hipStream_t s);
// residual of the fine level (N cells a side, right-hand side g) restricted
// to the next level in one pass, with R d into rd (frozen copy) and dc
// (start value of the coarse increment)
void launch_df_restrict_residual(int frn, const float* d, const float* xb, const float* g, float* rres, float* rd,
                                 float* dc, int N, const DfCoef& c, hipStream_t s);
void launch_restrict8_32(const float* fine, float* out1, float* out2, int nc, hipStream_t s);
void launch_df_rhs(int frn, const float* rd, const float* xb, const float* rres, float* f, int n, const DfCoef& c,
                   hipStream_t s);
void launch_df_correct(int frn, const float* dH, const float* rd, float* dfine, const float* xbfine, int nc,
                       hipStream_t s);
void launch_df_coarsest(int frn, float* d, const float* xb, const float* rhs, int n, int sweeps, const DfCoef& c,
                        hipStream_t s);
// the inner V-cycle from level lv[0] (<= 16^3 cells) to the coarsest lv[nlev-1]
// in one single-workgroup launch; false if nlev is out of range
struct TailLevelIn {
    int n;
    float *d, *xb, *f, *rres, *rd;
    DfCoef dc;
};
bool launch_df_tail(int frn, int nlev, const TailLevelIn* lv, hipStream_t s);

}  // namespace mgs
