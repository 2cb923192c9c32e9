/*
 * spk.h — C ABI of the MI355X spatial-interpolation kernels (libspk.so).
 *
 * Plain pointers and sizes, row-major C-contiguous arrays, int status codes.
 * Functions named after a reference cpdef take HOST pointers and keep the
 * reference's argument order and in-place semantics; the library copies to
 * the GPU, computes there and copies back on the context's stream.  The
 * spk_*_dev functions take DEVICE pointers and run asynchronously on the
 * context's stream.
 *
 * References are to faizan90/spinterps:  "pyx" = cyth/interpmthds.pyx.
 */
#ifndef SPK_H
#define SPK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define SPK_OK          0
#define SPK_EBADMODEL   1   /* unknown or malformed variogram string      */
#define SPK_ESHAPE      2   /* inconsistent sizes                         */
#define SPK_EHIP        3   /* HIP runtime error                          */
#define SPK_ENOTSPD     4   /* kriging system not positive definite       */
#define SPK_EARG        5   /* invalid argument (assert in the reference) */
#define SPK_ENOMEM      6
#define SPK_EINDEX      7   /* index out of range                         */

/* Text of the last error on the calling thread. */
const char *spk_last_error(void);
int spk_version(void);

/* ---- context: device, stream, workspaces ---------------------------- */
typedef struct spk_ctx spk_ctx;
int spk_ctx_create(int device, spk_ctx **out);
int spk_ctx_destroy(spk_ctx *ctx);
/* Use an existing hipStream_t (NULL = the null stream). */
int spk_ctx_set_stream(spk_ctx *ctx, void *hip_stream);
int spk_ctx_sync(spk_ctx *ctx);
/* Release cached device workspaces. */
int spk_ctx_trim(spk_ctx *ctx);

/* Captured launch sequences: the batched path runs the launch sequences
 * of its factorisation (Cholesky panels, condition estimate, dual vectors
 * and refinement: about 300 launches per batch of 2 000-row systems) as
 * hipGraphs, captured at their first use on a context and replayed for
 * every later batch with the same buffers and sizes (on = 1, the
 * default); 0 launches them one by one.  Same kernels, same results.
 * Replaces the per-step loop of interp/steps.py:315-401. */
int spk_ctx_set_graphs(spk_ctx *ctx, int on);

/* ---- tile autotuner --------------------------------------------------
 * Kernel families with several tile variants that write identical
 * results: "gemm_plain_f64" (the plain FP64 contraction lambda = K^-1 r
 * of interp/steps.py:416, k depth 8, 16 or 32 per LDS stage: variants
 * 0, 1, 2) and "gemm_tri_bf16" (the split-BF16 variance contraction, the
 * next chunk of delta loaded after (0) or while (1) a chunk is
 * multiplied).  With SPK_TUNE_AUTO a context times every variant on the
 * first launch of each size class (the system's rows in classes of 1024)
 * on the real operands and keeps the fastest; SPK_TUNE_OFF launches the
 * library's defaults (1 and 0).  The reference's per-step loop
 * (interp/steps.py:315-401) has no counterpart: its tiles are numpy's. */
#define SPK_TUNE_OFF   0
#define SPK_TUNE_AUTO  1   /* default */
int spk_ctx_set_tuning(spk_ctx *ctx, int mode);
/* Launch variant v of a family on this context (v < 0: tune or default
 * again).  SPK_EARG for an unknown family or variant. */
int spk_ctx_force_variant(spk_ctx *ctx, const char *family, int v);
/* The picked variants as text, one line per (family, size class):
 * "family class variant ms_0 ms_1 ...", NUL-terminated within cap bytes.
 * *len (optional) receives the full length. */
int spk_ctx_tuning_report(spk_ctx *ctx, char *buf, int64_t cap,
                          int64_t *len);

/* ---- variogram programs (pyx:86-95, pyx:174-184) -------------------- */
#define SPK_VG_MAX 8
enum { SPK_VG_RNG = 0, SPK_VG_NUG, SPK_VG_SPH, SPK_VG_EXP, SPK_VG_LIN,
       SPK_VG_GAU, SPK_VG_POW, SPK_VG_HOL };
typedef struct spk_vg {
  int32_t n;                    /* number of components                  */
  int32_t type[SPK_VG_MAX];     /* SPK_VG_*                              */
  double sill[SPK_VG_MAX];
  double range[SPK_VG_MAX];     /* clamped to >= 1e-5 (pyx:183)          */
  double sill_sum;              /* sum of sills (misc.py:1093)           */
  double range_max_raw;         /* max unclamped range (misc.py:1094)    */
} spk_vg;

/* Parse one variogram string as fill_vg_var_arr does (pyx:174-184). */
int spk_vg_parse(const char *vg_str, spk_vg *out);
/* misc.py:1074-1105 check_full_nuggetness. */
int spk_vg_full_nuggetness(const char *vg_str, double min_vg_val, int *out);

/* ---- host-pointer mirrors of the cyth API (cyth/__init__.py:9-21) --- */
/* pyx:123-143 */
int spk_fill_dists_2d_mat(spk_ctx *ctx, const double *x1s, const double *y1s,
                          int64_t n1, const double *x2s, const double *y2s,
                          int64_t n2, double *dists);
/* pyx:146-226 */
int spk_fill_vg_var_arr(spk_ctx *ctx, const double *dists, int64_t rows,
                        int64_t cols, double *in_vars, int covar_flag,
                        int diag_mat_flag, const char *vg_models_str,
                        double min_vg_val);
/* pyx:229-248; subset_arr is [sub_rows x sub_cols], only the leading
 * [n_rows x n_cols] block is written. */
int spk_copy_2d_arr_at_idxs(spk_ctx *ctx, const double *arr, int64_t arr_rows,
                            int64_t arr_cols, const int64_t *row_idxs,
                            int64_t n_rows, const int64_t *col_idxs,
                            int64_t n_cols, double *subset_arr,
                            int64_t sub_rows, int64_t sub_cols);
/* pyx:768-781 */
int spk_fill_dists_one_pt(spk_ctx *ctx, double x, double y, const double *xs,
                          const double *ys, int64_t n, double *dists);
/* pyx:784-795; *wts_sum receives the return value of the cpdef. */
int spk_fill_wts_and_sum(spk_ctx *ctx, const double *dists, double *wts,
                         int64_t n, double idw_exp, double *wts_sum);
/* pyx:798-808 */
int spk_get_mults_sum(spk_ctx *ctx, const double *wts, const double *data,
                      int64_t n, double *out);
/* pyx:98-120: vg_arr += model(h, r, s), no range clamp. */
int spk_fill_theo_vg_vals(spk_ctx *ctx, const char *vg_name,
                          const double *h_arr, int64_t n, double r, double s,
                          double *vg_arr);
/* pyx:893-923: out[n(n-1)/2], condensed, i > j order. */
int spk_get_nd_dists(spk_ctx *ctx, const double *pts, int64_t n_pts,
                     int64_t n_dims, double *out);
/* pyx:811-890 for one destination point. */
int spk_sel_equidist_refs(spk_ctx *ctx, double dst_x, double dst_y,
                          const double *ref_xs, const double *ref_ys,
                          int64_t n_refs, uint64_t n_pies,
                          double min_dist_thresh, int64_t not_neb_flag,
                          double *dists, double *tem_ref_sel_dists,
                          int64_t *ref_sel_pie_idxs, uint64_t *ref_pie_idxs,
                          uint64_t *ref_pie_cts);

/* ---- class API (pyx:251-719) ----------------------------------------- */
/* kind 0: OrdinaryKriging (pyx:251-362), 1: SimpleKriging (pyx:365-471),
 * 2: ExternalDriftKriging_MD (pyx:588-719).  Host pointers; n input points
 * (xi, yi, zi), C output points (xk, yk); EDK_MD drifts si [m x n] and
 * sk [m x C].  P = n + 1 (OK), n (SK), n + 1 + m (EDK_MD).  Outputs (any
 * may be NULL except zk): zk [C]; est_vars [C] (OK est_vars, SK
 * est_covars); lambdas [C x n]; mus [C] (OK) or [(m+1) x C] (EDK_MD
 * mus_arr); rhss [C x P]; in_vars [P x P] (SK: in_covars); in_vars_inv
 * [P x P]; in_dists [n x n].  SPK_ENOTSPD when the system is singular. */
int spk_krige_class(spk_ctx *ctx, int kind, const char *model,
                    const double *xi, const double *yi, const double *zi,
                    int64_t n, const double *si, int m, const double *xk,
                    const double *yk, const double *sk, int64_t C,
                    double *zk, double *est_vars, double *lambdas,
                    double *mus, double *rhss, double *in_vars,
                    double *in_vars_inv, double *in_dists);

/* ---- neighbour selection for all cells (interp/grps.py:141-247) ----- */
/* nrst: out[n_dst x min(n_nebs, n_refs)] sorted indices of the nearest. */
int spk_neb_idxs_nrst(spk_ctx *ctx, const double *dst_xs, const double *dst_ys,
                      int64_t n_dst, const double *ref_xs,
                      const double *ref_ys, int64_t n_refs, int64_t n_nebs,
                      int64_t *out);
/* pie: out[n_dst x n_nebs] (interp/grps.py:168-247, min_dist_thresh -1). */
int spk_neb_idxs_pie(spk_ctx *ctx, const double *dst_xs, const double *dst_ys,
                     int64_t n_dst, const double *ref_xs, const double *ref_ys,
                     int64_t n_refs, int64_t n_nebs, int64_t n_pies,
                     int64_t *out);

/* ---- batched step interpolation (interp/steps.py:245-435) ----------- */
enum { SPK_M_OK = 0, SPK_M_SK = 1, SPK_M_EDK = 2, SPK_M_IDW = 3,
       SPK_M_NNB = 4 };
/* SPK_P_FP64 : every step on its own, FP64 end to end, the reference's
 *              per-cell lambda = K^-1 rhs as one FP64 MFMA GEMM per step.
 * SPK_P_MIXED: all steps of the call in batched launches; estimates in the
 *              dual form (FP64), variance by a contraction with FP64
 *              Cholesky/Schur terms whose precision follows the step's
 *              condition: well-conditioned steps on split-BF16 operands
 *              (three bfloat16 terms per value, six plane products on BF16
 *              MFMA, FP32 sums per 32 rows, FP64 across them), the others
 *              on exact int8 plane products or FP64.
 * SPK_P_FP64B: the batched dual-form path with an FP64 contraction.
 * SPK_P_MIXED_F32: SPK_P_MIXED with FP32 operands on FP32 MFMA for the
 *              well-conditioned steps (also used by SPK_P_MIXED for
 *              systems above 8 192 rows). */
enum { SPK_P_FP64 = 0, SPK_P_MIXED = 1, SPK_P_FP64B = 2,
       SPK_P_MIXED_F32 = 3 };

/* Per-step status written to info[t] (bit flags). */
#define SPK_ST_NOSTNS    1   /* no station: outputs untouched (NaN)        */
#define SPK_ST_MEANFILL  2   /* interp flag off or nugget-only: mean fill  */
#define SPK_ST_FAILED    4   /* no solution: nearest-station fill (the
                                Cholesky of the covariance block and the
                                pseudo-inverse fallback both failed)      */
#define SPK_ST_ONESTN    8   /* one station: its value everywhere          */
#define SPK_ST_ILLCOND  16   /* results are set by rounding at the 1e-6
                                level (as the reference's pinv's are):
                                condition estimate > 1e12, or a residual
                                check failed.  SPK_P_FP64: some cell's
                                |sum of weights - 1| > 1e-9 (the sum is
                                exactly one in exact arithmetic).
                                SPK_P_MIXED / FP64B: the FP64 refinement
                                step of the dual vectors changed some
                                cell's estimate by more than 1e-7 of its
                                magnitude (at least 1e-3 of the step's
                                mean |value|). */
#define SPK_ST_FP64VAR  32   /* SPK_P_MIXED: the variance contraction ran
                                in FP64: condition estimate above a
                                caller's int8 limit (spk_opts.int8var_kappa),
                                or a
                                block of cells whose nearest stations
                                exceed the local fit's staging            */
#define SPK_ST_INT8VAR  64   /* SPK_P_MIXED: condition estimate above the
                                FP32 limit (1e5, Gaussian models 1e3, or
                                spk_opts.fp64var_kappa): the variance
                                contraction used exact int8 plane
                                products (unless FP64VAR is also set)    */
#define SPK_ST_PINV    128   /* the Cholesky of the covariance block failed
                                (indefinite or numerically singular): the
                                step was solved as the reference solves
                                every step, lambda = pinv(K) rhs per cell
                                with K the bordered system, pinv from a
                                symmetric eigendecomposition with numpy's
                                cutoff 1e-15 x the largest |eigenvalue|
                                (interp/steps.py:350-371, 403-435); also
                                SPK_ST_ILLCOND                           */

typedef struct spk_problem {
  int64_t n_stns;               /* N                                      */
  int64_t n_cells;              /* C                                      */
  int32_t n_drifts;             /* m (EDK)                                */
  int32_t pad_;
  const double *stn_x, *stn_y;  /* device [N], metres                     */
  const double *cell_x, *cell_y;/* device [C], metres                     */
  const double *stn_drift;      /* device [N x m] or NULL                 */
  const double *cell_drift;     /* device [m x C] or NULL                 */
  double origin_x, origin_y;    /* subtracted before any FP32 arithmetic  */
} spk_problem;

typedef struct spk_opts {
  int32_t method;               /* SPK_M_*                                */
  int32_t precision;            /* SPK_P_*                                */
  int32_t want_var;             /* OK variance                            */
  int32_t sk_mean;              /* SPK_M_SK: 0 = the reference's SK
                                   (lambda^T z, nearest station where
                                   sum(lambda) is not ~1,
                                   interp/steps.py:415-429); 1 = simple
                                   kriging with the step mean as known
                                   mean: m + lambda^T (z - m)            */
  double idw_exp;
  double min_vg_val;            /* interp/data.py default 0.0             */
  double min_var_thr;           /* interp_steps_flags (steps.py:760-765)  */
  double fp64var_kappa;         /* SPK_P_MIXED: condition estimate above
                                   which a step's variance contraction
                                   leaves FP32, for every model; 0 = the
                                   library's policy (1e5, Gaussian 1e3)   */
  double int8var_kappa;         /* SPK_P_MIXED: condition estimate up to
                                   which such a step uses exact int8 plane
                                   products, above it FP64; 0 = no limit,
                                   negative = never                       */
} spk_opts;

/* Interpolate n_steps steps.  vals: device [n_steps x N], NaN = station
 * missing at that step.  vgs: device [n_steps] (kriging only).  est, var:
 * device [n_steps x C] float64 (var may be NULL).  info: device [n_steps]
 * SPK_ST_* bits.  kappa: device [n_steps] condition estimate
 * max diag(C^-1) * max diag(C) of each kriging system, or NULL.
 * Outputs of steps without stations are left untouched. */
int spk_interp_steps_dev(spk_ctx *ctx, const spk_problem *prob,
                         const spk_opts *opts, int64_t n_steps,
                         const double *vals, const spk_vg *vgs, double *est,
                         double *var, int32_t *info, double *kappa);

/* The fields as the reference keeps them, in host memory: interp/steps.py
 * clips each interpolated field to [min_var_cut, max_var_cut]
 * (_mod_min_max, interp/steps.py:466-476) and stores it as float32.
 * est32 [n_steps x C] (required) takes the clipped estimates (has_min /
 * has_max select the cuts), var32 [n_steps x C] (NULL: none) the OK
 * variance cast only.  HOST pointers; with page-locked memory
 * (hipHostMalloc) each batch's fields are copied while later batches
 * compute. */
typedef struct spk_f32_out {
  float *est32;
  float *var32;
  int32_t has_min, has_max;
  double min_cut, max_cut;
} spk_f32_out;

/* spk_interp_steps_dev that also delivers the float32 host fields of
 * out32, batch by batch on a third stream (est / var are still written).
 * Asynchronous like spk_interp_steps_dev: the host fields are complete
 * when the context's stream has finished the call (spk_ctx_sync).
 * Replaces the per-step float32 writes of interp/steps.py:466-476 and
 * 815-831 into interp_flds. */
int spk_interp_steps_f32_dev(spk_ctx *ctx, const spk_problem *prob,
                             const spk_opts *opts, int64_t n_steps,
                             const double *vals, const spk_vg *vgs,
                             double *est, double *var, int32_t *info,
                             double *kappa, const spk_f32_out *out32);

/* ---- neighbourhood interpolation ------------------------------------
 * Each cell uses only n_nebs of the step's active stations: its nearest
 * (SPK_NEB_NRST, interp/grps.py:147-166) or n_nebs taken rank by rank over
 * n_pies angular sectors (SPK_NEB_PIE, interp/grps.py:168-247).  Stations
 * that no cell selects with every station present are treated as missing
 * (interp/steps.py:592-608).  Cells with the same station set form a
 * group (interp/grps.py:103-139) whose flags, mean fill, nearest-station
 * and kriging system are those of interp/steps.py:732-831 over the set.
 * Replaces the per-group loop of interp/steps.py:732-831 with its
 * _get_interp calls (interp/steps.py:245-401). */
#define SPK_NEB_NRST 1
#define SPK_NEB_PIE  2
typedef struct spk_neb {
  int32_t sel;                  /* SPK_NEB_NRST or SPK_NEB_PIE            */
  int32_t n_nebs;               /* stations per cell, 1..64               */
  int32_t n_pies;               /* SPK_NEB_PIE: sectors, 1..n_nebs        */
  int32_t pad_;
} spk_neb;

/* spk_interp_steps_dev's arguments and outputs, FP64 throughout
 * (opts->precision is not used).  info[t]: SPK_ST_NOSTNS for a step
 * without stations, else the OR over the step's groups of SPK_ST_ONESTN
 * (a one-station set), SPK_ST_MEANFILL (a group filled with its mean),
 * SPK_ST_PINV (a group solved by the pseudo-inverse of its bordered
 * system: its Cholesky failed or its condition estimate exceeded 1e14),
 * SPK_ST_FAILED (that eigendecomposition failed too: nearest station) and
 * SPK_ST_ILLCOND (condition estimate above 1e12, or eigenvalues dropped).
 * kappa[t]: the largest condition estimate of the step's groups, or NULL.
 * SPK_ESHAPE for pie steps with 1 < active stations < n_nebs (the
 * reference's row assignment fails there); SPK_EINDEX for a sector index
 * out of range (pyx:868-870). */
int spk_interp_neb_steps_dev(spk_ctx *ctx, const spk_problem *prob,
                             const spk_opts *opts, const spk_neb *neb,
                             int64_t n_steps, const double *vals,
                             const spk_vg *vgs, double *est, double *var,
                             int32_t *info, double *kappa);

/* No counterpart in the reference.  spk_interp_steps_dev for SPK_M_IDW and
 * SPK_M_NNB that allocates nothing, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same kernels and results; its scratch (re-centred
 * coordinates, per-step compaction) is a static device array of the
 * library, about 23 MB per device, held until the process ends.  At most
 * 8192 stations and 2^20 cells (SPK_ESHAPE beyond).  Calls that overlap on
 * one device share that array and race.  kappa may be NULL. */
int spk_idw_steps_noalloc_dev(spk_ctx *ctx, const spk_problem *prob,
                              const spk_opts *opts, int64_t n_steps,
                              const double *vals, double *est, int32_t *info,
                              double *kappa);

/* interp/steps.py:466-476 + the float32 cast of steps.py:815-831:
 * out32[i] = (float)clip(in64[i]) on device; NaN passes through.
 * has_min/has_max select the cuts (None in the reference = 0). */
int spk_clip_cast_dev(spk_ctx *ctx, const double *in64, float *out32,
                      int64_t n, int has_min, double min_cut, int has_max,
                      double max_cut);

/* ---- variogram estimation and fit (variograms/vgs.py, fitvgs.py) ---- */
/* All station pairs (i > j) sorted by distance (stable).  pi, pj, pd:
 * device [N (N - 1) / 2].  vgs.py:345-391. */
int spk_vgfit_pairs_dev(spk_ctx *ctx, const double *x, const double *y,
                        int64_t N, int32_t *pi, int32_t *pj, double *pd);
/* Empirical variograms of T steps: vals device [T x N] (NaN = no value);
 * pairs with both values, the mdr distance clip (vgs.py:597-659), bins of nk
 * consecutive pairs with the robust (classic = 0) or classic estimator
 * (vgs.py:399-431, 495-520).  bin_h, bin_g, bin_n: device [T x max_bins];
 * nbins: device [T]. */
int spk_vgfit_evg_dev(spk_ctx *ctx, const int32_t *pi, const int32_t *pj,
                      const double *pd, int64_t M, int64_t N,
                      const double *vals, int64_t T, double mdr, int nk,
                      int classic, int max_bins, double *bin_h, double *bin_g,
                      int32_t *bin_n, int32_t *nbins);
/* Differential-evolution fit (vgs.py:661-764) of n_cands candidate models
 * to each step's bins.  cands: HOST [n_cands x 6] int32 = {n_comp,
 * type[4] (SPK_VG_*), first_fixed}; first_fixed pins the first component
 * to the first bin (vgs.py:735-739).  var: device [T] (the step variance
 * bounds the sills at 2 var).  params: device [T x n_cands x 8] as (range,
 * sill) per component; fun, nit: device [T x n_cands].  range_ub is the
 * upper range bound (1e9 in vgs.py:747); popmul 50 and tol 0.01 are the
 * reference's settings.  A candidate with D = 2 n_comp parameters evolves
 * min(popmul D^2, 800) members; one workgroup per (step, candidate), whose
 * LDS (16 bytes per member and parameter of the launch's largest
 * candidate) sets how many share a CU. */
int spk_vgfit_de_dev(spk_ctx *ctx, const double *bin_h, const double *bin_g,
                     const int32_t *nbins, int max_bins, const double *var,
                     int64_t T, const int32_t *cands, int n_cands, int maxiter,
                     int popmul, double tol, double range_ub, uint64_t seed,
                     double *params, double *fun, int32_t *nit);

/* spk_vgfit_de_dev with a seeded start population (the learned proposal's
 * refinement).  Replaces the Latin-hypercube start of the same
 * differential evolution (vgs.py:758-764, scipy 'latinhypercube' init):
 * member 0 of (step t, candidate c) is init[(t * n_cands + c) * 8 + 0..D-1]
 * (clipped to the bounds); the others stratify, per dimension, the
 * log-box init * exp([-spread, spread]) clipped to the bounds.  spread in
 * [0, 20]; pinned dimensions stay pinned. */
int spk_vgfit_de_seeded_dev(spk_ctx *ctx, const double *bin_h,
                            const double *bin_g, const int32_t *nbins,
                            int max_bins, const double *var, int64_t T,
                            const int32_t *cands, int n_cands, int maxiter,
                            int popmul, double tol, double range_ub,
                            uint64_t seed, const double *init, double spread,
                            double *params, double *fun, int32_t *nit);

/* ---- learned variogram proposal ----------------------------------------
 * No counterpart function in the reference: the training data generator,
 * features and network that propose a model for the fit of
 * variograms/fitvgs.py:143-278.  All pointers are device memory. */

/* Unconditional Gaussian fields: for each step t the stations whose
 * mask_vals[t, i] is not NaN get mean[t] + (L xi)_i with L L^T the
 * covariance of vgs[t] (sum over components of sill - gamma, the nugget
 * sill on the diagonal, plus 1e-6 of the sill sum as jitter), xi standard
 * normal from a counter-based generator keyed on (seed, t, station rank).
 * out [T x N]: NaN at masked stations.  info [T]: SPK_ST_* (FAILED when
 * the Cholesky failed; the step's row stays NaN).  Returns SPK_EARG on a
 * context that has run spk_vgfit_pairs_dev: that call order stalled on the
 * GPU with no cause found, so it is refused (simulate first, or use a
 * second context). */
int spk_simulate_fields_dev(spk_ctx *ctx, const spk_problem *prob, int64_t T,
                            const double *mask_vals, const spk_vg *vgs,
                            const double *means, uint64_t seed, double *out,
                            int32_t *info);

/* Lag-bin features of each step from the distance-sorted pairs of
 * spk_vgfit_pairs_dev: nlag equal-width bins over (0, mdr * largest active
 * pair distance]; feat [T x 2 nlag] float: classic semivariance
 * (half mean squared difference) / step variance, then nlag times the
 * share of pairs per bin; scale [T x 2]: (mdr * largest distance,
 * variance).  nlag <= 64. */
int spk_vg_features_dev(spk_ctx *ctx, const int32_t *pi, const int32_t *pj,
                        const double *pd, int64_t M, int64_t N,
                        const double *vals, int64_t T, double mdr, int nlag,
                        float *feat, double *scale);

/* Multilayer perceptron 40 -> 128 -> 128 -> 128 -> 6 (GELU hidden
 * activations), float32 parameters in one flat array
 * [W1 b1 W2 b2 W3 b3 W4 b4], W_l row-major [out x in].  Outputs: logits of
 * (Sph, Exp, Gau), log(sill / variance), log(range / lag scale), nugget
 * fraction. */
int64_t spk_mlp_n_params(void);
int spk_mlp_init_dev(spk_ctx *ctx, uint64_t seed, float *params);
/* X [B x 40] -> out [B x 6]. */
int spk_mlp_forward_dev(spk_ctx *ctx, const float *params, const float *X,
                        int B, float *out);
/* Mean over the batch of softmax cross-entropy (cls [B] in 0..2) plus
 * reg_w * squared error of the 3 regression outputs against reg [B x 3]
 * (NaN targets ignored); writes loss [1] and grads (layout of params). */
int spk_mlp_loss_grad_dev(spk_ctx *ctx, const float *params, const float *X,
                          int B, const int32_t *cls, const float *reg,
                          float reg_w, float *grads, float *loss);
/* Adam update, step counted from 1 (bias correction 1 - beta^step). */
int spk_mlp_adam_dev(spk_ctx *ctx, float *params, const float *grads,
                     float *m, float *v, int64_t step, float lr, float b1,
                     float b2, float eps);

/* Device timing of the last spk_interp_steps_dev call, per stage (ms). */
int spk_last_timings(spk_ctx *ctx, double *ms, int max_n, int *n_out);

/* Kernel timing with hipEvents around the library's heavy launches:
 * enable/disable, clear, and read the total ms and launch count of one
 * kernel family ("gemm_tri_f32", "kgen", ...). */
int spk_prof_enable(spk_ctx *ctx, int on);
int spk_prof_reset(spk_ctx *ctx);
int spk_prof_get(spk_ctx *ctx, const char *name, double *total_ms,
                 int64_t *count);
/* Executed work counted while profiling is on: "gemm_tri_f32_macs", the
 * multiply-adds the FP32 triangular contraction executed, and
 * "gemm_tri_i8_ops", the int8 operations of the plane products the int8
 * contraction executed (both skip the zero chunks of delta, so these are
 * below the algorithmic counts); "neb_chunks" and "neb_pinv_chunks", the
 * neighbourhood group chunks spk_interp_neb_steps_dev solved and those of
 * them solved by the pseudo-inverse. */
int spk_prof_work(spk_ctx *ctx, const char *name, double *value);

#ifdef __cplusplus
}
#endif
#endif /* SPK_H */
