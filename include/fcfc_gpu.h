/*
 * fcfc_gpu.h -- C ABI of the MI355X pair-count engine for FCFC.
 *
 * The engine replaces FCFC's count step
 *     int count_pairs(const void *tree1, const void *tree2, CF *cf,
 *                     COUNT *cnt, bool isauto, bool withwt);
 * (src/fcfc/2pt_box/count_func.h:54-59, src/fcfc/2pt/count_func.h:54-59),
 * which eval_pairs() calls once per PAIR_COUNT
 * (src/fcfc/2pt_box/eval_cf.c:119,142).  Instead of kd-trees it takes the
 * catalogues as structure-of-arrays, already rescaled by the binning setup
 * (src/fcfc/2pt_box/build_tree.c:121-131), and returns the same histogram:
 * index s_bin + ns * (mu_bin | pi_bin), auto pairs counted once (the caller
 * doubles them, eval_cf.c:125-133), counts ADDED to `cnt`.
 *
 * All functions return 0 on success or a negative FCFC_ERR_* code; none of
 * them exits the process.  A context is not thread-safe: use one per host
 * thread.  Pointers are borrowed for the duration of the call.
 */
#ifndef FCFC_GPU_H
#define FCFC_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes: same values as src/util/define_comm.h:218-234 -------- */
#define FCFC_OK               0
#define FCFC_ERR_MEMORY      (-1)
#define FCFC_ERR_ARG         (-2)
#define FCFC_ERR_FILE        (-3)
#define FCFC_ERR_CFG         (-4)
#define FCFC_ERR_AST         (-5)
#define FCFC_ERR_ASCII       (-6)
#define FCFC_ERR_BINARY      (-7)
#define FCFC_ERR_CONF        (-20)
#define FCFC_ERR_DATA        (-21)
#define FCFC_ERR_CNVT        (-22)
#define FCFC_ERR_TREE        (-23)
#define FCFC_ERR_CF          (-24)
#define FCFC_ERR_SAVE        (-25)
#define FCFC_ERR_UNKNOWN     (-99)
/* no counterpart in the reference: a HIP runtime or kernel launch failure */
#define FCFC_ERR_GPU         (-60)

/* ---- binning schemes: src/fcfc/2pt_box/define.h:46-48 ------------------ */
#define FCFC_BIN_ISO          0
#define FCFC_BIN_SMU          1
#define FCFC_BIN_SPI          2

/* ---- geometry ----------------------------------------------------------- */
#define FCFC_GEOM_BOX         0   /* FCFC_2PT_BOX: periodic, LOS = z axis   */
#define FCFC_GEOM_SURVEY      1   /* FCFC_2PT: pair line of sight r1 + r2   */

/* ---- lookup table types: define.h:53-54 --------------------------------- */
#define FCFC_LOOKUP_TYPE_INT     0
#define FCFC_LOOKUP_TYPE_HYBRID  1

/*
 * Separation bins after FCFC's setup (src/fcfc/2pt_box/setup_cf.c:385-531,
 * src/fcfc/2pt/setup_cf.c:433-502): edges rescaled by `rescale` so that the
 * squared edges become integers where possible, plus the lookup tables that
 * the kernels stage in LDS.  Filled by fcfc_bins_linear / fcfc_bins_edges,
 * released by fcfc_bins_free.
 */
typedef struct {
  int geometry;          /* FCFC_GEOM_BOX or FCFC_GEOM_SURVEY                */
  int bintype;           /* FCFC_BIN_ISO / SMU / SPI                          */
  int ns, nmu, np;       /* bins in s (or s_perp), mu, pi                    */
  int ntot;              /* ns, ns*nmu or ns*np                               */
  double rescale;        /* coordinate factor (setup_cf.c:308-329)           */
  double *sbin_raw;      /* ns+1 edges in input units                         */
  double *pbin_raw;      /* np+1 pi edges in input units (SPI) or NULL       */
  double *sbin;          /* ns+1 rescaled edges                               */
  double *s2bin;         /* ns+1 rescaled squared edges                       */
  double *pbin;          /* np+1 rescaled pi edges (box) / pi^2 edges (survey) */
  int tabtype;           /* FCFC_LOOKUP_TYPE_INT or _HYBRID                   */
  int stab_len, ptab_len;
  uint16_t *stab;        /* table over integer parts of s^2 (create_lut.c)   */
  uint16_t *ptab;        /* table over integer parts of pi (or pi^2)         */
  int soff, poff;        /* (int) s2min and (int) pmin                        */
} fcfc_bins;

/* Linear bins [smin,smax) of width ds (and pi bins for SPI, nmu mu bins for
 * SMU), as cf_setup does for SEP_BIN_MIN/MAX/SIZE (setup_cf.c:614).
 * For the box geometry `bsize` (3 side lengths, input units) is checked:
 * the maximum separation must be below half the box (load_conf.c:1000). */
int fcfc_bins_linear(fcfc_bins *bins, int geometry, int bintype,
    double smin, double smax, double ds, int nmu,
    double pmin, double pmax, double dpi, const double *bsize);

/* Arbitrary edges (SEP_BIN_FILE / PI_BIN_FILE, setup_cf.c:596-640):
 * sedges[ns+1], pedges[np+1] (SPI only), increasing, input units. */
int fcfc_bins_edges(fcfc_bins *bins, int geometry, int bintype,
    const double *sedges, int ns, int nmu, const double *pedges, int np,
    const double *bsize);

void fcfc_bins_free(fcfc_bins *bins);

/*
 * A catalogue as structure-of-arrays in FP64.  Coordinates must already be
 * multiplied by bins.rescale (fcfc_prepare_catalog does that in place).
 * `s` is |r|^2 for survey (s,mu)/(s_perp,pi) counts, computed as FCFC's
 * data_preprocess does (src/fcfc/2pt/build_tree.c:35-133); NULL lets the
 * engine compute it.  `w` = per-object weights or NULL (unit weights).
 * `on_device` != 0 means the pointers are HIP device pointers.
 */
typedef struct {
  int64_t n;
  const double *x, *y, *z;
  const double *s;
  const double *w;
  int on_device;
} fcfc_cat;

/* Rescale coordinates in place (build_tree.c:121-131) and, for survey
 * (s,mu)/(s_perp,pi) bins, fill s[i] = |r_i|^2 with FCFC's lane rule
 * (FMA for the first floor(n/8)*8 objects, unfused for the rest).
 * `s` may be NULL when not needed; `sumw` (may be NULL) receives sum(w) or n.
 * Host pointers only. */
int fcfc_prepare_catalog(const fcfc_bins *bins, int64_t n, double *x,
    double *y, double *z, double *s, const double *w, double *sumw);

/* ---- engine context ----------------------------------------------------- */
typedef struct fcfc_gpu_ctx fcfc_gpu_ctx;

int fcfc_gpu_init(fcfc_gpu_ctx **ctx, int device);
int fcfc_gpu_destroy(fcfc_gpu_ctx *ctx);
/* Human-readable text for the last error of this context. */
const char *fcfc_gpu_last_error(const fcfc_gpu_ctx *ctx);

/* flags for fcfc_gpu_count* */
#define FCFC_GPU_AUTO         0x01u  /* b is ignored: auto pairs of a, once */
#define FCFC_GPU_WEIGHTED     0x02u  /* FP64 weighted counts (double cnt[]) */
#define FCFC_GPU_PREC_FP64    0x00u  /* FP64 distances; mu bin from the
                                        reciprocal square root, settled
                                        exactly near its edges             */
#define FCFC_GPU_PREC_FP32    0x10u  /* FP32 distances + FP64 edge recheck
                                        (same counts as FP64; launches it
                                        cannot serve run the FP64 form:
                                        a lower s^2 (or pi) edge far above
                                        zero, windows wider than half a
                                        table cell, tables too long for
                                        LDS, (s,mu) with nmu > 126)         */
#define FCFC_GPU_PREC_PLAIN   0x80u  /* FP64, the reference's per-pair
                                        arithmetic transcribed directly (an
                                        IEEE division per (s,mu) pair); the
                                        baseline the other forms are timed
                                        against                            */
#define FCFC_GPU_CNT_DEVICE   0x20u  /* cnt is a device pointer             */
#define FCFC_GPU_CENSUS       0x40u  /* count pairs whose bin is rounding-path
                                        dependent in the reference (FP32 mode) */
#define FCFC_GPU_GRADIENT     0x100u /* fcfc_gpu_autotune: tune the weight-
                                        gradient kernel (fcfc_gpu_count_backward) */

/*
 * Count pairs.  Replaces count_pairs(tree1, tree2, cf, cnt, isauto, withwt).
 * bsize: rescaled box side lengths (box geometry), NULL for survey.
 * cnt:   int64_t[ntot] (unweighted) or double[ntot] (FCFC_GPU_WEIGHTED);
 *        results are added to it.
 * A catalogue with a NaN or infinite coordinate is rejected with
 * FCFC_ERR_DATA (found while its grid is built, before any count runs;
 * the reference does not check, and its tree code is undefined there).
 */
int fcfc_gpu_count(fcfc_gpu_ctx *ctx, const fcfc_cat *a, const fcfc_cat *b,
    const fcfc_bins *bins, const double *bsize, uint32_t flags, void *cnt);

/*
 * No-allocation count.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * fcfc_gpu_count, but no device memory is allocated or freed: every device
 * buffer of the call comes from a fixed scratch array of 128 MiB per GPU
 * (held by the library from its first use on that GPU until the process
 * ends).  Catalogues whose sorted copies and sort scratch cannot fit (about
 * 2 - 3e6 objects in total; fewer with host catalogues, weights or survey
 * |r|^2) are rejected with FCFC_ERR_MEMORY and a message, before any work is
 * queued or, for a grid or tile list that overflows the rest, before the
 * count launch.  All calls on one GPU share its scratch array: calls that
 * overlap on one GPU (several threads or contexts) race and must not be
 * made.  Host memory (the call's bookkeeping) is still used.
 */
int fcfc_gpu_count_noalloc(fcfc_gpu_ctx *ctx, const fcfc_cat *a,
    const fcfc_cat *b, const fcfc_bins *bins, const double *bsize,
    uint32_t flags, void *cnt);

/* Per-call statistics of the last fcfc_gpu_count / _batch / _backward. */
typedef struct {
  double evals;          /* candidate pair distances evaluated              */
  double rechecked;      /* pairs re-evaluated in FP64 (FP32 mode)          */
  double census;         /* rounding-path ambiguous pairs (FCFC_GPU_CENSUS) */
  double kernel_ms;      /* device time of the count kernel(s)              */
  double total_ms;       /* wall time of the call                            */
  int    cell_div;       /* cells per max separation used                   */
  int    tile;           /* objects per A tile                              */
} fcfc_gpu_stats;
int fcfc_gpu_last_stats(const fcfc_gpu_ctx *ctx, fcfc_gpu_stats *st);

/*
 * Batched counting: many independent (catalogue pair, histogram) jobs in one
 * persistent launch.  No counterpart in the reference (it counts one
 * PAIR_COUNT at a time); each job has the semantics of fcfc_gpu_count.
 * The histogram of job k is cnt + k * bins->ntot (int64 or double).
 */
typedef struct {
  fcfc_cat a, b;          /* b ignored for auto jobs                        */
  double bsize[3];        /* rescaled box (box geometry)                    */
  uint32_t flags;         /* FCFC_GPU_AUTO | FCFC_GPU_WEIGHTED              */
} fcfc_batch_job;
int fcfc_gpu_count_batch(fcfc_gpu_ctx *ctx, int njob,
    const fcfc_batch_job *jobs, const fcfc_bins *bins, uint32_t flags,
    void *cnt);

/*
 * Prepared batches (no counterpart in the reference).  fcfc_gpu_plan_create
 * sorts the catalogues of njob jobs into their grids and tiles them once,
 * then records the device work of one count -- the weights re-read from the
 * callers' device arrays (fcfc_cat.w of on_device catalogues), histogram and
 * queue clears, the persistent count launch -- as a HIP graph.  Each
 * fcfc_gpu_plan_run replays the graph and adds the njob histograms into cnt
 * (as fcfc_gpu_count_batch; FCFC_GPU_CNT_DEVICE in flags: device pointer).
 * For counts repeated on fixed positions with changing weights (weight
 * fitting).  The positions are those at creation; the context must outlive
 * the plan and runs of one context must not overlap.
 */
typedef struct fcfc_gpu_plan fcfc_gpu_plan;
int fcfc_gpu_plan_create(fcfc_gpu_ctx *ctx, int njob,
    const fcfc_batch_job *jobs, const fcfc_bins *bins, uint32_t flags,
    fcfc_gpu_plan **plan);
int fcfc_gpu_plan_run(fcfc_gpu_plan *plan, void *cnt);
int fcfc_gpu_plan_destroy(fcfc_gpu_plan *plan);

/*
 * Backward of the weighted counts with respect to per-object weights.
 * No counterpart in the reference.  Given G = dL/dcnt (double[ntot], for the
 * histogram as returned, i.e. before the caller's auto doubling), adds
 *   ga[i] += sum_j G[bin(i,j)] * w_b[j]     (cross: also gb[j] += ...)
 * For auto counts every unordered pair contributes to both members and gb
 * is unused; for cross counts gb may be NULL when only a's gradient is
 * wanted.  ga/gb are double[n] in the ORIGINAL object order.  The pairs are
 * enumerated and binned exactly as fcfc_gpu_count enumerates them.
 */
int fcfc_gpu_count_backward(fcfc_gpu_ctx *ctx, const fcfc_cat *a,
    const fcfc_cat *b, const fcfc_bins *bins, const double *bsize,
    uint32_t flags, const double *grad_cnt, double *ga, double *gb);

/* ---- tuning ------------------------------------------------------------- */
/* Override the grid and launch shape (0 = automatic / autotuned):
 * cell_div = cells per maximum separation in x and y, tile = A objects per
 * tile (at most 2048; above 1024 only the weighted FP64 box counts with
 * integer tables hold such tiles, four objects per lane, other counts cut
 * them to 1024), zcell_div = z-cells per x-cell, nblock = persistent workgroups. */
int fcfc_gpu_set_tuning(fcfc_gpu_ctx *ctx, int cell_div, int tile,
    int zcell_div, int nblock);

/* Autotuner (no counterpart in the reference): times candidate grid / tile
 * shapes for counts like (a, b, bins, flags) on an eighth (beyond 16384
 * tiles: a sixteenth) of their tile list, taken as 16 runs of consecutive
 * tiles spread over the whole list (counts of at most 4096 tiles whole):
 * every pair of cell_div {2, 3, 4, 6} and zcell_div {1, 2, 4, 8, 16} at its
 * automatic tile size, then the best three pairs at the other tile sizes of
 * {512, 1024} (and 2048 for weighted FP64 box counts with integer tables), and
 * keeps the fastest in the
 * context for this geometry, binning, precision and density bucket; later
 * counts whose shape is automatic use it.  With FCFC_GPU_GRADIENT in flags
 * it tunes fcfc_gpu_count_backward for such counts instead.  best (may be
 * NULL) receives {cell_div, zcell_div, tile, estimated full-count
 * microseconds}. */
int fcfc_gpu_autotune(fcfc_gpu_ctx *ctx, const fcfc_cat *a, const fcfc_cat *b,
    const fcfc_bins *bins, const double *bsize, uint32_t flags, int *best);

/* ---- pair-count files (src/fcfc/2pt_box/save_res.c:45-296) -------------- */
/* Write a pair-count file byte-compatible with FCFC's save_res:
 * ofmt 0 = binary, 1 = ASCII.  `cnt` are the FINAL counts (auto already
 * doubled); norm and ncnt are computed as eval_pairs does (eval_cf.c:117-167).
 * n1/n2: object numbers, wt1/wt2: weight sums (weighted only). */
int fcfc_save_pair_count(const char *fname, const fcfc_bins *bins, int ofmt,
    const double *bsize_rescaled, int isauto, int weighted, uint64_t n1,
    uint64_t n2, double wt1, double wt2, const void *cnt);

/* ---- survey coordinate conversion (src/fcfc/2pt/cnvt_coord.c) ---------- */
/* Cosmology of COORD_CONVERT (src/fcfc/2pt/load_conf.c check_cosmo): pass
 * NaN for an unset key.  omega_l unset: flat, 1 - omega_m; w unset: -1;
 * err unset: 1e-8.  fcfc_cosmology_check fills the defaults and omega_k and
 * validates (omega_m in (0,1], omega_l >= 0, w <= -1/3, H^2(z) > 0,
 * err >= DBL_EPSILON), returning FCFC_ERR_CFG otherwise. */
typedef struct {
  double omega_m;        /* OMEGA_M                                         */
  double omega_l;        /* OMEGA_LAMBDA                                    */
  double omega_k;        /* 1 - omega_m - omega_l (set by the check)        */
  double w;              /* DE_EOS_W                                        */
  double err;            /* CMVDST_ERR: relative accuracy of the integral   */
} fcfc_cosmology;
int fcfc_cosmology_check(fcfc_cosmology *cm);

/* Order of the Gauss-Legendre rule (4-32) that integrates c/H(z) to the
 * relative accuracy cm->err at 128 redshifts spread over [min z, max z]
 * (cnvt_coord.c:cnvt_legauss_order); FCFC_ERR_DATA for a negative redshift,
 * FCFC_ERR_CNVT when order 32 does not converge. */
int fcfc_cmvdst_order(const fcfc_cosmology *cm, int64_t n, const double *z,
    int *order);

/* In place: x = RA (deg), y = Dec (deg), z = redshift -> comoving Cartesian
 * coordinates in Mpc/h (cnvt_coord.c:cnvt_coord_integr), with the order of
 * fcfc_cmvdst_order (returned in order_used, may be NULL).  Host pointers. */
int fcfc_convert_coords(const fcfc_cosmology *cm, int64_t n, double *x,
    double *y, double *z, int *order_used);

/* The same with distances from a natural cubic spline through nsample
 * (redshift, distance) samples in any order (Z_CMVDST_FILE,
 * cnvt_coord.c:cnvt_coord_interp).  A redshift outside [min zs, max zs)
 * fails with FCFC_ERR_CNVT. */
int fcfc_convert_coords_table(const double *zs, const double *ds,
    int64_t nsample, int64_t n, double *x, double *y, double *z);

/* ---- deterministic catalogue generator (tests and benchmarks) ----------- */
/* Uniform box [0,L)^3. */
int fcfc_gen_uniform(uint64_t seed, int64_t n, double L, double *x,
    double *y, double *z);
/* Box with nunif uniform objects plus nhalo halos of nper Gaussian members
 * (sigma), periodic wrap; weights (if w != NULL) lognormal(0, wsig) rounded
 * to float32.  Output length nunif + nhalo*nper. */
int fcfc_gen_clustered(uint64_t seed, int64_t nunif, int64_t nhalo,
    int64_t nper, double sigma, double L, double wsig, double *x, double *y,
    double *z, double *w);
/* Survey cap: r in [rmin,rmax), |RA| <= 30 deg, |Dec| <= 30 deg, observer at
 * the origin; FKP weights w = 1/(1 + n(r) P) with n(r) a Gaussian in r. */
int fcfc_gen_survey(uint64_t seed, int64_t n, double rmin, double rmax,
    double nbar_peak, double r_peak, double r_sigma, double P0, double *x,
    double *y, double *z, double *w);
/* Log-uniform sizes in [lo,hi], rounded. */
int fcfc_gen_sizes(uint64_t seed, int n, double lo, double hi, int64_t *out);

#ifdef __cplusplus
}
#endif
#endif
