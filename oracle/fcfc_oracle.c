/*
 * fcfc_oracle.c -- CPU restatement of FCFC's FP64 pair-count arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * Python package, the drop-in executables) links, loads or calls this file.
 * It is used by tests/, by __graft_entry__.smoke() as the checker and by the
 * cpu_baseline leg of bench.py.
 *
 * What it restates (all paths relative to the FCFC v1.0.1 source tree):
 *   - per-pair distance arithmetic of the box leaf kernels
 *       src/fcfc/2pt_box/metric_common.c:140-225   (scalar)
 *       src/fcfc/2pt_box/metric_common.c:377-534   (AVX-512 / AVX2 vector)
 *       src/fcfc/2pt_box/metric_common.c:555-756   (periodic wrap, vector)
 *   - survey leaf arithmetic
 *       src/fcfc/2pt/metric_common.c:142-259       (scalar)
 *       src/fcfc/2pt/metric_common.c:283-460       (vector)
 *   - binning: table lookup == edge search (src/util/create_lut.c:58-140,
 *     src/fcfc/2pt_box/metric_common.c:56-120); mu table mutab[m] = isqrt(m)
 *     (src/fcfc/2pt_box/setup_cf.c:513-531)
 *   - periodic image choice: node-pair shift from centre differences
 *     (src/fcfc/2pt_box/metric_kdtree.c:53-68); for a pair inside the
 *     separation range the shift equals the per-pair minimum image.
 *   - histogram index s + ns * (mu|pi) (metric_common.c:99-120); auto pairs
 *     are counted once (src/fcfc/2pt_box/dual_tree.c:245-373).
 *
 * The reference's FP64 result for one pair depends on the code path its
 * kd-tree happens to take (which side is vectorised, FMA or not, the mu
 * division rounding).  This oracle evaluates one chosen variant per call and
 * can count the pairs whose bin differs between variants (the "ambiguity
 * census"): when that census is zero, every path of the reference gives the
 * same histogram as the chosen variant.
 *
 * Variants (argument `variant`):
 *   0  AVX-512, vectorised side a:   dx = (a+S)-b, FMA s^2, RZ mu division
 *   1  AVX-512, vectorised side b:   dx = (b-S)-a
 *   2  AVX-512 periodic wrap:        dx = a-b, then -L if > L/2, +L if < -L/2
 *   3  scalar (no FMA, -ffp-contract=off), a-side shift, mu = (dz2/s2)*nmu2
 *   4  scalar periodic wrap
 *   5  AVX2: FMA s^2, round-to-nearest mu division then floor
 * For the survey geometry the shift is irrelevant: 0/1/2/5 are the FMA vector
 * formulas (5 with RN division) and 3/4 the scalar ones.
 *
 * Build: see oracle/Makefile (gcc -ffp-contract=off so that the scalar
 * variants really are unfused).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define ORC_BIN_ISO 0
#define ORC_BIN_SMU 1
#define ORC_BIN_SPI 2
#define ORC_NVAR 6

typedef struct {
  int bintype;          /* 0 iso, 1 (s,mu), 2 (s_perp,pi)                   */
  int ns, nmu, np;      /* numbers of bins                                  */
  const double *s2bin;  /* ns+1 rescaled squared s (or s_perp) edges        */
  const double *pbin;   /* np+1 rescaled pi edges (box) or pi^2 edges (survey) */
} orc_bins;

typedef struct {
  int64_t n;
  const double *x, *y, *z;
  const double *s;      /* |r|^2 for the survey (smu/spi), else NULL         */
  const double *w;      /* weights or NULL                                  */
} orc_cat;

/* ------------------------------------------------------------------------ */
/* Bin search: largest k with edge[k] <= v, for v in [edge[0], edge[n]).      */
/* Equivalent to the integer / hybrid lookup tables of create_lut.c:58-140    */
/* followed by the downward refinement of FCFC_LOOKUP_HYBRID                   */
/* (metric_common.c:56-64).                                                    */
static inline int edge_search(const double *edge, int n, double v) {
  int lo = 0, hi = n;           /* edge[lo] <= v < edge[hi] */
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (v < edge[mid]) hi = mid; else lo = mid;
  }
  return lo;
}

/* integer square root; mutab[m] of setup_cf.c:513-531 equals isqrt(m). */
static inline int isqrt_i(int64_t m) {
  int64_t r = (int64_t) sqrt((double) m);
  while (r * r > m) r--;
  while ((r + 1) * (r + 1) <= m) r++;
  return (int) r;
}

/* floor of the exact quotient a/b for a >= 0, b > 0: the AVX-512 path
 * computes trunc(RZ(a/b)) (metric_common.c:484-486 with
 * _MM_FROUND_TO_ZERO then cvtt).  Round-to-nearest can only differ from RZ
 * when the RN result is an integer above the exact quotient.               */
static inline double floor_div_rz(double a, double b) {
  double q = a / b;
  double f = floor(q);
  if (f == q && q > 0) {
    if (fma(-q, b, a) < 0) f = q - 1;
  }
  return f;
}

/* Result of evaluating one pair: -1 if out of range, else histogram index. */
typedef struct { int idx; } pair_res;

/* ---------------------------- box geometry ------------------------------ */
static inline int box_pair(const double a[3], const double b[3],
    const double L[3], const double h[3], const orc_bins *bn, int var) {
  double dx, dy, dz;
  if (var == 2 || var == 4) {   /* per-pair periodic wrap */
    dx = a[0] - b[0]; if (dx > h[0]) dx -= L[0]; else if (dx < -h[0]) dx += L[0];
    dy = a[1] - b[1]; if (dy > h[1]) dy -= L[1]; else if (dy < -h[1]) dy += L[1];
    dz = a[2] - b[2]; if (dz > h[2]) dz -= L[2]; else if (dz < -h[2]) dz += L[2];
  }
  else {
    /* node-pair shift: S = -L if (a-b) > L/2, +L if (a-b) < -L/2 */
    double S[3];
    for (int k = 0; k < 3; k++) {
      double d = a[k] - b[k];
      S[k] = (d > h[k]) ? -L[k] : ((d < -h[k]) ? L[k] : 0);
    }
    if (var == 1) {   /* vectorised b side: x1 = b - S, d = x1 - a */
      dx = (b[0] - S[0]) - a[0];
      dy = (b[1] - S[1]) - a[1];
      dz = (b[2] - S[2]) - a[2];
    }
    else {
      dx = (a[0] + S[0]) - b[0];
      dy = (a[1] + S[1]) - b[1];
      dz = (a[2] + S[2]) - b[2];
    }
  }
  const int scalar = (var == 3 || var == 4);
  double s2, dz2 = 0, pi = 0;
  if (bn->bintype == ORC_BIN_SPI) {
    pi = fabs(dz);
    if (scalar) s2 = dx * dx + dy * dy;             /* :156 */
    else s2 = fma(dy, dy, dx * dx);                 /* :416-420 */
    if (!(pi < bn->pbin[bn->np]) || pi < bn->pbin[0]) return -1;
  }
  else {
    dz2 = dz * dz;
    if (scalar) s2 = dx * dx + dy * dy + dz2;       /* :160-166 */
    else s2 = fma(dy, dy, fma(dx, dx, dz2));        /* :426-430 */
  }
  if (!(s2 < bn->s2bin[bn->ns]) || s2 < bn->s2bin[0]) return -1;
  int sb = edge_search(bn->s2bin, bn->ns, s2);
  if (bn->bintype == ORC_BIN_ISO) return sb;
  if (bn->bintype == ORC_BIN_SPI) return sb + bn->ns * edge_search(bn->pbin, bn->np, pi);
  /* (s,mu) */
  const double nmu2 = (double) bn->nmu * bn->nmu;
  int64_t m;
  if (scalar) {               /* :184 */
    m = (s2 < DBL_EPSILON) ? 0 : (int64_t) ((dz2 / s2) * nmu2);
  }
  else if (var == 5) {        /* AVX2: floor(RN(RN(dz2*nmu2)/s2)) :473-477 */
    m = (s2 < DBL_EPSILON) ? 0 : (int64_t) floor((dz2 * nmu2) / s2);
  }
  else {                      /* AVX-512: trunc(RZ(RN(dz2*nmu2)/s2)) :484-486 */
    m = (s2 < DBL_EPSILON) ? 0 : (int64_t) floor_div_rz(dz2 * nmu2, s2);
  }
  if (m >= (int64_t) nmu2) return -1;                 /* mu = 1 dropped */
  return sb + bn->ns * isqrt_i(m);
}

/* --------------------------- survey geometry ---------------------------- */
static inline int survey_pair(const double a[4], const double b[4],
    const orc_bins *bn, int var) {
  const int scalar = (var == 3 || var == 4);
  if (bn->bintype == ORC_BIN_ISO) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    double s2;
    if (scalar) s2 = dx * dx + dy * dy + dz * dz;               /* :165-168 */
    else s2 = fma(dz, dz, fma(dy, dy, dx * dx));                /* :327-333 */
    if (!(s2 < bn->s2bin[bn->ns]) || s2 < bn->s2bin[0]) return -1;
    return edge_search(bn->s2bin, bn->ns, s2);
  }
  double t;
  if (scalar) t = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * 2;  /* :170 */
  else t = fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0])) * 2.0;   /* :335-347 */
  double s = a[3] + b[3];
  double d = a[3] - b[3];
  if (bn->bintype == ORC_BIN_SMU) {
    double s2 = s - t;
    if (!(s2 < bn->s2bin[bn->ns]) || s2 < bn->s2bin[0]) return -1;
    double p2 = (d * d) / (s + t);          /* scalar: d*d/(s+t) :181 */
    if (!scalar) p2 = (d * d) / (t + s);    /* vector: z2 = x2 + y2 = t + s */
    int sb = edge_search(bn->s2bin, bn->ns, s2);
    const double nmu2 = (double) bn->nmu * bn->nmu;
    int64_t m;
    if (scalar) m = (s2 < DBL_EPSILON) ? 0 : (int64_t) ((p2 / s2) * nmu2);
    else if (var == 5) m = (s2 < DBL_EPSILON) ? 0 : (int64_t) floor((p2 * nmu2) / s2);
    else m = (s2 < DBL_EPSILON) ? 0 : (int64_t) floor_div_rz(p2 * nmu2, s2);
    if (m >= (int64_t) nmu2) return -1;
    return sb + bn->ns * isqrt_i(m);
  }
  /* (s_perp, pi): pi^2 then s_perp^2 = (s - t) - pi^2 */
  double p2 = scalar ? (d * d) / (s + t) : (d * d) / (t + s);
  if (!(p2 < bn->pbin[bn->np]) || p2 < bn->pbin[0]) return -1;
  double sp2 = (s - t) - p2;
  if (!(sp2 < bn->s2bin[bn->ns])) return -1;
  if (bn->s2bin[0] != 0 && sp2 < bn->s2bin[0]) return -1;
  /* with s2min == 0 the reference has no lower check: a slightly negative
   * sp2 truncates to table index 0 (metric_common.c:207-212). */
  int sb = (sp2 < bn->s2bin[0]) ? 0 : edge_search(bn->s2bin, bn->ns, sp2);
  return sb + bn->ns * edge_search(bn->pbin, bn->np, p2);
}

/* ------------------------- neighbour cell list -------------------------- */
typedef struct {
  int nc[3];
  double lo[3], cw[3];
  int64_t *start;       /* ncell+1 */
  int64_t *idx;         /* n */
} cgrid;

static int grid_build(cgrid *g, const orc_cat *c, const double lo[3],
    const double hi[3], double rmax) {
  int64_t ncell = 1;
  for (int k = 0; k < 3; k++) {
    double ext = hi[k] - lo[k];
    int n = (int) (ext / rmax);
    if (n < 1) n = 1;
    if (n > 256) n = 256;
    g->nc[k] = n;
    g->lo[k] = lo[k];
    g->cw[k] = ext / n;
    ncell *= n;
  }
  g->start = calloc(ncell + 1, sizeof(int64_t));
  g->idx = malloc(sizeof(int64_t) * (c->n ? c->n : 1));
  int *cid = malloc(sizeof(int) * (c->n ? c->n : 1));
  if (!g->start || !g->idx || !cid) return -1;
  const double *p[3] = {c->x, c->y, c->z};
  for (int64_t i = 0; i < c->n; i++) {
    int id = 0;
    for (int k = 0; k < 3; k++) {
      int q = (int) floor((p[k][i] - g->lo[k]) / g->cw[k]);
      if (q < 0) q = 0;
      if (q >= g->nc[k]) q = g->nc[k] - 1;
      id = id * g->nc[k] + q;
    }
    cid[i] = id;
    g->start[id + 1]++;
  }
  for (int64_t i = 0; i < ncell; i++) g->start[i + 1] += g->start[i];
  int64_t *fill = malloc(sizeof(int64_t) * ncell);
  memcpy(fill, g->start, sizeof(int64_t) * ncell);
  for (int64_t i = 0; i < c->n; i++) g->idx[fill[cid[i]]++] = i;
  free(fill);
  free(cid);
  return 0;
}

static void grid_free(cgrid *g) { free(g->start); free(g->idx); }

/*
 * Core loop.  For every pair in range (each unordered pair once for auto
 * counts) evaluate variant `var`, accumulate into icnt/dcnt, and optionally
 * count the pairs whose result differs between any two variants of
 * `census_mask` (bit v = include variant v).  If `grad` is non-NULL the
 * weight gradient g_a[i] += G[bin] * w_b[j] (and g_b[j] += G[bin] * w_a[i])
 * is accumulated with the canonical variant's bins.
 */
static int run_pairs(int geom, const double *bsize, const orc_cat *A,
    const orc_cat *B, int isauto, const orc_bins *bn, int var,
    int census_mask, int64_t *icnt, double *dcnt, int64_t *census,
    const double *G, double *ga, double *gb, int nthread) {
  const int64_t ntot = (int64_t) bn->ns *
      (bn->bintype == ORC_BIN_ISO ? 1 : (bn->bintype == ORC_BIN_SMU ? bn->nmu : bn->np));
  /* maximum 3-D separation of interest */
  double r2max = bn->s2bin[bn->ns];
  if (bn->bintype == ORC_BIN_SPI) {
    double pm = bn->pbin[bn->np];
    r2max += (geom == 0) ? pm * pm : pm;     /* survey pbin holds pi^2 */
  }
  /* survey smu/spi: the 4-D formula has absolute rounding error of order
   * 1e-10; pad the neighbour radius generously (pruning only). */
  double rmax = sqrt(r2max) * (1 + 1e-6) + 1e-6;
  double lo[3], hi[3];
  if (geom == 0) {
    for (int k = 0; k < 3; k++) { lo[k] = 0; hi[k] = bsize[k]; }
  }
  else {
    for (int k = 0; k < 3; k++) { lo[k] = 1e300; hi[k] = -1e300; }
    const orc_cat *cs[2] = {A, B};
    for (int c = 0; c < 2; c++) {
      const double *p[3] = {cs[c]->x, cs[c]->y, cs[c]->z};
      for (int64_t i = 0; i < cs[c]->n; i++)
        for (int k = 0; k < 3; k++) {
          if (p[k][i] < lo[k]) lo[k] = p[k][i];
          if (p[k][i] > hi[k]) hi[k] = p[k][i];
        }
    }
    for (int k = 0; k < 3; k++) { hi[k] += 1e-9 * (1 + fabs(hi[k])); if (hi[k] <= lo[k]) hi[k] = lo[k] + 1; }
  }
  /* cells of half the neighbour radius (5^3 cells searched instead of 3^3
   * of the full radius: 2.5 times fewer candidates) */
  cgrid g;
  if (grid_build(&g, B, lo, hi, 0.5 * rmax)) return -1;
  /* candidates beyond the padded radius are out of range on every rounding
   * path (the paths differ by a few ulps of s^2; the survey formula by
   * ~1e-10 absolute): skipped before any variant is evaluated */
  const double r2pad = r2max * (1 + 1e-6) + 1e-6;
  double L[3] = {0, 0, 0}, h[3] = {0, 0, 0};
  if (geom == 0) for (int k = 0; k < 3; k++) { L[k] = bsize[k]; h[k] = bsize[k] * 0.5; }
  const int withwt = (A->w != NULL) || (B->w != NULL);
  int64_t cens_tot = 0;
  int reach[3];
  for (int k = 0; k < 3; k++) {
    reach[k] = (int) ceil(rmax / g.cw[k]);
    if (geom == 0 && 2 * reach[k] + 1 > g.nc[k]) reach[k] = -1;   /* all cells */
  }

#pragma omp parallel num_threads(nthread) reduction(+:cens_tot)
  {
    int64_t *ic = calloc(ntot, sizeof(int64_t));
    double *dc = calloc(ntot, sizeof(double));
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < A->n; i++) {
      double a[4] = {A->x[i], A->y[i], A->z[i], A->s ? A->s[i] : 0};
      const double wa = A->w ? A->w[i] : 1.0;
      double gacc = 0;
      int cc[3];
      for (int k = 0; k < 3; k++) {
        int q = (int) floor((a[k] - g.lo[k]) / g.cw[k]);
        if (q < 0) q = 0;
        if (q >= g.nc[k]) q = g.nc[k] - 1;
        cc[k] = q;
      }
      int r0[3], r1[3];
      for (int k = 0; k < 3; k++) {
        if (reach[k] < 0) { r0[k] = 0; r1[k] = g.nc[k] - 1; }
        else if (geom == 0) { r0[k] = cc[k] - reach[k]; r1[k] = cc[k] + reach[k]; }
        else {
          r0[k] = cc[k] - reach[k]; r1[k] = cc[k] + reach[k];
          if (r0[k] < 0) r0[k] = 0;
          if (r1[k] >= g.nc[k]) r1[k] = g.nc[k] - 1;
        }
      }
      for (int ox = r0[0]; ox <= r1[0]; ox++) {
        int qx = reach[0] < 0 ? ox : ((ox % g.nc[0]) + g.nc[0]) % g.nc[0];
        for (int oy = r0[1]; oy <= r1[1]; oy++) {
          int qy = reach[1] < 0 ? oy : ((oy % g.nc[1]) + g.nc[1]) % g.nc[1];
          for (int oz = r0[2]; oz <= r1[2]; oz++) {
            int qz = reach[2] < 0 ? oz : ((oz % g.nc[2]) + g.nc[2]) % g.nc[2];
            int64_t cid = ((int64_t) qx * g.nc[1] + qy) * g.nc[2] + qz;
            for (int64_t t = g.start[cid]; t < g.start[cid + 1]; t++) {
              int64_t j = g.idx[t];
              if (isauto && j <= i) continue;
              double b[4] = {B->x[j], B->y[j], B->z[j], B->s ? B->s[j] : 0};
              {
                double q = 0;
                for (int k = 0; k < 3; k++) {
                  double d = a[k] - b[k];
                  if (geom == 0) { if (d > h[k]) d -= L[k]; else if (d < -h[k]) d += L[k]; }
                  q += d * d;
                }
                if (q > r2pad) continue;
              }
              int r = (geom == 0) ? box_pair(a, b, L, h, bn, var)
                                  : survey_pair(a, b, bn, var);
              if (census_mask) {
                int amb = 0;
                for (int v = 0; v < ORC_NVAR; v++) {
                  if (!(census_mask & (1 << v)) || v == var) continue;
                  int rv = (geom == 0) ? box_pair(a, b, L, h, bn, v)
                                       : survey_pair(a, b, bn, v);
                  if (rv != r) { amb = 1; break; }
                }
                cens_tot += amb;
              }
              if (r < 0) continue;
              const double wb = B->w ? B->w[j] : 1.0;
              if (withwt) dc[r] += wa * wb; else ic[r] += 1;
              if (G) {
                gacc += G[r] * wb;
                if (gb) {
#pragma omp atomic
                  gb[j] += G[r] * wa;
                }
              }
            }
          }
        }
      }
      if (G && ga) {
        /* atomic: for auto gradients ga and gb are the same array, which
         * other threads update at index i as their pair partner */
#pragma omp atomic
        ga[i] += gacc;
      }
    }
#pragma omp critical
    {
      for (int64_t k = 0; k < ntot; k++) {
        if (icnt) icnt[k] += ic[k];
        if (dcnt) dcnt[k] += dc[k];
      }
    }
    free(ic); free(dc);
  }
  grid_free(&g);
  if (census) *census += cens_tot;
  return 0;
}

/* Public entry points (ctypes). ----------------------------------------- */

/* Pair counts with the reference arithmetic variant `var`.  geom 0 = periodic
 * box (bsize = rescaled side lengths), 1 = survey (bsize ignored).  Counts are
 * ADDED to icnt (unweighted) or dcnt (weighted); auto counts are unordered
 * pairs (the caller doubles them, eval_cf.c:125-133). */
int orc_count(int geom, const double *bsize, const orc_cat *A,
    const orc_cat *B, int isauto, const orc_bins *bn, int var,
    int census_mask, int64_t *icnt, double *dcnt, int64_t *census,
    int nthread) {
  return run_pairs(geom, bsize, A, isauto ? A : B, isauto, bn, var,
      census_mask, icnt, dcnt, census, NULL, NULL, NULL, nthread);
}

/* FP64 weight gradient of sum_k G[k] * cnt[k] with cnt the weighted counts
 * (canonical variant 0), i.e. ga[i] += sum_j G[bin(i,j)] w_b[j] and, for
 * cross counts, gb[j] += sum_i G[bin(i,j)] w_a[i].  For auto counts every
 * unordered pair contributes to both of its members. */
int orc_weight_grad(int geom, const double *bsize, const orc_cat *A,
    const orc_cat *B, int isauto, const orc_bins *bn, const double *G,
    double *ga, double *gb, int nthread) {
  if (isauto) {
    /* both members of an unordered pair get a contribution: use gb = ga */
    return run_pairs(geom, bsize, A, A, 1, bn, 0, 0, NULL, NULL, NULL, G,
        ga, ga, nthread);
  }
  return run_pairs(geom, bsize, A, B, 0, bn, 0, 0, NULL, NULL, NULL, G, ga,
      gb, nthread);
}

/* Survey |r|^2 exactly as data_preprocess (src/fcfc/2pt/build_tree.c:35-133)
 * for an AVX-512 build (lanes = 8) or AVX2 (lanes = 4) or scalar (lanes = 0):
 * FMA for the first floor(n/lanes)*lanes points, unfused for the rest. */
void orc_survey_norm2(int64_t n, const double *x, const double *y,
    const double *z, int lanes, double *s) {
  int64_t nv = lanes ? (n / lanes) * lanes : 0;
  for (int64_t i = 0; i < nv; i++) s[i] = fma(z[i], z[i], fma(y[i], y[i], x[i] * x[i]));
  for (int64_t i = nv; i < n; i++) s[i] = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
}
