/*
 * fcfc_host.c -- host side of the pair-count engine (C99):
 *   - separation-bin setup and lookup tables, following FCFC's cf_setup
 *     (src/fcfc/2pt_box/setup_cf.c:298-531, src/fcfc/2pt/setup_cf.c:433-502,
 *     src/util/create_lut.c:58-140);
 *   - catalogue preparation: rescaling (src/fcfc/2pt_box/build_tree.c:121-156)
 *     and survey |r|^2 (src/fcfc/2pt/build_tree.c:35-133);
 *   - pair-count files in FCFC's binary / ASCII layout
 *     (src/fcfc/2pt_box/save_res.c:45-296).
 * Compiled with -ffp-contract=off: every floating-point operation here is
 * rounded exactly as written, like FCFC's own -std=c99 build.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fcfc_gpu.h"

#define BIN_MAX_DIGIT_TO_INT    4        /* define.h:87 */
#define LOOKUP_INT_MAX_SIZE     40960    /* define.h:80 */
#define LOOKUP_HYBRID_MAX_SIZE  32768    /* define.h:81 */
#define LOOKUP_HYBRID_MIN_SIZE  8192     /* define.h:82 */
#define REAL_TOL                1e-10    /* define_comm.h DOUBLE_TOL */
#define REAL_MAXINT             0x1p+53
#define MAX_NSBIN               10000    /* define_comm.h:147 */
#define MAX_NMU                 255      /* define_comm.h:148 */

/* --------------------------------------------------------------------- */
/* Rescaling factors: the least factor turning the bin parameters into
 * integers (setup_cf.c:298-376).                                          */

static size_t gcd_sz(size_t a, size_t b) {
  while (b) { size_t t = b; b = a % b; a = t; }
  return a;
}

static double least_fac2(double a, double b) {
  if (a < 0 || a > REAL_MAXINT || b < 0 || b > REAL_MAXINT) return 0;
  if (a == 0) return 1 / b;
  if (b == 0) return 1 / a;
  size_t i = 0, ifac = 1;
  do {
    if (fabs(round(a) - a) < REAL_TOL && fabs(round(b) - b) < REAL_TOL) break;
    a *= 10; b *= 10; ifac *= 10; i++;
  } while (i <= BIN_MAX_DIGIT_TO_INT);
  if (i == BIN_MAX_DIGIT_TO_INT + 1 || a > REAL_MAXINT || b > REAL_MAXINT)
    return 0;
  size_t g = gcd_sz((size_t) round(a), (size_t) round(b));
  return ifac / (double) g;
}

static double least_fac4(double a, double b, double c, double d) {
  if (a < 0 || a > REAL_MAXINT || b < 0 || b > REAL_MAXINT ||
      c < 0 || c > REAL_MAXINT || d < 0 || d > REAL_MAXINT) return 0;
  size_t i = 0, ifac = 1;
  do {
    if (fabs(round(a) - a) < REAL_TOL && fabs(round(b) - b) < REAL_TOL &&
        fabs(round(c) - c) < REAL_TOL && fabs(round(d) - d) < REAL_TOL) break;
    a *= 10; b *= 10; c *= 10; d *= 10; ifac *= 10; i++;
  } while (i <= BIN_MAX_DIGIT_TO_INT);
  if (i == BIN_MAX_DIGIT_TO_INT + 1 || a > REAL_MAXINT || b > REAL_MAXINT ||
      c > REAL_MAXINT || d > REAL_MAXINT) return 0;
  /* the reference mixes rounded and truncated operands here
   * (setup_cf.c:367-373); keep the same conversions */
  size_t g = 0;
  if (a != 0) g = (size_t) round(a);
  if (b != 0) g = g ? gcd_sz(g, (size_t) b) : (size_t) round(b);
  if (c != 0) g = g ? gcd_sz(g, (size_t) c) : (size_t) round(c);
  if (d != 0) g = g ? gcd_sz(g, (size_t) d) : (size_t) round(d);
  if (g == 0) g = 1;
  return ifac / (double) g;
}

/* Truncate the fraction of a positive number to a power of 2
 * (define_comm.h REAL_TRUNC_FRAC: pow(FLT_RADIX, logb(x))). */
static double trunc_frac(double x) { return pow(2.0, logb(x)); }

/* Integer lookup table over [edge[0], edge[n]) (create_lut.c:58-93):
 * tab[k] = bin of the integer k + edge[0]. */
static uint16_t *lut_int(const double *edge, int n, int *len) {
  const long lo = (long) edge[0], hi = (long) edge[n];
  const long ntab = hi - lo;
  uint16_t *tab = calloc((size_t) (ntab > 0 ? ntab : 1), sizeof *tab);
  if (!tab) return NULL;
  int b = 1;
  for (long i = 0; i < ntab; i++) {
    if (i + lo < edge[b]) tab[i] = (uint16_t) (b - 1);
    else {
      if (++b > n) { free(tab); return NULL; }
      i--;
    }
  }
  *len = (int) ntab;
  return tab;
}

/* Hybrid table (create_lut.c:95-140): entries below n are exact bins,
 * entries n + k mean "bin k or lower, refine against the edges". */
static uint16_t *lut_hybrid(const double *edge, int n, int *len) {
  const long lo = (long) edge[0];
  const long hi = (long) ceil(edge[n]);
  const long ntab = hi - lo;
  uint16_t *tab = calloc((size_t) (ntab > 0 ? ntab : 1), sizeof *tab);
  if (!tab) return NULL;
  int b = 1;
  long e = (long) edge[b];
  for (long i = 0; i < ntab; i++) {
    if (i + lo < e) tab[i] = (uint16_t) (b - 1);
    else {
      while (++b <= n) {
        e = (long) edge[b];
        if (i + lo < e) { tab[i] = (uint16_t) (n + b - 1); break; }
      }
      if (b > n) { for (; i < ntab; i++) tab[i] = (uint16_t) (n * 2); break; }
    }
  }
  *len = (int) ntab;
  return tab;
}

void fcfc_bins_free(fcfc_bins *b) {
  if (!b) return;
  free(b->sbin_raw); free(b->pbin_raw); free(b->sbin); free(b->s2bin);
  free(b->pbin); free(b->stab); free(b->ptab);
  memset(b, 0, sizeof *b);
}

/* Shared part of the two setup routes.  `lin` != 0: linear bins given by
 * (smin, ds[, pmin, dpi]) so that the integer-table route may apply. */
static int bins_finish(fcfc_bins *b, int lin, double smin, double ds,
    double pmin, double dpi) {
  const int ns = b->ns, np = b->np;
  if (b->bintype == FCFC_BIN_SPI) {
    const int survey = (b->geometry == FCFC_GEOM_SURVEY);
    if (lin) {          /* create_tab_sp_pi, integer route (setup_cf.c:439-478) */
      const double fac = least_fac4(smin, ds, pmin, dpi);
      if (fac != 0) {
        double s1 = b->sbin_raw[0] * fac; s1 = round(s1 * s1);
        double s2 = b->sbin_raw[ns] * fac; s2 = round(s2 * s2);
        if (s1 <= REAL_MAXINT && s2 <= REAL_MAXINT) {
          /* the length check uses the unsquared pi range in both geometries
           * (2pt/setup_cf.c:443-447), although the survey table spans pi^2 */
          double q1 = b->pbin_raw[0] * fac, q2 = b->pbin_raw[np] * fac;
          size_t nts = (size_t) s2 - (size_t) s1;
          size_t ntp = (size_t) q2 - (size_t) q1;
          if (nts <= LOOKUP_INT_MAX_SIZE && ntp <= LOOKUP_INT_MAX_SIZE) {
            b->rescale = fac;
            for (int i = 0; i <= ns; i++) {
              b->sbin[i] = round(b->sbin_raw[i] * fac);
              b->s2bin[i] = b->sbin[i] * b->sbin[i];
            }
            for (int i = 0; i <= np; i++) {
              double p = round(b->pbin_raw[i] * fac);
              b->pbin[i] = survey ? p * p : p;
            }
            b->tabtype = FCFC_LOOKUP_TYPE_INT;
            b->stab = lut_int(b->s2bin, ns, &b->stab_len);
            b->ptab = lut_int(b->pbin, np, &b->ptab_len);
            return (b->stab && b->ptab) ? 0 : FCFC_ERR_MEMORY;
          }
        }
      }
    }
    /* hybrid route (setup_cf.c:480-503) */
    double smax = b->sbin_raw[ns]; smax *= smax;
    double pmax = b->pbin_raw[np];
    if (survey) pmax *= pmax;
    double larger = (smax >= pmax) ? smax : pmax;
    double smaller = (smax >= pmax) ? pmax : smax;
    double fac = LOOKUP_HYBRID_MAX_SIZE / larger;
    fac = trunc_frac(fac);
    while (fac * smaller > LOOKUP_HYBRID_MIN_SIZE * 2) fac *= 0.5;
    b->rescale = fac;
    for (int i = 0; i <= ns; i++) {
      b->sbin[i] = b->sbin_raw[i] * fac;
      b->s2bin[i] = b->sbin[i] * b->sbin[i];
    }
    for (int i = 0; i <= np; i++) {
      double p = b->pbin_raw[i] * fac;
      b->pbin[i] = survey ? p * p : p;
    }
    b->tabtype = FCFC_LOOKUP_TYPE_HYBRID;
    b->stab = lut_hybrid(b->s2bin, ns, &b->stab_len);
    b->ptab = lut_hybrid(b->pbin, np, &b->ptab_len);
    return (b->stab && b->ptab) ? 0 : FCFC_ERR_MEMORY;
  }
  /* iso / smu: create_tab_sbin (setup_cf.c:385-430) */
  if (lin) {
    const double fac = least_fac2(smin, ds);
    if (fac != 0) {
      double s1 = b->sbin_raw[0] * fac; s1 = round(s1 * s1);
      double s2 = b->sbin_raw[ns] * fac; s2 = round(s2 * s2);
      if (s1 <= REAL_MAXINT && s2 <= REAL_MAXINT) {
        size_t nt = (size_t) s2 - (size_t) s1;
        if (nt <= LOOKUP_INT_MAX_SIZE) {
          b->rescale = fac;
          for (int i = 0; i <= ns; i++) {
            b->sbin[i] = round(b->sbin_raw[i] * fac);
            b->s2bin[i] = b->sbin[i] * b->sbin[i];
          }
          b->tabtype = FCFC_LOOKUP_TYPE_INT;
          b->stab = lut_int(b->s2bin, ns, &b->stab_len);
          return b->stab ? 0 : FCFC_ERR_MEMORY;
        }
      }
    }
  }
  double smax = b->sbin_raw[ns]; smax *= smax;
  double fac = LOOKUP_HYBRID_MAX_SIZE / smax;
  fac = trunc_frac(fac);
  while (fac * smax > LOOKUP_HYBRID_MIN_SIZE * 2) fac *= 0.5;
  b->rescale = fac;
  for (int i = 0; i <= ns; i++) {
    b->sbin[i] = b->sbin_raw[i] * fac;
    b->s2bin[i] = b->sbin[i] * b->sbin[i];
  }
  b->tabtype = FCFC_LOOKUP_TYPE_HYBRID;
  b->stab = lut_hybrid(b->s2bin, ns, &b->stab_len);
  return b->stab ? 0 : FCFC_ERR_MEMORY;
}

static int bins_alloc(fcfc_bins *b, int geometry, int bintype, int ns,
    int nmu, int np) {
  memset(b, 0, sizeof *b);
  if (geometry != FCFC_GEOM_BOX && geometry != FCFC_GEOM_SURVEY) return FCFC_ERR_ARG;
  if (bintype < FCFC_BIN_ISO || bintype > FCFC_BIN_SPI) return FCFC_ERR_ARG;
  if (ns < 1 || ns > MAX_NSBIN) return FCFC_ERR_CFG;
  if (bintype == FCFC_BIN_SMU && (nmu <= 1 || nmu > MAX_NMU)) return FCFC_ERR_CFG;
  if (bintype == FCFC_BIN_SPI && (np < 1 || np > MAX_NSBIN)) return FCFC_ERR_CFG;
  b->geometry = geometry;
  b->bintype = bintype;
  b->ns = ns;
  b->nmu = (bintype == FCFC_BIN_SMU) ? nmu : 1;
  b->np = (bintype == FCFC_BIN_SPI) ? np : 0;
  b->ntot = ns * (bintype == FCFC_BIN_SMU ? b->nmu : (bintype == FCFC_BIN_SPI ? np : 1));
  if (!(b->sbin_raw = malloc(sizeof(double) * (ns + 1))) ||
      !(b->sbin = malloc(sizeof(double) * (ns + 1))) ||
      !(b->s2bin = malloc(sizeof(double) * (ns + 1)))) return FCFC_ERR_MEMORY;
  if (bintype == FCFC_BIN_SPI) {
    if (!(b->pbin_raw = malloc(sizeof(double) * (np + 1))) ||
        !(b->pbin = malloc(sizeof(double) * (np + 1)))) return FCFC_ERR_MEMORY;
  }
  return 0;
}

/* Number of linear bins as conf_verify counts them (load_conf.c:981-1005):
 * step from min by width until max - tol. */
static int count_linear(double lo, double hi, double w, double *top) {
  if (!isfinite(w) || w <= 0) return -1;
  if (lo + w > hi + REAL_TOL) return -1;
  double m = lo;
  int n = 0;
  while (m < hi - REAL_TOL) {
    m += w;
    if (++n > MAX_NSBIN) return -1;
  }
  *top = m;
  return n;
}

int fcfc_bins_linear(fcfc_bins *b, int geometry, int bintype, double smin,
    double smax, double ds, int nmu, double pmin, double pmax, double dpi,
    const double *bsize) {
  if (!b) return FCFC_ERR_ARG;
  double stop = 0, ptop = 0;
  int ns = count_linear(smin, smax, ds, &stop);
  if (ns < 1 || smin < 0) return FCFC_ERR_CFG;
  int np = 0;
  if (bintype == FCFC_BIN_SPI) {
    np = count_linear(pmin, pmax, dpi, &ptop);
    if (np < 1 || pmin < 0) return FCFC_ERR_CFG;
  }
  if (geometry == FCFC_GEOM_BOX) {
    if (!bsize) return FCFC_ERR_ARG;
    for (int k = 0; k < 3; k++) {
      if (stop >= bsize[k] * 0.5) return FCFC_ERR_CFG;
      if (bintype == FCFC_BIN_SPI && ptop >= bsize[k] * 0.5) return FCFC_ERR_CFG;
    }
  }
  int e = bins_alloc(b, geometry, bintype, ns, nmu, np);
  if (e) { fcfc_bins_free(b); return e; }
  for (int i = 0; i <= ns; i++) b->sbin_raw[i] = smin + ds * i;   /* setup_cf.c:614 */
  if (bintype == FCFC_BIN_SPI)
    for (int i = 0; i <= b->np; i++) b->pbin_raw[i] = pmin + dpi * i;
  if ((e = bins_finish(b, 1, smin, ds, pmin, dpi))) { fcfc_bins_free(b); return e; }
  b->soff = (int) b->s2bin[0];
  b->poff = (b->bintype == FCFC_BIN_SPI) ? (int) b->pbin[0] : 0;
  return 0;
}

int fcfc_bins_edges(fcfc_bins *b, int geometry, int bintype,
    const double *sedges, int ns, int nmu, const double *pedges, int np,
    const double *bsize) {
  if (!b || !sedges) return FCFC_ERR_ARG;
  if (bintype == FCFC_BIN_SPI && !pedges) return FCFC_ERR_ARG;
  for (int i = 0; i < ns; i++) if (!(sedges[i + 1] > sedges[i])) return FCFC_ERR_CFG;
  if (sedges[0] < 0) return FCFC_ERR_CFG;
  if (bintype == FCFC_BIN_SPI) {
    for (int i = 0; i < np; i++) if (!(pedges[i + 1] > pedges[i])) return FCFC_ERR_CFG;
    if (pedges[0] < 0) return FCFC_ERR_CFG;
  }
  if (geometry == FCFC_GEOM_BOX) {
    if (!bsize) return FCFC_ERR_ARG;
    for (int k = 0; k < 3; k++) {
      if (sedges[ns] >= bsize[k] * 0.5) return FCFC_ERR_CFG;
      if (bintype == FCFC_BIN_SPI && pedges[np] >= bsize[k] * 0.5) return FCFC_ERR_CFG;
    }
  }
  int e = bins_alloc(b, geometry, bintype, ns, nmu, np);
  if (e) { fcfc_bins_free(b); return e; }
  memcpy(b->sbin_raw, sedges, sizeof(double) * (ns + 1));
  if (bintype == FCFC_BIN_SPI) memcpy(b->pbin_raw, pedges, sizeof(double) * (np + 1));
  if ((e = bins_finish(b, 0, 0, 0, 0, 0))) { fcfc_bins_free(b); return e; }
  b->soff = (int) b->s2bin[0];
  b->poff = (b->bintype == FCFC_BIN_SPI) ? (int) b->pbin[0] : 0;
  return 0;
}

/* --------------------------------------------------------------------- */
int fcfc_prepare_catalog(const fcfc_bins *b, int64_t n, double *x, double *y,
    double *z, double *s, const double *w, double *sumw) {
  if (!b || n < 0 || (n && (!x || !y || !z))) return FCFC_ERR_ARG;
  const double f = b->rescale;
  if (f != 1) {
    for (int64_t i = 0; i < n; i++) { x[i] *= f; y[i] *= f; z[i] *= f; }
  }
  if (s) {
    /* data_preprocess lane rule for an AVX-512 build: FMA in 8-wide blocks,
     * unfused for the n % 8 tail (2pt/build_tree.c:59-92) */
    const int64_t nv = (n / 8) * 8;
    for (int64_t i = 0; i < nv; i++)
      s[i] = fma(z[i], z[i], fma(y[i], y[i], x[i] * x[i]));
    for (int64_t i = nv; i < n; i++)
      s[i] = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
  }
  if (sumw) {
    if (w) { double t = 0; for (int64_t i = 0; i < n; i++) t += w[i]; *sumw = t; }
    else *sumw = (double) n;
  }
  return 0;
}

/* --------------------------------------------------------------------- */
/* Pair-count files: save_res.c:45-296.  The writer emits the same header
 * line, the same Fortran-style blocks [u64 size][payload][u64 size] and the
 * same "%.10lg" text. */

static int wblock(FILE *fp, const void *p, uint64_t size) {
  if (fwrite(&size, sizeof size, 1, fp) != 1) return FCFC_ERR_FILE;
  if (size && fwrite(p, size, 1, fp) != 1) return FCFC_ERR_FILE;
  if (fwrite(&size, sizeof size, 1, fp) != 1) return FCFC_ERR_FILE;
  return 0;
}

#define OFMT "%.10lg"

int fcfc_save_pair_count(const char *fname, const fcfc_bins *b, int ofmt,
    const double *bsize, int isauto, int weighted, uint64_t n1, uint64_t n2,
    double wt1, double wt2, const void *cnt) {
  if (!fname || !b || !cnt) return FCFC_ERR_ARG;
  const int box = (b->geometry == FCFC_GEOM_BOX);
  if (box && !bsize) return FCFC_ERR_ARG;
  const char *code = box ? "FCFC_2PT_BOX" : "FCFC_2PT";
  const size_t ntot = (size_t) b->ntot;
  double norm;
  if (isauto) norm = weighted ? wt1 * (wt1 - 1) : (double) n1 * (n1 - 1);
  else norm = weighted ? wt1 * wt2 : (double) n1 * n2;
  double *ncnt = malloc(sizeof(double) * (ntot ? ntot : 1));
  if (!ncnt) return FCFC_ERR_MEMORY;
  for (size_t k = 0; k < ntot; k++)
    ncnt[k] = (weighted ? ((const double *) cnt)[k]
                        : (double) ((const int64_t *) cnt)[k]) / norm;
  FILE *fp = fopen(fname, ofmt == 0 ? "wb" : "w");
  if (!fp) { free(ncnt); return FCFC_ERR_FILE; }
  int e = 0;
  fprintf(fp, "# Created by %s v1.0.1 (2) ; format = %d\n", code, ofmt);
  if (ofmt == 0) {
    int32_t spec[8];
    spec[0] = 1;                              /* double precision */
    spec[1] = 0;                              /* mu = 1 excluded */
    spec[2] = box ? 1 : 0;                    /* periodic metric */
    spec[3] = isauto ? 0 : 1;                 /* FCFC_COUNT_AUTO/CROSS */
    spec[4] = weighted ? 1 : 0;
    spec[5] = b->bintype;
    spec[6] = b->ns;
    spec[7] = (b->bintype == FCFC_BIN_SMU) ? b->nmu : b->np;
    uint64_t sz = (b->bintype == FCFC_BIN_ISO) ? 7 * sizeof(int32_t) : 8 * sizeof(int32_t);
    if (!e) e = wblock(fp, spec, sz);
    if (box && !e) e = wblock(fp, bsize, 3 * sizeof(double));
    if (!e) e = wblock(fp, b->sbin_raw, (uint64_t) (b->ns + 1) * sizeof(double));
    if (b->bintype == FCFC_BIN_SPI && !e)
      e = wblock(fp, b->pbin_raw, (uint64_t) (b->np + 1) * sizeof(double));
    if (!e) {
      if (isauto) {
        if (weighted) e = wblock(fp, &wt1, sizeof(double));
        else { uint64_t v = n1; e = wblock(fp, &v, sizeof v); }
      }
      else {
        if (weighted) { double v[2] = {wt1, wt2}; e = wblock(fp, v, sizeof v); }
        else { uint64_t v[2] = {n1, n2}; e = wblock(fp, v, sizeof v); }
      }
    }
    if (!e) e = wblock(fp, &norm, sizeof(double));
    if (!e) e = wblock(fp, ncnt, ntot * sizeof(double));
    if (!e) e = wblock(fp, cnt, ntot * 8);
  }
  else {
    if (isauto) {
      fprintf(fp, "# Number of tracers: %llu (auto pair counts", (unsigned long long) n1);
      if (weighted) fprintf(fp, ", weighted number: " OFMT " ", wt1);
      fprintf(fp, ")\n");
    }
    else {
      fprintf(fp, "# Numbers of tracers: %llu and %llu", (unsigned long long) n1,
          (unsigned long long) n2);
      if (weighted) fprintf(fp, " (weighted numbers: " OFMT " and " OFMT " )", wt1, wt2);
      fprintf(fp, "\n");
    }
    fprintf(fp, "# Normalization (total number of pairs): " OFMT "\n", norm);
    if (b->bintype == FCFC_BIN_SMU) {
      fprintf(fp, "# Columns: s_min(1) s_max(2) mu_min(3) mu_max(4) normalized_pair_count(5)\n");
      for (int j = 0; j < b->nmu; j++)
        for (int i = 0; i < b->ns; i++)
          fprintf(fp, OFMT " " OFMT " " OFMT " " OFMT " " OFMT "\n", b->sbin_raw[i],
              b->sbin_raw[i + 1], j / (double) b->nmu, (j + 1) / (double) b->nmu,
              ncnt[i + (size_t) j * b->ns]);
    }
    else if (b->bintype == FCFC_BIN_SPI) {
      fprintf(fp, "# Columns: s_perp_min(1) s_perp_max(2) pi_min(3) pi_max(4) normalized_pair_count(5)\n");
      for (int j = 0; j < b->np; j++)
        for (int i = 0; i < b->ns; i++)
          fprintf(fp, OFMT " " OFMT " " OFMT " " OFMT " " OFMT "\n", b->sbin_raw[i],
              b->sbin_raw[i + 1], b->pbin_raw[j], b->pbin_raw[j + 1],
              ncnt[i + (size_t) j * b->ns]);
    }
    else {
      fprintf(fp, "# Columns: s_min(1) s_max(2) normalized_pair_count(3)\n");
      for (int i = 0; i < b->ns; i++)
        fprintf(fp, OFMT " " OFMT " " OFMT "\n", b->sbin_raw[i], b->sbin_raw[i + 1], ncnt[i]);
    }
  }
  if (ferror(fp)) e = FCFC_ERR_FILE;
  if (fclose(fp)) e = FCFC_ERR_FILE;
  free(ncnt);
  return e;
}

/* --------------------------------------------------------------------- */
/* Deterministic generator.  xoshiro256** (Blackman & Vigna 2018) seeded by
 * splitmix64; uniform doubles (u >> 11) * 2^-53.  log and exp are written
 * here from their textbook series with only + - * / so that catalogues are
 * bit-identical on every host (this file is built with -ffp-contract=off):
 *   log(x) = k ln2 + 2 atanh((m-1)/(m+1)),  m in [sqrt(1/2), sqrt(2))
 *   exp(x) = 2^k * sum_{i<=24} r^i / i!,    |r| <= ln2 / 2                  */

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static void rng_seed(rng_t *r, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rng_next(rng_t *r) {
  uint64_t *s = r->s;
  const uint64_t res = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
  s[2] ^= t; s[3] = rotl(s[3], 45);
  return res;
}
static inline double rng_u01(rng_t *r) { return (rng_next(r) >> 11) * 0x1p-53; }

static const double LN2_HI = 0x1.62e42fefa39efp-1;

static double gen_log(double x) {        /* x > 0, finite */
  int k;
  double m = frexp(x, &k);                /* x = m 2^k, m in [0.5, 1) */
  if (m < 0x1.6a09e667f3bcdp-1) { m *= 2; k -= 1; }   /* m in [sqrt(.5), sqrt(2)) */
  double f = (m - 1) / (m + 1);
  double f2 = f * f, term = f, sum = 0;
  for (int i = 1; i <= 41; i += 2) { sum += term / i; term *= f2; }
  return k * LN2_HI + 2 * sum;
}

static double gen_exp(double x) {
  double kd = floor(x / LN2_HI + 0.5);
  double r = x - kd * LN2_HI;
  double sum = 1, term = 1;
  for (int i = 1; i <= 24; i++) { term = term * r / i; sum += term; }
  return ldexp(sum, (int) kd);
}

/* Marsaglia polar method. */
static double gen_gauss(rng_t *r, int *have, double *spare) {
  if (*have) { *have = 0; return *spare; }
  double u, v, s;
  do {
    u = 2 * rng_u01(r) - 1;
    v = 2 * rng_u01(r) - 1;
    s = u * u + v * v;
  } while (s >= 1 || s == 0);
  double f = sqrt(-2 * gen_log(s) / s);
  *spare = v * f; *have = 1;
  return u * f;
}

static inline double wrap_box(double x, double L) {
  x = x - L * floor(x / L);
  if (x >= L || x < 0) x = 0;
  return x;
}

int fcfc_gen_uniform(uint64_t seed, int64_t n, double L, double *x,
    double *y, double *z) {
  if (n < 0 || !(L > 0)) return FCFC_ERR_ARG;
  rng_t r; rng_seed(&r, seed);
  for (int64_t i = 0; i < n; i++) {
    x[i] = wrap_box(rng_u01(&r) * L, L);
    y[i] = wrap_box(rng_u01(&r) * L, L);
    z[i] = wrap_box(rng_u01(&r) * L, L);
  }
  return 0;
}

int fcfc_gen_clustered(uint64_t seed, int64_t nunif, int64_t nhalo,
    int64_t nper, double sigma, double L, double wsig, double *x, double *y,
    double *z, double *w) {
  if (nunif < 0 || nhalo < 0 || nper < 0 || !(L > 0)) return FCFC_ERR_ARG;
  rng_t r; rng_seed(&r, seed);
  int have = 0; double spare = 0;
  int64_t i = 0;
  for (; i < nunif; i++) {
    x[i] = wrap_box(rng_u01(&r) * L, L);
    y[i] = wrap_box(rng_u01(&r) * L, L);
    z[i] = wrap_box(rng_u01(&r) * L, L);
  }
  for (int64_t h = 0; h < nhalo; h++) {
    double cx = rng_u01(&r) * L, cy = rng_u01(&r) * L, cz = rng_u01(&r) * L;
    for (int64_t k = 0; k < nper; k++, i++) {
      x[i] = wrap_box(cx + sigma * gen_gauss(&r, &have, &spare), L);
      y[i] = wrap_box(cy + sigma * gen_gauss(&r, &have, &spare), L);
      z[i] = wrap_box(cz + sigma * gen_gauss(&r, &have, &spare), L);
    }
  }
  if (w) {
    rng_t rw; rng_seed(&rw, seed ^ 0x5eed5eed5eed5eedULL);
    int hw = 0; double sw = 0;
    for (int64_t k = 0; k < i; k++)
      w[k] = (double) (float) gen_exp(wsig * gen_gauss(&rw, &hw, &sw));
  }
  return 0;
}

int fcfc_gen_survey(uint64_t seed, int64_t n, double rmin, double rmax,
    double nbar_peak, double r_peak, double r_sigma, double P0, double *x,
    double *y, double *z, double *w) {
  if (n < 0 || !(rmax > rmin) || rmin < 0) return FCFC_ERR_ARG;
  rng_t r; rng_seed(&r, seed);
  const double r2lo = rmin * rmin, r2hi = rmax * rmax;
  const double ymax = rmax * 0.57735026918962584;   /* tan(30 deg) */
  const double zmax = rmax * 0.5;                   /* sin(30 deg) */
  for (int64_t i = 0; i < n; i++) {
    double px, py, pz, r2;
    for (;;) {
      px = rng_u01(&r) * rmax;
      py = (2 * rng_u01(&r) - 1) * ymax;
      pz = (2 * rng_u01(&r) - 1) * zmax;
      r2 = px * px + py * py + pz * pz;
      if (r2 < r2lo || r2 >= r2hi) continue;
      if (3 * py * py > px * px) continue;          /* |RA| <= 30 deg  */
      if (4 * pz * pz > r2) continue;               /* |Dec| <= 30 deg */
      break;
    }
    x[i] = px; y[i] = py; z[i] = pz;
    if (w) {
      double rr = sqrt(r2), d = rr - r_peak;
      double nb = nbar_peak * gen_exp(-d * d / (2 * r_sigma * r_sigma));
      w[i] = (double) (float) (1 / (1 + nb * P0));
    }
  }
  return 0;
}

int fcfc_gen_sizes(uint64_t seed, int n, double lo, double hi, int64_t *out) {
  if (n < 0 || !(lo > 0) || !(hi >= lo)) return FCFC_ERR_ARG;
  rng_t r; rng_seed(&r, seed);
  const double a = gen_log(lo), b = gen_log(hi);
  for (int i = 0; i < n; i++) out[i] = (int64_t) floor(gen_exp(a + (b - a) * rng_u01(&r)) + 0.5);
  return 0;
}

/* --------------------------------------------------------------------- */
/* Survey coordinate conversion: (RA, Dec) in degrees and redshift ->     */
/* comoving Cartesian coordinates in Mpc/h (src/fcfc/2pt/cnvt_coord.c).   */
/* The comoving distance is the integral of c/H(z) (H0 = 100 h km/s/Mpc)  */
/* by Gauss-Legendre quadrature of an order chosen by a convergence test  */
/* over the catalogue's redshift range, or a natural cubic spline through */
/* (z, distance) samples.  The arithmetic follows the reference's         */
/* expression order (same libm calls, no contraction), so converted       */
/* coordinates are bit-identical on the same C library.                   */
#include "fcfc_legauss.h"

#define CM_SPEED_OF_LIGHT 299792.458            /* define_comm.h:124 */
#define CM_DEG2RAD        0x1.1df46a2529d39p-6   /* pi / 180, define_comm.h:137 */
#define CM_NUM_ZSAMPLE    128                    /* 2pt/define.h:99 */

int fcfc_cosmology_check(fcfc_cosmology *cm) {
  if (!cm) return FCFC_ERR_ARG;
  if (!(cm->omega_m > 0 && cm->omega_m <= 1)) return FCFC_ERR_CFG;   /* load_conf.c:700-705 */
  if (isnan(cm->omega_l)) { cm->omega_l = 1 - cm->omega_m; cm->omega_k = 0; }
  else if (cm->omega_l < 0) return FCFC_ERR_CFG;
  else cm->omega_k = 1 - cm->omega_m - cm->omega_l;
  if (isnan(cm->w)) cm->w = -1;
  else if (cm->w > -1 / (double) 3) return FCFC_ERR_CFG;
  /* H^2(z) > 0 for all z (load_conf.c:726-733) */
  const double w3 = cm->w * 3, widx = w3 + 1;
  if (cm->omega_k * pow(cm->omega_l * (-widx), widx / w3) <=
      cm->omega_l * w3 * pow(cm->omega_m, widx / w3)) return FCFC_ERR_CFG;
  if (isnan(cm->err)) cm->err = 1e-8;
  else if (cm->err < 0x1p-52) return FCFC_ERR_CFG;
  return 0;
}

/* c / H(z) (cnvt_coord.c:cnvt_integrand) */
static double cm_integrand(const fcfc_cosmology *cm, double widx, double z) {
  const double z1 = z + 1;
  const double z2 = z1 * z1;
  double d = cm->omega_m * z2 * z1;
  if (cm->omega_k) d += cm->omega_k * z2;
  if (widx) d += cm->omega_l * pow(z1, widx);
  else d += cm->omega_l;
  return CM_SPEED_OF_LIGHT * 0.01 / sqrt(d);
}

/* integral from 0 to z with the rule of the given order, abscissas from
 * the outermost in and the centre last (cnvt_coord.c:cnvt_legauss) */
static double cm_legauss(const fcfc_cosmology *cm, double widx, int order, double z) {
  const fcfc_lg_rule *q = &fcfc_lg_rules[order - FCFC_LG_MIN_ORDER];
  const double zp = z * 0.5;
  double sum = 0;
  for (int i = 0; i < order / 2; i++) {
    const double z1 = zp * (1 + q->x[i]);
    const double z2 = zp * (1 - q->x[i]);
    sum += q->w[i] * (cm_integrand(cm, widx, z1) + cm_integrand(cm, widx, z2));
  }
  if (order & 1) sum += q->mid * cm_integrand(cm, widx, zp);
  return sum * zp;
}

int fcfc_cmvdst_order(const fcfc_cosmology *cm, int64_t n, const double *z, int *order) {
  if (!cm || !order || n < 0 || (n && !z)) return FCFC_ERR_ARG;
  double zmin = 1.79769313486231570815e308, zmax = -1.79769313486231570815e308;
  for (int64_t i = 0; i < n; i++) {
    if (z[i] < 0) return FCFC_ERR_DATA;                 /* cnvt_coord.c:cnvt_z_sample */
    if (zmax < z[i]) zmax = z[i];
    if (zmin > z[i]) zmin = z[i];
  }
  if (zmin > zmax) return FCFC_ERR_DATA;
  const double widx = (cm->w == -1) ? 0 : 3 * (1 + cm->w);
  int best = 0;
  for (int i = 0; i < CM_NUM_ZSAMPLE; i++) {
    const double zs = zmin + i * (zmax - zmin) / (CM_NUM_ZSAMPLE - 1);
    double oint, nint = 0;
    int k = FCFC_LG_MIN_ORDER - 1;
    do {
      /* the reference would read past its table at order 33: here the
       * test fails instead */
      if (k + 1 > FCFC_LG_MAX_ORDER) return FCFC_ERR_CNVT;
      oint = nint;
      nint = cm_legauss(cm, widx, ++k, zs);
    } while (fabs(nint - oint) > nint * cm->err);
    if (best < k) best = k;
  }
  *order = best;
  return 0;
}

static void cm_to_cartesian(double dist, double *x, double *y, double *z) {
  const double ra = *x * CM_DEG2RAD;
  const double dec = *y * CM_DEG2RAD;
  *x = dist * cos(dec) * cos(ra);
  *y = dist * cos(dec) * sin(ra);
  *z = dist * sin(dec);
}

int fcfc_convert_coords(const fcfc_cosmology *cm, int64_t n, double *x, double *y,
    double *z, int *order_used) {
  if (!cm || n < 0 || (n && (!x || !y || !z))) return FCFC_ERR_ARG;
  int order = 0;
  int e = fcfc_cmvdst_order(cm, n, z, &order);
  if (e) return e;
  const double widx = (cm->w == -1) ? 0 : 3 * (1 + cm->w);
  for (int64_t i = 0; i < n; i++)
    cm_to_cartesian(cm_legauss(cm, widx, order, z[i]), x + i, y + i, z + i);
  if (order_used) *order_used = order;
  return 0;
}

/* Natural cubic spline through (zs[i], ds[i]), zs increasing: second
 * derivatives by the tridiagonal sweep (forward elimination, back
 * substitution), evaluated with the usual cubic form -- the operation
 * order of src/math/cspline.c, so interpolated distances match bit for bit. */
typedef struct { int64_t n; double *z, *d, *ypp; } cm_spline;

static int cm_cmp_pair(const void *a, const void *b) {
  const double *p = (const double *) a, *q = (const double *) b;
  if (p[0] < q[0]) return -1;
  if (p[0] > q[0]) return 1;
  return (p[2] < q[2]) ? -1 : (p[2] > q[2]);      /* stable: by input position */
}

static int cm_spline_init(cm_spline *s, const double *zs, const double *ds, int64_t n) {
  s->n = n;
  s->z = malloc(sizeof(double) * n);
  s->d = malloc(sizeof(double) * n);
  s->ypp = malloc(sizeof(double) * n);
  double *tmp = malloc(sizeof(double) * 3 * n), *cp = malloc(sizeof(double) * n);
  if (!s->z || !s->d || !s->ypp || !tmp || !cp) { free(tmp); free(cp); return FCFC_ERR_MEMORY; }
  for (int64_t i = 0; i < n; i++) { tmp[3 * i] = zs[i]; tmp[3 * i + 1] = ds[i]; tmp[3 * i + 2] = (double) i; }
  qsort(tmp, (size_t) n, 3 * sizeof(double), cm_cmp_pair);
  for (int64_t i = 0; i < n; i++) { s->z[i] = tmp[3 * i]; s->d[i] = tmp[3 * i + 1]; }
  free(tmp);
  const double *x = s->z, *y = s->d;
  double *ypp = s->ypp;
  double xn = x[1], yn = y[1], h = x[1] - x[0];
  double slope = (y[1] - y[0]) / h;
  cp[0] = cp[n - 1] = ypp[0] = ypp[n - 1] = 0;
  int64_t j = 1;
  for (; j < n - 1; j++) {
    const double xo = xn, yo = yn, hprev = h, so = slope;
    xn = x[j + 1];
    yn = y[j + 1];
    h = xn - xo;
    slope = (yn - yo) / h;
    const double diag = (h + hprev) * 2;
    const double inv = 1 / (diag - hprev * cp[j - 1]);
    const double rhs = (slope - so) * 6;
    ypp[j] = (rhs - hprev * ypp[j - 1]) * inv;
    cp[j] = h * inv;
  }
  while (j) {
    j -= 1;
    ypp[j] -= cp[j] * ypp[j + 1];
  }
  free(cp);
  return 0;
}

static void cm_spline_free(cm_spline *s) { free(s->z); free(s->d); free(s->ypp); }

/* distance at zv, z[0] <= zv < z[n-1]: the interval z[i] <= zv < z[i+1] */
static double cm_spline_eval(const cm_spline *s, double zv) {
  int64_t lo = 0, hi = s->n - 1;                /* z[lo] <= zv < z[hi] */
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (s->z[mid] <= zv) lo = mid; else hi = mid;
  }
  const int64_t i = lo, k = i + 1;
  const double *x = s->z, *y = s->d, *ypp = s->ypp;
  const double ba = x[k] - x[i];
  const double xa = zv - x[i];
  const double bx = x[k] - zv;
  const double ba2 = ba * ba;
  const double lower = xa * y[k] + bx * y[i];
  const double c = (xa * xa - ba2) * xa * ypp[k];
  const double d = (bx * bx - ba2) * bx * ypp[i];
  return (lower + 0x1.5555555555555p-3 * (c + d)) / ba;      /* 1/6 */
}

int fcfc_convert_coords_table(const double *zs, const double *ds, int64_t nsample,
    int64_t n, double *x, double *y, double *z) {
  if (nsample < 2 || !zs || !ds || n < 0 || (n && (!x || !y || !z))) return FCFC_ERR_ARG;
  cm_spline s;
  int e = cm_spline_init(&s, zs, ds, nsample);
  if (e) { cm_spline_free(&s); return e; }
  /* a redshift outside the samples has no distance: the reference would
   * carry an infinite coordinate on; here the conversion fails */
  for (int64_t i = 0; i < n; i++)
    if (!(z[i] >= s.z[0] && z[i] < s.z[s.n - 1])) { cm_spline_free(&s); return FCFC_ERR_CNVT; }
  for (int64_t i = 0; i < n; i++) cm_to_cartesian(cm_spline_eval(&s, z[i]), x + i, y + i, z + i);
  cm_spline_free(&s);
  return 0;
}
