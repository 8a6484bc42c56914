/*
 * fcfc_kernels.h -- device-side data layout and pair arithmetic shared by the
 * gfx950 kernels of the pair-count engine.
 *
 * Layout in HBM (one catalogue, after the grid build):
 *   objects sorted by cell id  cid = (ix * ncy + iy) * ncz + iz,
 *   FP64 SoA x[], y[], z[] (+ s[] = |r|^2 for survey (s,mu)/(s_perp,pi),
 *   + w[] weights), idx[] = original object index, cell_start[ncell + 1].
 * A "row" is the set of cells sharing (ix, iy): its objects are contiguous and
 * ordered by z-cell, so every z-window of a row is one contiguous range.
 *
 * Work unit: an A tile = up to NT*R consecutive objects of one row whose
 * z-cells span a bounded range.  For each xy offset (a "strip") the B objects
 * that can be within the maximum separation form at most three contiguous
 * segments (three periodic images in z).  A persistent workgroup takes tiles
 * from a global queue, packs the segments of a batch of strips into LDS
 * chunks, and every lane evaluates its A objects against each staged B object.
 *
 * Arithmetic contract (FP64 path), after FCFC's AVX-512 leaf kernels
 * (src/fcfc/2pt_box/metric_common.c:377-534):  B objects are staged as
 *   b' = fl(b + D)  with D the periodic image offset in {-L, 0, +L},
 * and every pair computes  dx = fl(b'.x - a.x), dz2 = fl(dz*dz),
 *   s2 = fma(dy, dy, fma(dx, dx, dz2)),
 * which is the reference's "vectorise node b" block (x1 = b - shift,
 * metric_common.c:1154-1163).  The mu index is floor of the EXACT quotient
 * fl(dz2*nmu^2)/s2, i.e. trunc(RZ(.)) as in metric_common.c:484-486.
 */
#ifndef FCFC_KERNELS_H
#define FCFC_KERNELS_H

#include <stdint.h>

#define FK_NT        512       /* threads per workgroup (8 wave64)          */
#define FK_NSB       64        /* strips per batch = one wavefront          */
#define FK_MAXSEG    3         /* z images per strip                        */
#define FK_CH        256       /* B objects per LDS chunk                   */
#define FK_OUT       0xFFFFu   /* table sentinel: outside [s2min, s2max)    */
#define FK_QCAP      128       /* FP32 re-evaluation queue per wavefront    */

/* Binning constants passed by value to the kernels. */
struct FkBins {
  int bintype, geometry;
  int ns, nmu, np, ntot;
  int nsh, ntoh;             /* LDS histogram row stride (odd) and slots    */
  int tabtype;               /* 0 integer table, 1 hybrid table             */
  int soff, poff;            /* (int) s2min, (int) pmin (or pi^2 min)       */
  int stab_len, ptab_len;    /* table lengths without sentinels             */
  int ssearch, psearch;      /* hybrid lookups: tables too long for LDS, bin
                                by binary search over the edges instead    */
  int scap32, pcap32;        /* last entry of the zero-based FP32 tables    */
  double s2min, s2max;       /* squared s (or s_perp) range                 */
  double pmin, pmax;         /* pi range (box) or pi^2 range (survey)       */
  double nmu2;               /* nmu * nmu as double                          */
  double nmu_d;              /* nmu as double                                */
  double mu_c;               /* 2^36 + 1 - mu_sh: the fixed-point mu value  */
  double mu_cmax;            /* 2^36 + nmu + 1.5: its ceiling (survey)       */
  unsigned mu_t16, mu_pad;   /* mu bin settled exactly when the fraction     */
                             /* (low 16 bits) of that value is >= mu_t16     */
  double s2_exact;           /* ... or s^2 is below this value               */
  double q_c;                /* survey (s_perp, pi), FP64 integer tables,   */
  double q_caps, q_capp;     /* pi^2 and s_perp^2 in fixed point (fk_sv64): */
  double q_bmin;             /* addend, ceilings, least |a+b|^2 of the fast */
  unsigned q_t16, q_pad;     /* path, fraction threshold of edge cells      */
  double s2skip, pskip;      /* box: a pair at s^2 >= s2skip or pi >= pskip  */
  float s2skip32, pskip32;   /* is out of range (FP64 / FP32 values)         */
  float win2s, win2p;        /* FP32 tables: window 2 e of the edge cells,   */
                             /* the largest over the jobs of a launch        */
  const uint16_t *stab;      /* device tables (global), staged to LDS       */
  const uint16_t *ptab;
  const double *s2bin;       /* edges for the hybrid refinement              */
  const double *pbin;
};

/* Grid geometry (shared by A and B of one job). */
struct FkGrid {
  int ncx, ncy, ncz;
  int periodic;
  double ox, oy, oz;         /* origin                                       */
  double wx, wy, wz;         /* cell widths                                  */
  double Lx, Ly, Lz;         /* box sizes (periodic)                         */
  double R2;                 /* squared maximum 3-D separation (padded)     */
  double Rxy2;               /* squared maximum xy separation (padded)      */
  double hz_const;           /* >= 0: fixed z half-window (box s_perp,pi)   */
};

struct FkCat {
  int n;
  const double *x, *y, *z, *s, *w;
  const int *idx;
  const int *cell_start;
};

/* A work item: the A objects [p0, p1) of one row (z-cells iz0..iz1) against
 * the strips [s0, s1) of its job; a tile with many strips is split into
 * several items so that the queue stays long (items of one tile add into
 * the same histogram; only the item with s0 == 0 counts the pairs inside
 * the tile). */
struct FkTile { int row, p0, p1, iz0, iz1, job, s0, s1; };

/* Floor of the exact quotient a / b for a >= 0, b > 0 (b not tiny):
 * approximate quotient from the hardware reciprocal, then one-step
 * corrections with exact-sign fma residuals. */
__device__ __forceinline__ double fk_floor_div(double a, double b) {
  double q = a * __builtin_amdgcn_rcp(b);
  double m = __builtin_floor(q);
  if (__builtin_fma(-m, b, a) < 0.0) m -= 1.0;
  else if (__builtin_fma(-(m + 1.0), b, a) >= 0.0) m += 1.0;
  return m;
}

/* integer square root of 0 <= m < 2^22 (mu table, setup_cf.c:513-531):
 * sqrt(m + 1/2) is at least 1/(4(n+1)) away from any integer. */
__device__ __forceinline__ int fk_isqrt(int m) {
  return (int) __builtin_sqrtf((float) m + 0.5f);
}

#endif
