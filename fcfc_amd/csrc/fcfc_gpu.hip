/*
 * fcfc_gpu.hip -- gfx950 pair-count engine for FCFC: grid build, tiling,
 * the persistent counting kernels (FP64, FP32 + FP64 edge recheck, weight
 * backward) and the C ABI declared in include/fcfc_gpu.h.
 *
 * Build (see fcfc_amd/build.py):
 *   hipcc --offload-arch=gfx950 -O3 -ffp-contract=off -fPIC -shared ...
 * -ffp-contract=off matters: the FP64 pair arithmetic must round exactly as
 * the reference's (every fma below is explicit).
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <type_traits>
#include <vector>

#include "fcfc_gpu.h"
#include "fcfc_kernels.h"

#define MODE_C64  0   /* FP64 counts                                    */
#define MODE_C32  1   /* FP32 distances + FP64 recheck near bin edges   */
#define MODE_GRAD 2   /* weight gradient (FP64 bins)                    */
#define MODE_C32C 3   /* MODE_C32 with the rounding census of the rechecked pairs */
#define MODE_P64  4   /* FP64, the reference's per-pair arithmetic as written */

/* ======================================================================== */
/*                          grid build kernels                              */
/* ======================================================================== */

__device__ __forceinline__ int fk_cell1(double v, double o, double w, int nc) {
  int q = (int) floor((v - o) / w);
  return q < 0 ? 0 : (q >= nc ? nc - 1 : q);
}

__global__ void k_fill_int(int *p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_zero_u64(unsigned long long *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

/* Exclusive scan of n ints into out[0..n] (out[n] = total), one workgroup
 * of 1024 threads, each scanning a contiguous slice. */
__global__ void __launch_bounds__(1024) k_scan1(const int *__restrict__ in,
    int *__restrict__ out, int n) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = t * per, e = min(n, b + per);
  int s = 0;
  for (int i = b; i < e; i++) s += in[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int i = b; i < e; i++) { int v = in[i]; out[i] = run; run += v; }
  if (t == 1023) out[n] = part[1023];
}

/* ---- grid build: every catalogue of a call sorted into its job's grid
 * in one pass per stage (cell counting sort), so that a batch of hundreds
 * of small catalogues costs a handful of launches ---- */
struct FkStage {
  int n, obj0, cell0, cs0;    /* objects; offsets in the concatenated object,
                                 count and cell-start index spaces */
  int need_s;                 /* survey |r|^2 wanted (computed when s is null) */
  FkGrid g;
  const double *x, *y, *z, *s, *w;      /* inputs; s, w may be null */
  double *ox, *oy, *oz, *os, *ow;       /* sorted copies; os, ow null when unused */
  int *oidx, *ocs;                      /* original index, cell starts [ncell + 1] */
};

/* request owning index v of the concatenated space whose starts are off */
__device__ __forceinline__ int fk_stage_find(const FkStage *__restrict__ st, int nst, int v,
    int FkStage::*off) {
  int lo = 0, hi = nst - 1;                   /* last request with start <= v */
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (st[mid].*off <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_stage_assign(int nobj, int nst, const FkStage *__restrict__ st,
    int *__restrict__ cid, int *__restrict__ rank, int *__restrict__ count, int *bad) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < nobj; o += gridDim.x * blockDim.x) {
    const FkStage &q = st[fk_stage_find(st, nst, o, &FkStage::obj0)];
    const int i = o - q.obj0;
    /* a non-finite coordinate has no cell and no bin: the call fails */
    if (!(__builtin_isfinite(q.x[i]) && __builtin_isfinite(q.y[i]) && __builtin_isfinite(q.z[i])))
      atomicOr(bad, 1);
    const int ix = fk_cell1(q.x[i], q.g.ox, q.g.wx, q.g.ncx);
    const int iy = fk_cell1(q.y[i], q.g.oy, q.g.wy, q.g.ncy);
    const int iz = fk_cell1(q.z[i], q.g.oz, q.g.wz, q.g.ncz);
    const int c = (ix * q.g.ncy + iy) * q.g.ncz + iz;
    cid[o] = c;
    rank[o] = atomicAdd(count + q.cell0 + c, 1);
  }
}

/* cell starts of each catalogue from the scan of all counts */
__global__ void k_stage_starts(int ncs, int nst, const FkStage *__restrict__ st,
    const int *__restrict__ gstart) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ncs; t += gridDim.x * blockDim.x) {
    const FkStage &q = st[fk_stage_find(st, nst, t, &FkStage::cs0)];
    q.ocs[t - q.cs0] = gstart[q.cell0 + (t - q.cs0)] - gstart[q.cell0];
  }
}

/* sorted copies; |r|^2 with FCFC's data_preprocess lane rule
 * (2pt/build_tree.c:59-92: FMA for the first floor(n/8)*8 objects, unfused
 * for the tail), unit weights where a weighted count has none
 * (build_tree.c:143-156) */
__global__ void k_stage_scatter(int nobj, int nst, const FkStage *__restrict__ st,
    const int *__restrict__ cid, const int *__restrict__ rank, const int *__restrict__ gstart) {
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < nobj; o += gridDim.x * blockDim.x) {
    const FkStage &q = st[fk_stage_find(st, nst, o, &FkStage::obj0)];
    const int i = o - q.obj0;
    const int d = gstart[q.cell0 + cid[o]] - gstart[q.cell0] + rank[o];
    const double a = q.x[i], b = q.y[i], c = q.z[i];
    q.ox[d] = a; q.oy[d] = b; q.oz[d] = c;
    if (q.os) {
      double v;
      if (q.s) v = q.s[i];
      else if (i < (q.n / 8) * 8) v = __builtin_fma(c, c, __builtin_fma(b, b, a * a));
      else { double t = a * a; t = t + b * b; v = t + c * c; }
      q.os[d] = v;
    }
    if (q.ow) q.ow[d] = q.w ? q.w[i] : 1.0;
    q.oidx[d] = i;
  }
}

/* Work items of the tile list: tile ti becomes S items (each a share of the
 * strip list) for ti < tt and St > S finer items from tt on, so that the
 * items the persistent workgroups take last are short and the launch ends
 * evenly.  ncap tiles in all: of each group of G consecutive tiles of the
 * tiling, its first `run` tiles.  A full count has one group (G = run =
 * the tile count); the autotuner's samples take runs spread over the whole
 * catalogue, each keeping the neighbourhood, and so the cache reuse, of a
 * full count. */
struct FkSplit { int S, tt, St, ncap, G, run; };

/* list position of tile ti of the tiling, or -1 when it is not sampled */
__host__ __device__ inline int split_sample(const FkSplit &s, int ti) {
  const int g = ti / s.G, p = ti - g * s.G;
  return p < s.run ? g * s.run + p : -1;
}

__host__ __device__ inline long long split_pos(const FkSplit &s, int ti) {
  return ti < s.tt ? (long long) s.S * ti
                   : (long long) s.S * s.tt + (long long) s.St * (ti - s.tt);
}

/* Tiles of one row: consecutive cells, at most T objects, at most zmaxc
 * z-cells; a cell with more than T objects is split into T-sized tiles.
 * t0: index of the row's first tile in the tile list. */
__device__ int row_tiles(const int *__restrict__ cs, int ncz, int T, int zmaxc,
    int row, int job, int nstrips, const FkSplit &sp, int t0, FkTile *out) {
  int nt = 0, open = 0, tp0 = 0, tcnt = 0, tz0 = 0, tz1 = 0;
  /* emit tile (p0, p1, iz0, iz1) as items over the strip list */
  auto emit = [&](int p0, int p1, int iz0, int iz1) {
    const int ti = t0 + nt;
    const int si = out ? split_sample(sp, ti) : -1;
    if (si >= 0 && si < sp.ncap) {
      const int S = si < sp.tt ? sp.S : sp.St;
      FkTile *o = out + split_pos(sp, si);
      for (int q = 0; q < S; q++)
        o[q] = FkTile{row, p0, p1, iz0, iz1, job,
                      (int) ((long long) nstrips * q / S),
                      (int) ((long long) nstrips * (q + 1) / S)};
    }
    nt++;
  };
  for (int iz = 0; iz < ncz; iz++) {
    const int c = cs[iz + 1] - cs[iz];
    if (c == 0) continue;
    if (open && (tcnt + c > T || iz - tz0 + 1 > zmaxc)) {
      emit(tp0, tp0 + tcnt, tz0, tz1);
      open = 0;
    }
    if (c > T) {
      for (int k = 0; k < c; k += T) emit(cs[iz] + k, cs[iz] + min(k + T, c), iz, iz);
      continue;
    }
    if (!open) { open = 1; tp0 = cs[iz]; tcnt = 0; tz0 = iz; }
    tcnt += c; tz1 = iz;
  }
  if (open) emit(tp0, tp0 + tcnt, tz0, tz1);
  return nt;
}

/* Tiling of all jobs of a launch in one pass: job k owns rows row0 ..
 * row0 + ncx ncy - 1 of the launch's row space (row0 ascending). */
struct FkTileJob { int row0, ncz, T, zmaxc, nstrips; const int *cs; };

__device__ __forceinline__ int fk_tile_job(const FkTileJob *__restrict__ tj, int njob, int gr) {
  int lo = 0, hi = njob - 1;                  /* last job with row0 <= gr */
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tj[mid].row0 <= gr) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_tile_count(int nrow, int njob, const FkTileJob *__restrict__ tj,
    int *__restrict__ ntile) {
  for (int gr = blockIdx.x * blockDim.x + threadIdx.x; gr < nrow; gr += gridDim.x * blockDim.x) {
    const FkTileJob t = tj[fk_tile_job(tj, njob, gr)];
    const int r = gr - t.row0;
    ntile[gr] = row_tiles(t.cs + (size_t) r * t.ncz, t.ncz, t.T, t.zmaxc, r, 0, 0,
                          FkSplit{1, 0, 1, 0, 1, 1}, 0, nullptr);
  }
}
/* the tiles of row gr are numbered from toff[gr] */
__global__ void k_tile_emit(int nrow, int njob, const FkTileJob *__restrict__ tj,
    const int *__restrict__ toff, FkSplit sp, FkTile *__restrict__ tiles) {
  for (int gr = blockIdx.x * blockDim.x + threadIdx.x; gr < nrow; gr += gridDim.x * blockDim.x) {
    const int k = fk_tile_job(tj, njob, gr);
    const FkTileJob t = tj[k];
    const int r = gr - t.row0;
    row_tiles(t.cs + (size_t) r * t.ncz, t.ncz, t.T, t.zmaxc, r, k, t.nstrips, sp,
              toff[gr], tiles);
  }
}

/* ======================================================================== */
/*                             count kernel                                 */
/* ======================================================================== */

/* v of lane (lane with the DPP pattern CTRL applied), both halves moved */
template <int CTRL>
__device__ __forceinline__ double fk_dpp(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int) (unsigned) u, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int) (unsigned) (u >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((unsigned long long) (unsigned) hi << 32) | (unsigned) lo);
}

/* Sum of v over each row of 16 lanes, in every lane of the row: quad
 * butterflies (lane ^ 1, lane ^ 2), then the row's half mirror (lane i of
 * 8 with 7 - i) and mirror (lane i of 16 with 15 - i).  Every lane of the
 * wavefront must be active. */
__device__ __forceinline__ double fk_row_sum(double v) {
  v += fk_dpp<0xB1>(v);       /* quad_perm [1, 0, 3, 2] */
  v += fk_dpp<0x4E>(v);       /* quad_perm [2, 3, 0, 1] */
  v += fk_dpp<0x141>(v);      /* row_half_mirror */
  v += fk_dpp<0x140>(v);      /* row_mirror */
  return v;
}

/* FP32 exactness windows (see fp32_windows on the host). */
struct FkWin {
  float e_s2, e_s2x2;         /* absolute window on s^2 (or s_perp^2), 2 e_s2 */
  float e_pi, e_pix2;         /* absolute window on pi, 2 e_pi                */
  float idle_x, idle_z;       /* FP32 coordinates of idle lanes             */
  float mu_sh;                /* error bound of the FP32 mu value t         */
  float mu_c;                 /* 129 - sh: the fixed-point t (fk_mu32)      */
  unsigned mu_t16;            /* mu bin certain if t's fraction is below it */
  float s2_floor;             /* pairs below this s^2 are re-evaluated      */
  float nmu_f;
  float m2_floor;             /* survey: pairs with |a + b|^2 below this    */
};                            /* are re-evaluated                          */

/* A value that is the same in every lane (read from the job or tile held in
 * LDS), moved to a scalar register: the compiler cannot tell that it is
 * uniform and would otherwise keep it in a vector register. */
__device__ __forceinline__ float fk_uni(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ FkWin fk_uni(const FkWin &w) {
  FkWin u;
  u.e_s2 = fk_uni(w.e_s2); u.e_s2x2 = fk_uni(w.e_s2x2);
  u.e_pi = fk_uni(w.e_pi); u.e_pix2 = fk_uni(w.e_pix2);
  u.idle_x = fk_uni(w.idle_x); u.idle_z = fk_uni(w.idle_z);
  u.mu_sh = fk_uni(w.mu_sh); u.mu_c = fk_uni(w.mu_c);
  u.mu_t16 = (unsigned) __builtin_amdgcn_readfirstlane((int) w.mu_t16);
  u.s2_floor = fk_uni(w.s2_floor); u.nmu_f = fk_uni(w.nmu_f);
  u.m2_floor = fk_uni(w.m2_floor);
  return u;
}

/* Flags of the integer cells of the FP32 tables, set while they are staged
 * (the LDS cells then carry a float threshold in place of the flags). */
#define FK_F_EDGE 0x10000u    /* a bin edge at the lower end of the integer cell */
#define FK_F_ALW  0x20000u    /* always re-evaluate (with FK_F_EDGE set too)     */
#define FK_F_ALWAYS 0x4000u   /* FP32 threshold (high half) of FK_F_ALW cells: 2 */
#define FK_F_NEVER  0xBF80u   /* ... of cells without flags: -1                  */

/* One counting job: a catalogue pair and its histogram.  A launch counts the
 * tiles of one or many jobs (batched mocks); tiles carry their job index. */
struct FkJob {
  FkCat A, B;
  FkGrid g;
  const int *strips;          /* (dx, dy, diag) triples */
  int nstrips;
  int autoc;                  /* COUNT: A == B, half shell + self tile;
                                 GRAD: A == B, exclude j == i          */
  int wt;                     /* weighted histogram (double)          */
  long long hist_off;         /* histogram of this job: gcnt + hist_off */
  double *gA;                 /* GRAD: output for A objects, original order (or null) */
  double *gB;                 /* GRAD: output for B objects, original order (or null;
                                 auto counts: the same array as gA)   */
  FkWin win;                  /* FP32 error model (MODE_C32)          */
};

struct FkArgs {
  const FkJob *jobs;
  int njobs;
  FkBins bins;
  const FkTile *tiles;
  int ntiles;
  unsigned int *queue;
  unsigned long long *gcnt_u;
  double *gcnt_d;
  unsigned long long *stats;  /* evals, rechecked, census */
  int nh;                     /* LDS histogram copies */
  /* LDS byte offsets */
  int off_hist, off_stab, off_ptab, off_edges, off_strip, off_chunk, off_queue;
  int stab_c0, ptab_c0;       /* table entries before cell 0 */
  int census;
  const double *G;            /* GRAD: dL/dcnt */
};

/* LDS strip table */
struct FkStripLds {
  int off[FK_NSB + 1];
  int b0[FK_NSB * FK_MAXSEG];
  int len[FK_NSB * FK_MAXSEG];
  int code[FK_NSB * FK_MAXSEG];
};

/* Compute the (up to three) B segments of one strip for tile t. */
__device__ void fk_strip_segments(const FkJob &J, const FkTile &t, int dx,
    int dy, int diag, int *b0, int *len, int *code) {
  const FkGrid &g = J.g;
  b0[0] = b0[1] = b0[2] = 0;
  len[0] = len[1] = len[2] = 0;
  code[0] = code[1] = code[2] = 13;
  const int ix = t.row / g.ncy, iy = t.row - (t.row / g.ncy) * g.ncy;
  int bx = ix + dx, by = iy + dy, cx = 0, cy = 0;
  if (g.periodic) {
    if (bx < 0) { bx += g.ncx; cx = -1; } else if (bx >= g.ncx) { bx -= g.ncx; cx = 1; }
    if (by < 0) { by += g.ncy; cy = -1; } else if (by >= g.ncy) { by -= g.ncy; cy = 1; }
  }
  else if (bx < 0 || bx >= g.ncx || by < 0 || by >= g.ncy) return;
  const double gx = (double) max(abs(dx) - 1, 0) * g.wx;
  const double gy = (double) max(abs(dy) - 1, 0) * g.wy;
  const double d2 = gx * gx + gy * gy;
  if (d2 >= g.Rxy2) return;
  const double hz = (g.hz_const >= 0) ? g.hz_const : sqrt(fmax(g.R2 - d2, 0.0));
  const double zc0 = (double) t.iz0 - hz / g.wz - 1e-6;
  const double zc1 = (double) (t.iz1 + 1) + hz / g.wz + 1e-6;
  int izlo = (int) floor(zc0), izhi = (int) floor(zc1);
  const int *cs = J.B.cell_start + (size_t) (bx * g.ncy + by) * g.ncz;
  const int cxy = (cx + 1) + 3 * (cy + 1);
  if (g.periodic) {
    if (izhi - izlo + 1 > g.ncz) {   /* cannot happen with the tile limits */
      izlo = 0; izhi = g.ncz - 1;
    }
    if (izlo < 0) {                  /* image below z = 0: B shifted by -Lz */
      int c0 = izlo + g.ncz, c1 = g.ncz - 1;
      b0[0] = cs[c0]; len[0] = cs[c1 + 1] - cs[c0]; code[0] = cxy + 9 * 0;
    }
    {
      int c0 = max(izlo, 0), c1 = min(izhi, g.ncz - 1);
      if (diag) {
        /* auto counts, same row: only objects after the tile (j > i) */
        int e = cs[c1 + 1];
        b0[1] = t.p1; len[1] = max(e - t.p1, 0); code[1] = cxy + 9 * 1;
      }
      else if (c1 >= c0) { b0[1] = cs[c0]; len[1] = cs[c1 + 1] - cs[c0]; code[1] = cxy + 9 * 1; }
    }
    if (izhi >= g.ncz && !diag) {    /* image above: B shifted by +Lz */
      int c0 = 0, c1 = izhi - g.ncz;
      b0[2] = cs[c0]; len[2] = cs[c1 + 1] - cs[c0]; code[2] = cxy + 9 * 2;
    }
  }
  else {
    int c0 = max(izlo, 0), c1 = min(izhi, g.ncz - 1);
    if (diag) {
      int e = (c1 >= 0) ? cs[c1 + 1] : t.p1;
      b0[0] = t.p1; len[0] = max(e - t.p1, 0); code[0] = 13;
    }
    else if (c1 >= c0) { b0[0] = cs[c0]; len[0] = cs[c1 + 1] - cs[c0]; code[0] = 13; }
  }
}

/* ---------------- FP64 pair evaluation (the arithmetic contract) -------- */

/* Pair functions return an LDS histogram index sb + nsh * mb; every value
 * >= ntoh = nsh * rows means "not counted".  The row stride nsh is ns
 * rounded up to an odd number: lanes of one histogram atomic that share an
 * s bin but differ in mu (or pi) then fall on different LDS banks (an even
 * stride such as 40 folds them onto a few).  Table sentinels are FK_OUT =
 * 0xFFFF >= ntoh (the histogram lives in LDS, so ntoh < 40960), hence
 * sb + nsh * mb needs no range test of its own: an out-of-range s (or pi)
 * gives sb >= ntoh, a dropped mu gives mb = nmu and sb + nsh * nmu >= ntoh.
 * fk_hslot maps an LDS slot to the reference layout sb + ns * mb. */
__device__ __forceinline__ int fk_hslot(int i, int ns, int nsh) {
  const int row = i / nsh, col = i - row * nsh;
  return col < ns ? col + ns * row : -1;
}
#define FK_NONE 0xFFFFFFFFu
#define FK_NMU32 126          /* most mu bins of the FP32 (s, mu) form      */
#define FK_DBL_EPS 2.220446049250313e-16
#define FK_FAR 1.0e30          /* coordinates of idle lanes: every pair out */

/* s^2 (or s_perp^2) bin from the LDS table, FK_OUT outside [s2min, s2max).
 * Integer tables (setup_cf.c:385-410): cell (int) s2 - (int) s2min, exactly
 * the reference's cvtt lookup (metric_common.c:500-510).  The LDS copy is
 * zero-based: entry k belongs to (int) s2 = k, the entries below (int) s2min
 * and the one at (int) s2min + len are FK_OUT, so with s2 >= 0 the range
 * check is one min.  Hybrid tables (pointer at cell 0) refine downwards
 * against the edges (FCFC_LOOKUP_HYBRID, metric_common.c:56-64): the result
 * is the last edge at or below s2 (bin 0 below the first edge).  Tables too
 * long for LDS are replaced by a binary search for that same edge.  MK: the
 * table is an FP32 one, whose cells hold the bin in their low half. */
__device__ __forceinline__ unsigned fk_edge_search(const double *e, int n, double v) {
  int lo = 0, hi = n - 1;                 /* last i in [0, n) with e[i] <= v, else 0 */
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (e[mid] <= v) lo = mid; else hi = mid - 1;
  }
  return (unsigned) lo;
}

template <bool HYB, bool MK = false>
__device__ __forceinline__ unsigned fk_sbin(double s2, const FkBins &bn,
    const unsigned *stab, const double *s2e) {
  if (HYB) {
    if (!(s2 < bn.s2max) || (bn.s2min != 0 && s2 < bn.s2min)) return FK_OUT;
    if (bn.ssearch) return fk_edge_search(s2e, bn.ns, s2);
    int k = (int) s2 - bn.soff;
    k = k < 0 ? 0 : (k >= bn.stab_len ? bn.stab_len - 1 : k);
    unsigned sb = MK ? (stab[k] & 0xFFFFu) : stab[k];
    if (sb >= (unsigned) bn.ns) {
      sb -= bn.ns;
      while (sb != 0 && s2 < s2e[sb]) sb--;
    }
    return sb;
  }
  /* unsigned: a value below -1 (a survey s_perp^2 from an infinite pi^2)
   * reads the "out" sentinel, as for values above the range */
  const unsigned k = min((unsigned) (int) s2, (unsigned) (bn.soff + bn.stab_len));
  return MK ? (stab[k] & 0xFFFFu) : stab[k];
}

template <bool HYB, bool MK = false>
__device__ __forceinline__ unsigned fk_pbin(double p, const FkBins &bn,
    const unsigned *ptab, const double *pe) {
  if (HYB) {
    if (!(p < bn.pmax) || p < bn.pmin) return FK_OUT;
    if (bn.psearch) return fk_edge_search(pe, bn.np, p);
    int k = (int) p - bn.poff;
    k = k < 0 ? 0 : (k >= bn.ptab_len ? bn.ptab_len - 1 : k);
    unsigned pb = MK ? (ptab[k] & 0xFFFFu) : ptab[k];
    if (pb >= (unsigned) bn.np) {
      pb -= bn.np;
      while (pb != 0 && p < pe[pb]) pb--;
    }
    return pb;
  }
  const unsigned k = min((unsigned) (int) p, (unsigned) (bn.poff + bn.ptab_len));
  return MK ? (ptab[k] & 0xFFFFu) : ptab[k];
}

/* mu bin of (s, mu) counts.  The reference takes m = floor of the exact
 * quotient a / s2 with a = fl(num * nmu^2) (trunc of the RZ division,
 * metric_common.c:484-486), sets m = 0 when s2 < DBL_EPSILON, drops the pair
 * when m >= nmu^2 and bins it in mutab[m] = isqrt(m) (setup_cf.c:513-531).
 * Since isqrt(floor(q)) = floor(sqrt(q)) for q >= 0, the bin is floor(t),
 * t = sqrt(a / s2), and the drop rule is floor(t) >= nmu.
 *
 * Fast path: x = nmu |dz| rsq(s2) (hardware reciprocal square root; |x - t|
 * < D = nmu 2^-20 including the products) goes into one fma with the addend
 * mu_c = 2^36 + 1 - sh, sh a multiple of 2^-16 with sh >= D + 2^-16.  The
 * sum lies in [2^36, 2^37), where doubles are spaced 2^-16, so the fma
 * rounds it once onto that grid and the low word of the result is the
 * fixed-point v = round((x - sh + 1) 2^16): its high half is K =
 * floor(x - sh) + 1 and its low half f the fraction.  When f < mu_t16 =
 * floor(2^16 (1 - 2 sh)), x lies in [K - 1 + sh - 2^-17, K - 1 + 1 - sh) and
 * the exact t strictly inside (K - 1, K): the bin is K - 1, and the
 * histogram index plus nsh is K nsh + sb, one 16-bit multiply-add of the
 * high half (fk_mad_hi16).  A t within sh of an integer (K = 0 included)
 * fails the fraction test.  Those pairs, and all with s2 below s2_exact =
 * s2max 2^-12 (which covers s2 < DBL_EPSILON and bounds t by nmu (1 +
 * 2^-21) on the fast path), are settled exactly: t_sh = v 2^-16 - 1 (exact:
 * the fma result minus 2^36 + 1) is within 1/2 of t, so k = rint(t_sh) is
 * the bin if a - k^2 s2 >= 0 (one fma, exact sign; k^2 is exact), else
 * k - 1. */
#define FK_MU_C0 68719476737.0          /* 2^36 + 1 */

/* hi16(v) * m + c (m < 2^16) in one instruction: the 16-bit multiply-add
 * reading the high half of its first operand */
__device__ __forceinline__ unsigned fk_mad_hi16(unsigned v, unsigned m, unsigned c) {
  unsigned r;
  asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,0,0,0]" : "=v"(r) : "v"(v), "s"(m), "v"(c));
  return r;
}

/* hi16(v) * 4 + c, c uniform: the LDS byte address of entry hi16(v) of a
 * table of 32-bit entries at byte address c */
__device__ __forceinline__ unsigned fk_mad4_hi16(unsigned v, unsigned c) {
  unsigned r;
  asm("v_mad_u32_u16 %0, %1, 4, %2 op_sel:[1,0,0,0]" : "=v"(r) : "v"(v), "s"(c));
  return r;
}

/* lo16(a) * m + c in one instruction (the high half of a is ignored) */
__device__ __forceinline__ unsigned fk_mad_lo16(unsigned a, unsigned m, unsigned c) {
  unsigned r;
  asm("v_mad_u32_u16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(m), "v"(c));
  return r;
}

/* min / max of x and a uniform bound, one v_min_f64 / v_max_f64 (fmin and
 * fmax would also canonicalise the bound on every call); a NaN x gives the
 * bound */
__device__ __forceinline__ double fk_min64(double x, double c) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(x), "s"(c));
  return r;
}
__device__ __forceinline__ double fk_max64(double x, double c) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(x), "s"(c));
  return r;
}

/* LDS byte address of a pointer into the workgroup's LDS, and a 32-bit
 * LDS read at such an address (an ordinary load: no asm) */
__device__ __forceinline__ unsigned fk_lds_addr(const void *p) {
  return (unsigned) (uintptr_t) (const __attribute__((address_space(3))) char *) p;
}
__device__ __forceinline__ unsigned fk_lds_u32(unsigned addr) {
  return *(const __attribute__((address_space(3))) unsigned *) (uintptr_t) addr;
}

__device__ __forceinline__ unsigned fk_mu_exact(double a, double s2, double t_sh,
    const FkBins &bn) {
  if (!(s2 >= FK_DBL_EPS)) return 0u;
  double k = __builtin_rint(t_sh);
  if (__builtin_fma(-(k * k), s2, a) < 0.0) k -= 1.0;
  return (unsigned) (int) __builtin_fmin(k, bn.nmu_d);
}

/* The same mu bin transcribed directly from the reference's arithmetic:
 * the correctly rounded quotient q = fl(a / s2), floored, with the one
 * correction that turns round-to-nearest into truncation of the exact
 * quotient (q can only round up onto an integer k, and then a - k s2 < 0),
 * and the isqrt table mapping m to the bin. */
__device__ __forceinline__ unsigned fk_mu_plain(double num, double s2, const FkBins &bn) {
  if (!(s2 >= FK_DBL_EPS)) return 0u;
  const double a = num * bn.nmu2;
  const double q = a / s2;
  double m = __builtin_floor(q);
  if (m == q && __builtin_fma(-m, s2, a) < 0.0) m -= 1.0;
  if (m >= bn.nmu2) return (unsigned) bn.nmu;
  return (unsigned) fk_isqrt((int) m);
}

/* The (s, mu) index plus nsh from x = anmu r = nmu |dz| rsq(s2) >= 0; fa()
 * gives a = dz^2 nmu^2 for the exact path (evaluated there only).  Every
 * index that leaves the fast path is at least nsh (K >= 1).  CAP: survey
 * pairs, whose x has no bound near a = -b (|a + b|^2 cancels) and may be
 * NaN there, have the value capped at K = nmu + 1 (not counted, as the
 * reference drops mu >= 1).  MUT: sb is the raw cell of an integer table
 * whose high half holds the fraction threshold, mu_t16, or 0 in the cells
 * below s2_exact (whose pairs all take the exact path), so one compare of
 * the two low halves replaces the threshold and s2_exact compares; the
 * threshold then rides in the index too, which grows by mu_t16 << 16. */
template <bool CAP, bool MUT, class FA>
__device__ __forceinline__ unsigned fk_smu_index(double anmu, double r, FA fa, double s2,
    unsigned sb, const FkBins &bn) {
  double tm = __builtin_fma(anmu, r, bn.mu_c);
  if (CAP) tm = fk_min64(tm, bn.mu_cmax);
  const unsigned v = (unsigned) __builtin_bit_cast(unsigned long long, tm);
  unsigned idx = fk_mad_hi16(v, (unsigned) bn.nsh, sb);
  const bool exact = MUT ? !((v & 0xFFFFu) < (sb >> 16))
                         : (!((v & 0xFFFFu) < bn.mu_t16) || s2 < bn.s2_exact);
  if (exact)
    idx = (sb & 0xFFFFu) + __umul24((unsigned) bn.nsh, fk_mu_exact(fa(), s2, tm - FK_MU_C0, bn) + 1u)
          + (MUT ? bn.mu_t16 << 16 : 0u);
  return idx;
}

/* Box pair, B already at its periodic image (b' = fl(b + D)).  Returns the
 * histogram index, plus nsh (and mu_t16 << 16 with MUT) for (s, mu) bins
 * outside the plain form (fk_smu_index); >= ntot (+ that): not counted. */
template <int BIN, bool HYB, bool MK = false, bool PLAIN = false, bool MUT = false>
__device__ __forceinline__ unsigned fk_box64(double dx, double dy, double dz,
    const FkBins &bn, const unsigned *stab, const unsigned *ptab,
    const double *s2e, const double *pe) {
  if (BIN == FCFC_BIN_SPI) {
    const double pi = __builtin_fabs(dz);
    const double s2 = __builtin_fma(dy, dy, dx * dx);          /* :416-420 */
    const unsigned sb = fk_sbin<HYB, MK>(s2, bn, stab, s2e);
    const unsigned pb = fk_pbin<HYB, MK>(pi, bn, ptab, pe);
    return sb + __umul24((unsigned) bn.nsh, pb);
  }
  const double dz2 = dz * dz;
  const double s2 = __builtin_fma(dy, dy, __builtin_fma(dx, dx, dz2));  /* :426-430 */
  const unsigned sb = fk_sbin<HYB, MK>(s2, bn, stab, s2e);
  if (BIN == FCFC_BIN_ISO) return sb;
  if (PLAIN) return sb + __umul24((unsigned) bn.nsh, fk_mu_plain(dz2, s2, bn));
  return fk_smu_index<false, MUT>(__builtin_fabs(dz) * bn.nmu_d, __builtin_amdgcn_rsq(s2),
                      [&] { return dz2 * bn.nmu2; }, s2, sb, bn);
}

/* Survey pair (src/fcfc/2pt/metric_common.c:283-460, FMA vector formulas):
 * t = 2 a.b, s = |a|^2 + |b|^2, d = |a|^2 - |b|^2, pi^2 = d^2 / (t + s),
 * s^2 = s - t, s_perp^2 = (s - t) - pi^2.  (s, mu) bins need pi^2 only in
 * the exact branch of the mu bin: sqrt(pi^2) = |d| / sqrt(t + s).  MUT:
 * integer tables whose cells carry a threshold in their high half ((s, mu):
 * the mu fraction threshold; (s_perp, pi): fixed-point lookups, below);
 * os (op) is the s (pi) table's LDS byte address for those lookups. */
template <int BIN, bool HYB, bool PLAIN = false, bool MK = false, bool MUT = false>
__device__ __forceinline__ unsigned fk_sv64(double ax, double ay, double az,
    double as, double bx, double by, double bz, double bs, const FkBins &bn,
    const unsigned *stab, const unsigned *ptab, const double *s2e, const double *pe,
    unsigned os = 0u, unsigned op = 0u) {
  if (BIN == FCFC_BIN_ISO) {
    const double dx = ax - bx, dy = ay - by, dz = az - bz;
    const double s2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
    return fk_sbin<HYB, MK>(s2, bn, stab, s2e);
  }
  const double t = __builtin_fma(az, bz, __builtin_fma(ay, by, ax * bx)) * 2.0;
  const double s = as + bs;
  const double d = as - bs;
  if (BIN == FCFC_BIN_SMU) {
    const double s2 = s - t;
    const unsigned sb = fk_sbin<HYB, MK>(s2, bn, stab, s2e);
    if (PLAIN) return sb + __umul24((unsigned) bn.nsh, fk_mu_plain((d * d) / (t + s), s2, bn));
    return fk_smu_index<true, MUT>(__builtin_fabs(d) * bn.nmu_d,
                        __builtin_amdgcn_rsq((t + s) * s2),
                        [&] { return ((d * d) / (t + s)) * bn.nmu2; }, s2, sb, bn);
  }
  /* (s_perp, pi): pbin holds pi^2 edges; with s2min == 0 the reference has
   * no lower bound on s_perp^2 (a rounding-negative value lands in bin 0) */
  if (MUT) {
    /* Integer tables in the FP64 forms: no division on the fast path.
     * With y = rcp(B) refined once (B = t + s, |y - 1/B| <= 2^-45.9 / B for
     * the documented 2^-23 of v_rcp_f64 and B >= q_bmin), q = n y is within
     * 2^-29.8 of the reference's correctly rounded n / B below 2^16, and
     * s_perp^2 = (s - t) - q within 2^-29.7 of the reference's.  Each is
     * read in fixed point as the mu bin is (fk_smu_index): x + q_c, q_c =
     * 2^36 + 2 - h (h = 2^-15), lies in [2^36, 2^37) (clamped to it) and is
     * rounded once onto the grid 2^-16, so the low word holds K = floor(x -
     * h) + 2 in its high half and the fraction f in its low half.  The
     * tables (staged by k_count, qt_entry) are indexed by K: entry K holds
     * the bin of (int) x for x in (K - 2, K - 1) (entry 0: out, entry 1:
     * the cell of (-1, 0)) and in its high half the largest f for which x
     * is certainly in that interval (q_t16) when the bin changes at K - 1,
     * 0xFFFF when it does not.  Anything else -- f above it, B below q_bmin (zero,
     * negative, NaN) -- is binned with the reference's division. */
    const double n = d * d, bb = t + s;
    const double y0 = __builtin_amdgcn_rcp(bb);
    const double q = n * __builtin_fma(y0, __builtin_fma(-bb, y0, 1.0), y0);
    const double sp = (s - t) - q;
    const unsigned vq = (unsigned) __builtin_bit_cast(unsigned long long,
        fk_min64(q + bn.q_c, bn.q_capp));
    const unsigned vs = (unsigned) __builtin_bit_cast(unsigned long long,
        fk_min64(fk_max64(sp + bn.q_c, 68719476736.0), bn.q_caps));
    const unsigned eq = fk_lds_u32(fk_mad4_hi16(vq, op));
    const unsigned es = fk_lds_u32(fk_mad4_hi16(vs, os));
    unsigned idx = fk_mad_lo16(eq, (unsigned) bn.nsh, es & 0xFFFFu);
    const bool clear = (bb >= bn.q_bmin) & ((vq & 0xFFFFu) <= (eq >> 16)) &
                       ((vs & 0xFFFFu) <= (es >> 16));
    if (!clear) {
      const double p2 = n / bb;
      const double sp2 = (s - t) - p2;
      const unsigned ks = min((unsigned) (int) sp2, (unsigned) (bn.soff + bn.stab_len)) + 2u;
      const unsigned kp = min((unsigned) (int) p2, (unsigned) (bn.poff + bn.ptab_len)) + 2u;
      idx = (stab[ks] & 0xFFFFu) + __umul24((unsigned) bn.nsh, ptab[kp] & 0xFFFFu);
    }
    return idx;
  }
  const double p2 = (d * d) / (t + s);
  const double sp2 = (s - t) - p2;
  const unsigned sb = fk_sbin<HYB, MK>(sp2, bn, stab, s2e);
  const unsigned pb = fk_pbin<HYB, MK>(p2, bn, ptab, pe);
  return sb + __umul24((unsigned) bn.nsh, pb);
}

/* The other arithmetic paths of the reference, for the rounding census
 * (box only): vectorised a-side ((a+S)-b, var 0), per-pair periodic wrap
 * (var 2) and the unfused scalar remainder path (var 3). */
template <int BIN, bool HYB>
__device__ unsigned fk_box_variant(int var, const double a[3], const double bo[3],
    const int dcode[3], const FkGrid &g, const FkBins &bn, const unsigned *stab,
    const unsigned *ptab, const double *s2e, const double *pe) {
  const double L[3] = {g.Lx, g.Ly, g.Lz};
  double d[3];
  for (int k = 0; k < 3; k++) {
    const double D = dcode[k] * L[k];        /* b image = b + D, S = -D */
    if (var == 0 || var == 3) d[k] = (a[k] - D) - bo[k];
    else {
      double v = a[k] - bo[k];
      const double h = L[k] * 0.5;
      if (v > h) v = v - L[k]; else if (v < -h) v = v + L[k];
      d[k] = v;
    }
  }
  const int scalar = (var == 3);
  if (BIN == FCFC_BIN_SPI) {
    const double pi = fabs(d[2]);
    const double s2 = scalar ? (d[0] * d[0] + d[1] * d[1]) : __builtin_fma(d[1], d[1], d[0] * d[0]);
    const unsigned sb = fk_sbin<HYB, true>(s2, bn, stab, s2e);
    const unsigned pb = fk_pbin<HYB, true>(pi, bn, ptab, pe);
    return sb + __umul24((unsigned) bn.nsh, pb);
  }
  const double dz2 = d[2] * d[2];
  double s2;
  if (scalar) { double t = d[0] * d[0]; t = t + d[1] * d[1]; s2 = t + dz2; }
  else s2 = __builtin_fma(d[1], d[1], __builtin_fma(d[0], d[0], dz2));
  const unsigned sb = fk_sbin<HYB, true>(s2, bn, stab, s2e);
  if (BIN == FCFC_BIN_ISO || sb == FK_OUT) return sb;
  int m = 0;
  if (s2 >= FK_DBL_EPS) {
    if (scalar) {
      const double q = (dz2 / s2) * bn.nmu2;           /* :184 */
      if (q >= bn.nmu2) return FK_NONE;
      m = (int) q;
    }
    else {
      const double mq = fk_floor_div(dz2 * bn.nmu2, s2);
      if (mq >= bn.nmu2) return FK_NONE;
      m = (int) mq;
    }
  }
  return sb + __umul24((unsigned) bn.nsh, (unsigned) fk_isqrt(m));
}

/* ---------------- FP32 pair evaluation with exactness windows ----------- */
/* Coordinates relative to the A tile centre in FP32.  Returns the histogram
 * index (>= ntot: not counted; (s, mu) bins: the index plus 0x4301 nsh, see
 * fk_mu32) and sets ambig when the pair lies within the
 * rounding bound of a bin edge that matters (it is then re-evaluated in
 * FP64 and the index is ignored).  Window test of s^2 (and pi): with
 * u = s2 + e, s2 is within e of an integer k exactly when fract(u) < 2e,
 * and then k = (int) u; otherwise (int) u = (int) s2.  So the table entry
 * at (int) u is the cell of s2 when s2 is clear of the integers, and the
 * cell above the nearby integer otherwise, whose FK_F_EDGE flag says
 * whether that integer is a bin edge (if not, both cells give the same
 * bin).  A cell of the FP32 tables holds the bin in its low 16 bits and in
 * its high 16 bits those of a float threshold: -1 without flags, the window
 * 2e rounded up (the largest of the launch's jobs) with FK_F_EDGE, 2 with
 * FK_F_ALW.  Read as a float, the cell lies within 2^-7 relative of its
 * threshold, on the side away from zero (thresholds are normal numbers:
 * comparisons with subnormals are slow), so the test is one compare of the
 * cell itself: fract(u) < cell.  A pair whose s (or pi) is certainly out of
 * range reads a flagless sentinel. */
/* FP32 mu bin floor(t) from x = anmu r (|x - t| < D above s2_floor, the
 * cells below s2_floor carry FK_F_ALW: every pair there is re-evaluated;
 * t <= nmu (1 + 2^-20) <= 126), in fixed point as in fk_mu_exact: the fma
 * with mu_c = 129 - sh (sh a multiple of 2^-16, sh >= D + 2^-16) rounds
 * once onto the grid 2^-16 of [128, 256), so the float's bits are
 * 0x43000000 + round((x - sh + 1) 2^16): the high half is 0x4300 + K, K =
 * floor(x - sh) + 1, the low half the fraction f.  f < mu_t16 = floor(2^16
 * (1 - 2 sh)) puts the exact t strictly inside (K - 1, K); otherwise
 * ambig_mu.  Returns the index plus 0x4301 nsh: (0x4300 + K) nsh + sb. */
__device__ __forceinline__ unsigned fk_mu32(float anmu, float r, unsigned sb,
    const FkBins &bn, const FkWin &wn, bool &ambig_mu) {
  const unsigned v = __float_as_uint(__builtin_fmaf(anmu, r, wn.mu_c));
  ambig_mu = !((v & 0xFFFFu) < wn.mu_t16);
  return fk_mad_hi16(v, (unsigned) bn.nsh, sb);
}

template <int BIN>
__device__ __forceinline__ unsigned fk_box32(float dx, float dy, float dz,
    const FkBins &bn, const FkWin &wn, const unsigned *stab,
    const unsigned *ptab, bool &ambig, bool &ambig_mu) {
  ambig_mu = false;
  float s2;
  if (BIN == FCFC_BIN_SPI) s2 = __builtin_fmaf(dy, dy, dx * dx);
  else s2 = __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, dz * dz));
  /* u >= 0: truncation is the floor.  The zero-based FP32 tables run to
   * (int) s2min + len + 1: the entry after the last cell flags the edge at
   * s2max, the last one is a plain FK_OUT for pairs certainly out of range
   * (as are the entries below s2min) */
  const float u = s2 + wn.e_s2;
  const unsigned e = stab[min((int) u, bn.scap32)];
  const bool amb = __builtin_amdgcn_fractf(u) < __uint_as_float(e);
  const unsigned sb = e & 0xFFFFu;
  if (BIN == FCFC_BIN_ISO) { ambig = amb; return sb; }
  if (BIN == FCFC_BIN_SPI) {
    const float up = __builtin_fabsf(dz) + wn.e_pi;
    const unsigned ep = ptab[min((int) up, bn.pcap32)];
    const bool ambp = __builtin_amdgcn_fractf(up) < __uint_as_float(ep);
    const unsigned pb = ep & 0xFFFFu;
    const bool out = ((sb == FK_OUT) & !amb) | ((pb == FK_OUT) & !ambp);
    ambig = (amb | ambp) & !out;
    return sb + __umul24((unsigned) bn.nsh, pb);
  }
  /* mu bin = floor(t), t = nmu |dz| / s (fk_mu32).  A pair out of range in
   * s whose t is near an integer is re-evaluated too (rare, and cheaper than
   * testing sb here); idle lanes sit where t = 1/2. */
  ambig = amb;                                 /* the caller ORs the two */
  return fk_mu32(wn.nmu_f * __builtin_fabsf(dz), __builtin_amdgcn_rsqf(s2), sb, bn, wn, ambig_mu);
}

/* FP32 survey pair.  d = b - a from coordinates relative to the A tile
 * centre (as for box pairs) and m = a + b = 2 a + d (m* = 2 a in FP32), so
 * that s^2 = |d|^2, pi = d.m / |m| and mu = pi / s carry the FP32 error of
 * differences, not the cancellation of |a|^2 + |b|^2 - 2 a.b that the
 * reference's FP64 formulas have (metric_common.c:338-364): its values
 * differ from the exact ones by its own rounding, which the windows include
 * (fp32_windows).  Pairs with |m|^2 below m2_floor (both objects near the
 * observer, on opposite sides) are always re-evaluated, as are pairs of
 * cells flagged for it (the three conditions come back apart, so that each
 * ballot is its compare's own mask).  (s_perp, pi): the tables hold pi^2 and s_perp^2 =
 * s^2 - pi^2, whose FP32 value may come out slightly negative (index 0, as
 * the reference's truncation gives); lanes whose indices come out below -1
 * (pairs below m2_floor) read the "out" sentinel. */
template <int BIN>
__device__ __forceinline__ unsigned fk_sv32(float dx, float dy, float dz, float mx, float my,
    float mz, const FkBins &bn, const FkWin &wn, const unsigned *stab, const unsigned *ptab,
    bool &ambig, bool &ambig_mu, bool &ambig_m) {
  ambig_mu = false;
  ambig_m = false;
  const float s2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
  if (BIN == FCFC_BIN_ISO) {
    const float u = s2 + wn.e_s2;
    const unsigned e = stab[min((int) u, bn.scap32)];
    ambig = __builtin_amdgcn_fractf(u) < __uint_as_float(e);
    return e & 0xFFFFu;
  }
  const float p = __builtin_fmaf(dz, mz, __builtin_fmaf(dy, my, dx * mx));
  const float m2 = __builtin_fmaf(mz, mz, __builtin_fmaf(my, my, mx * mx));
  ambig_m = m2 < wn.m2_floor;          /* near the observer: re-evaluate */
  if (BIN == FCFC_BIN_SPI) {
    const float p2 = (p * p) * __builtin_amdgcn_rcpf(m2);
    const float sp2 = s2 - p2;
    const float u = sp2 + wn.e_s2;
    const unsigned e = stab[min((unsigned) (int) u, (unsigned) bn.scap32)];
    const bool amb = __builtin_amdgcn_fractf(u) < __uint_as_float(e);
    const float up = p2 + wn.e_pi;
    const unsigned ep = ptab[min((unsigned) (int) up, (unsigned) bn.pcap32)];
    const bool ambp = __builtin_amdgcn_fractf(up) < __uint_as_float(ep);
    ambig = amb;
    ambig_mu = ambp;
    return (e & 0xFFFFu) + __umul24((unsigned) bn.nsh, ep & 0xFFFFu);
  }
  const float u = s2 + wn.e_s2;
  const unsigned e = stab[min((int) u, bn.scap32)];
  ambig = __builtin_amdgcn_fractf(u) < __uint_as_float(e);
  return fk_mu32(wn.nmu_f * __builtin_fabsf(p),
                 __builtin_amdgcn_rsqf(m2 * s2), e & 0xFFFFu, bn, wn,
                 ambig_mu);
}

/* Exact FP64 re-evaluation of a pair flagged by the FP32 windows (rare: the
 * pairs are queued and re-evaluated together after each chunk).  The tables
 * are the flagged FP32 ones, read through their bin bits.  With census
 * enabled the other rounding paths of the reference are evaluated as well
 * and pairs whose bin depends on the path are counted. */
template <int BIN, bool HYB>
__device__ __forceinline__ bool fk_census(const FkBins &bn, const FkJob &J, double ax,
    double ay, double az, const double bo[3], const int c3[3], unsigned idx,
    const unsigned *stab, const unsigned *ptab, const double *s2e, const double *pe) {
  const unsigned nt = (unsigned) bn.ntoh;
  const double av3[3] = {ax, ay, az};
  bool differ = false;
  for (int v = 0; v < 4; v++) {
    if (v == 1) continue;
    unsigned iv = fk_box_variant<BIN, HYB>(v, av3, bo, c3, J.g, bn, stab, ptab, s2e, pe);
    if (iv >= nt) iv = FK_NONE;
    differ = differ || (iv != idx);
  }
  return differ;
}

template <int BIN, bool HYB, bool CEN>
__device__ __forceinline__ unsigned fk_recheck(const FkBins &bn, const FkJob &J,
    double ax, double ay, double az, int jg, int code, const unsigned *stab,
    const unsigned *ptab, const double *s2e, const double *pe, unsigned int *ncensus) {
  const int c3[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
  const double bo[3] = {J.B.x[jg], J.B.y[jg], J.B.z[jg]};
  const double ex = bo[0] + c3[0] * J.g.Lx;
  const double ey = bo[1] + c3[1] * J.g.Ly;
  const double ez = bo[2] + c3[2] * J.g.Lz;
  unsigned idx = fk_box64<BIN, HYB, true>(ex - ax, ey - ay, ez - az, bn, stab, ptab, s2e, pe);
  if (BIN == FCFC_BIN_SMU) idx -= (unsigned) bn.nsh;
  if (idx >= (unsigned) bn.ntoh) idx = FK_NONE;
  if (CEN && fk_census<BIN, HYB>(bn, J, ax, ay, az, bo, c3, idx, stab, ptab, s2e, pe))
    (*ncensus)++;
  return idx;
}

/* The same for a survey pair: A object p, B object jg, with the reference's
 * FP64 formulas (fk_sv64). */
template <int BIN, bool HYB>
__device__ __forceinline__ unsigned fk_recheck_sv(const FkBins &bn, const FkJob &J, int p,
    int jg, const unsigned *stab, const unsigned *ptab, const double *s2e, const double *pe) {
  const double as = (BIN == FCFC_BIN_ISO) ? 0.0 : J.A.s[p];
  const double bs = (BIN == FCFC_BIN_ISO) ? 0.0 : J.B.s[jg];
  unsigned idx = fk_sv64<BIN, HYB, false, true>(J.A.x[p], J.A.y[p], J.A.z[p], as, J.B.x[jg],
                                                J.B.y[jg], J.B.z[jg], bs, bn, stab, ptab, s2e, pe);
  if (BIN == FCFC_BIN_SMU) idx -= (unsigned) bn.nsh;
  return idx >= (unsigned) bn.ntoh ? FK_NONE : idx;
}

/* ---------------------------- the kernel --------------------------------- */

/* One persistent kernel for counts (FP64 or FP32 + recheck) and for the
 * weight gradient.  Template parameters: binning, geometry, weighted
 * histogram support, mode, hybrid tables, A objects per lane.  Unweighted
 * kernels are held to 64 VGPRs (8 waves per SIMD, 4 workgroups per CU):
 * occupancy pays more than the FP32 form's spills, which sit in its rare
 * re-evaluation paths.  Weighted ones get 80 (6 waves per SIMD), as their
 * FP64 LDS histograms limit them to 3 workgroups per CU anyway. */
template <int BIN, int GEOM, bool WT, int MODE, bool HYB, int R>
__global__ void __launch_bounds__(FK_NT, WT ? 6 : 8)
k_count(const FkArgs a) {
  extern __shared__ __attribute__((aligned(16))) char lds[];
  constexpr bool F32 = (MODE == MODE_C32 || MODE == MODE_C32C);
  constexpr bool CENS = (MODE == MODE_C32C);
  constexpr bool PLAIN = (MODE == MODE_P64);
  constexpr bool GRAD = (MODE == MODE_GRAD);
  constexpr bool SV32 = (GEOM == FCFC_GEOM_SURVEY && (MODE == MODE_C32 || MODE == MODE_C32C));
  /* |r|^2 of the objects: survey (s, mu) and (s_perp, pi) in FP64 (the FP32
   * form reads them from the catalogue when it re-evaluates a pair) */
  constexpr bool NEED_S = (GEOM == FCFC_GEOM_SURVEY && BIN != FCFC_BIN_ISO && !SV32);
  constexpr bool NEED_W = WT || GRAD;
  constexpr bool NEED_C = F32;
  const int tid = threadIdx.x;
  /* the wavefront index, in a scalar register (with everything derived
   * from it: the wavefront's histogram copy and queue) */
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const FkBins &bn = a.bins;
  const unsigned ntot = (unsigned) bn.ntoh;   /* LDS histogram slots */
  /* the pair functions' (s, mu) indices carry an offset (fk_mu_exact,
   * fk_mu32); indices below it are never counted.  MUT: the FP64 integer
   * table cells carry the mu fraction threshold (fk_smu_index) */
  constexpr bool MUT = (BIN == FCFC_BIN_SMU && !F32 && !PLAIN && !HYB);
  /* QT: survey (s_perp, pi) in the FP64 forms with integer tables, pi^2
   * and s_perp^2 read in fixed point (fk_sv64); the tables start at floor
   * cell -2 and carry fraction thresholds (qt_entry below) */
  constexpr bool QT = (GEOM == FCFC_GEOM_SURVEY && BIN == FCFC_BIN_SPI && !F32 && !PLAIN && !HYB);
  const unsigned ioff = (BIN != FCFC_BIN_SMU || PLAIN) ? 0u
                        : F32 ? 0x4301u * (unsigned) bn.nsh
                        : (unsigned) bn.nsh + (MUT ? bn.mu_t16 << 16 : 0u);
  const unsigned ntot_o = ntot + ioff;

  /* ---- LDS carve-up (lds_plan on the host computes the same offsets) ---- */
  unsigned int *hist_u = (unsigned int *) (lds + a.off_hist);
  double *hist_d = (double *) (lds + a.off_hist);
  /* tables: raw[c0 + c] is integer cell c, c0 = a.stab_c0 (a.ptab_c0)
   * entries precede cell 0 and two follow the last one.  Integer tables and
   * all FP32 ones are zero-based (c0 = (int) s2min: entry k is the one of
   * (int) s2 = k); hybrid ones for the FP64 forms may have c0 = 1.  Hybrid
   * lookups index cells from raw + c0 */
  unsigned *stab_raw = (unsigned *) (lds + a.off_stab);
  unsigned *ptab_raw = (unsigned *) (lds + a.off_ptab);
  const unsigned *stab = HYB ? stab_raw + a.stab_c0 : stab_raw;
  const unsigned *ptab = HYB ? ptab_raw + a.ptab_c0 : ptab_raw;
  const unsigned os_lds = fk_lds_addr(stab_raw), op_lds = fk_lds_addr(ptab_raw);
  /* FP32 re-evaluation queue of this wavefront: (A slot in the wavefront
   * << 8) | chunk index */
  uint16_t *wq = (uint16_t *) (lds + a.off_queue) + wave * FK_QCAP;
  double *s2e = (double *) (lds + a.off_edges);
  double *pe = s2e + (bn.ns + 1);
  FkStripLds *st = (FkStripLds *) (lds + a.off_strip);
  char *cp = lds + a.off_chunk;
  double *cx = (double *) cp, *cy = cx + FK_CH, *cz = cy + FK_CH;
  float *fx = (float *) cp, *fy = fx + FK_CH, *fz = fy + FK_CH;
  cp += (F32 ? 12 : 24) * FK_CH;
  double *cs_ = (double *) cp;                 /* survey |r|^2 */
  if (NEED_S) cp += 8 * FK_CH;
  double *cw = (double *) cp;                  /* weights */
  if (NEED_W) cp += 8 * FK_CH;
  double *cb = (double *) cp;                  /* GRAD: gradient share of each B object */
  if (GRAD) cp += 8 * FK_CH;
  int *cj = (int *) cp;                        /* sorted index */
  cp += 4 * FK_CH;
  int *cc = (int *) cp;                        /* periodic image code */
  __shared__ int s_tile;

  /* ---- stage tables ---- */
  const int hist_elems = a.nh * bn.ntoh;
  if (GRAD) {
    for (int i = tid; i < bn.ntoh; i += FK_NT) {
      const int g = fk_hslot(i, bn.ns, bn.nsh);
      hist_d[i] = g >= 0 ? a.G[g] : 0.0;
    }
  }
  else if (WT) { for (int i = tid; i < hist_elems; i += FK_NT) hist_d[i] = 0.0; }   /* covers the u32 view */
  else { for (int i = tid; i < hist_elems; i += FK_NT) hist_u[i] = 0u; }
  {
    /* cell c of a table of len cells; FP32 entries carry the edge flags */
    auto cell = [](const uint16_t *t, int len, int c) -> unsigned {
      return (c >= 0 && c < len) ? (unsigned) t[c] : FK_OUT;
    };
    /* alw_below > 0: cells whose lower end lies below alw_below + 1 are
     * always re-evaluated (the extra cell covers pairs just below the first
     * other cell, which the window test reads through that cell).  Hybrid
     * tables (lut_hybrid, setup_cf.c) give the cell holding a non-integer
     * lower range edge the first bin and leave values below the edge to the
     * lookup's explicit range check, so that cell is always re-evaluated as
     * well (refinement cells, v >= nb, always are). */
    auto entry = [&](const uint16_t *t, int len, int nb, int c, double off, double alw_below,
                     double lo_edge) -> unsigned {
      const unsigned v = cell(t, len, c);
      if (!F32) return v;
      unsigned f = 0;
      if (cell(t, len, c - 1) != v) f |= FK_F_EDGE;
      /* the cell after one holding a non-integer lower edge: a value just
       * below that edge may land at its lower end (same bin, no edge of the
       * table there) */
      if (HYB && c == 1 && off != lo_edge) f |= FK_F_EDGE;
      if (v != FK_OUT && (v >= (unsigned) nb || (alw_below > 0 && off + c < alw_below + 1.0) ||
                          (HYB && c == 0 && off != lo_edge)))
        f |= FK_F_ALW | FK_F_EDGE;
      return v | f;
    };
    /* FP32 tables: bin and the high half of the threshold of fract(u),
     * the window rounded up */
    auto put = [&](unsigned *raw, int i, unsigned v, float win2) {
      if (!F32) { raw[i] = v; return; }
      const unsigned th = (v & FK_F_ALW) ? FK_F_ALWAYS
                          : ((v & FK_F_EDGE) ? (__float_as_uint(win2) + 0xFFFFu) >> 16 : FK_F_NEVER);
      raw[i] = (th << 16) | (v & 0xFFFFu);
    };
    /* QT tables: entry K holds the bin of (int) x for x in (K - 2, K - 1)
     * (entry 0: out; entry 1, x in (-1, 0): cell 0) and in its high half
     * the fraction threshold of fk_sv64's fixed-point lookup: q_t16 where
     * the bin changes at x = K - 1, 0xFFFF where it does not; off + len + 4
     * entries (the last one out, also for the ceilings q_caps / q_capp) */
    auto qt_entry = [&](const uint16_t *t, int len, int off, int k) -> unsigned {
      auto low = [&](int e) -> unsigned { return e == 0 ? FK_OUT : cell(t, len, max(e - 2, 0) - off); };
      const unsigned v = low(k);
      return ((v != low(k + 1) ? bn.q_t16 : 0xFFFFu) << 16) | v;
    };
    if (QT) {
      for (int i = tid; i < bn.soff + bn.stab_len + 4; i += FK_NT)
        stab_raw[i] = qt_entry(bn.stab, bn.stab_len, bn.soff, i);
      for (int i = tid; i < bn.poff + bn.ptab_len + 4; i += FK_NT)
        ptab_raw[i] = qt_entry(bn.ptab, bn.ptab_len, bn.poff, i);
    }
    const double s2f = (F32 && BIN == FCFC_BIN_SMU) ? (double) a.jobs[0].win.s2_floor : 0.0;
    if (!QT)
    for (int i = tid; i < a.stab_c0 + bn.stab_len + 2; i += FK_NT) {
      unsigned v = entry(bn.stab, bn.stab_len, bn.ns, i - a.stab_c0, bn.soff, s2f, bn.s2min);
      /* MUT (zero-based: entry i holds (int) s2 = i): the mu fraction
       * threshold, 0 where the cell reaches below s2_exact */
      if (MUT && !((double) i < bn.s2_exact)) v |= bn.mu_t16 << 16;
      put(stab_raw, i, v, bn.win2s);
    }
    if (BIN == FCFC_BIN_SPI && !QT)
      for (int i = tid; i < a.ptab_c0 + bn.ptab_len + 2; i += FK_NT)
        put(ptab_raw, i, entry(bn.ptab, bn.ptab_len, bn.np, i - a.ptab_c0, bn.poff, 0.0, bn.pmin),
            bn.win2p);
  }
  if (HYB) {
    for (int i = tid; i <= bn.ns; i += FK_NT) s2e[i] = bn.s2bin[i];
    if (BIN == FCFC_BIN_SPI) for (int i = tid; i <= bn.np; i += FK_NT) pe[i] = bn.pbin[i];
  }

  double evals = 0;
  unsigned int nrecheck = 0, ncensus = 0;
  double since_flush = 0;
  unsigned int *myhist_u = hist_u + (wave % a.nh) * bn.ntoh;
  double *myhist_d = hist_d + (wave % a.nh) * bn.ntoh;
  /* byte offsets in lds of these histograms, minus the pair functions'
   * index offset times the slot size: lds + (o + idx * size) is slot
   * idx - ioff (unsigned arithmetic, modulo 2^32; in range for every
   * counted index) */
  const unsigned ohist_u = (unsigned) a.off_hist + 4u * ((unsigned) (wave % a.nh) * ntot - ioff);
  const unsigned ohist_d = (unsigned) a.off_hist + 8u * ((unsigned) (wave % a.nh) * ntot - ioff);
  const unsigned ohist_g = (unsigned) a.off_hist - 8u * ioff;
  __shared__ FkJob sJ;
  const FkJob &J = sJ;
  int cur = -1;

  /* add the LDS histograms of the current job to its global histogram */
  auto flush = [&]() {
    __syncthreads();
    if (cur >= 0) {
      const bool w = WT && J.wt;
      for (int i = tid; i < bn.ntoh; i += FK_NT) {
        const int g = fk_hslot(i, bn.ns, bn.nsh);    /* padding slots stay 0 */
        if (g < 0) continue;
        if (w) {
          double v = 0;
          for (int c = 0; c < a.nh; c++) { v += hist_d[c * bn.ntoh + i]; hist_d[c * bn.ntoh + i] = 0.0; }
          if (v != 0) unsafeAtomicAdd(a.gcnt_d + J.hist_off + g, v);
        }
        else {
          unsigned long long v = 0;
          for (int c = 0; c < a.nh; c++) { v += hist_u[c * bn.ntoh + i]; hist_u[c * bn.ntoh + i] = 0u; }
          if (v) atomicAdd(a.gcnt_u + J.hist_off + g, v);
        }
      }
    }
    __syncthreads();
  };

  for (;;) {
    __syncthreads();
    if (tid == 0) s_tile = (int) atomicAdd(a.queue, 1u);
    __syncthreads();
    const int tix = s_tile;
    if (tix >= a.ntiles) break;
    const FkTile t = a.tiles[tix];
    if (t.job != cur) {
      /* switch job: flush the histogram of the previous one, stage the new
       * job descriptor in LDS */
      if (!GRAD) flush();
      __syncthreads();
      {
        const int *src = (const int *) (a.jobs + t.job);
        int *dst = (int *) &sJ;
        for (int i = tid; i < (int) (sizeof(FkJob) / sizeof(int)); i += FK_NT) dst[i] = src[i];
      }
      __syncthreads();
      cur = t.job;
      since_flush = 0;
    }
    const int ix = t.row / J.g.ncy, iy = t.row - ix * J.g.ncy;
    /* weights exist for this job (a weighted-capable kernel may run
     * unweighted jobs of a batch, whose catalogues carry no weights) */
    const bool jw = GRAD || (WT && J.wt);

    /* A objects of this lane; idle lanes sit far away, so that every pair
     * they form is out of range */
    double ax[R], ay[R], az[R], as[R], aw[R], gacc[R];
    float fax[R], fay[R], faz[R];
    int ai[R];
    const double tcx = J.g.ox + (ix + 0.5) * J.g.wx;
    const double tcy = J.g.oy + (iy + 0.5) * J.g.wy;
    const double tcz = J.g.oz + 0.5 * (t.iz0 + t.iz1 + 1) * J.g.wz;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int p = t.p0 + tid * R + r;     /* a wavefront holds one run of 64 R objects */
      const bool v = p < t.p1;
      ai[r] = v ? p : 0x7fffffff;
      ax[r] = v ? J.A.x[p] : FK_FAR;
      ay[r] = v ? J.A.y[p] : FK_FAR;
      az[r] = v ? J.A.z[p] : FK_FAR;
      as[r] = NEED_S ? (v ? J.A.s[p] : 3.0 * FK_FAR * FK_FAR) : 0.0;
      aw[r] = (jw && v) ? J.A.w[p] : 1.0;
      gacc[r] = 0.0;
      /* idle lanes: |a| = 1e18 and |dz| / s = 1 / (2 nmu), so that the mu
       * estimate of their pairs is 1/2, clear of the windows */
      fax[r] = v ? (float) (ax[r] - tcx) : J.win.idle_x;
      fay[r] = v ? (float) (ay[r] - tcy) : 0.0f;
      faz[r] = v ? (float) (az[r] - tcz) : J.win.idle_z;
    }
    /* survey FP32: m = a + b = 2 a' + d + 2 c (a' = a - c, c = tile centre).
     * Idle lanes (a' = (10^6, 0, 0): s^2 and pi^2 far out of range) may
     * look ambiguous in mu: their queue bits are masked off */
    const float c2x = fk_uni((float) (2.0 * tcx)), c2y = fk_uni((float) (2.0 * tcy)),
                c2z = fk_uni((float) (2.0 * tcz));
    /* the FP32 windows of the job, in scalar registers for the pair loop */
    const FkWin wn = F32 ? fk_uni(J.win) : J.win;
    unsigned long long actm[R];
#pragma unroll
    for (int r = 0; r < R; r++) actm[r] = __builtin_amdgcn_ballot_w64(t.p0 + tid * R + r < t.p1);
    const int nva = min(t.p1 - t.p0, R * FK_NT);
    const bool jwt = WT && J.wt;

    /* FP32: FP64 re-evaluation of the n queued pairs of this wavefront
     * (entries: A slot << 8 | chunk index), while the chunk is staged */
    const int lane = tid & 63;
    auto drain = [&](int n) {
      for (int e0 = 0; e0 < n; e0 += 64) {
        const unsigned v = (e0 + lane < n) ? (unsigned) wq[e0 + lane] : 0xFFFFu;
        const int p = t.p0 + wave * 64 * R + (int) (v >> 8);
        /* entries of idle lanes (possible only when the windows cover
         * everything) are dropped */
        if (e0 + lane < n && p < t.p1) {
          const int j = (int) (v & 255u);
          unsigned idx;
          if (GEOM == FCFC_GEOM_SURVEY)
            idx = fk_recheck_sv<BIN, HYB>(bn, J, p, cj[j], stab, ptab, s2e, pe);
          else
            idx = fk_recheck<BIN, HYB, CENS>(bn, J, J.A.x[p], J.A.y[p], J.A.z[p], cj[j], cc[j],
                                             stab, ptab, s2e, pe, &ncensus);
          nrecheck++;
          if (idx < ntot) {
            if (WT && jwt) atomicAdd(myhist_d + idx, J.A.w[p] * cw[j]);
            else atomicAdd(myhist_u + idx, 1u);
          }
        }
      }
    };

    /* one staged chunk of cnt B objects against the lane's A objects */
    auto run_chunk = [&](auto selfc, int cnt) {
      constexpr bool SELF = decltype(selfc)::value;
      int qn = 0;                 /* queued pairs of this wavefront */
#pragma unroll 2
      for (int j = 0; j < cnt; j++) {
        double bx = 0, by = 0, bz = 0, bs = 0;
        float gx = 0, gy = 0, gz = 0;
        if (F32) { gx = fx[j]; gy = fy[j]; gz = fz[j]; }
        else { bx = cx[j]; by = cy[j]; bz = cz[j]; }
        if (NEED_S) bs = cs_[j];
        const double wj = NEED_W ? cw[j] : 1.0;
        const int jj = (SELF || F32) ? cj[j] : 0;
        if (GEOM == FCFC_GEOM_BOX) {
          /* the wavefront's A objects are one run of consecutive objects
           * of one cell column: when this B object is out of range of all
           * of them, skip the binning (same expressions as the pair
           * functions, so the compiler shares them).  Every form skips, so
           * the plain form differs from the FP64 one in the per-pair mu
           * arithmetic only */
          bool any = false;
#pragma unroll
          for (int r = 0; r < R; r++) {
            if (F32) {
              const float ex = gx - fax[r], ey = gy - fay[r], ez = gz - faz[r];
              if (BIN == FCFC_BIN_SPI)
                any |= (__builtin_fmaf(ey, ey, ex * ex) < bn.s2skip32) &
                       (__builtin_fabsf(ez) < bn.pskip32);
              else any |= __builtin_fmaf(ey, ey, __builtin_fmaf(ex, ex, ez * ez)) < bn.s2skip32;
            }
            else {
              const double ex = bx - ax[r], ey = by - ay[r], ez = bz - az[r];
              if (BIN == FCFC_BIN_SPI)
                any |= (__builtin_fma(ey, ey, ex * ex) < bn.s2skip) &
                       (__builtin_fabs(ez) < bn.pskip);
              else any |= __builtin_fma(ey, ey, __builtin_fma(ex, ex, ez * ez)) < bn.s2skip;
            }
          }
          if (__builtin_amdgcn_ballot_w64(any) == 0) continue;
        }
        if (GEOM == FCFC_GEOM_SURVEY) {
          /* the same skip for survey pairs: |a - b|^2 (iso) or s - t
           * (= s^2 = s_perp^2 + pi^2) beyond the largest counted value */
          bool any = false;
#pragma unroll
          for (int r = 0; r < R; r++) {
            if (SV32) {
              const float ex = gx - fax[r], ey = gy - fay[r], ez = gz - faz[r];
              any |= __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex)) < bn.s2skip32;
            }
            else if (BIN == FCFC_BIN_ISO) {
              const double ex = ax[r] - bx, ey = ay[r] - by, ez = az[r] - bz;
              any |= __builtin_fma(ez, ez, __builtin_fma(ey, ey, ex * ex)) < bn.s2skip;
            }
            else {
              const double tt = __builtin_fma(az[r], bz, __builtin_fma(ay[r], by, ax[r] * bx)) * 2.0;
              any |= (as[r] + bs) - tt < bn.s2skip;
            }
          }
          if (__builtin_amdgcn_ballot_w64(any) == 0) continue;
        }
        double bsum = 0.0;          /* GRAD: sum of G[bin] w_a over this lane's pairs */
#pragma unroll
        for (int r = 0; r < R; r++) {
          unsigned idx;
          bool queued = false;      /* FP32: left to the FP64 re-evaluation */
          if (F32) {
            bool amb, ambm;
            bool ambx = false;
            if (SV32) {
              const float ex = gx - fax[r], ey = gy - fay[r], ez = gz - faz[r];
              idx = fk_sv32<BIN>(ex, ey, ez, __builtin_fmaf(2.0f, fax[r], ex) + c2x,
                                 __builtin_fmaf(2.0f, fay[r], ey) + c2y,
                                 __builtin_fmaf(2.0f, faz[r], ez) + c2z, bn, wn, stab_raw,
                                 ptab_raw, amb, ambm, ambx);
            }
            else
              idx = fk_box32<BIN>(gx - fax[r], gy - fay[r], gz - faz[r], bn, wn, stab_raw,
                                  ptab_raw, amb, ambm);
            if (SELF) {
              const bool keep = jj > ai[r];
              amb = amb & keep; ambm = ambm & keep; idx = keep ? idx : FK_NONE;
            }
            /* pairs near an edge: queue (A slot, chunk entry) for FP64.  The
             * two conditions are balloted apart (each ballot is its compare's
             * own mask) and the lane's bit is read back from the mask inside
             * the rare branch, so no lane value is kept across it */
            if (SELF) ambx = ambx & (jj > ai[r]);
            unsigned long long m = __builtin_amdgcn_ballot_w64(amb) |
                __builtin_amdgcn_ballot_w64(ambm) | __builtin_amdgcn_ballot_w64(ambx);
            if (SV32) m &= actm[r];
            if (m) {
              if ((m >> lane) & 1ull) {
                const int pos = qn + (int) __builtin_amdgcn_mbcnt_hi((unsigned) (m >> 32),
                                           __builtin_amdgcn_mbcnt_lo((unsigned) m, 0u));
                wq[pos] = (uint16_t) (((unsigned) (lane * R + r) << 8) | (unsigned) j);
              }
              qn += __popcll(m);
              if (qn > FK_QCAP - 64) { drain(qn); qn = 0; }
            }
            queued = amb | ambm | ambx;
          }
          else if (GEOM == FCFC_GEOM_BOX)
            idx = fk_box64<BIN, HYB, false, PLAIN, MUT>(bx - ax[r], by - ay[r], bz - az[r], bn,
                                                         stab, ptab, s2e, pe);
          else
            idx = fk_sv64<BIN, HYB, PLAIN, false, MUT || QT>(ax[r], ay[r], az[r], as[r], bx, by,
                                                             bz, bs, bn, stab, ptab, s2e, pe,
                                                             os_lds, op_lds);
          if (!F32 && SELF && !(jj > ai[r])) idx = FK_NONE;
          if (idx < ntot_o && !queued) {
            if (GRAD) {
              const double gk = *(const double *) (lds + (ohist_g + 8u * idx));
              gacc[r] += gk * wj;
              bsum += gk * aw[r];
            }
            else if (WT && jwt) atomicAdd((double *) (lds + (ohist_d + 8u * idx)), aw[r] * wj);
            else atomicAdd((unsigned int *) (lds + (ohist_u + 4u * idx)), 1u);
          }
        }
        if (GRAD && J.gB) {
          /* the B object's share over the wavefront's pairs: one sum per
           * row of 16 lanes, added into the chunk's LDS accumulator */
          const double v = fk_row_sum(bsum);
          if ((lane & 15) == 0 && v != 0.0) atomicAdd(cb + j, v);
        }
      }
      if (F32 && qn) drain(qn);
    };

    /* GRAD: after a chunk, the B objects' shares go to the output */
    auto grad_flush = [&](int cnt) {
      if (!GRAD || !J.gB) return;
      __syncthreads();
      for (int q = tid; q < cnt; q += FK_NT) {
        const double v = cb[q];
        if (v != 0.0) unsafeAtomicAdd(J.gB + J.B.idx[cj[q]], v);
      }
    };

    auto guard = [&](int cnt) {
      if (GRAD || jwt) return;
      since_flush += (double) cnt * nva;
      if (since_flush > 2.0e9) { flush(); since_flush = 0; }
    };

    /* ---- auto counts: pairs inside the tile (j > i), first item only ---- */
    if (J.autoc && t.s0 == 0) {
      for (int c0 = t.p0; c0 < t.p1; c0 += FK_CH) {
        const int cnt = min(FK_CH, t.p1 - c0);
        __syncthreads();
        for (int q = tid; q < cnt; q += FK_NT) {
          const int jg = c0 + q;
          const double x = J.A.x[jg], y = J.A.y[jg], z = J.A.z[jg];
          if (F32) { fx[q] = (float) (x - tcx); fy[q] = (float) (y - tcy); fz[q] = (float) (z - tcz); cc[q] = 13; }
          else { cx[q] = x; cy[q] = y; cz[q] = z; }
          if (NEED_S) cs_[q] = J.A.s[jg];
          if (NEED_W) cw[q] = jw ? J.A.w[jg] : 1.0;
          if (GRAD) cb[q] = 0.0;
          cj[q] = jg;
        }
        __syncthreads();
        run_chunk(std::integral_constant<bool, true>(), cnt);
        grad_flush(cnt);
        evals += (double) cnt * nva * 0.5;
        guard(cnt);
      }
    }

    /* ---- strips in batches of one wavefront ---- */
    for (int sb0 = t.s0; sb0 < t.s1; sb0 += FK_NSB) {
      __syncthreads();
      if (wave == 0) {
        const int l = tid;
        int tot = 0;
        int b0[3], len[3], code[3];
        if (sb0 + l < t.s1) {
          const int *sp = J.strips + 3 * (sb0 + l);
          fk_strip_segments(J, t, sp[0], sp[1], sp[2], b0, len, code);
          tot = len[0] + len[1] + len[2];
        }
        else { b0[0] = b0[1] = b0[2] = 0; len[0] = len[1] = len[2] = 0; code[0] = code[1] = code[2] = 13; }
        for (int k = 0; k < 3; k++) { st->b0[3 * l + k] = b0[k]; st->len[3 * l + k] = len[k]; st->code[3 * l + k] = code[k]; }
        int inc = tot;
        for (int o = 1; o < 64; o <<= 1) {
          int v = __shfl_up(inc, o, 64);
          if (l >= o) inc += v;
        }
        st->off[l] = inc - tot;
        if (l == 63) st->off[64] = inc;
      }
      __syncthreads();
      const int total = st->off[FK_NSB];
      for (int c0 = 0; c0 < total; c0 += FK_CH) {
        const int cnt = min(FK_CH, total - c0);
        if (c0 > 0) __syncthreads();
        for (int q = tid; q < cnt; q += FK_NT) {
          const int pos = c0 + q;
          int lo = 0, hi = FK_NSB;          /* off[lo] <= pos < off[hi] */
          while (hi - lo > 1) { int mid = (lo + hi) >> 1; if (st->off[mid] <= pos) lo = mid; else hi = mid; }
          int rem = pos - st->off[lo];
          int sgi = 3 * lo;
          if (rem >= st->len[sgi]) { rem -= st->len[sgi]; sgi++; if (rem >= st->len[sgi]) { rem -= st->len[sgi]; sgi++; } }
          const int jg = st->b0[sgi] + rem;
          const int code = st->code[sgi];
          double x = J.B.x[jg], y = J.B.y[jg], z = J.B.z[jg];
          if (GEOM == FCFC_GEOM_BOX && code != 13) {
            x = x + (double) (code % 3 - 1) * J.g.Lx;
            y = y + (double) ((code / 3) % 3 - 1) * J.g.Ly;
            z = z + (double) (code / 9 - 1) * J.g.Lz;
          }
          if (F32) { fx[q] = (float) (x - tcx); fy[q] = (float) (y - tcy); fz[q] = (float) (z - tcz); cc[q] = code; }
          else { cx[q] = x; cy[q] = y; cz[q] = z; }
          if (NEED_S) cs_[q] = J.B.s[jg];
          if (NEED_W) cw[q] = jw ? J.B.w[jg] : 1.0;
          if (GRAD) cb[q] = 0.0;
          cj[q] = jg;
        }
        __syncthreads();
        run_chunk(std::integral_constant<bool, false>(), cnt);
        grad_flush(cnt);
        evals += (double) cnt * nva;
        guard(cnt);
      }
    }

    if (GRAD && J.gA) {
#pragma unroll
      for (int r = 0; r < R; r++)
        if (ai[r] != 0x7fffffff && gacc[r] != 0.0) unsafeAtomicAdd(J.gA + J.A.idx[ai[r]], gacc[r]);
    }
  }

  if (!GRAD) flush();
  if (tid == 0) atomicAdd(a.stats + 0, (unsigned long long) evals);
  if (F32) {
    if (nrecheck) atomicAdd(a.stats + 1, (unsigned long long) nrecheck);
    if (ncensus) atomicAdd(a.stats + 2, (unsigned long long) ncensus);
  }
}

/* add a device histogram into the caller's device array */
__global__ void k_add_u64(long long *dst, const unsigned long long *src, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] += (long long) src[i];
}
__global__ void k_add_f64(double *dst, const double *src, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] += src[i];
}

/* ======================================================================== */
/*                                host side                                 */
/* ======================================================================== */

/* Device scratch of the no-allocation count (fcfc_gpu_count_noalloc): every
 * device buffer of such a call -- uploads of host catalogues, sorted copies,
 * cell and tiling arrays, histograms, queue and counters -- is carved from
 * this array.  Each GPU the library runs on holds its own copy until the
 * process ends.  In host code the name is only a stand-in: fcfc_gpu_init
 * resolves the address of the current GPU's copy with hipGetSymbolAddress
 * (once per device) and every host use goes through that address. */
#define FK_SCRATCH_BYTES ((size_t) 128 << 20)
#define FK_MAXDEV 64
static __device__ __align__(256) unsigned char scratch[FK_SCRATCH_BYTES];
static void *g_scratch_ptr[FK_MAXDEV];
static std::mutex g_scratch_mu;

struct DevBuf { void *p = nullptr; size_t cap = 0; };

struct fcfc_gpu_ctx {
  int device = 0;
  int ncu = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  char err[512] = {0};
  fcfc_gpu_stats stats{};
  int t_cell_div = 0, t_tile = 0, t_zdiv = 0, t_nblock = 0;
  int t_nh = 0;                /* LDS histogram copies (0: by occupancy) */
  int census = 0;
  int tile_frac = 0;           /* autotuner: count about 1/tile_frac of the
                                  tiles (16 runs spread over the list) */
  /* tiles (before their split into work items) of the last launch built */
  int last_ntiles = 0;
  int last_nsampled = 0;       /* ... and how many of them it counted */
  struct Tuned { int geom, bintype, mode, autoc, wt, bucket, cd, zd, T; };
  std::vector<Tuned> tuned;    /* autotuner results per density bucket */
  std::vector<DevBuf> bufs;
  void *scratch = nullptr;     /* device address of this GPU's scratch array */
  char scratch_err[128] = {0};
};

static void set_err(fcfc_gpu_ctx *c, const char *fmt, ...) {
  if (!c) return;
  va_list ap; va_start(ap, fmt);
  vsnprintf(c->err, sizeof c->err, fmt, ap);
  va_end(ap);
}

#define HIPCHK(ctx, call) do { hipError_t _e = (call); if (_e != hipSuccess) { \
    set_err(ctx, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, __LINE__); \
    return FCFC_ERR_GPU; } } while (0)

/* workspace slots */
enum {
  WS_IN_X, WS_IN_Y, WS_IN_Z, WS_IN_S, WS_IN_W,          /* uploads, catalogue A */
  WS_IN2_X, WS_IN2_Y, WS_IN2_Z, WS_IN2_S, WS_IN2_W,     /* uploads, catalogue B */
  WS_A_X, WS_A_Y, WS_A_Z, WS_A_S, WS_A_W, WS_A_IDX, WS_A_CS,
  WS_B_X, WS_B_Y, WS_B_Z, WS_B_S, WS_B_W, WS_B_IDX, WS_B_CS,
  WS_CID, WS_RANK, WS_COUNT, WS_GSTART, WS_STAGE, WS_TJOBS, WS_ROWN, WS_ROWOFF, WS_TILES,
  WS_STRIPS, WS_STAB, WS_PTAB, WS_S2E, WS_PE, WS_HIST, WS_MISC, WS_G, WS_GA, WS_GB, WS_JOBS,
  WS_NSLOT
};

static void *ws(fcfc_gpu_ctx *c, int slot, size_t bytes, int *err) {
  if ((int) c->bufs.size() < WS_NSLOT) c->bufs.resize(WS_NSLOT);
  DevBuf &b = c->bufs[slot];
  if (bytes == 0) bytes = 8;
  if (b.cap < bytes) {
    if (b.p) { (void) hipStreamSynchronize(c->stream); (void) hipFree(b.p); b.p = nullptr; b.cap = 0; }
    size_t nb = bytes + bytes / 8 + 256;
    if (hipMalloc(&b.p, nb) != hipSuccess) {
      set_err(c, "hipMalloc of %zu bytes failed", nb);
      *err = FCFC_ERR_MEMORY;
      return nullptr;
    }
    b.cap = nb;
  }
  return b.p;
}

extern "C" int fcfc_gpu_init(fcfc_gpu_ctx **out, int device) {
  if (!out) return FCFC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FCFC_ERR_GPU;
  if (device < 0 || device >= ndev) return FCFC_ERR_ARG;
  fcfc_gpu_ctx *c = new (std::nothrow) fcfc_gpu_ctx();
  if (!c) return FCFC_ERR_MEMORY;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return FCFC_ERR_GPU; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return FCFC_ERR_GPU; }
  c->ncu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c; return FCFC_ERR_GPU;
  }
  /* address of this GPU's scratch array (no-allocation counts), resolved
   * once per device with the device current; a failure leaves the context
   * usable and makes fcfc_gpu_count_noalloc report it */
  if (device < FK_MAXDEV) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e == hipSuccess && p) g_scratch_ptr[device] = p;
      else snprintf(c->scratch_err, sizeof c->scratch_err, "hipGetSymbolAddress: %s",
                    hipGetErrorString(e));
    }
    c->scratch = g_scratch_ptr[device];
  }
  else snprintf(c->scratch_err, sizeof c->scratch_err, "device index beyond %d", FK_MAXDEV - 1);
  *out = c;
  return 0;
}

extern "C" int fcfc_gpu_destroy(fcfc_gpu_ctx *c) {
  if (!c) return 0;
  (void) hipSetDevice(c->device);
  if (c->stream) (void) hipStreamSynchronize(c->stream);
  for (auto &b : c->bufs) if (b.p) (void) hipFree(b.p);
  if (c->ev0) (void) hipEventDestroy(c->ev0);
  if (c->ev1) (void) hipEventDestroy(c->ev1);
  if (c->stream) (void) hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" const char *fcfc_gpu_last_error(const fcfc_gpu_ctx *c) {
  return c ? c->err : "no context";
}

extern "C" int fcfc_gpu_last_stats(const fcfc_gpu_ctx *c, fcfc_gpu_stats *st) {
  if (!c || !st) return FCFC_ERR_ARG;
  *st = c->stats;
  return 0;
}

extern "C" int fcfc_gpu_set_tuning(fcfc_gpu_ctx *c, int cell_div, int tile,
    int zdiv, int nblock) {
  if (!c) return FCFC_ERR_ARG;
  if (cell_div < 0 || cell_div > 32 || tile < 0 || tile > 4 * FK_NT || zdiv < 0 || zdiv > 64 || nblock < 0)
    return FCFC_ERR_ARG;
  c->t_cell_div = cell_div; c->t_tile = tile; c->t_zdiv = zdiv; c->t_nblock = nblock;
  return 0;
}

/* ---- per-job plan ------------------------------------------------------ */

struct JobPlan {
  int bucket;       /* density bucket for the autotuner cache */
  FkGrid g;
  int T;            /* objects per tile */
  int zmaxc;        /* max z-cells per tile */
  int R;            /* objects per lane */
  int T_auto;       /* T chosen by the plan (not by the caller's tuning) */
  std::vector<int> strips;
  int nstrips;
};

static double max_sep2(const fcfc_bins *b, double *rxy2, double *hzc) {
  double s2max = b->s2bin[b->ns];
  *hzc = -1;
  if (b->bintype == FCFC_BIN_SPI) {
    double pm = b->pbin[b->np];
    if (b->geometry == FCFC_GEOM_BOX) {
      *rxy2 = s2max;
      *hzc = pm;
      return s2max + pm * pm;
    }
    /* survey: s^2 = s_perp^2 + pi^2 with pbin holding pi^2 */
    *rxy2 = s2max + pm;
    return s2max + pm;
  }
  *rxy2 = s2max;
  return s2max;
}

/* Build the grid geometry and strip list.  lo/hi: bounding box (survey). */
static int make_plan(fcfc_gpu_ctx *c, const fcfc_bins *b, const double *bsize,
    const double lo[3], const double hi[3], int autoc, int wt, int mode, int64_t nA,
    int64_t nB, JobPlan *P) {
  double rxy2, hzc;
  double r2 = max_sep2(b, &rxy2, &hzc);
  /* pad: the survey formula carries ~1e-10 absolute rounding in s^2 */
  const double pad = (b->geometry == FCFC_GEOM_SURVEY) ? 1e-6 : 1e-9;
  r2 = r2 * (1 + pad) + pad;
  rxy2 = rxy2 * (1 + pad) + pad;
  if (hzc >= 0) hzc = hzc * (1 + pad) + pad;
  const double R = sqrt(r2), Rxy = sqrt(rxy2);
  FkGrid &g = P->g;
  memset(&g, 0, sizeof g);
  g.R2 = r2; g.Rxy2 = rxy2; g.hz_const = hzc;
  /* objects per unit volume decides the default cell division */
  int cd = c->t_cell_div;
  int zd = c->t_zdiv;
  double vol;
  double ext[3];
  if (b->geometry == FCFC_GEOM_BOX) {
    for (int k = 0; k < 3; k++) ext[k] = bsize[k];
  }
  else {
    for (int k = 0; k < 3; k++) { ext[k] = hi[k] - lo[k]; if (!(ext[k] > 0)) ext[k] = 1e-6 * (1 + fabs(hi[k])); }
  }
  vol = ext[0] * ext[1] * ext[2];
  const double nden = (double) std::max<int64_t>(nB, 1) / vol;
  /* shapes picked by fcfc_gpu_autotune for this kind of job */
  /* density buckets (half octaves of objects per separation volume) of
   * B and of A: a tuned shape serves counts of the same kind only */
  const double ndenA = (double) std::max<int64_t>(nA, 1) / vol;
  const double rv = sqrt(rxy2) * sqrt(rxy2) * sqrt(r2);
  const int bucket = 4096 * (int) floor(2 * log2(nden * rv + 1e-300)) +
                     (int) floor(2 * log2(ndenA * rv + 1e-300));
  int tT = c->t_tile;
  if (cd <= 0 || zd <= 0 || tT <= 0) {
    for (const auto &t : c->tuned)
      if (t.geom == b->geometry && t.bintype == b->bintype && t.mode == mode &&
          t.autoc == autoc && t.wt == wt && t.bucket == bucket) {
        if (cd <= 0) cd = t.cd;
        if (zd <= 0) zd = t.zd;
        if (tT <= 0) tT = t.T;
        break;
      }
  }
  P->bucket = bucket;
  if (cd <= 0) {
    /* three cells per maximum separation (measured best for dense boxes);
     * fewer for sparse catalogues, so that columns keep enough objects */
    double objs_per_col = nden * Rxy * Rxy;   /* per unit z, for cd = 1 */
    cd = 3;
    while (cd > 1 && objs_per_col / (cd * cd) < 4.0) cd--;
  }
  if (zd <= 0) zd = 8;
  /* cells of Rxy / cd, but never more cells than about 4 per object
   * (between 2^18 and 2^25): a sparse catalogue in a large volume gets
   * larger cells instead of a grid of mostly empty ones */
  const double cap = std::min(std::max(4.0 * (double) (nA + nB), 262144.0), 33554432.0);
  const double w_cap = cbrt(ext[0] * ext[1] * ext[2] * zd / cap);
  double w = std::max(Rxy / cd, w_cap);
  if (b->geometry == FCFC_GEOM_BOX) {
    g.periodic = 1;
    g.Lx = bsize[0]; g.Ly = bsize[1]; g.Lz = bsize[2];
    int n[2];
    for (int k = 0; k < 2; k++) {
      int nc = (int) floor(ext[k] / w);
      if (nc < 1) nc = 1;
      for (;;) {
        double wk = ext[k] / nc;
        int reach = (int) ceil(Rxy / wk);
        if (2 * reach + 1 <= nc) break;
        nc++;
        if (nc > 4096) { set_err(c, "separation range too close to half the box"); return FCFC_ERR_CFG; }
      }
      n[k] = nc;
    }
    g.ncx = n[0]; g.ncy = n[1];
    g.wx = ext[0] / g.ncx; g.wy = ext[1] / g.ncy;
    double wzt = std::min(g.wx, g.wy) / zd;
    /* z windows must stay shorter than the box: Lz - 2 R >= 4 wz */
    double zroom = ext[2] - 2 * (hzc >= 0 ? hzc : R);
    if (zroom <= 0) { set_err(c, "separation range not below half the box"); return FCFC_ERR_CFG; }
    if (wzt > zroom / 6) wzt = zroom / 6;
    g.ncz = std::max(1, (int) floor(ext[2] / wzt));
    g.wz = ext[2] / g.ncz;
    g.ox = g.oy = g.oz = 0;
    P->zmaxc = (int) floor(zroom / g.wz) - 4;
    if (P->zmaxc < 1) P->zmaxc = 1;
    P->zmaxc = std::min(P->zmaxc, std::max(1, (int) ceil(2 * zd * 1.0)));
  }
  else {
    g.periodic = 0;
    g.ncx = std::max(1, (int) ceil(ext[0] / w));
    g.ncy = std::max(1, (int) ceil(ext[1] / w));
    g.wx = w * (1 + 1e-12); g.wy = w * (1 + 1e-12);
    g.wz = w / zd;
    g.ncz = std::max(1, (int) ceil(ext[2] / g.wz) + 1);
    g.ox = lo[0]; g.oy = lo[1]; g.oz = lo[2];
    P->zmaxc = std::max(1, 2 * zd);
    if ((double) g.ncx * g.ncy * g.ncz > 2.0e8) {
      set_err(c, "grid too large (%d x %d x %d cells)", g.ncx, g.ncy, g.ncz);
      return FCFC_ERR_CFG;
    }
  }
  /* objects per tile: two per lane when the rows are long enough to fill
   * such tiles, one per lane otherwise (and always for the gradient) */
  P->T_auto = (tT <= 0);
  if (tT > 0) P->T = std::min(tT, 4 * FK_NT);
  else {
    /* two objects per lane only when the z extent allowed to a tile holds
     * enough objects to fill such tiles (idle lanes cost full issue slots) */
    const double per_window = (double) nA / ((double) g.ncx * g.ncy) * P->zmaxc / g.ncz;
    P->T = (per_window >= 2.5 * FK_NT) ? 2 * FK_NT : FK_NT;
  }
  P->R = (P->T + FK_NT - 1) / FK_NT;
  /* strips */
  P->strips.clear();
  const int kx = (int) ceil(Rxy / g.wx), ky = (int) ceil(Rxy / g.wy);
  for (int dx = -kx; dx <= kx; dx++) {
    for (int dy = -ky; dy <= ky; dy++) {
      double gx = std::max(abs(dx) - 1, 0) * g.wx, gy = std::max(abs(dy) - 1, 0) * g.wy;
      if (gx * gx + gy * gy >= rxy2) continue;
      int diag = 0;
      if (autoc) {
        /* auto counts and their gradient: half plane, and on the own row
         * only the objects after the tile */
        if (dx < 0 || (dx == 0 && dy < 0)) continue;
        if (dx == 0 && dy == 0) diag = 1;
      }
      P->strips.push_back(dx); P->strips.push_back(dy); P->strips.push_back(diag);
    }
  }
  P->nstrips = (int) P->strips.size() / 3;
  return 0;
}

/* sorted catalogue on the device */
struct DevCat {
  int n = 0;
  double *x = nullptr, *y = nullptr, *z = nullptr, *s = nullptr, *w = nullptr;
  int *idx = nullptr, *cs = nullptr;
};

static inline int nblk(int64_t n, int bs) {
  int64_t b = (n + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (int) b;
}

/* Device memory of one call: the context's reusable workspace slots for a
 * single job, or allocations owned by the call (freed at its end) when many
 * jobs need their own sorted catalogues. */
struct Alloc {
  fcfc_gpu_ctx *c;
  bool own;
  std::vector<void *> blocks;
  char *cur = nullptr;         /* own: buffers are carved from 64 MiB chunks */
  size_t left = 0;             /* (a batch of hundreds of small catalogues   */
  int err = 0;                 /* would otherwise pay a hipMalloc per array) */
  char *arena = nullptr;       /* no-allocation calls: every buffer, the     */
  size_t arena_cap = 0;        /* temporaries included, is carved from the   */
  size_t arena_used = 0;       /* device scratch array; nothing is allocated */
  Alloc(fcfc_gpu_ctx *c_, bool own_) : c(c_), own(own_) {}
  ~Alloc() { for (void *p : blocks) (void) hipFree(p); }
  void use_arena(void *p, size_t cap) { arena = (char *) p; arena_cap = cap; arena_used = 0; own = false; }
  void *carve(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t) 255;
    if (bytes > arena_cap - arena_used) {
      if (!err) set_err(c, "no-allocation count: its %zu-byte device scratch is exhausted "
                        "(%zu bytes in use, %zu more requested); count these catalogues "
                        "with fcfc_gpu_count", arena_cap, arena_used, bytes);
      err = FCFC_ERR_MEMORY;
      return nullptr;
    }
    void *r = arena + arena_used;
    arena_used += bytes;
    return r;
  }
  /* temporaries of one call (sort, tiling, histograms): the context's
   * reusable slots, or the arena */
  void *tmp(int slot, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (arena) return carve(bytes);
    return ws(c, slot, bytes, &err);
  }
  void *get(int slot, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (arena) return carve(bytes);
    if (!own) return ws(c, slot, bytes, &err);
    bytes = (bytes + 255) & ~(size_t) 255;
    if (bytes > left) {
      const size_t chunk = std::max(bytes, (size_t) 64 << 20);
      void *p = nullptr;
      if (hipMalloc(&p, chunk) != hipSuccess) {
        set_err(c, "hipMalloc of %zu bytes failed", chunk);
        err = FCFC_ERR_MEMORY;
        return nullptr;
      }
      blocks.push_back(p);
      cur = (char *) p;
      left = chunk;
    }
    void *r = cur;
    cur += bytes;
    left -= bytes;
    return r;
  }
};

__global__ void k_gather(int n, const int *__restrict__ idx, const double *__restrict__ src,
    double *__restrict__ dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[idx[i]];
}

__global__ void k_fill_f64(double *p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_minmax(int n, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, double *out) {
  __shared__ double sm[6][256];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    v[0] = fmin(v[0], x[i]); v[1] = fmin(v[1], y[i]); v[2] = fmin(v[2], z[i]);
    v[3] = fmax(v[3], x[i]); v[4] = fmax(v[4], y[i]); v[5] = fmax(v[5], z[i]);
  }
  for (int k = 0; k < 6; k++) sm[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int) threadIdx.x < s)
      for (int k = 0; k < 6; k++)
        sm[k][threadIdx.x] = (k < 3) ? fmin(sm[k][threadIdx.x], sm[k][threadIdx.x + s])
                                     : fmax(sm[k][threadIdx.x], sm[k][threadIdx.x + s]);
    __syncthreads();
  }
  /* out[8 + 6 b + q]: block b's minima (q < 3) and maxima (q >= 3); the
   * first 8 slots are not written */
  if (threadIdx.x < 6) out[8 + blockIdx.x * 6 + threadIdx.x] = sm[threadIdx.x][0];
}



static int bbox(fcfc_gpu_ctx *c, Alloc &al, const fcfc_cat *k, double lo[3], double hi[3]) {
  if (k->n <= 0) return 0;
  if (!k->on_device) {
    for (int64_t i = 0; i < k->n; i++) {
      const double v[3] = {k->x[i], k->y[i], k->z[i]};
      for (int d = 0; d < 3; d++) { lo[d] = fmin(lo[d], v[d]); hi[d] = fmax(hi[d], v[d]); }
    }
    return 0;
  }
  const int nbk = 64;
  double *d = (double *) al.tmp(WS_MISC, sizeof(double) * (8 + 6 * nbk));
  if (al.err) return al.err;
  hipLaunchKernelGGL(k_minmax, dim3(nbk), dim3(256), 0, c->stream, (int) k->n, k->x, k->y, k->z, d);
  std::vector<double> h(8 + 6 * nbk);
  HIPCHK(c, hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int bk = 0; bk < nbk; bk++)
    for (int q = 0; q < 3; q++) {
      lo[q] = fmin(lo[q], h[8 + bk * 6 + q]);
      hi[q] = fmax(hi[q], h[8 + bk * 6 + 3 + q]);
    }
  return 0;
}

/* Prepare the sort of one catalogue into the grid of a job: upload host
 * inputs, allocate the sorted copies (so the job's descriptors can point
 * at them at once) and queue the work for stage_run.  slot: first of 7
 * consecutive workspace slots (single-job calls). */
static int stage_prep(fcfc_gpu_ctx *c, Alloc &al, const fcfc_cat *k, int need_s,
    int wt, const FkGrid &g, int slot, int upslot, DevCat *out, std::vector<FkStage> &st) {
  const int n = (int) k->n;
  const size_t nb = sizeof(double) * (size_t) (n > 0 ? n : 1);
  const int ncell = g.ncx * g.ncy * g.ncz;
  /* inputs on the device */
  const double *x = k->x, *y = k->y, *z = k->z, *s = need_s ? k->s : nullptr, *w = wt ? k->w : nullptr;
  if (!k->on_device && n > 0) {
    double *dx = (double *) al.get(upslot + 0, nb), *dy = (double *) al.get(upslot + 1, nb),
           *dz = (double *) al.get(upslot + 2, nb);
    if (al.err) return al.err;
    HIPCHK(c, hipMemcpyAsync(dx, k->x, nb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dy, k->y, nb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dz, k->z, nb, hipMemcpyHostToDevice, c->stream));
    x = dx; y = dy; z = dz;
    if (s) {
      double *ds = (double *) al.get(upslot + 3, nb);
      if (al.err) return al.err;
      HIPCHK(c, hipMemcpyAsync(ds, k->s, nb, hipMemcpyHostToDevice, c->stream));
      s = ds;
    }
    if (w) {
      double *dw = (double *) al.get(upslot + 4, nb);
      if (al.err) return al.err;
      HIPCHK(c, hipMemcpyAsync(dw, k->w, nb, hipMemcpyHostToDevice, c->stream));
      w = dw;
    }
  }
  out->n = n;
  out->x = (double *) al.get(slot + 0, nb);
  out->y = (double *) al.get(slot + 1, nb);
  out->z = (double *) al.get(slot + 2, nb);
  out->s = need_s ? (double *) al.get(slot + 3, nb) : nullptr;
  out->w = wt ? (double *) al.get(slot + 4, nb) : nullptr;
  out->idx = (int *) al.get(slot + 5, sizeof(int) * (size_t) (n > 0 ? n : 1));
  out->cs = (int *) al.get(slot + 6, sizeof(int) * ((size_t) ncell + 1));
  if (al.err) return al.err;
  FkStage q;
  memset(&q, 0, sizeof q);
  q.n = n; q.need_s = need_s; q.g = g;
  q.x = x; q.y = y; q.z = z; q.s = s; q.w = w;
  q.ox = out->x; q.oy = out->y; q.oz = out->z; q.os = out->s; q.ow = out->w;
  q.oidx = out->idx; q.ocs = out->cs;
  st.push_back(q);
  return 0;
}

/* Run the queued sorts of a call: one launch per stage for all of them. */
static int stage_run(fcfc_gpu_ctx *c, Alloc &al, std::vector<FkStage> &st) {
  if (st.empty()) return 0;
  long long nobj = 0, ncell = 0, ncs = 0;
  for (FkStage &q : st) {
    q.obj0 = (int) nobj; q.cell0 = (int) ncell; q.cs0 = (int) ncs;
    const long long nc = (long long) q.g.ncx * q.g.ncy * q.g.ncz;
    nobj += q.n; ncell += nc; ncs += nc + 1;
    if (nobj > INT_MAX / 2 || ncs > INT_MAX / 2) { set_err(c, "catalogues of one call too large"); return FCFC_ERR_CFG; }
  }
  const int nst = (int) st.size();
  FkStage *dst = (FkStage *) al.tmp(WS_STAGE, sizeof(FkStage) * st.size());
  int *cid = (int *) al.tmp(WS_CID, sizeof(int) * (size_t) (nobj > 0 ? nobj : 1));
  int *rank = (int *) al.tmp(WS_RANK, sizeof(int) * (size_t) (nobj > 0 ? nobj : 1));
  /* count[ncell]: set when an object has a non-finite coordinate */
  int *count = (int *) al.tmp(WS_COUNT, sizeof(int) * ((size_t) ncell + 1));
  int *gstart = (int *) al.tmp(WS_GSTART, sizeof(int) * ((size_t) ncell + 1));
  if (al.err) return al.err;
  HIPCHK(c, hipMemcpyAsync(dst, st.data(), sizeof(FkStage) * st.size(), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_fill_int, dim3(nblk(ncell + 1, 256)), dim3(256), 0, c->stream, count,
                     (int) ncell + 1, 0);
  if (nobj > 0)
    hipLaunchKernelGGL(k_stage_assign, dim3(nblk(nobj, 256)), dim3(256), 0, c->stream, (int) nobj, nst, dst,
                       cid, rank, count, count + ncell);
  hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, c->stream, count, gstart, (int) ncell);
  hipLaunchKernelGGL(k_stage_starts, dim3(nblk(ncs, 256)), dim3(256), 0, c->stream, (int) ncs, nst, dst, gstart);
  if (nobj > 0)
    hipLaunchKernelGGL(k_stage_scatter, dim3(nblk(nobj, 256)), dim3(256), 0, c->stream, (int) nobj, nst, dst,
                       cid, rank, gstart);
  HIPCHK(c, hipGetLastError());
  int bad = 0;
  HIPCHK(c, hipMemcpyAsync(&bad, count + ncell, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  /* the descriptors are read by the kernels above: keep the host copy until
   * the copy has completed */
  HIPCHK(c, hipStreamSynchronize(c->stream));
  st.clear();
  if (bad) { set_err(c, "a catalogue has a non-finite coordinate"); return FCFC_ERR_DATA; }
  return 0;
}

/* LDS histogram geometry: row stride ns rounded up to an odd number (rows
 * of mu or pi bins), see fk_hslot */
static int hist_stride(const fcfc_bins *b) {
  return b->bintype == FCFC_BIN_ISO ? b->ns : (b->ns | 1);
}
static int hist_slots(const fcfc_bins *b) {
  return hist_stride(b) * (b->ntot / b->ns);
}

struct LdsPlan { int off_hist, off_stab, off_ptab, off_edges, off_strip, off_chunk, off_queue, total, nh, stab_c0, ptab_c0; };

static int align16(int v) { return (v + 15) & ~15; }

/* Table entries staged before cell 0: the integer tables and every FP32
 * table are zero-based (entry k is the one of (int) s2 = k, so the lookup
 * needs no lower range check), the hybrid tables of the FP64 forms keep one
 * sentinel.  A zero-based table costs 4 bytes per entry below s2min; tables
 * that start beyond FK_ZB_MAXOFF entries run through the hybrid lookups
 * instead, which check the range against the edges and give the same bins
 * for integer edges. */
#define FK_ZB_MAXOFF 4096
static bool zb_fits(const fcfc_bins *b) {
  return b->soff <= FK_ZB_MAXOFF && (b->bintype != FCFC_BIN_SPI || b->poff <= FK_ZB_MAXOFF);
}

static LdsPlan lds_plan(const fcfc_bins *b, int mode, int wt, int need_s, int nh, int hyb,
    int ssearch = 0, int psearch = 0) {
  LdsPlan L;
  int o = 0;
  L.nh = nh;
  L.off_hist = o;
  if (mode == MODE_GRAD) o += hist_slots(b) * 8;
  else o += nh * hist_slots(b) * (wt ? 8 : 4);
  o = align16(o);
  L.off_edges = o;
  if (hyb) o = align16(o + 8 * (b->ns + 1 + (b->bintype == FCFC_BIN_SPI ? b->np + 1 : 0)));
  L.off_strip = o; o = align16(o + (int) sizeof(FkStripLds));
  L.off_chunk = o;
  /* per staged B object: coordinates, |r|^2 (survey), weight, index, image code */
  int per = (mode == MODE_C32) ? 12 : 24;
  if (need_s && mode != MODE_C32) per += 8;
  if (wt || mode == MODE_GRAD) per += 8;
  if (mode == MODE_GRAD) per += 8;             /* gradient share of each B object */
  per += 4;
  if (mode == MODE_C32) per += 4;
  o = align16(o + FK_CH * per);
  L.off_queue = o;
  if (mode == MODE_C32) o += (FK_NT / 64) * FK_QCAP * 2;
  o = align16(o);
  const int zero_based = (!hyb || mode == MODE_C32);
  L.stab_c0 = zero_based ? b->soff : 1;
  L.ptab_c0 = zero_based ? b->poff : 1;
  /* two entries after the last cell (four for the survey (s_perp, pi)
   * tables of the FP64 forms, which start at floor cell -2) */
  L.off_stab = o; o = align16(o + (L.stab_c0 + (ssearch ? 0 : b->stab_len) + 4) * 4);
  L.off_ptab = o;
  if (b->bintype == FCFC_BIN_SPI)
    o = align16(o + (L.ptab_c0 + (psearch ? 0 : b->ptab_len) + 4) * 4);
  L.total = o;
  return L;
}

typedef void (*kfn_t)(const FkArgs);

/* Kernel variants: binning x geometry x weights x mode x table type x A
 * objects per lane (R = 1: tiles of FK_NT objects, R = 2: 2 FK_NT, R = 4:
 * 4 FK_NT).  Hybrid tables use one R (2; the gradient 1); R = 4 exists for
 * the weighted FP64 box counts with integer tables only, whose LDS-bound
 * histogram updates gain from reading each staged B object for more A
 * objects (fk_lanes_max gives the largest R of a kernel family).  The
 * survey kernels hold four more FP64 values per object and spill at R = 4
 * (weighted (s_perp, pi) DR: 950 ms against 830 at R = 2). */
static int fk_lanes_max(int geom, int mode, int wt, int hyb) {
  if (hyb) return mode == MODE_GRAD ? 1 : 2;
  return (wt && mode == MODE_C64 && geom == FCFC_GEOM_BOX) ? 4 : 2;
}

template <int BIN, int GEOM, bool WT, int MODE>
static kfn_t pick5(int hyb, int R) {
  if constexpr (MODE == MODE_GRAD) {
    if (hyb) return k_count<BIN, GEOM, WT, MODE, true, 1>;
    return R == 1 ? k_count<BIN, GEOM, WT, MODE, false, 1> : k_count<BIN, GEOM, WT, MODE, false, 2>;
  }
  else {
    if (hyb) return k_count<BIN, GEOM, WT, MODE, true, 2>;
    if constexpr (WT && MODE == MODE_C64 && GEOM == FCFC_GEOM_BOX)
      if (R == 4) return k_count<BIN, GEOM, WT, MODE, false, 4>;
    return R == 1 ? k_count<BIN, GEOM, WT, MODE, false, 1> : k_count<BIN, GEOM, WT, MODE, false, 2>;
  }
}

template <int BIN, int GEOM>
static kfn_t pick3(int wt, int mode, int hyb, int R) {
  if (mode == MODE_GRAD) return pick5<BIN, GEOM, true, MODE_GRAD>(hyb, R);
  if (mode == MODE_C32)
    return wt ? pick5<BIN, GEOM, true, MODE_C32>(hyb, R) : pick5<BIN, GEOM, false, MODE_C32>(hyb, R);
  if constexpr (GEOM == FCFC_GEOM_BOX) {
    if (mode == MODE_C32C)
      return wt ? pick5<BIN, GEOM, true, MODE_C32C>(hyb, R) : pick5<BIN, GEOM, false, MODE_C32C>(hyb, R);
  }
  if (mode == MODE_P64)
    return wt ? pick5<BIN, GEOM, true, MODE_P64>(hyb, R) : pick5<BIN, GEOM, false, MODE_P64>(hyb, R);
  return wt ? pick5<BIN, GEOM, true, MODE_C64>(hyb, R) : pick5<BIN, GEOM, false, MODE_C64>(hyb, R);
}

static kfn_t pick_kernel(int bintype, int geom, int wt, int mode, int hyb, int R) {
  if (geom == FCFC_GEOM_BOX) {
    if (bintype == FCFC_BIN_ISO) return pick3<FCFC_BIN_ISO, FCFC_GEOM_BOX>(wt, mode, hyb, R);
    if (bintype == FCFC_BIN_SMU) return pick3<FCFC_BIN_SMU, FCFC_GEOM_BOX>(wt, mode, hyb, R);
    return pick3<FCFC_BIN_SPI, FCFC_GEOM_BOX>(wt, mode, hyb, R);
  }
  if (bintype == FCFC_BIN_ISO) return pick3<FCFC_BIN_ISO, FCFC_GEOM_SURVEY>(wt, mode, hyb, R);
  if (bintype == FCFC_BIN_SMU) return pick3<FCFC_BIN_SMU, FCFC_GEOM_SURVEY>(wt, mode, hyb, R);
  return pick3<FCFC_BIN_SPI, FCFC_GEOM_SURVEY>(wt, mode, hyb, R);
}

/* FP32 exactness windows.  With coordinates relative to the A tile centre
 * c, each FP32 difference satisfies |d32 - d| <= u (2|d| + E) + tiny
 * (u = 2^-24, E = tile extent), so
 *   |s2_32 - s2| <= 7 u s2 + 2 sqrt(3) u E s + 3 (u (2 s + E))^2,
 * evaluated at the largest in-range s^2 (plus the first integer above it:
 * beyond that the table answers "outside" without a window).  The mu bin
 * floor(t), t = nmu |dz| / s, has |t32 - t| <= t 10u + (sqrt(3) u E t +
 * nmu u E) / s; with s^2 >= s2_floor (fp32_s2_floor) and t <= nmu this is
 * a constant D, and the bin is certain when |fract(t32) - 1/2| <= 1/2 - D.
 * All bounds are inflated by 25%; pairs below s2_floor are always
 * re-evaluated.
 *
 * Survey pairs (fk_sv32) add two things.  The reference's own values carry
 * the rounding of its FP64 formulas: s^2 = (|a|^2 + |b|^2) - 2 a.b is off
 * the exact value by at most about 8 eps R^2 (R = largest |r| of the grid,
 * eps = 2^-53, |r|^2 given with 2 eps), and its pi = (|a|^2 - |b|^2) / |m|
 * by 5 eps R^2 / |m| (m = a + b).  And the FP32 m = (2 a' + d) + 2 c (c the
 * tile centre, a' = a - c) has error |dm| <= 2u|c| + 2 sqrt(3) ek + 2u|m| +
 * 2uE (ek = u (2 s + E) + tiny, 2|c| <= |m| + s + E), so p = d.m is off by
 * at most |m| (sqrt(3) ek + 6 u s) + s (u (s + 3E) + 2 sqrt(3) ek) and pi
 * by e_pi = sqrt(3) ek + 6 u s + s (u (s + 3E) + 2 sqrt(3) ek) / |m|.  With |m|^2 >= m2_floor (smaller |m| is re-evaluated) these are
 * constants: the pi^2 window is 2 pi e_pi + e_pi^2 + pi^2 (relative error of
 * |m|^2, 4u of the reciprocal, 2u of the products) plus the reference's
 * terms, the s_perp^2 window the sum of the s^2 and pi^2 windows plus the
 * rounding of the difference, and the mu window (t = nmu |p| / (|m| s))
 * gets e_pi / s plus half the relative errors of s^2 and |m|^2 and 8u for
 * the reciprocal square root of |m|^2 s^2 and that product. */
static double fp32_extent(const FkGrid &g, int zmaxc) {
  return std::max(g.wx, std::max(g.wy, (zmaxc + 1) * g.wz));
}

static void fp32_windows(const fcfc_bins *b, const FkGrid &g, int zmaxc, double s2f,
    FkWin *win);

/* s^2 below which a job's pairs are always re-evaluated (the LDS table flags
 * those cells once per launch, so a launch uses the largest over its jobs).
 * For (s, mu) bins the floor trades the pairs below it -- about
 * ((floor + 1) / s2max)^(3/2) of a uniform catalogue's in-range pairs,
 * counting the flagged cell above it -- against the mu windows, 2 D of the
 * pairs above it with D growing as 1 / sqrt(floor): the candidate with the
 * fewest expected re-evaluations is taken. */
static double fp32_s2_floor(const fcfc_bins *b, const FkGrid &g, int zmaxc) {
  const double u = ldexp(1.0, -24);
  const double s2max = b->s2bin[b->ns];
  const double lo = pow(1e4 * u * fp32_extent(g, zmaxc), 2);
  if (b->bintype != FCFC_BIN_SMU) return std::max(s2max / 64.0, lo);
  double best = std::max(s2max / 64.0, lo), bcost = 1e300;
  for (double x = 1.0 / 64; x >= 1.0 / 16384; x *= 0.5) {
    const double s2f = std::max(x * s2max, lo);
    FkWin w;
    fp32_windows(b, g, zmaxc, s2f, &w);
    const double cost = pow(std::min((s2f + 1) / s2max, 1.0), 1.5) + 2.0 * w.mu_sh;
    if (cost < bcost) { bcost = cost; best = s2f; }
  }
  return best;
}

/* largest |r| of a survey grid (its corners) */
static double grid_rmax(const FkGrid &g) {
  double r2 = 0;
  for (int k = 0; k < 8; k++) {
    const double x = (k & 1) ? g.ox + g.ncx * g.wx : g.ox;
    const double y = (k & 2) ? g.oy + g.ncy * g.wy : g.oy;
    const double z = (k & 4) ? g.oz + g.ncz * g.wz : g.oz;
    r2 = std::max(r2, x * x + y * y + z * z);
  }
  return sqrt(r2);
}

static void fp32_windows(const fcfc_bins *b, const FkGrid &g, int zmaxc, double s2f,
    FkWin *win) {
  const double u = ldexp(1.0, -24), eps = ldexp(1.0, -53);
  const bool sv = (b->geometry == FCFC_GEOM_SURVEY);
  const double s2max = b->s2bin[b->ns];
  /* box (s_perp, pi): pbin holds pi edges; survey: pi^2 edges */
  const double pmx = (b->bintype == FCFC_BIN_SPI) ? b->pbin[b->np] : 0;
  const double S2 = s2max + (sv ? pmx : pmx * pmx) + 2.0;
  const double s = sqrt(S2);
  const double E = fp32_extent(g, zmaxc);
  const double R = sv ? grid_rmax(g) : 0.0;
  const double tiny = ldexp(1.0, -44) * std::max(std::max(g.Lx, std::max(g.Ly, g.Lz)), R) * 4;
  const double ek = u * (2 * s + E) + tiny;
  double es2 = 7 * u * S2 + 2 * sqrt(3.0) * u * E * s + 3 * ek * ek + 6 * ek * tiny;
  double epi = u * (2 * pmx + E) + tiny;          /* box: window on pi */
  const double m2f = sv ? S2 : 0.0;
  const double mf = sqrt(m2f);
  double dmu = 0;                                 /* survey mu window terms */
  if (sv) {
    es2 += 8 * eps * R * R + 4 * eps * S2;        /* the reference's s^2 */
    const double epa = sqrt(3.0) * ek + 6 * u * s + s * (u * (s + 3 * E) + 2 * sqrt(3.0) * ek) / mf;
    const double m2rel = u * (9 + 2 * (s + 3 * E) / mf) + 4 * sqrt(3.0) * ek / mf;
    if (b->bintype == FCFC_BIN_SPI) {
      const double pi = sqrt(pmx);
      const double ep2 = 2 * pi * epa + epa * epa + pmx * (m2rel + 6 * u)
                         + 2 * pi * 5 * eps * R * R / mf + pmx * 6 * eps * R * R / m2f + 4 * eps * pmx;
      epi = ep2;                                  /* window on pi^2 */
      es2 = es2 + ep2 + u * S2;                   /* window on s_perp^2 */
    }
    else {
      const double sf = sqrt(std::max(s2f, 1e-300));
      const double es2rel = 7 * u + 2 * sqrt(3.0) * u * E / sf + 3 * pow(u * (2 + E / sf), 2);
      dmu = (sqrt(3.0) * u * (2 + E / sf) + sqrt(3.0) * tiny / sf + 5 * u + u * s / mf
             + sqrt(3.0) * ek / mf) + 0.5 * es2rel + 0.5 * m2rel + 8 * u + 3 * u
            + 5 * eps * R * R / (mf * sf) + 4 * eps * R * R / (sf * sf);
    }
  }
  FkWin &w = *win;
  w.e_s2 = (float) (es2 * 1.25 + 1e-7);
  w.e_pi = (float) (epi * 1.25 + 1e-7);
  w.e_s2x2 = 2.0f * w.e_s2;
  w.e_pix2 = 2.0f * w.e_pi;
  const double cz = 0.5 / std::max(b->nmu, 1);
  if (sv) { w.idle_x = 1.0e6f; w.idle_z = 0.0f; }
  else {
    w.idle_z = (float) (1e18 * cz);
    w.idle_x = (float) (1e18 * sqrt(1.0 - cz * cz));
  }
  const double nmu = b->nmu;
  double D = 1.25 * (nmu * 10 * u + (sqrt(3.0) * u * E * nmu + nmu * (u * E + tiny)) / sqrt(s2f))
             + 4 * nmu * u;
  if (sv) D = 1.25 * nmu * dmu + 4 * nmu * u;
  /* fixed-point mu (fk_mu32): sh >= D + 2^-16 on the grid 2^-16; the
   * fraction threshold is 0 (every pair re-evaluated) when the windows
   * cover everything */
  const double sh = ceil((std::min(D, 0.5) + ldexp(1.0, -16)) * 65536.0) / 65536.0;
  w.mu_sh = (float) sh;
  w.mu_c = (float) (129.0 - sh);
  w.mu_t16 = (unsigned) std::max(0.0, floor(65536.0 * (1.0 - 2.0 * sh)));
  w.s2_floor = (float) s2f;
  w.nmu_f = (float) b->nmu;
  w.m2_floor = (float) m2f;
}

/* One counting (or gradient) job on device-resident, sorted catalogues. */

/* A job ready to count: plan, sorted catalogues, output. */
struct PJob {
  JobPlan P;
  DevCat A, B;
  int autoc = 0, wt = 0;
  double *gA = nullptr, *gB = nullptr;
  long long hist_off = 0;
};

static int cat_ok(const fcfc_cat *k) {
  return k && k->n >= 0 && k->n < 0x7fffffff && (k->n == 0 || (k->x && k->y && k->z));
}

static int bins_ok(const fcfc_bins *b) {
  return b && b->ns > 0 && b->s2bin && b->stab && b->ntot > 0 &&
      (b->bintype != FCFC_BIN_SPI || (b->pbin && b->ptab));
}

/* Plan a job and stage its catalogues.  Sorted copies are shared between
 * jobs of one call when every array a copy is built from (coordinates,
 * |r|^2, weights of weighted jobs) and the grid coincide: two catalogues on
 * the same positions with different weights (jackknife or bootstrap
 * weights) get copies of their own. */
struct CopyKey {
  const double *x, *y, *z, *s, *w; int64_t n; int wt; FkGrid g; DevCat d;
  bool same(const fcfc_cat *k, int wt_, int need_s, const FkGrid &g_) const;
};

static bool same_grid(const FkGrid &a, const FkGrid &b) {
  return a.ncx == b.ncx && a.ncy == b.ncy && a.ncz == b.ncz && a.ox == b.ox && a.oy == b.oy &&
      a.oz == b.oz && a.wx == b.wx && a.wy == b.wy && a.wz == b.wz;
}

bool CopyKey::same(const fcfc_cat *k, int wt_, int need_s, const FkGrid &g_) const {
  return x == k->x && y == k->y && z == k->z && n == k->n && wt == wt_ &&
      (!need_s || s == k->s) && (!wt_ || w == k->w) && same_grid(g, g_);
}

/* weights of a sorted catalogue re-read from the caller's device array */
struct Gather { const double *src; const int *idx; double *dst; int n; };

static int setup_job(fcfc_gpu_ctx *c, Alloc &al, std::vector<CopyKey> &copies,
    const fcfc_cat *a, const fcfc_cat *b, const fcfc_bins *bins, const double *bsize,
    int autoc, int wt, int mode, PJob *J, std::vector<FkStage> &stages,
    std::vector<Gather> *gathers = nullptr) {
  const int need_s = (bins->geometry == FCFC_GEOM_SURVEY && bins->bintype != FCFC_BIN_ISO);
  /* cross counts are symmetric in the two catalogues: the larger one goes to
   * the lanes (fuller tiles, tighter windows and fewer lanes sharing one
   * histogram bin when the other is clustered); the gradient keeps the
   * caller's orientation */
  if (!autoc && mode != MODE_GRAD && b->n > a->n) std::swap(a, b);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  int e;
  if (bins->geometry == FCFC_GEOM_SURVEY) {
    if ((e = bbox(c, al, a, lo, hi))) return e;
    if (!autoc && (e = bbox(c, al, b, lo, hi))) return e;
    if (!(lo[0] <= hi[0])) { for (int k = 0; k < 3; k++) { lo[k] = 0; hi[k] = 1; } }
  }
  if ((e = make_plan(c, bins, bsize, lo, hi, autoc, wt, mode, a->n, autoc ? a->n : b->n, &J->P))) return e;
  J->autoc = autoc;
  J->wt = wt;
  const fcfc_cat *cs[2] = {a, autoc ? a : b};
  DevCat *dst[2] = {&J->A, &J->B};
  for (int q = 0; q < (autoc ? 1 : 2); q++) {
    const fcfc_cat *k = cs[q];
    bool found = false;
    for (auto &cp : copies)
      if (cp.same(k, wt, need_s, J->P.g)) { *dst[q] = cp.d; found = true; break; }
    if (found) continue;
    const int slot = (q == 0) ? WS_A_X : WS_B_X;
    const int upslot = (q == 0) ? WS_IN_X : WS_IN2_X;
    if ((e = stage_prep(c, al, k, need_s, wt, J->P.g, slot, upslot, dst[q], stages))) return e;
    if (al.own) copies.push_back(CopyKey{k->x, k->y, k->z, k->s, k->w, k->n, wt, J->P.g, *dst[q]});
    if (gathers && wt && k->w && k->on_device && k->n > 0)
      gathers->push_back(Gather{k->w, dst[q]->idx, dst[q]->w, (int) k->n});
  }
  if (autoc) J->B = J->A;
  return 0;
}

/* A recorded count launch: kernel, grid, LDS bytes and arguments, and the
 * host copies of the strip triples and job descriptors its uploads read.
 * They live as long as the launch, so that the asynchronous uploads do not
 * depend on how the runtime stages pageable host memory. */
struct Launch {
  kfn_t fn = nullptr;
  int nblock = 0, lds = 0, ntiles = 0;
  FkArgs a;
  std::vector<int> hstr;
  std::vector<FkJob> hj;
};

/* Tile all jobs and prepare the persistent launch that counts (or
 * differentiates) them; every buffer comes from al. */
static int build_launch(fcfc_gpu_ctx *c, Alloc &al, const fcfc_bins *b,
    std::vector<PJob> &jobs, int mode, const double *G, void *dhist,
    unsigned long long *dstats, unsigned int *dqueue, Launch *out) {
  const int geom = b->geometry;
  int m = mode;
  const bool zb = zb_fits(b);
  if (m == MODE_C32 && !zb) m = MODE_C64;                      /* FP32 tables are zero-based */
  /* the FP32 fixed-point mu value holds t <= nmu (1 + 2^-20) below 127 */
  if (m == MODE_C32 && b->bintype == FCFC_BIN_SMU && b->nmu > FK_NMU32) m = MODE_C64;
  int hyb = (b->tabtype || !zb) ? 1 : 0;
  const int need_s = (geom == FCFC_GEOM_SURVEY && b->bintype != FCFC_BIN_ISO);
  int anywt = 0;
  for (auto &j : jobs) anywt |= j.wt;
  if (m == MODE_GRAD) anywt = 1;
  const int lds_max = 160 * 1024 - 2048;       /* room for the static LDS */
  /* the FP32 window test reads one table cell and the edge at its lower end:
   * windows must stay well inside a cell (wide only for a few objects in a
   * huge volume, whose grid cells, and so tiles, are large) */
  if (m == MODE_C32) {
    double s2f = 0;
    for (const PJob &J : jobs) s2f = std::max(s2f, fp32_s2_floor(b, J.P.g, J.P.zmaxc));
    for (const PJob &J : jobs) {
      FkWin w;
      fp32_windows(b, J.P.g, J.P.zmaxc, s2f, &w);
      if (!(w.e_s2x2 < 0.5f) || !(w.e_pix2 < 0.5f)) { m = MODE_C64; break; }
    }
  }
  /* tables too long for LDS (hybrid binnings rescaled to large integers):
   * hybrid lookups that binary-search the edges, in the FP64 forms */
  int ssearch = 0, psearch = 0;
  if (lds_plan(b, m, anywt, need_s, 1, hyb).total > lds_max) {
    if (m == MODE_C32) m = MODE_C64;
    hyb = 1;
    ssearch = 1;
    if (lds_plan(b, m, anywt, need_s, 1, hyb, 1, 0).total > lds_max) psearch = 1;
  }
  /* one kernel serves every job of the launch: tiles of the longest kind.
   * A lane holds R objects (1, 2 or 4), so no tile may exceed R FK_NT
   * objects: R is cut to what the kernel family provides (fk_lanes_max),
   * and a tile size set by the caller or the autotuner is clamped to what
   * the kernel covers */
  {
    int R = 1;
    for (auto &j : jobs) R = std::max(R, (j.P.T + FK_NT - 1) / FK_NT);
    if (R > 2) R = 4;
    R = std::min(R, fk_lanes_max(geom, m, anywt, hyb));
    for (auto &j : jobs) {
      j.P.R = R;
      if (j.P.T_auto || j.P.T > R * FK_NT) j.P.T = R * FK_NT;
    }
  }
  /* ---- kernel and LDS plan ---- */
  const int R = jobs.empty() ? 1 : jobs[0].P.R;
  auto occ_of = [&](kfn_t f, const LdsPlan &P, int *occ) -> int {
    HIPCHK(c, hipFuncSetAttribute((const void *) f, hipFuncAttributeMaxDynamicSharedMemorySize, P.total));
    HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, (const void *) f, FK_NT, P.total));
    return 0;
  };
  /* the rounding census compares the box arithmetic paths */
  const int km = (m == MODE_C32 && c->census && geom == FCFC_GEOM_BOX) ? MODE_C32C : m;
  kfn_t fn = pick_kernel(b->bintype, geom, anywt, km, hyb, R);
  /* histogram copies: as many as keep the occupancy of a single copy */
  LdsPlan L = lds_plan(b, m, anywt, need_s, 1, hyb, ssearch, psearch);
  if (L.total > lds_max) {
    set_err(c, "binning tables need %d bytes of LDS (limit %d)", L.total, lds_max);
    return FCFC_ERR_CFG;
  }
  if (m != MODE_GRAD) {
    int occ1 = 0;
    if (int e = occ_of(fn, L, &occ1)) return e;
    const int nhmax = c->t_nh > 0 ? c->t_nh : 8;
    for (int nh = 2; nh <= nhmax; nh *= 2) {
      LdsPlan L2 = lds_plan(b, m, anywt, need_s, nh, hyb, ssearch, psearch);
      int occ = 0;
      if (L2.total > lds_max) break;
      if (int e = occ_of(fn, L2, &occ)) return e;
      if (occ < occ1 && c->t_nh <= 0) break;
      L = L2;
    }
  }
  HIPCHK(c, hipFuncSetAttribute((const void *) fn, hipFuncAttributeMaxDynamicSharedMemorySize, L.total));
  /* ---- tiles of all jobs ---- */
  std::vector<int> rowoff(jobs.size() + 1, 0);
  for (size_t k = 0; k < jobs.size(); k++) rowoff[k + 1] = rowoff[k] + jobs[k].P.g.ncx * jobs[k].P.g.ncy;
  const int nrow = rowoff.back();
  int *rown = (int *) al.get(WS_ROWN, sizeof(int) * (size_t) nrow);
  int *rowscan = (int *) al.get(WS_ROWOFF, sizeof(int) * ((size_t) nrow + 1));
  FkTileJob *dtj = (FkTileJob *) al.get(WS_TJOBS, sizeof(FkTileJob) * jobs.size());
  if (al.err) return al.err;
  std::vector<FkTileJob> htj(jobs.size());
  for (size_t k = 0; k < jobs.size(); k++) {
    const PJob &J = jobs[k];
    htj[k] = FkTileJob{rowoff[k], J.P.g.ncz, J.P.T, J.P.zmaxc, J.P.nstrips, J.A.cs};
  }
  HIPCHK(c, hipMemcpyAsync(dtj, htj.data(), sizeof(FkTileJob) * jobs.size(), hipMemcpyHostToDevice, c->stream));
  const int njob = (int) jobs.size();
  if (nrow > 0)
    hipLaunchKernelGGL(k_tile_count, dim3(nblk(nrow, 64)), dim3(64), 0, c->stream, nrow, njob, dtj, rown);
  hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, c->stream, rown, rowscan, nrow);
  int ntile0 = 0;
  HIPCHK(c, hipMemcpyAsync(&ntile0, rowscan + nrow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  /* persistent workgroups, and work items per tile: enough items that the
   * last ones are short against the whole (16 per workgroup) */
  int occ = 0;
  HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *) fn, FK_NT, L.total));
  if (occ < 1) occ = 1;
  const int nwg = c->t_nblock > 0 ? c->t_nblock : occ * c->ncu;
  /* the autotuner's samples count about 1/tile_frac of the tiles, in 16
   * runs of consecutive tiles spread evenly over the list, split into work
   * items by the same rule as a full count (about 16 per workgroup), so
   * that a sample ends as evenly */
  int grp = std::max(ntile0, 1), run = grp;
  if (c->tile_frac > 1 && ntile0 > 0) {
    grp = (ntile0 + 15) / 16;
    run = std::max(1, grp / c->tile_frac);
  }
  const int glast = ntile0 > 0 ? (ntile0 - 1) / grp : 0;
  const int ncap = ntile0 > 0 ? glast * run + std::min(run, ntile0 - glast * grp) : 0;
  FkSplit sp{1, ncap, 1, ncap, grp, run};
  if (ntile0 > 0) {
    int minstr = 1 << 30;
    for (auto &j : jobs) minstr = std::min(minstr, std::max(j.P.nstrips, 1));
    const int smax = std::min(minstr, 64);
    sp.S = (int) std::min<long long>(std::max<long long>(1, (16LL * nwg + ncap - 1) / ncap),
                                     smax);
    /* the tiles of about two items per workgroup at the end of the list
     * are split four times finer */
    sp.St = std::min(4 * sp.S, smax);
    sp.tt = (int) std::max<long long>(0, ncap - (2LL * nwg + sp.S - 1) / sp.S);
    if (sp.St <= sp.S) { sp.St = sp.S; sp.tt = ncap; }
    if (split_pos(sp, ncap) > (1LL << 30)) sp = FkSplit{1, ncap, 1, ncap, grp, run};
  }
  const int ntiles = (int) split_pos(sp, ncap);
  FkTile *tiles = (FkTile *) al.get(WS_TILES, sizeof(FkTile) * (size_t) (ntiles > 0 ? ntiles : 1));
  if (al.err) return al.err;
  if (nrow > 0)
    hipLaunchKernelGGL(k_tile_emit, dim3(nblk(nrow, 64)), dim3(64), 0, c->stream, nrow, njob, dtj,
                       rowscan, sp, tiles);
  /* ---- tables, strips, job descriptors ---- */
  size_t nstr = 0;
  for (auto &j : jobs) nstr += j.P.strips.size();
  uint16_t *dstab = (uint16_t *) al.get(WS_STAB, sizeof(uint16_t) * (b->stab_len + 1));
  uint16_t *dptab = (uint16_t *) al.get(WS_PTAB, sizeof(uint16_t) * (b->ptab_len + 1));
  double *ds2e = (double *) al.get(WS_S2E, sizeof(double) * (b->ns + 1));
  double *dpe = (double *) al.get(WS_PE, sizeof(double) * (b->np + 1));
  int *dstrips = (int *) al.get(WS_STRIPS, sizeof(int) * (nstr ? nstr : 1));
  FkJob *djobs = nullptr;
  if (al.err) return al.err;
  std::vector<int> &hstr = out->hstr;
  hstr.clear();
  hstr.reserve(nstr);
  std::vector<FkJob> &hj = out->hj;
  hj.assign(jobs.size(), FkJob{});
  double s2f = 0;
  float win2s = 0.0f, win2p = 0.0f;
  if (m == MODE_C32)
    for (const PJob &J : jobs) s2f = std::max(s2f, fp32_s2_floor(b, J.P.g, J.P.zmaxc));
  for (size_t k = 0; k < jobs.size(); k++) {
    const PJob &J = jobs[k];
    FkJob &f = hj[k];
    memset(&f, 0, sizeof f);
    f.A = FkCat{J.A.n, J.A.x, J.A.y, J.A.z, J.A.s, J.A.w, J.A.idx, J.A.cs};
    f.B = FkCat{J.B.n, J.B.x, J.B.y, J.B.z, J.B.s, J.B.w, J.B.idx, J.B.cs};
    f.g = J.P.g;
    f.strips = dstrips + hstr.size();
    f.nstrips = J.P.nstrips;
    f.autoc = J.autoc;
    f.wt = J.wt;
    f.hist_off = J.hist_off;
    f.gA = J.gA;
    f.gB = J.gB;
    if (m == MODE_C32) {
      fp32_windows(b, J.P.g, J.P.zmaxc, s2f, &f.win);
      win2s = std::max(win2s, f.win.e_s2x2);
      win2p = std::max(win2p, f.win.e_pix2);
    }
    hstr.insert(hstr.end(), J.P.strips.begin(), J.P.strips.end());
  }
  djobs = (FkJob *) al.get(WS_JOBS, sizeof(FkJob) * jobs.size());
  if (al.err) return al.err;
  HIPCHK(c, hipMemcpyAsync(dstab, b->stab, sizeof(uint16_t) * b->stab_len, hipMemcpyHostToDevice, c->stream));
  if (b->bintype == FCFC_BIN_SPI)
    HIPCHK(c, hipMemcpyAsync(dptab, b->ptab, sizeof(uint16_t) * b->ptab_len, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(ds2e, b->s2bin, sizeof(double) * (b->ns + 1), hipMemcpyHostToDevice, c->stream));
  if (b->bintype == FCFC_BIN_SPI)
    HIPCHK(c, hipMemcpyAsync(dpe, b->pbin, sizeof(double) * (b->np + 1), hipMemcpyHostToDevice, c->stream));
  if (nstr) HIPCHK(c, hipMemcpyAsync(dstrips, hstr.data(), sizeof(int) * nstr, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(djobs, hj.data(), sizeof(FkJob) * jobs.size(), hipMemcpyHostToDevice, c->stream));

  FkArgs a;
  memset(&a, 0, sizeof a);
  a.jobs = djobs;
  a.njobs = (int) jobs.size();
  FkBins &fb = a.bins;
  fb.bintype = b->bintype; fb.geometry = geom;
  fb.ns = b->ns; fb.nmu = b->nmu; fb.np = b->np; fb.ntot = b->ntot;
  fb.nsh = hist_stride(b); fb.ntoh = hist_slots(b);
  fb.scap32 = b->soff + b->stab_len + 1; fb.pcap32 = b->poff + b->ptab_len + 1;
  fb.tabtype = b->tabtype; fb.soff = b->soff; fb.poff = b->poff;
  /* searched tables are not staged to LDS (length 0: sentinels only) */
  fb.stab_len = ssearch ? 0 : b->stab_len; fb.ptab_len = psearch ? 0 : b->ptab_len;
  fb.ssearch = ssearch; fb.psearch = psearch;
  fb.s2min = b->s2bin[0]; fb.s2max = b->s2bin[b->ns];
  fb.pmin = (b->bintype == FCFC_BIN_SPI) ? b->pbin[0] : 0;
  fb.pmax = (b->bintype == FCFC_BIN_SPI) ? b->pbin[b->np] : 0;
  fb.nmu2 = (double) b->nmu * b->nmu;
  fb.nmu_d = (double) b->nmu;
  fb.s2_exact = std::max(b->s2bin[b->ns] * ldexp(1.0, -12), 4 * FK_DBL_EPS);
  /* no pair at or beyond these is counted (margins cover the other rounding
   * orientation of the gradient and the FP32 windows) */
  /* survey (s_perp, pi): s^2 = s_perp^2 + pi^2 (pbin holds pi^2 edges) */
  fb.s2skip = (geom == FCFC_GEOM_SURVEY && b->bintype == FCFC_BIN_SPI)
      ? (fb.s2max + fb.pmax) * (1 + ldexp(1.0, -30)) + 1e-6 * (fb.s2max + fb.pmax + 1)
      : fb.s2max * (1 + ldexp(1.0, -30));
  fb.pskip = (b->bintype == FCFC_BIN_SPI) ? fb.pmax * (1 + ldexp(1.0, -30)) : 1e300;
  fb.s2skip32 = (float) (fb.s2max + (geom == FCFC_GEOM_SURVEY ? fb.pmax : 0.0) + 2.0);
  fb.win2s = win2s;
  fb.win2p = win2p;
  fb.pskip32 = (b->bintype == FCFC_BIN_SPI) ? (float) (fb.pmax + 2.0) : 3.0e38f;
  {
    /* fixed-point mu (fk_mu_exact): sh >= nmu 2^-20 + 2^-16 on the grid
     * 2^-16, the addend 2^36 + 1 - sh exact */
    const double sh = ceil((ldexp((double) b->nmu, -20) + ldexp(1.0, -16)) * 65536.0) / 65536.0;
    fb.mu_c = ldexp(1.0, 36) + 1.0 - sh;
    fb.mu_cmax = ldexp(1.0, 36) + b->nmu + 1.5;
    fb.mu_t16 = (unsigned) std::max(0.0, floor(65536.0 * (1.0 - 2.0 * sh)));
    /* the FP64 integer-table kernels add mu_t16 << 16 to their (s, mu)
     * indices (fk_smu_index): keep the largest, (nmu + 1) nsh + 0xFFFF +
     * (mu_t16 << 16), below 2^32 (a lower threshold only sends more pairs
     * to the exact path) */
    const unsigned long long room = (0xFFFF0000ull - (unsigned long long) (b->nmu + 1) * fb.nsh) >> 16;
    fb.mu_t16 = (unsigned) std::min<unsigned long long>(fb.mu_t16, room);
    fb.mu_pad = 0;
  }
  {
    /* survey (s_perp, pi) with integer tables (fk_sv64): x + q_c rounds
     * onto the grid 2^-16 of [2^36, 2^37) with K = floor(x - h) + 2, h =
     * 2^-15 >= 2^-16 plus the 2^-29.7 error bound of the fast pi^2 and
     * s_perp^2; a fraction f <= q_t16 = 65532 puts x strictly inside
     * (K - 2, K - 1) (f/2^16 + h + 2^-17 + 2^-29.7 < 1).  The ceilings are
     * the last table entries, off + len + 3 (< 2^16: the tables fit in LDS) */
    fb.q_c = ldexp(1.0, 36) + 2.0 - ldexp(1.0, -15);
    fb.q_caps = ldexp(1.0, 36) + (double) (b->soff + b->stab_len + 3);
    fb.q_capp = ldexp(1.0, 36) + (double) (b->poff + b->ptab_len + 3);
    fb.q_bmin = ldexp(1.0, -1000);
    fb.q_t16 = 65532u;
    fb.q_pad = 0;
  }
  fb.stab = dstab; fb.ptab = dptab; fb.s2bin = ds2e; fb.pbin = dpe;
  a.tiles = tiles; a.ntiles = ntiles;
  a.queue = dqueue;
  a.gcnt_u = (unsigned long long *) dhist; a.gcnt_d = (double *) dhist;
  a.stats = dstats;
  a.nh = L.nh;
  a.off_hist = L.off_hist; a.off_stab = L.off_stab; a.off_ptab = L.off_ptab;
  a.stab_c0 = L.stab_c0; a.ptab_c0 = L.ptab_c0;
  a.off_edges = L.off_edges; a.off_strip = L.off_strip;
  a.off_chunk = L.off_chunk; a.off_queue = L.off_queue;
  a.census = c->census;
  a.G = G;
  int nblock = nwg;
  if (nblock > a.ntiles) nblock = std::max(a.ntiles, 1);
  c->last_ntiles = ntile0;
  c->last_nsampled = ncap;
  c->stats.tile = jobs.empty() ? 0 : jobs[0].P.T;
  c->stats.cell_div = jobs.empty() ? 0 : (int) lround(sqrt(jobs[0].P.g.Rxy2) / jobs[0].P.g.wx);
  out->fn = fn; out->nblock = nblock; out->lds = L.total; out->ntiles = a.ntiles; out->a = a;
  return 0;
}

/* Waits for the stream when it goes out of scope: declared after the
 * launches of a call (destroyed before them), it keeps their host copies
 * alive until the stream has drained, on the error returns as well. */
struct StreamDrain {
  fcfc_gpu_ctx *c;
  ~StreamDrain() { (void) hipStreamSynchronize(c->stream); }
};

static int launch(fcfc_gpu_ctx *c, const Launch &Lc) {
  if (Lc.ntiles > 0)
    hipLaunchKernelGGL(Lc.fn, dim3(Lc.nblock), dim3(FK_NT), Lc.lds, c->stream, Lc.a);
  HIPCHK(c, hipGetLastError());
  return 0;
}

/* Tile all jobs, then count (or differentiate) them in one persistent launch. */
static int run_jobs(fcfc_gpu_ctx *c, Alloc &al, const fcfc_bins *b,
    std::vector<PJob> &jobs, int mode, const double *G, void *dhist,
    unsigned long long *dstats, unsigned int *dqueue, float *kms) {
  Launch Lc;
  StreamDrain drain{c};
  int e = build_launch(c, al, b, jobs, mode, G, dhist, dstats, dqueue, &Lc);
  if (e) return e;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  if ((e = launch(c, Lc))) return e;
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipEventElapsedTime(kms, c->ev0, c->ev1));
  return 0;
}

/* Jobs of a batch grouped for their launches: one kernel variant per
 * launch, so weighted and unweighted jobs, and jobs whose tiles want one or
 * two objects per lane, go to separate persistent launches. */
static std::vector<std::vector<PJob>> launch_groups(const std::vector<PJob> &jobs) {
  std::vector<std::vector<PJob>> g;
  std::vector<int> key;
  for (const PJob &j : jobs) {
    const int k = 2 * (j.wt ? 1 : 0) + ((j.P.T > FK_NT) ? 1 : 0);
    size_t q = 0;
    while (q < key.size() && key[q] != k) q++;
    if (q == key.size()) { key.push_back(k); g.emplace_back(); }
    g[q].push_back(j);
  }
  return g;
}

static int stats_from(fcfc_gpu_ctx *c, unsigned long long *dstats) {
  unsigned long long st[3] = {0, 0, 0};
  HIPCHK(c, hipMemcpyAsync(st, dstats, sizeof st, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->stats.evals = (double) st[0];
  c->stats.rechecked = (double) st[1];
  c->stats.census = (double) st[2];
  return 0;
}

/* add the device histograms of the jobs (8 bytes per bin: int64 or
 * double) into the caller's cnt (device or host) */
static int deliver(fcfc_gpu_ctx *c, const std::vector<PJob> &jobs, int ntot,
    const void *dh, uint32_t flags, void *cnt) {
  const int njob = (int) jobs.size();
  const int n = ntot * njob;
  if (flags & FCFC_GPU_CNT_DEVICE) {
    for (int k = 0; k < njob; k++) {
      char *dst = (char *) cnt + (size_t) k * ntot * 8;
      const char *src = (const char *) dh + (size_t) k * ntot * 8;
      if (jobs[k].wt) hipLaunchKernelGGL(k_add_f64, dim3(nblk(ntot, 256)), dim3(256), 0, c->stream, (double *) dst, (const double *) src, ntot);
      else hipLaunchKernelGGL(k_add_u64, dim3(nblk(ntot, 256)), dim3(256), 0, c->stream, (long long *) dst, (const unsigned long long *) src, ntot);
    }
    HIPCHK(c, hipGetLastError());
  }
  else {
    std::vector<unsigned long long> h(n);
    HIPCHK(c, hipMemcpyAsync(h.data(), dh, 8 * (size_t) n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < njob; k++) {
      const size_t o = (size_t) k * ntot;
      if (jobs[k].wt) { double *dst = (double *) cnt + o; const double *hd = (const double *) h.data() + o; for (int i = 0; i < ntot; i++) dst[i] += hd[i]; }
      else { int64_t *dst = (int64_t *) cnt + o; for (int i = 0; i < ntot; i++) dst[i] += (int64_t) h[o + i]; }
    }
  }
  return 0;
}

/* counts of njob jobs; job k's histogram lands in cnt + k * ntot */
static int count_jobs(fcfc_gpu_ctx *c, int njob, const fcfc_cat *const *as,
    const fcfc_cat *const *bs, const double *const *bsz, const uint32_t *fl,
    const fcfc_bins *bins, uint32_t flags, void *cnt, void *arena = nullptr,
    size_t arena_cap = 0) {
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(c, hipSetDevice(c->device));
  memset(&c->stats, 0, sizeof c->stats);
  c->census = (flags & FCFC_GPU_CENSUS) ? 1 : 0;
  const int mode = (flags & FCFC_GPU_PREC_FP32) ? MODE_C32 :
                   ((flags & FCFC_GPU_PREC_PLAIN) ? MODE_P64 : MODE_C64);
  Alloc al(c, njob > 1);
  if (arena) al.use_arena(arena, arena_cap);
  std::vector<CopyKey> copies;
  std::vector<FkStage> stages;
  std::vector<PJob> jobs(njob);
  for (int k = 0; k < njob; k++) {
    const int autoc = (fl[k] & FCFC_GPU_AUTO) ? 1 : 0;
    const int wt = (fl[k] & FCFC_GPU_WEIGHTED) ? 1 : 0;
    if (!cat_ok(as[k]) || (!autoc && !cat_ok(bs[k]))) { set_err(c, "invalid catalogue in job %d", k); return FCFC_ERR_ARG; }
    if (bins->geometry == FCFC_GEOM_BOX && !bsz[k]) { set_err(c, "box size required"); return FCFC_ERR_ARG; }
    int e = setup_job(c, al, copies, as[k], bs[k], bins, bsz[k], autoc, wt, mode, &jobs[k], stages);
    if (e) return e;
    jobs[k].hist_off = (long long) k * bins->ntot;
  }
  if (int e = stage_run(c, al, stages)) return e;
  const size_t hbytes = 8 * (size_t) bins->ntot * njob;
  void *dh = al.tmp(WS_HIST, hbytes);
  unsigned long long *misc = (unsigned long long *) al.tmp(WS_MISC, 64 * 8);
  if (al.err) return al.err;
  hipLaunchKernelGGL(k_zero_u64, dim3(nblk((int64_t) bins->ntot * njob, 256)), dim3(256), 0, c->stream,
      (unsigned long long *) dh, bins->ntot * njob);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, misc, 64);
  /* one persistent launch per group (each with its own queue word) */
  std::vector<std::vector<PJob>> groups = launch_groups(jobs);
  std::vector<Launch> L(groups.size());
  StreamDrain drain{c};
  for (size_t q = 0; q < groups.size(); q++) {
    int e = build_launch(c, al, bins, groups[q], mode, nullptr, dh, misc + 8,
                         (unsigned int *) misc + q, &L[q]);
    if (e) return e;
  }
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (const Launch &Lq : L) { int e = launch(c, Lq); if (e) return e; }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  int e = deliver(c, jobs, bins->ntot, dh, flags, cnt);
  if (e) return e;
  if ((e = stats_from(c, misc + 8))) return e;
  c->stats.kernel_ms = ms;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

extern "C" int fcfc_gpu_count(fcfc_gpu_ctx *c, const fcfc_cat *a,
    const fcfc_cat *b, const fcfc_bins *bins, const double *bsize,
    uint32_t flags, void *cnt) {
  if (!c) return FCFC_ERR_ARG;
  if (!bins_ok(bins) || !cnt) { set_err(c, "invalid argument"); return FCFC_ERR_ARG; }
  const fcfc_cat *as[1] = {a}, *bs[1] = {b};
  const double *bz[1] = {bsize};
  uint32_t fl[1] = {flags & (FCFC_GPU_AUTO | FCFC_GPU_WEIGHTED)};
  return count_jobs(c, 1, as, bs, bz, fl, bins, flags, cnt);
}

/* Device bytes a no-allocation count needs at least (uploads, sorted copies,
 * per-object sort scratch); cells, tiles and histograms come on top. */
static size_t noalloc_floor(const fcfc_cat *k, const fcfc_bins *bins, int wt) {
  if (!k || k->n <= 0) return 0;
  const int need_s = (bins->geometry == FCFC_GEOM_SURVEY && bins->bintype != FCFC_BIN_ISO);
  size_t per = 24 + 4 + 8 + (wt ? 8 : 0) + (need_s ? 8 : 0);     /* sorted copy + sort */
  if (!k->on_device) per += 24 + (wt && k->w ? 8 : 0) + (need_s && k->s ? 8 : 0);
  return per * (size_t) k->n;
}

extern "C" int fcfc_gpu_count_noalloc(fcfc_gpu_ctx *c, const fcfc_cat *a,
    const fcfc_cat *b, const fcfc_bins *bins, const double *bsize,
    uint32_t flags, void *cnt) {
  if (!c) return FCFC_ERR_ARG;
  if (!bins_ok(bins) || !cnt) { set_err(c, "invalid argument"); return FCFC_ERR_ARG; }
  if (!c->scratch) {
    set_err(c, "no-allocation count: the device scratch of GPU %d is unavailable (%s)",
            c->device, c->scratch_err);
    return FCFC_ERR_GPU;
  }
  const int autoc = (flags & FCFC_GPU_AUTO) ? 1 : 0;
  const int wt = (flags & FCFC_GPU_WEIGHTED) ? 1 : 0;
  if (!cat_ok(a) || (!autoc && !cat_ok(b))) { set_err(c, "invalid catalogue"); return FCFC_ERR_ARG; }
  const size_t need = noalloc_floor(a, bins, wt) + (autoc ? 0 : noalloc_floor(b, bins, wt));
  if (need > FK_SCRATCH_BYTES) {
    set_err(c, "no-allocation count: at most %zu bytes of device scratch, these catalogues "
            "need more than %zu; count them with fcfc_gpu_count", FK_SCRATCH_BYTES, need);
    return FCFC_ERR_MEMORY;
  }
  const fcfc_cat *as[1] = {a}, *bs[1] = {b};
  const double *bz[1] = {bsize};
  uint32_t fl[1] = {flags & (FCFC_GPU_AUTO | FCFC_GPU_WEIGHTED)};
  return count_jobs(c, 1, as, bs, bz, fl, bins, flags, cnt, c->scratch, FK_SCRATCH_BYTES);
}

extern "C" int fcfc_gpu_count_batch(fcfc_gpu_ctx *c, int njob,
    const fcfc_batch_job *jobs, const fcfc_bins *bins, uint32_t flags,
    void *cnt) {
  if (!c || njob < 0 || (njob && !jobs) || !cnt) return FCFC_ERR_ARG;
  if (!bins_ok(bins)) { set_err(c, "invalid bins"); return FCFC_ERR_ARG; }
  if (njob == 0) return 0;
  std::vector<const fcfc_cat *> as(njob), bs(njob);
  std::vector<const double *> bz(njob);
  std::vector<uint32_t> fl(njob);
  for (int k = 0; k < njob; k++) {
    as[k] = &jobs[k].a; bs[k] = &jobs[k].b; bz[k] = jobs[k].bsize;
    fl[k] = jobs[k].flags & (FCFC_GPU_AUTO | FCFC_GPU_WEIGHTED);
  }
  return count_jobs(c, njob, as.data(), bs.data(), bz.data(), fl.data(), bins,
      flags & ~(FCFC_GPU_AUTO | FCFC_GPU_WEIGHTED), cnt);
}

extern "C" int fcfc_gpu_count_backward(fcfc_gpu_ctx *c, const fcfc_cat *a,
    const fcfc_cat *b, const fcfc_bins *bins, const double *bsize,
    uint32_t flags, const double *grad_cnt, double *ga, double *gb) {
  if (!c) return FCFC_ERR_ARG;
  auto t0 = std::chrono::steady_clock::now();
  const int autoc = (flags & FCFC_GPU_AUTO) ? 1 : 0;
  if (!cat_ok(a) || (!autoc && !cat_ok(b)) || !bins_ok(bins) || !grad_cnt || !ga) {
    set_err(c, "invalid argument"); return FCFC_ERR_ARG;
  }
  if (bins->geometry == FCFC_GEOM_BOX && !bsize) { set_err(c, "box size required"); return FCFC_ERR_ARG; }
  HIPCHK(c, hipSetDevice(c->device));
  memset(&c->stats, 0, sizeof c->stats);
  c->census = 0;
  const int dev = (flags & FCFC_GPU_CNT_DEVICE) ? 1 : 0;
  /* one pass over the pairs exactly as the count enumerates them (auto: each
   * unordered pair once; cross: the larger catalogue in the lanes), so every
   * pair is binned as in the count; each pair adds G[bin] w_b to its A object
   * and G[bin] w_a to its B object */
  const int sw = (!autoc && b->n > a->n) ? 1 : 0;
  const fcfc_cat *pa = sw ? b : a, *pb = autoc ? a : (sw ? a : b);
  Alloc al(c, false);
  std::vector<CopyKey> copies;
  std::vector<FkStage> stages;
  std::vector<PJob> jobs(1);
  StreamDrain drain{c};
  if (int e = setup_job(c, al, copies, pa, pb, bins, bsize, autoc, 1, MODE_GRAD, &jobs[0], stages))
    return e;
  double *dga = (double *) al.get(WS_GA, sizeof(double) * (size_t) (a->n > 0 ? a->n : 1));
  double *dgb = (!autoc && gb) ? (double *) al.get(WS_GB, sizeof(double) * (size_t) (b->n > 0 ? b->n : 1))
                               : nullptr;
  if (al.err) return al.err;
  hipLaunchKernelGGL(k_fill_f64, dim3(nblk(a->n, 256)), dim3(256), 0, c->stream, dga, (int64_t) a->n, 0.0);
  if (dgb)
    hipLaunchKernelGGL(k_fill_f64, dim3(nblk(b->n, 256)), dim3(256), 0, c->stream, dgb, (int64_t) b->n, 0.0);
  jobs[0].gA = sw ? dgb : dga;
  jobs[0].gB = autoc ? dga : (sw ? dga : dgb);
  if (int e = stage_run(c, al, stages)) return e;
  double *dG = (double *) al.tmp(WS_G, sizeof(double) * bins->ntot);
  unsigned long long *misc = (unsigned long long *) al.tmp(WS_MISC, 64 * 8);
  if (al.err) return al.err;
  HIPCHK(c, hipMemcpyAsync(dG, grad_cnt, sizeof(double) * bins->ntot,
      dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, misc, 64);
  float ms = 0;
  int e = run_jobs(c, al, bins, jobs, MODE_GRAD, dG, nullptr, misc + 8, (unsigned int *) misc, &ms);
  if (e) return e;
  for (int p = 0; p < 2; p++) {
    const fcfc_cat *k = p ? b : a;
    double *gout = p ? gb : ga, *gdev = p ? dgb : dga;
    if (!gdev) continue;
    if (dev) {
      hipLaunchKernelGGL(k_add_f64, dim3(nblk(k->n, 256)), dim3(256), 0, c->stream, gout, gdev, (int) k->n);
      HIPCHK(c, hipGetLastError());
    }
    else {
      std::vector<double> h((size_t) k->n);
      HIPCHK(c, hipMemcpyAsync(h.data(), gdev, sizeof(double) * (size_t) k->n, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (int64_t i = 0; i < k->n; i++) gout[i] += h[(size_t) i];
    }
  }
  if ((e = stats_from(c, misc + 8))) return e;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->stats.kernel_ms = ms;
  c->stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

/* ---- prepared batches replayed as a HIP graph -------------------------- */

struct fcfc_gpu_plan {
  fcfc_gpu_ctx *c;
  Alloc al;                          /* every device buffer of the plan */
  std::vector<PJob> jobs;
  std::vector<Gather> gathers;
  int ntot = 0;
  uint32_t flags = 0;
  std::vector<Launch> L;             /* one per launch group */
  void *dh = nullptr;                /* njob x ntot histograms (8 bytes per bin) */
  unsigned long long *misc = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  explicit fcfc_gpu_plan(fcfc_gpu_ctx *c_) : c(c_), al(c_, true) {}
  ~fcfc_gpu_plan() {
    if (c && c->stream) (void) hipStreamSynchronize(c->stream);
    if (exec) (void) hipGraphExecDestroy(exec);
    if (graph) (void) hipGraphDestroy(graph);
  }
};

/* Sort and tile the catalogues of njob jobs once; record the launch
 * sequence of one count -- weight gathers from the callers' device arrays,
 * histogram and queue clears (fill kernels), the persistent count kernel --
 * as a HIP graph. */
extern "C" int fcfc_gpu_plan_create(fcfc_gpu_ctx *c, int njob, const fcfc_batch_job *jobs,
    const fcfc_bins *bins, uint32_t flags, fcfc_gpu_plan **out) {
  if (!c || !out || njob < 1 || !jobs || !bins_ok(bins)) return FCFC_ERR_ARG;
  *out = nullptr;
  HIPCHK(c, hipSetDevice(c->device));
  memset(&c->stats, 0, sizeof c->stats);
  c->census = 0;
  const int mode = (flags & FCFC_GPU_PREC_FP32) ? MODE_C32 :
                   ((flags & FCFC_GPU_PREC_PLAIN) ? MODE_P64 : MODE_C64);
  std::unique_ptr<fcfc_gpu_plan> P(new fcfc_gpu_plan(c));
  P->ntot = bins->ntot;
  P->flags = flags;
  P->jobs.resize(njob);
  std::vector<CopyKey> copies;
  std::vector<FkStage> stages;
  StreamDrain drain{c};
  for (int k = 0; k < njob; k++) {
    const fcfc_batch_job &j = jobs[k];
    const int autoc = (j.flags & FCFC_GPU_AUTO) ? 1 : 0;
    const int wt = (j.flags & FCFC_GPU_WEIGHTED) ? 1 : 0;
    if (!cat_ok(&j.a) || (!autoc && !cat_ok(&j.b))) { set_err(c, "invalid catalogue in job %d", k); return FCFC_ERR_ARG; }
    int e = setup_job(c, P->al, copies, &j.a, &j.b, bins, j.bsize, autoc, wt, mode, &P->jobs[k], stages,
                      &P->gathers);
    if (e) return e;
    P->jobs[k].hist_off = (long long) k * bins->ntot;
  }
  if (int e = stage_run(c, P->al, stages)) return e;
  P->dh = P->al.get(WS_HIST, 8 * (size_t) bins->ntot * njob);
  P->misc = (unsigned long long *) P->al.get(WS_MISC, 64 * 8);
  if (P->al.err) return P->al.err;
  std::vector<std::vector<PJob>> groups = launch_groups(P->jobs);
  P->L.resize(groups.size());
  int e = 0;
  for (size_t q = 0; q < groups.size(); q++)
    if ((e = build_launch(c, P->al, bins, groups[q], mode, nullptr, P->dh, P->misc + 8,
                          (unsigned int *) P->misc + q, &P->L[q]))) return e;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (const Gather &g : P->gathers)
    hipLaunchKernelGGL(k_gather, dim3(nblk(g.n, 256)), dim3(256), 0, c->stream, g.n, g.idx, g.src, g.dst);
  hipLaunchKernelGGL(k_zero_u64, dim3(nblk((int64_t) bins->ntot * njob, 256)), dim3(256), 0, c->stream,
      (unsigned long long *) P->dh, bins->ntot * njob);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, P->misc, 64);
  for (const Launch &Lq : P->L) if ((e = launch(c, Lq))) break;
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
  if (e) { if (graph) (void) hipGraphDestroy(graph); return e; }
  if (ce != hipSuccess) { set_err(c, "graph capture failed: %s", hipGetErrorString(ce)); return FCFC_ERR_GPU; }
  P->graph = graph;
  HIPCHK(c, hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
  *out = P.release();
  return 0;
}

/* One count of every job of the plan, histograms added into cnt (njob x
 * ntot; FCFC_GPU_CNT_DEVICE in the plan's flags: a device pointer). */
extern "C" int fcfc_gpu_plan_run(fcfc_gpu_plan *P, void *cnt) {
  if (!P || !cnt) return FCFC_ERR_ARG;
  fcfc_gpu_ctx *c = P->c;
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  HIPCHK(c, hipGraphLaunch(P->exec, c->stream));
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  int e = deliver(c, P->jobs, P->ntot, P->dh, P->flags, cnt);
  if (e) return e;
  if ((e = stats_from(c, P->misc + 8))) return e;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->stats.kernel_ms = ms;
  c->stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

extern "C" int fcfc_gpu_plan_destroy(fcfc_gpu_plan *P) {
  delete P;
  return 0;
}

/* Project-written autotuner: for jobs like (a, b, bins, flags) try grid and
 * tile shapes (cells per maximum separation, z-cells per cell, A objects
 * per tile), timing each on the first part of the tile list (a slab of the
 * box at the true density: an eighth, a sixteenth beyond 16384 tiles) and
 * extrapolating to all tiles (counts of at most 4096 tiles are timed
 * whole).  Every pair of cell_div and zdiv is timed (the best zdiv depends
 * on cell_div: a descent over one, then the other, missed the best pair of
 * a cross count), then the tile size.  The winner is cached in the
 * context for this geometry / binning / precision and density bucket and
 * used by later counts whose shape is left automatic. */
extern "C" int fcfc_gpu_autotune(fcfc_gpu_ctx *c, const fcfc_cat *a,
    const fcfc_cat *b, const fcfc_bins *bins, const double *bsize,
    uint32_t flags, int *best) {
  if (!c || !bins_ok(bins)) return FCFC_ERR_ARG;
  const int autoc = (flags & FCFC_GPU_AUTO) ? 1 : 0;
  if (!cat_ok(a) || (!autoc && !cat_ok(b))) { set_err(c, "invalid argument"); return FCFC_ERR_ARG; }
  const int grad = (flags & FCFC_GPU_GRADIENT) ? 1 : 0;
  const int mode = grad ? MODE_GRAD : (flags & FCFC_GPU_PREC_FP32) ? MODE_C32 :
                   ((flags & FCFC_GPU_PREC_PLAIN) ? MODE_P64 : MODE_C64);
  const int save_cd = c->t_cell_div, save_zd = c->t_zdiv, save_t = c->t_tile;
  std::vector<unsigned char> hcnt(8 * (size_t) bins->ntot);
  /* gradient timing: dL/dcnt = 1 in every bin (every pair contributes),
   * gradient of a only, as weight fitting asks for it */
  std::vector<double> gone(grad ? (size_t) bins->ntot : 0, 1.0);
  std::vector<double> gout(grad ? (size_t) std::max<int64_t>(a->n, 1) : 0);
  const uint32_t cflags = flags & ~(FCFC_GPU_CNT_DEVICE | FCFC_GPU_GRADIENT);
  auto run = [&]() -> int {
    if (grad)
      return fcfc_gpu_count_backward(c, a, b, bins, bsize, flags & FCFC_GPU_AUTO, gone.data(),
                                     gout.data(), nullptr);
    memset(hcnt.data(), 0, hcnt.size());
    return fcfc_gpu_count(c, a, b, bins, bsize, cflags, hcnt.data());
  };
  double bestms = 1e300;
  int bcd = 3, bzd = 8, bT = 0;
  int last_T = 0;
  /* FCFC_GPU_TUNE_LOG set and not "0": every candidate's timing on stderr */
  const char *tl = getenv("FCFC_GPU_TUNE_LOG");
  const bool tune_log = tl && strcmp(tl, "0") != 0;
  auto measure = [&](int cd, int zd, int T, double *est) -> int {
    c->t_cell_div = cd; c->t_zdiv = zd; c->t_tile = T;
    /* first pass: learn the number of tiles of this shape, counting 16 */
    c->tile_frac = 1 << 30;
    int e = run();
    if (e) return e;
    const int nt = c->last_ntiles;
    /* small counts (a few thousand tiles, milliseconds) are timed whole: an
     * eighth of them is too few tiles to rank shapes */
    c->tile_frac = (nt <= 4096) ? 1 : (nt > 16384 ? 16 : 8);
    e = run();
    if (e) return e;
    *est = c->stats.kernel_ms * (double) nt / (double) std::max(1, c->last_nsampled);
    last_T = c->stats.tile;
    if (tune_log)
      fprintf(stderr, "fcfc_gpu_autotune: cell_div %d zcell_div %d tile %d: %d of %d tiles "
              "in %.3f ms, estimate %.3f ms\n", cd, zd, last_T, c->last_nsampled, nt,
              c->stats.kernel_ms, *est);
    return 0;
  };
  int e = 0;
  const int cds[] = {2, 3, 4, 6};
  const int zds[] = {1, 2, 4, 8, 16};
  struct Cand { double est; int cd, zd, T; };
  std::vector<Cand> grid;
  for (int cd : cds) {
    for (int zd : zds) {
      double est;
      if ((e = measure(cd, zd, 0, &est))) break;
      grid.push_back({est, cd, zd, last_T});
      if (est < bestms) { bestms = est; bcd = cd; bzd = zd; bT = last_T; }
    }
    if (e) break;
  }
  /* the grid runs each shape at its automatic tile size (FK_NT objects for
   * small cells); the best three shapes are also timed at the other sizes
   * the kernel family holds (fk_lanes_max: 4 FK_NT for weighted FP64 box
   * counts with integer tables, FK_NT for the hybrid gradient kernel) */
  std::sort(grid.begin(), grid.end(), [](const Cand &x, const Cand &y) { return x.est < y.est; });
  const int lmax = fk_lanes_max(bins->geometry, mode, grad || (flags & FCFC_GPU_WEIGHTED) ? 1 : 0,
                                (bins->tabtype || !zb_fits(bins)) ? 1 : 0);
  const int Ts[] = {FK_NT, 2 * FK_NT, 4 * FK_NT};
  for (size_t k = 0; k < grid.size() && k < 3 && !e; k++) {
    for (int T : Ts) {
      double est;
      if (T == grid[k].T || T > lmax * FK_NT) continue;
      if ((e = measure(grid[k].cd, grid[k].zd, T, &est))) break;
      if (est < bestms) { bestms = est; bcd = grid[k].cd; bzd = grid[k].zd; bT = last_T; }
    }
  }
  c->tile_frac = 0;
  c->t_cell_div = save_cd; c->t_zdiv = save_zd; c->t_tile = save_t;
  if (e) return e;
  /* cache for the density bucket of this job, keyed as the count (or the
   * gradient) plans it: the larger catalogue of a cross count is A */
  {
    JobPlan P;
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    const fcfc_cat *ka = a, *kb = autoc ? a : b;
    if (!autoc && kb->n > ka->n) std::swap(ka, kb);
    if (bins->geometry == FCFC_GEOM_SURVEY) {
      for (int k = 0; k < 3; k++) { lo[k] = 1e300; hi[k] = -1e300; }
      Alloc tal(c, false);
      bbox(c, tal, ka, lo, hi);
      if (!autoc) bbox(c, tal, kb, lo, hi);
    }
    const int wt = (grad || (flags & FCFC_GPU_WEIGHTED)) ? 1 : 0;
    make_plan(c, bins, bsize, lo, hi, autoc, wt, mode, ka->n, kb->n, &P);
    bool found = false;
    for (auto &t : c->tuned)
      if (t.geom == bins->geometry && t.bintype == bins->bintype && t.mode == mode &&
          t.autoc == autoc && t.wt == wt && t.bucket == P.bucket) {
        t.cd = bcd; t.zd = bzd; t.T = bT; found = true;
      }
    if (!found) c->tuned.push_back({bins->geometry, bins->bintype, mode, autoc, wt,
                                    P.bucket, bcd, bzd, bT});
  }
  if (best) { best[0] = bcd; best[1] = bzd; best[2] = bT; best[3] = (int) lround(bestms * 1000); }
  return 0;
}
