"""fcfc_amd -- MI355X pair-count engine for FCFC (Python side).

The compute path is the HIP library fcfc_amd/lib/libfcfc_gpu.so (C ABI in
include/fcfc_gpu.h).  This module mirrors FCFC's count step
    count_pairs(tree1, tree2, cf, cnt, isauto, withwt)
(src/fcfc/2pt_box/count_func.h:54-59) over structure-of-arrays catalogues:

    bins = Bins.linear(GEOM_BOX, BIN_SMU, smin=0, smax=200, ds=5, nmu=100,
                       bsize=[1000]*3)
    d = Catalog(bins, x, y, z, w)              # rescaled as FCFC does
    with Engine() as eng:
        cnt = eng.count_pairs(d, d, bins, isauto=True, withwt=True)

`cnt` holds unordered auto pairs (FCFC's caller doubles them,
src/fcfc/2pt_box/eval_cf.c:125-133); `pair_counts()` does the doubling and
normalisation of eval_pairs.  There is no CPU fallback: without the library
every call raises.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import (FCFC_BIN_ISO as BIN_ISO, FCFC_BIN_SMU as BIN_SMU,
                   FCFC_BIN_SPI as BIN_SPI, FCFC_GEOM_BOX as GEOM_BOX,
                   FCFC_GEOM_SURVEY as GEOM_SURVEY)

__all__ = ["Bins", "Catalog", "Engine", "FcfcError", "BIN_ISO", "BIN_SMU",
           "BIN_SPI", "GEOM_BOX", "GEOM_SURVEY", "pair_counts", "gen_uniform",
           "gen_clustered", "gen_survey", "gen_sizes", "save_pair_count",
           "convert_coords", "convert_coords_table", "lib"]


def lib():
    return L.load()


class FcfcError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"FCFC error {code}: {msg}")
        self.code = code


def _prec_flag(precision):
    """'fp64': FP64 distances with the fast exact mu bin; 'fp32': FP32
    distances with FP64 re-evaluation near bin edges; 'plain': the
    reference's per-pair FP64 arithmetic transcribed directly."""
    try:
        return {"fp64": L.FCFC_GPU_PREC_FP64, "fp32": L.FCFC_GPU_PREC_FP32,
                "plain": L.FCFC_GPU_PREC_PLAIN}[precision]
    except KeyError:
        raise ValueError("precision must be 'fp64', 'fp32' or 'plain'") from None


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Bins:
    """Separation bins and lookup tables (FCFC's cf_setup binning,
    src/fcfc/2pt_box/setup_cf.c:385-531)."""

    def __init__(self):
        self._b = L.fcfc_bins()
        self._live = False
        self.bsize = None            # box sizes in input units (box geometry)

    @classmethod
    def linear(cls, geometry, bintype, smin, smax, ds, nmu=1, pmin=0.0,
               pmax=0.0, dpi=1.0, bsize=None):
        self = cls()
        bs = None
        if bsize is not None:
            bsize = np.broadcast_to(np.asarray(bsize, dtype=np.float64), (3,)).copy()
            bs = _dptr(bsize)
            self.bsize = bsize
        e = lib().fcfc_bins_linear(C.byref(self._b), geometry, bintype, smin, smax,
                                   ds, nmu, pmin, pmax, dpi, bs)
        if e:
            raise FcfcError(e, "fcfc_bins_linear")
        self._live = True
        return self

    @classmethod
    def edges(cls, geometry, bintype, sedges, nmu=1, pedges=None, bsize=None):
        self = cls()
        se = _f64(sedges)
        pe = _f64(pedges) if pedges is not None else None
        bs = None
        if bsize is not None:
            bsize = np.broadcast_to(np.asarray(bsize, dtype=np.float64), (3,)).copy()
            bs = _dptr(bsize)
            self.bsize = bsize
        e = lib().fcfc_bins_edges(C.byref(self._b), geometry, bintype, _dptr(se),
                                  len(se) - 1, nmu, _dptr(pe),
                                  (len(pe) - 1) if pe is not None else 0, bs)
        if e:
            raise FcfcError(e, "fcfc_bins_edges")
        self._live = True
        return self

    def close(self):
        if self._live:
            lib().fcfc_bins_free(C.byref(self._b))
            self._live = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- views ---------------------------------------------------------------
    def _arr(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if ptr else None

    geometry = property(lambda s: s._b.geometry)
    bintype = property(lambda s: s._b.bintype)
    ns = property(lambda s: s._b.ns)
    nmu = property(lambda s: s._b.nmu)
    np = property(lambda s: s._b.np)
    ntot = property(lambda s: s._b.ntot)
    rescale = property(lambda s: s._b.rescale)
    tabtype = property(lambda s: s._b.tabtype)
    sbin_raw = property(lambda s: s._arr(s._b.sbin_raw, s._b.ns + 1))
    pbin_raw = property(lambda s: s._arr(s._b.pbin_raw, s._b.np + 1))
    s2bin = property(lambda s: s._arr(s._b.s2bin, s._b.ns + 1))
    pbin = property(lambda s: s._arr(s._b.pbin, s._b.np + 1))

    @property
    def stab(self):
        return np.ctypeslib.as_array(self._b.stab, shape=(self._b.stab_len,)).copy()

    @property
    def bsize_rescaled(self):
        if self.bsize is None:
            return None
        return self.bsize * self.rescale

    @property
    def shape(self):
        if self.bintype == BIN_SMU:
            return (self.nmu, self.ns)
        if self.bintype == BIN_SPI:
            return (self.np, self.ns)
        return (self.ns,)


class Catalog:
    """Structure-of-arrays catalogue, rescaled by the binning factor.

    Host catalogues are numpy float64 arrays (copied and rescaled here, as
    FCFC's tree_create does, build_tree.c:121-156).  Device catalogues are
    torch CUDA float64 tensors that the caller already rescaled; for survey
    (s,mu)/(s_perp,pi) bins `s` may be left None and is computed on the GPU.
    """

    def __init__(self, bins, x, y, z, w=None, s=None, rescaled=False):
        self.device = hasattr(x, "is_cuda") and x.is_cuda
        if self.device:
            import torch
            for t in (x, y, z) + tuple(v for v in (w, s) if v is not None):
                if t.dtype != torch.float64 or not t.is_contiguous():
                    raise ValueError("device catalogues must be contiguous float64 tensors")
            self.x, self.y, self.z, self.w, self.s = x, y, z, w, s
            self.n = int(x.numel())
            self.sumw = float(w.sum().item()) if w is not None else float(self.n)
            return
        self.x = _f64(x).copy()
        self.y = _f64(y).copy()
        self.z = _f64(z).copy()
        self.w = _f64(w).copy() if w is not None else None
        self.n = len(self.x)
        need_s = bins.geometry == GEOM_SURVEY and bins.bintype != BIN_ISO
        self.s = np.empty(self.n) if need_s and s is None else (_f64(s) if s is not None else None)
        sumw = C.c_double(0)
        if rescaled:
            b = Bins.__new__(Bins)
            lb = L.fcfc_bins()
            C.memmove(C.byref(lb), C.byref(bins._b), C.sizeof(lb))
            lb.rescale = 1.0
            e = lib().fcfc_prepare_catalog(C.byref(lb), self.n, _dptr(self.x), _dptr(self.y),
                                           _dptr(self.z), _dptr(self.s) if (need_s and s is None) else None,
                                           _dptr(self.w), C.byref(sumw))
        else:
            e = lib().fcfc_prepare_catalog(C.byref(bins._b), self.n, _dptr(self.x),
                                           _dptr(self.y), _dptr(self.z),
                                           _dptr(self.s) if (need_s and s is None) else None,
                                           _dptr(self.w), C.byref(sumw))
        if e:
            raise FcfcError(e, "fcfc_prepare_catalog")
        self.sumw = sumw.value

    def _c(self):
        c = L.fcfc_cat()
        c.n = self.n
        if self.device:
            c.x, c.y, c.z = self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr()
            c.s = self.s.data_ptr() if self.s is not None else None
            c.w = self.w.data_ptr() if self.w is not None else None
            c.on_device = 1
        else:
            c.x, c.y, c.z = (a.ctypes.data for a in (self.x, self.y, self.z))
            c.s = self.s.ctypes.data if self.s is not None else None
            c.w = self.w.ctypes.data if self.w is not None else None
            c.on_device = 0
        return c


def _job_array(jobs, bins):
    arr = (L.fcfc_batch_job * len(jobs))()
    bs = bins.bsize_rescaled
    for k, (a, b, isauto, withwt) in enumerate(jobs):
        arr[k].a = a._c()
        arr[k].b = (b if b is not None else a)._c()
        if bs is not None:
            for i in range(3):
                arr[k].bsize[i] = bs[i]
        arr[k].flags = (L.FCFC_GPU_AUTO if isauto else 0) | (L.FCFC_GPU_WEIGHTED if withwt else 0)
    return arr


class Engine:
    """One HIP device context (fcfc_gpu_init)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        e = lib().fcfc_gpu_init(C.byref(self._h), device)
        if e:
            raise FcfcError(e, "fcfc_gpu_init: no usable GPU")

    def close(self):
        if self._h:
            lib().fcfc_gpu_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, e, what):
        if e:
            msg = lib().fcfc_gpu_last_error(self._h)
            raise FcfcError(e, f"{what}: {msg.decode() if msg else ''}")

    @staticmethod
    def _sync_torch(*cats):
        if any(c is not None and c.device for c in cats):
            import torch
            torch.cuda.synchronize()

    def set_tuning(self, cell_div=0, tile=0, zdiv=0, nblock=0):
        self._check(lib().fcfc_gpu_set_tuning(self._h, cell_div, tile, zdiv, nblock),
                    "fcfc_gpu_set_tuning")

    def autotune(self, a, b, bins, isauto=False, withwt=False, precision="fp64",
                 gradient=False):
        """Pick grid/tile shapes for counts like this one (cached in the
        engine per density bucket), or with gradient=True for the weight
        gradient of such counts (count_backward).  Returns (cell_div, zdiv,
        tile, est_ms)."""
        flags = (L.FCFC_GPU_AUTO if isauto else 0) | (L.FCFC_GPU_WEIGHTED if withwt else 0)
        flags |= _prec_flag(precision)
        if gradient:
            flags |= L.FCFC_GPU_GRADIENT
        best = (C.c_int * 4)()
        bs = bins.bsize_rescaled
        ca = a._c()
        cb = (b if b is not None else a)._c()
        self._sync_torch(a, b)
        e = lib().fcfc_gpu_autotune(self._h, C.byref(ca), C.byref(cb), C.byref(bins._b),
                                    _dptr(bs) if bs is not None else None, flags, best)
        self._check(e, "fcfc_gpu_autotune")
        return best[0], best[1], best[2], best[3] / 1000.0

    def count_pairs(self, a, b, bins, isauto=False, withwt=False,
                    precision="fp64", census=False, out=None):
        """Histogram of unordered (auto) or all (cross) pairs; the result is
        added to `out` when given (FCFC accumulates into cnt)."""
        flags = (L.FCFC_GPU_AUTO if isauto else 0) | (L.FCFC_GPU_WEIGHTED if withwt else 0)
        flags |= _prec_flag(precision)
        if census:
            flags |= L.FCFC_GPU_CENSUS
        if out is None:
            out = np.zeros(bins.ntot, dtype=np.float64 if withwt else np.int64)
        bs = bins.bsize_rescaled
        bsp = _dptr(bs) if bs is not None else None
        ca = a._c()
        cb = (b if b is not None else a)._c()
        self._sync_torch(a, b)
        e = lib().fcfc_gpu_count(self._h, C.byref(ca), C.byref(cb), C.byref(bins._b), bsp,
                                 flags, out.ctypes.data)
        self._check(e, "fcfc_gpu_count")
        return out

    def count_backward(self, a, b, bins, grad_cnt, isauto=False, need_b=True):
        """d(sum_k G_k cnt_k)/dw for the objects of a (and of b for cross
        counts unless need_b is False, which skips accumulating b's terms).
        One pass over the pairs as the count enumerates them."""
        g = _f64(grad_cnt)
        ga = np.zeros(a.n)
        gb = np.zeros((b if b is not None else a).n)
        bs = bins.bsize_rescaled
        ca = a._c()
        cb = (b if b is not None else a)._c()
        self._sync_torch(a, b)
        e = lib().fcfc_gpu_count_backward(self._h, C.byref(ca), C.byref(cb), C.byref(bins._b),
                                          _dptr(bs) if bs is not None else None,
                                          L.FCFC_GPU_AUTO if isauto else 0, _dptr(g),
                                          _dptr(ga), None if (isauto or not need_b) else _dptr(gb))
        self._check(e, "fcfc_gpu_count_backward")
        return (ga, None) if (isauto or not need_b) else (ga, gb)

    def plan(self, jobs, bins, precision="fp64"):
        """A prepared batch (see fcfc_gpu_plan_create): jobs as for
        count_batch, catalogues on the device.  plan.run() recounts them,
        re-reading the weights from the catalogues' device arrays."""
        return Plan(self, jobs, bins, precision)

    def count_batch(self, jobs, bins, precision="fp64"):
        """jobs: list of (a, b_or_None, isauto, withwt).  Returns one
        histogram per job (float64 for weighted jobs, int64 otherwise); the
        histograms are rows of one array allocated for the call."""
        n = len(jobs)
        arr = _job_array(jobs, bins)
        raw = np.zeros((max(n, 1), bins.ntot), dtype=np.int64)
        flags = _prec_flag(precision)
        self._sync_torch(*[c for j in jobs for c in j[:2] if c is not None])
        e = lib().fcfc_gpu_count_batch(self._h, n, arr, C.byref(bins._b), flags, raw.ctypes.data)
        self._check(e, "fcfc_gpu_count_batch")
        out = []
        for k, (_, _, _, withwt) in enumerate(jobs):
            out.append(raw[k].view(np.float64) if withwt else raw[k])
        return out

    def stats(self):
        s = L.fcfc_gpu_stats()
        self._check(lib().fcfc_gpu_last_stats(self._h, C.byref(s)), "fcfc_gpu_last_stats")
        return {k: getattr(s, k) for k, _ in s._fields_}



class Plan:
    """Prepared batch replayed as a HIP graph (fcfc_gpu_plan_*).  The jobs'
    catalogues must stay alive (their device arrays are re-read by run())."""

    def __init__(self, eng, jobs, bins, precision="fp64"):
        self._eng = eng
        self._jobs = list(jobs)
        self._keep = [c for j in self._jobs for c in j[:2] if c is not None]
        self.bins = bins
        arr = _job_array(self._jobs, bins)
        eng._sync_torch(*self._keep)
        h = C.c_void_p()
        e = lib().fcfc_gpu_plan_create(eng._h, len(self._jobs), arr, C.byref(bins._b),
                                       _prec_flag(precision), C.byref(h))
        eng._check(e, "fcfc_gpu_plan_create")
        self._h = h

    def run(self):
        """Count every job again; returns one histogram per job."""
        n = len(self._jobs)
        raw = np.zeros((n, self.bins.ntot), dtype=np.int64)
        self._eng._sync_torch(*self._keep)
        e = lib().fcfc_gpu_plan_run(self._h, raw.ctypes.data)
        self._eng._check(e, "fcfc_gpu_plan_run")
        return [raw[k].view(np.float64).copy() if j[3] else raw[k].copy()
                for k, j in enumerate(self._jobs)]

    def close(self):
        if getattr(self, "_h", None):
            lib().fcfc_gpu_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

def pair_counts(cnt, a, b, isauto, withwt):
    """The caller conventions of FCFC's eval_pairs (eval_cf.c:117-167):
    auto counts doubled; norm = n(n-1) or W(W-1) (auto), n1 n2 or W1 W2
    (cross); returns (counts, norm, normalised counts)."""
    c = np.array(cnt, copy=True)
    if isauto:
        c = c * 2
        norm = a.sumw * (a.sumw - 1) if withwt else float(a.n) * (a.n - 1)
    else:
        norm = a.sumw * b.sumw if withwt else float(a.n) * b.n
    return c, norm, c / norm


def save_pair_count(fname, bins, cnt, a, b, isauto, withwt, ofmt=0):
    """Write an FCFC pair-count file (save_res.c:177-296 binary, :45-164 ASCII)
    from final (doubled) counts."""
    bs = bins.bsize_rescaled
    c = np.ascontiguousarray(cnt, dtype=np.float64 if withwt else np.int64)
    e = lib().fcfc_save_pair_count(fname.encode(), C.byref(bins._b), ofmt,
                                   _dptr(bs) if bs is not None else None, int(isauto),
                                   int(withwt), a.n, (b if b is not None else a).n,
                                   a.sumw, (b if b is not None else a).sumw, c.ctypes.data)
    if e:
        raise FcfcError(e, "fcfc_save_pair_count")


# -- survey coordinate conversion (src/fcfc/2pt/cnvt_coord.c) -----------------
def convert_coords(ra, dec, z, omega_m, omega_l=None, w=None, err=None):
    """(RA, Dec) in degrees and redshift -> comoving Cartesian coordinates
    (Mpc/h), integrating c/H(z) by Gauss-Legendre quadrature of the order
    FCFC's convergence test picks; returns (x, y, z, order)."""
    nan = float("nan")
    cm = L.fcfc_cosmology(omega_m, nan if omega_l is None else omega_l, 0.0,
                          nan if w is None else w, nan if err is None else err)
    e = lib().fcfc_cosmology_check(C.byref(cm))
    if e:
        raise FcfcError(e, "fcfc_cosmology_check: invalid cosmology")
    x, y, zz = (np.array(a, dtype=np.float64, copy=True) for a in (ra, dec, z))
    order = C.c_int(0)
    e = lib().fcfc_convert_coords(C.byref(cm), len(x), _dptr(x), _dptr(y), _dptr(zz),
                                  C.byref(order))
    if e:
        raise FcfcError(e, "fcfc_convert_coords")
    return x, y, zz, order.value


def convert_coords_table(zs, ds, ra, dec, z):
    """The same with distances from a natural cubic spline through the
    (redshift, distance) samples (Z_CMVDST_FILE); returns (x, y, z)."""
    zs, ds = _f64(zs), _f64(ds)
    x, y, zz = (np.array(a, dtype=np.float64, copy=True) for a in (ra, dec, z))
    e = lib().fcfc_convert_coords_table(_dptr(zs), _dptr(ds), len(zs), len(x), _dptr(x),
                                        _dptr(y), _dptr(zz))
    if e:
        raise FcfcError(e, "fcfc_convert_coords_table")
    return x, y, zz


# -- deterministic generator ---------------------------------------------------
def gen_uniform(seed, n, L_):
    x, y, z = (np.empty(n) for _ in range(3))
    e = lib().fcfc_gen_uniform(seed, n, L_, _dptr(x), _dptr(y), _dptr(z))
    if e:
        raise FcfcError(e, "fcfc_gen_uniform")
    return x, y, z


def gen_clustered(seed, nunif, nhalo, nper, sigma, L_, wsig=None):
    n = nunif + nhalo * nper
    x, y, z = (np.empty(n) for _ in range(3))
    w = np.empty(n) if wsig is not None else None
    e = lib().fcfc_gen_clustered(seed, nunif, nhalo, nper, sigma, L_,
                                 wsig if wsig is not None else 0.0, _dptr(x), _dptr(y),
                                 _dptr(z), _dptr(w))
    if e:
        raise FcfcError(e, "fcfc_gen_clustered")
    return x, y, z, w


def gen_survey(seed, n, rmin=1000.0, rmax=2000.0, nbar=3e-4, rpeak=1500.0,
               rsig=200.0, P0=1e4):
    x, y, z, w = (np.empty(n) for _ in range(4))
    e = lib().fcfc_gen_survey(seed, n, rmin, rmax, nbar, rpeak, rsig, P0, _dptr(x),
                              _dptr(y), _dptr(z), _dptr(w))
    if e:
        raise FcfcError(e, "fcfc_gen_survey")
    return x, y, z, w


def gen_sizes(seed, n, lo, hi):
    out = np.empty(n, dtype=np.int64)
    e = lib().fcfc_gen_sizes(seed, n, lo, hi, out.ctypes.data_as(C.POINTER(C.c_int64)))
    if e:
        raise FcfcError(e, "fcfc_gen_sizes")
    return out
