"""Correlation functions from pair counts, and weight gradients.

* Post-count formulas of FCFC's eval_cf.c: normalisation (:117-167),
  analytic RR for periodic boxes (:216-242), multipoles with midpoint mu and
  Legendre polynomials (:326-354, src/math/legpoly.h:47-61), projected
  w_p (:365-388).  numpy versions for files and checks, torch versions for
  differentiable pipelines.
* WeightedPairCount: a torch.autograd.Function whose forward is the GPU
  weighted pair count and whose backward is the GPU weight-gradient kernel
  (d cnt / d w, contracted with the incoming gradient).
"""
import functools

import numpy as np

from . import Catalog


def legpoly(ell, x):
    """Legendre P_ell(x) for ell <= 6, in the same polynomial forms as
    src/math/legpoly.h:47-61."""
    x2 = x * x
    if ell == 0:
        return x * 0 + 1
    if ell == 1:
        return x
    if ell == 2:
        return 1.5 * x2 - 0.5
    if ell == 3:
        return 2.5 * x * (x2 - 0.6)
    if ell == 4:
        return 4.375 * x2 * x2 - 3.75 * x2 + 0.375
    if ell == 5:
        return 8.75 * x * (0.9 * x2 * x2 - x2 + 0.21428571428571427)
    if ell == 6:
        return 14.4375 * x2 * x2 * x2 - 19.6875 * x2 * x2 + 6.5625 * x2 - 0.3125
    raise ValueError("ell must be <= 6")


@functools.lru_cache(maxsize=32)
def _pole_weights(nmu, poles):
    mu = (np.arange(nmu) + 0.5) / nmu
    w = np.stack([(2 * ell + 1) / nmu * legpoly(ell, mu) for ell in poles])
    w.setflags(write=False)
    return w


def multipoles(xi_smu, nmu, poles=(0, 2, 4)):
    """xi_ell(s) = (2 ell + 1)/nmu sum_j xi(s, mu_j) P_ell(mu_j), mu_j the bin
    midpoints (eval_cf.c:326-354).  xi_smu: (..., nmu, ns); returns
    (..., len(poles), ns), one matrix product with the cached weights."""
    with np.errstate(invalid="ignore", over="ignore"):      # NaN bins propagate
        return np.matmul(_pole_weights(nmu, tuple(poles)), xi_smu)


def wp(xi_spi, pbin_raw):
    """w_p(s_perp) = sum_j 2 xi(s_perp, pi_j) dpi_j (eval_cf.c:365-388);
    xi_spi: (np, ns)."""
    dpi = np.diff(pbin_raw)
    return 2 * np.tensordot(dpi, xi_spi, axes=([0], [0]))


def analytic_rr(bins, bsize):
    """Normalised analytic RR of a periodic box (eval_cf.c:216-242), raw
    (unrescaled) edges; shape (nmu|np, ns) or (ns,)."""
    V = float(np.prod(np.broadcast_to(bsize, (3,))))
    s = bins.sbin_raw
    if bins.bintype == 2:
        p = bins.pbin_raw
        fac = 2 * np.pi / V
        return fac * np.outer(np.diff(p), s[1:] ** 2 - s[:-1] ** 2)
    fac = 4 * np.pi / (3 * bins.nmu * V)
    rr = fac * (s[1:] ** 3 - s[:-1] ** 3)
    return np.tile(rr, (bins.nmu, 1)) if bins.bintype == 1 else rr


def landy_szalay(dd, dr, rr):
    """(DD - 2 DR + RR) / RR on normalised counts."""
    return (dd - 2 * dr + rr) / rr


# --------------------------------------------------------------------------
def _torch():
    import torch
    return torch


class WeightedPairCount:
    """Differentiable weighted pair counts.

        cnt = WeightedPairCount.apply(wa, wb, eng, A, B, bins, isauto)

    wa (wb): per-object weights of catalogue A (B), 1-D torch tensors on the
    GPU (any float dtype; counted in FP64).  A, B: device Catalogs holding
    the rescaled coordinates (their own weights are ignored).  Returns the
    FP64 histogram of bins.ntot entries (auto pairs counted once, as FCFC's
    count_pairs).  Backward: d(sum_k g_k cnt_k)/dw from the GPU gradient
    kernel.  For auto counts pass wb = None.
    """

    _fn = None

    @classmethod
    def apply(cls, wa, wb, eng, A, B, bins, isauto, precision="fp64"):
        if cls._fn is None:
            cls._fn = _make_fn()
        return cls._fn.apply(wa, wb, eng, A, B, bins, isauto, precision)


def _with_weights(cat, bins, w):
    torch = _torch()
    wd = w.detach().to(torch.float64).contiguous()
    return Catalog(bins, cat.x, cat.y, cat.z, wd, cat.s)


def _make_fn():
    torch = _torch()

    class _Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, wa, wb, eng, A, B, bins, isauto, precision):
            Aw = _with_weights(A, bins, wa)
            Bw = None if isauto else _with_weights(B, bins, wb)
            cnt = eng.count_pairs(Aw, Bw, bins, isauto=isauto, withwt=True, precision=precision)
            ctx.stuff = (eng, Aw, Bw, bins, isauto, wa.dtype, None if wb is None else wb.dtype)
            return torch.from_numpy(cnt).to(wa.device)

        @staticmethod
        def backward(ctx, g):
            eng, Aw, Bw, bins, isauto, da, db = ctx.stuff
            gn = g.detach().to(torch.float64).cpu().numpy()
            ga, gb = eng.count_backward(Aw, Bw, bins, gn, isauto=isauto)
            dev = Aw.x.device
            ga_t = torch.from_numpy(ga).to(dev).to(da)
            gb_t = None if (isauto or gb is None) else torch.from_numpy(gb).to(dev).to(db)
            return ga_t, gb_t, None, None, None, None, None, None

    return _Fn


class WeightFitCounts:
    """Differentiable DD and DR of fixed data and random positions for
    changing data weights, for fitting loops: the counts come from a
    prepared plan (catalogues sorted and tiled once, each forward a HIP graph
    replay that re-reads the weights), the gradients from the GPU backward
    kernel.

        fit = WeightFitCounts(eng, D, R, bins)      # D, R: device Catalogs
        dd, dr = fit(wd)                            # wd: (D.n,) tensor
    R keeps the weights it was created with (none: unit weights).
    """

    def __init__(self, eng, D, R, bins, precision="fp64"):
        torch = _torch()
        self.eng, self.bins = eng, bins
        self._wd = torch.ones(D.n, dtype=torch.float64, device=D.x.device)
        wr = R.w if R.w is not None else torch.ones(R.n, dtype=torch.float64, device=R.x.device)
        self.D = Catalog(bins, D.x, D.y, D.z, self._wd, D.s)
        self.R = Catalog(bins, R.x, R.y, R.z, wr.contiguous(), R.s)
        self.plan = eng.plan([(self.D, None, True, True), (self.D, self.R, False, True)],
                             bins, precision)
        self._fn = _make_fit_fn()

    def __call__(self, wd):
        return self._fn.apply(wd, self)

    def close(self):
        self.plan.close()


def _make_fit_fn():
    torch = _torch()

    class _Fit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, wd, fit):
            fit._wd.copy_(wd.detach().to(torch.float64))
            dd, dr = fit.plan.run()
            ctx.fit = fit
            ctx.wdtype = wd.dtype
            dev = wd.device
            return torch.from_numpy(dd).to(dev), torch.from_numpy(dr).to(dev)

        @staticmethod
        def backward(ctx, gdd, gdr):
            fit = ctx.fit
            g = np.zeros(fit.D.n)
            if gdd is not None:
                ga, _ = fit.eng.count_backward(fit.D, None, fit.bins,
                                               gdd.detach().to(torch.float64).cpu().numpy(),
                                               isauto=True)
                g += ga
            if gdr is not None:
                ga, _ = fit.eng.count_backward(fit.D, fit.R, fit.bins,
                                               gdr.detach().to(torch.float64).cpu().numpy(),
                                               isauto=False, need_b=False)
                g += ga
            return torch.from_numpy(g).to(fit._wd.device).to(ctx.wdtype), None

    return _Fit


def xi_multipoles_torch(wd, wr, eng, D, R, bins, rr_cnt=None, poles=(0, 2), fit=None):
    """Differentiable Landy-Szalay xi_ell(s) from (s,mu) counts: DD and DR
    through WeightedPairCount (or through `fit`, a WeightFitCounts of the
    same catalogues, in fitting loops), RR (independent of wd) counted once
    and passed in as rr_cnt (unordered, FP64) when available."""
    torch = _torch()
    ns, nmu = bins.ns, bins.nmu
    if fit is not None:
        dd, dr = fit(wd)
        dd = 2 * dd
    else:
        dd = 2 * WeightedPairCount.apply(wd, None, eng, D, None, bins, True)
        dr = WeightedPairCount.apply(wd, wr, eng, D, R, bins, False)
    if rr_cnt is None:
        rr_cnt = WeightedPairCount.apply(wr, None, eng, R, None, bins, True).detach()
    rr = 2 * rr_cnt
    Wd = wd.to(torch.float64).sum()
    Wr = wr.to(torch.float64).sum()
    ddn = dd / (Wd * (Wd - 1))
    drn = dr / (Wd * Wr)
    rrn = rr / (Wr * (Wr - 1))
    # bins without random pairs carry no information (FCFC would print inf/nan)
    safe = torch.where(rrn > 0, rrn, torch.ones_like(rrn))
    xi = torch.where(rrn > 0, (ddn - 2 * drn + rrn) / safe, torch.zeros_like(rrn)).reshape(nmu, ns)
    mu = (torch.arange(nmu, dtype=torch.float64, device=xi.device) + 0.5) / nmu
    out = []
    for ell in poles:
        w = (2 * ell + 1) / nmu * legpoly(ell, mu)
        out.append((xi * w[:, None]).sum(0))
    return torch.stack(out)
