"""Weight gradients: the GPU backward kernel against the FP64 CPU gradient of
the oracle, against central finite differences of the GPU counts, and the
torch.autograd wrapper (gradcheck) on a differentiable xi_ell pipeline."""
import numpy as np
import pytest

import fcfc_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu
L = 1000.0


@pytest.fixture(scope="module")
def eng():
    with F.Engine() as e:
        yield e


@pytest.fixture(scope="module")
def setup():
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=10, bsize=L)
    x, y, z, w = F.gen_clustered(31, 1500, 20, 30, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(32, 2500, L)
    wr = np.linspace(0.6, 1.4, len(xr))
    D = F.Catalog(b, x, y, z, w)
    R = F.Catalog(b, xr, yr, zr, wr)
    G = np.random.default_rng(3).normal(size=b.ntot)
    return b, D, R, G


@pytest.mark.parametrize("isauto", [True, False])
def test_backward_matches_fp64_oracle(eng, setup, isauto):
    b, D, R, G = setup
    ob = O.oracle_bins(1, b.s2bin, b.nmu)
    print(f"case backward isauto={isauto}", flush=True)
    ga, gb = eng.count_backward(D, None if isauto else R, b, G, isauto=isauto)
    ra, rb = O.weight_grad(0, b.bsize_rescaled, (D.x, D.y, D.z, None, D.w),
                           None if isauto else (R.x, R.y, R.z, None, R.w), isauto, ob, G)
    np.testing.assert_allclose(ga, ra, rtol=1e-6, atol=1e-9 * np.abs(ra).max())
    if not isauto:
        np.testing.assert_allclose(gb, rb, rtol=1e-6, atol=1e-9 * np.abs(rb).max())


@pytest.mark.parametrize("bt", [F.BIN_SMU, F.BIN_SPI])
def test_backward_hybrid_and_spi(eng, bt):
    """The gradient kernel with hybrid (non-integer) tables and with
    (s_perp, pi) bins, cross and auto, against the FP64 oracle."""
    b = F.Bins.linear(F.GEOM_BOX, bt, 0.5, 146.5, 7.3, nmu=7 if bt == F.BIN_SMU else 1,
                      pmin=1.5, pmax=96.5, dpi=3.7, bsize=L)
    assert b.tabtype == 1
    x, y, z, w = F.gen_clustered(41, 1200, 15, 30, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(42, 2000, L)
    wr = np.linspace(0.7, 1.3, len(xr))
    D = F.Catalog(b, x, y, z, w)
    R = F.Catalog(b, xr, yr, zr, wr)
    G = np.random.default_rng(5).normal(size=b.ntot)
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu, b.pbin if bt == F.BIN_SPI else None)
    for isauto in (True, False):
        print(f"case backward hybrid bt={bt} isauto={isauto}", flush=True)
        ga, gb = eng.count_backward(D, None if isauto else R, b, G, isauto=isauto)
        ra, rb = O.weight_grad(0, b.bsize_rescaled, (D.x, D.y, D.z, None, D.w),
                               None if isauto else (R.x, R.y, R.z, None, R.w), isauto, ob, G)
        np.testing.assert_allclose(ga, ra, rtol=1e-6, atol=1e-9 * np.abs(ra).max())
        if not isauto:
            np.testing.assert_allclose(gb, rb, rtol=1e-6, atol=1e-9 * np.abs(rb).max())


@pytest.mark.parametrize("bt", [F.BIN_SMU, F.BIN_SPI])
def test_backward_survey(eng, bt):
    """Survey geometry: auto gradient, and cross gradients with the smaller
    catalogue first (the pass puts the other one in the lanes and reduces
    the first one's terms over wavefronts) and with the larger first and
    only its own gradient asked for, against the FP64 oracle."""
    b = F.Bins.linear(F.GEOM_SURVEY, bt, 0, 200, 5, nmu=10 if bt == F.BIN_SMU else 1,
                      pmin=0, pmax=200 if bt == F.BIN_SPI else 0, dpi=5)
    x, y, z, w = F.gen_survey(51, 3000)
    xr, yr, zr, wr = F.gen_survey(52, 5000)
    D = F.Catalog(b, x, y, z, w)
    R = F.Catalog(b, xr, yr, zr, wr)
    G = np.random.default_rng(9).normal(size=b.ntot)
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu, b.pbin if bt == F.BIN_SPI else None)
    tD = (D.x, D.y, D.z, D.s, D.w)
    tR = (R.x, R.y, R.z, R.s, R.w)

    def close(g, r):
        np.testing.assert_allclose(g, r, rtol=1e-6, atol=1e-9 * np.abs(r).max())

    print(f"case backward survey bt={bt} auto", flush=True)
    ga, _ = eng.count_backward(D, None, b, G, isauto=True)
    ra, _ = O.weight_grad(1, None, tD, None, True, ob, G)
    close(ga, ra)
    print(f"case backward survey bt={bt} cross", flush=True)
    ga, gb = eng.count_backward(D, R, b, G)
    ra, rb = O.weight_grad(1, None, tD, tR, False, ob, G)
    close(ga, ra)
    close(gb, rb)
    print(f"case backward survey bt={bt} cross, larger first, one side", flush=True)
    gr, none = eng.count_backward(R, D, b, G, need_b=False)
    assert none is None
    close(gr, rb)


def test_backward_finite_difference(eng, setup):
    b, D, R, G = setup
    ga, gb = eng.count_backward(D, R, b, G, isauto=False)
    for i in (3, 800, 1999):
        vals = []
        for sgn in (1, -1):
            w2 = D.w.copy()
            w2[i] += sgn * 1e-3
            D2 = F.Catalog.__new__(F.Catalog)
            D2.__dict__.update(D.__dict__)
            D2.w = w2
            c = eng.count_pairs(D2, R, b, isauto=False, withwt=True)
            vals.append(G @ c)
        fd = (vals[0] - vals[1]) / 2e-3
        assert ga[i] == pytest.approx(fd, rel=1e-7, abs=1e-9)


def test_autograd_xi_multipoles():
    import torch
    from fcfc_amd.xi import WeightedPairCount, xi_multipoles_torch
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=10, bsize=L)
    x, y, z, w = F.gen_clustered(41, 400, 8, 25, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(42, 900, L)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a * b.rescale)).cuda()
    D = F.Catalog(b, dev(x), dev(y), dev(z))
    R = F.Catalog(b, dev(xr), dev(yr), dev(zr))
    wd = torch.tensor(w, dtype=torch.float64, device="cuda", requires_grad=True)
    wr = torch.ones(len(xr), dtype=torch.float64, device="cuda")
    with F.Engine() as eng:
        rr = WeightedPairCount.apply(wr, None, eng, R, None, b, True).detach()
        target = torch.zeros(2, b.ns, dtype=torch.float64, device="cuda")

        def loss_fn(wd_):
            xi = xi_multipoles_torch(wd_, wr, eng, D, R, b, rr_cnt=rr, poles=(0, 2))
            return ((xi - target) ** 2).sum()

        loss = loss_fn(wd)
        loss.backward()
        g = wd.grad.detach().cpu().numpy()
        # central differences of the full pipeline on a few weights
        for i in (0, 17, 333):
            e = torch.zeros_like(wd)
            e[i] = 1e-4
            with torch.no_grad():
                fd = (loss_fn(wd + e) - loss_fn(wd - e)).item() / 2e-4
            assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-8 * abs(g).max())
        # float32 weights (the differentiable-weights configuration)
        w32 = torch.tensor(w, dtype=torch.float32, device="cuda", requires_grad=True)
        cnt = WeightedPairCount.apply(w32, None, eng, D, None, b, True)
        cnt.sum().backward()
        assert w32.grad.dtype == torch.float32 and torch.isfinite(w32.grad).all()


def test_weight_fit_plan_matches_eager():
    """The plan-based fitting counts (HIP graph replays) give the same loss
    and gradient as the eager autograd path, over several weight updates."""
    import torch
    from fcfc_amd.xi import WeightFitCounts, WeightedPairCount, xi_multipoles_torch
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=10, bsize=L)
    x, y, z, w = F.gen_clustered(51, 500, 8, 25, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(52, 1200, L)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a * b.rescale)).cuda()
    D = F.Catalog(b, dev(x), dev(y), dev(z))
    R = F.Catalog(b, dev(xr), dev(yr), dev(zr))
    wr = torch.ones(len(xr), dtype=torch.float64, device="cuda")
    with F.Engine() as eng:
        rr = WeightedPairCount.apply(wr, None, eng, R, None, b, True).detach()
        fit = WeightFitCounts(eng, D, R, b)
        try:
            wd = torch.tensor(w, dtype=torch.float64, device="cuda")
            for step in range(3):
                print(f"case weight fit step={step}", flush=True)
                w1 = wd.clone().requires_grad_(True)
                l1 = (xi_multipoles_torch(w1, wr, eng, D, R, b, rr_cnt=rr, fit=fit) ** 2).sum()
                l1.backward()
                w2 = wd.clone().requires_grad_(True)
                l2 = (xi_multipoles_torch(w2, wr, eng, D, R, b, rr_cnt=rr) ** 2).sum()
                l2.backward()
                assert l1.item() == pytest.approx(l2.item(), rel=1e-12)
                np.testing.assert_allclose(w1.grad.cpu().numpy(), w2.grad.cpu().numpy(),
                                           rtol=1e-9, atol=1e-12 * float(w2.grad.abs().max()))
                wd = wd * (1.0 + 0.05 * torch.sin(torch.arange(len(w), device="cuda") * 0.1))
        finally:
            fit.close()


def test_backward_hybrid_tiles_beyond_one_per_lane(eng):
    """Hybrid binning holds one A object per lane in the gradient kernel.
    Dense halos put more than 1024 objects in a cell, so tiles are as large
    as the tile size allows: with 1024-object tiles set by the caller, and
    with the autotuner's shape, every object's gradient must still match
    the FP64 oracle (auto and cross)."""
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0.5, 146.5, 7.3, nmu=7, bsize=L)
    assert b.tabtype == 1
    x, y, z, w = F.gen_clustered(71, 1000, 4, 3000, 1.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(72, 3000, L)
    wr = np.linspace(0.7, 1.3, len(xr))
    D = F.Catalog(b, x, y, z, w)
    R = F.Catalog(b, xr, yr, zr, wr)
    G = np.random.default_rng(7).normal(size=b.ntot)
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu)
    refs = {True: O.weight_grad(0, b.bsize_rescaled, (D.x, D.y, D.z, None, D.w), None, True, ob, G),
            False: O.weight_grad(0, b.bsize_rescaled, (D.x, D.y, D.z, None, D.w),
                                 (R.x, R.y, R.z, None, R.w), False, ob, G)}

    def check(tag):
        for isauto in (True, False):
            print(f"case backward hybrid dense {tag} isauto={isauto}", flush=True)
            ga, gb = eng.count_backward(D, None if isauto else R, b, G, isauto=isauto)
            ra, rb = refs[isauto]
            np.testing.assert_allclose(ga, ra, rtol=1e-6, atol=1e-9 * np.abs(ra).max())
            if not isauto:
                np.testing.assert_allclose(gb, rb, rtol=1e-6, atol=1e-9 * np.abs(rb).max())

    try:
        eng.set_tuning(tile=1024)
        check("tile=1024")
    finally:
        eng.set_tuning()
    for isauto in (True, False):
        eng.autotune(D, None if isauto else R, b, isauto=isauto, withwt=True, gradient=True)
    check("autotuned")
