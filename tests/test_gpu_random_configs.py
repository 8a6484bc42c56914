"""GPU parity over seeded random configurations: geometry, binning type,
integer or hybrid bin edges, number of mu bins, lower pi edge, box shape,
catalogue sizes and clustering, auto or cross (in both catalogue orders),
every form, single and batched counts, prepared plans (HIP graph replays)
and the weight gradient -- against
the oracle (counts identical when the ambiguity census is zero, weighted
counts within 1e-12 relative, gradients within 1e-6)."""
import numpy as np
import pytest

import fcfc_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu
WT_RTOL = 1e-12


@pytest.fixture(scope="module")
def eng():
    with F.Engine() as e:
        yield e


def _config(seed, lfrac=0.4):
    """Configuration of one seed; separations up to lfrac of the shortest
    box side (the reference requires less than half)."""
    rng = np.random.default_rng(1000 + seed)
    geom = F.GEOM_BOX if seed % 3 else F.GEOM_SURVEY
    bt = (F.BIN_ISO, F.BIN_SMU, F.BIN_SPI)[(seed // 3) % 3]
    if geom == F.GEOM_BOX:
        Lb = rng.choice([1000.0, 600.0, 850.0], 3) if rng.random() < 0.3 else np.full(3, 1000.0)
        lmax = lfrac * Lb.min()
    else:
        Lb, lmax = None, 300.0
    if rng.random() < 0.5:            # integer-friendly edges
        ds = float(rng.choice([2.0, 4.0, 5.0, 10.0]))
        smin = float(rng.choice([0.0, ds, 2 * ds]))
    else:                             # edges the rescaling cannot make integer: hybrid tables
        ds = float(np.round(rng.uniform(1.3, 9.7), 2))
        smin = float(np.round(rng.uniform(0.0, 3.0 * ds), 2))
    ns = int(rng.integers(3, 41))
    smax = min(smin + ns * ds, lmax)
    ns = max(1, int((smax - smin) // ds))
    smax = smin + ns * ds
    nmu = int(rng.choice([2, 3, 10, 37, 100])) if bt == F.BIN_SMU else 1
    if bt == F.BIN_SPI:
        dpi = float(rng.choice([2.0, 5.0, 7.5]))
        pmin = float(rng.choice([0.0, 0.0, dpi, 3 * dpi]))
        pmax = pmin + dpi * int(rng.integers(2, 30))
        pmax = min(pmax, lmax)
        pmax = pmin + dpi * max(1, int((pmax - pmin) // dpi))
    else:
        dpi, pmin, pmax = 1.0, 0.0, 0.0
    b = F.Bins.linear(geom, bt, smin, smax, ds, nmu=nmu, pmin=pmin, pmax=pmax, dpi=dpi,
                      bsize=Lb)
    n1 = int(rng.integers(1, 4000))
    n2 = int(rng.integers(1, 6000))
    if geom == F.GEOM_BOX:
        def cat(n, clustered):
            if clustered:
                x, y, z, w = F.gen_clustered(int(rng.integers(1 << 30)), n // 2, max(1, n // 100),
                                             max(1, (n - n // 2) // max(1, n // 100)), 2.0,
                                             1000.0, 0.3)
                x, y, z = x * Lb[0] / 1000.0, y * Lb[1] / 1000.0, z * Lb[2] / 1000.0
            else:
                x, y, z = (rng.random(n) * Lb[k] for k in range(3))
                w = rng.lognormal(0.0, 0.3, n)
            return F.Catalog(b, x, y, z, w)
        A = cat(n1, rng.random() < 0.5)
        B = cat(n2, False)
    else:
        x, y, z, w = F.gen_survey(int(rng.integers(1 << 30)), n1)
        A = F.Catalog(b, x, y, z, w)
        x, y, z, w = F.gen_survey(int(rng.integers(1 << 30)), n2)
        B = F.Catalog(b, x, y, z, w)
    return geom, bt, b, A, B


# seed 8: hybrid edges whose table does not fit in LDS; seed 127: a hybrid
# lower edge just below an integer, inside the FP32 window of the next cell
@pytest.mark.parametrize("seed", list(range(48)) + [127])
def test_random_config(eng, seed, lfrac=0.4):
    geom, bt, b, A, B = _config(seed, lfrac)
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu, b.pbin if bt == F.BIN_SPI else None)
    og, bs, var, cmask = ((0, b.bsize_rescaled, 1, 0b11111) if geom == F.GEOM_BOX
                          else (1, None, 0, 0b111111))
    forms = ("fp64", "fp32", "plain")
    tA = (A.x, A.y, A.z, A.s, None)
    tB = (B.x, B.y, B.z, B.s, None)
    wA = (A.x, A.y, A.z, A.s, A.w)
    wB = (B.x, B.y, B.z, B.s, B.w)
    print(f"case random seed={seed} geom={geom} bt={bt} tab={b.tabtype} ns={b.ns} "
          f"nmu={b.nmu} nA={A.n} nB={B.n}", flush=True)
    ref_a, cen_a = O.count(og, bs, tA, None, True, ob, variant=var, census_mask=cmask)
    ref_x, cen_x = O.count(og, bs, tA, tB, False, ob, variant=var, census_mask=cmask)
    ref_aw, _ = O.count(og, bs, wA, None, True, ob, variant=var, weighted=True)
    ref_xw, _ = O.count(og, bs, wA, wB, False, ob, variant=var, weighted=True)

    def same(g, r, cen):
        diff = np.abs(g.astype(np.int64) - r.astype(np.int64)).sum()
        assert diff <= 2 * cen, f"counts differ by {diff} (census {cen})"
        if cen == 0:
            np.testing.assert_array_equal(g, r)

    def close(g, r):
        scale = np.maximum(np.abs(r), 1e-300)
        rel = np.max(np.abs(g - r) / scale) if len(r) else 0.0
        assert rel <= WT_RTOL, f"weighted counts differ: max rel {rel}"

    for prec in forms:
        same(eng.count_pairs(A, None, b, isauto=True, precision=prec), ref_a, cen_a)
        same(eng.count_pairs(A, B, b, precision=prec), ref_x, cen_x)
        same(eng.count_pairs(B, A, b, precision=prec), ref_x, cen_x)
        close(eng.count_pairs(A, None, b, isauto=True, withwt=True, precision=prec), ref_aw)
        close(eng.count_pairs(A, B, b, withwt=True, precision=prec), ref_xw)
    # one batched launch of mixed jobs, in every form
    for prec in forms:
        ba, bx, bxw = eng.count_batch([(A, None, True, False), (A, B, False, False),
                                       (A, B, False, True)], b, precision=prec)
        same(ba, ref_a, cen_a)
        same(bx, ref_x, cen_x)
        close(bxw, ref_xw)
    # a prepared plan on device copies of the catalogues, replayed twice, in
    # one form per seed
    import torch
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
    dA = F.Catalog(b, dev(A.x), dev(A.y), dev(A.z), dev(A.w))
    dB = F.Catalog(b, dev(B.x), dev(B.y), dev(B.z), dev(B.w))
    plan = eng.plan([(dA, None, True, False), (dA, dB, False, True)], b,
                    precision=forms[seed % 3])
    try:
        for _ in range(2):
            pa, pxw = plan.run()
            same(pa, ref_a, cen_a)
            close(pxw, ref_xw)
    finally:
        plan.close()
    # weight gradients (pairs binned as the count bins them: equal bins
    # when the census is zero)
    if cen_a == 0 and cen_x == 0:
        G = np.random.default_rng(seed).normal(size=b.ntot)

        def grad_close(g, r):
            np.testing.assert_allclose(g, r, rtol=1e-6, atol=1e-9 * max(np.abs(r).max(), 1e-300))

        ga, _ = eng.count_backward(A, None, b, G, isauto=True)
        ra, _ = O.weight_grad(og, bs, wA, None, True, ob, G)
        grad_close(ga, ra)
        ga, gb = eng.count_backward(A, B, b, G)
        ra, rb = O.weight_grad(og, bs, wA, wB, False, ob, G)
        grad_close(ga, ra)
        grad_close(gb, rb)


@pytest.mark.parametrize("seed", range(12))
def test_random_config_near_half_box(eng, seed):
    """Separations up to 0.49 of the shortest box side: periodic images of
    the strips as close to aliasing as the reference allows."""
    test_random_config(eng, seed, lfrac=0.49)


def _large_config(seed):
    rng = np.random.default_rng(5000 + seed)
    bt = (F.BIN_ISO, F.BIN_SMU, F.BIN_SPI)[seed % 3]
    L = 1000.0
    if seed % 2 == 0:                 # integer-friendly edges; odd seeds: hybrid tables
        ds = float(rng.choice([2.0, 4.0, 5.0]))
        smin = float(rng.choice([0.0, ds]))
    else:
        ds = float(np.round(rng.uniform(1.3, 6.0), 2))
        smin = float(np.round(rng.uniform(0.0, 2.0 * ds), 2))
    ns = max(1, int((min(smin + int(rng.integers(5, 41)) * ds, 120.0) - smin) // ds))
    nmu = int(rng.choice([5, 40, 100])) if bt == F.BIN_SMU else 1
    dpi, pmax = (5.0, 5.0 * int(rng.integers(4, 21))) if bt == F.BIN_SPI else (1.0, 0.0)
    b = F.Bins.linear(F.GEOM_BOX, bt, smin, smin + ns * ds, ds, nmu=nmu, pmin=0.0, pmax=pmax,
                      dpi=dpi, bsize=L)
    na = int(rng.integers(30000, 80000))
    x, y, z, w = F.gen_clustered(int(rng.integers(1 << 30)), na // 2, na // 400, 200, 2.0, L, 0.3)
    A = F.Catalog(b, x, y, z, w)
    xr, yr, zr = F.gen_uniform(int(rng.integers(1 << 30)), int(rng.integers(30000, 100000)), L)
    B = F.Catalog(b, xr, yr, zr)
    shape = (int(rng.choice([1, 2, 3, 4, 6])), int(rng.choice([64, 256, 512, 1024])),
             int(rng.choice([1, 2, 4, 8, 16])), int(rng.choice([0, 7, 64, 300])))
    return bt, b, A, B, shape


@pytest.mark.parametrize("seed", range(6))
def test_random_large(eng, seed):
    """Larger catalogues (tiles of two objects per lane, many work items per
    tile, both regions of the end-of-launch split) under random grid, tile
    and workgroup shapes, every form and the weight gradient, against the
    oracle."""
    bt, b, A, B, (cd, tile, zd, nb) = _large_config(seed)
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu, b.pbin if bt == F.BIN_SPI else None)
    bs = b.bsize_rescaled
    print(f"case random large seed={seed} bt={bt} tab={b.tabtype} ns={b.ns} nA={A.n} nB={B.n} "
          f"shape={(cd, tile, zd, nb)}", flush=True)
    ref_a, cen_a = O.count(0, bs, (A.x, A.y, A.z, None, None), None, True, ob, variant=1,
                           census_mask=0b11111)
    ref_x, cen_x = O.count(0, bs, (A.x, A.y, A.z, None, None), (B.x, B.y, B.z, None, None),
                           False, ob, variant=1, census_mask=0b11111)
    ref_xw, _ = O.count(0, bs, (A.x, A.y, A.z, None, A.w), (B.x, B.y, B.z, None, None), False,
                        ob, variant=1, weighted=True)
    try:
        eng.set_tuning(cd, tile, zd, nb)
        for prec in ("fp64", "fp32", "plain"):
            g = eng.count_pairs(A, None, b, isauto=True, precision=prec)
            assert np.abs(g - ref_a).sum() <= 2 * cen_a
            if cen_a == 0:
                np.testing.assert_array_equal(g, ref_a)
            g = eng.count_pairs(A, B, b, precision=prec)
            assert np.abs(g - ref_x).sum() <= 2 * cen_x
            if cen_x == 0:
                np.testing.assert_array_equal(g, ref_x)
            g = eng.count_pairs(A, B, b, withwt=True, precision=prec)
            scale = np.maximum(np.abs(ref_xw), 1e-300)
            assert np.max(np.abs(g - ref_xw) / scale) <= WT_RTOL
        if cen_a == 0 and cen_x == 0:
            # weight gradients of A (auto and cross, A's catalogue first
            # and second) under the same shapes
            G = np.random.default_rng(100 + seed).normal(size=b.ntot)
            ra, _ = O.weight_grad(0, bs, (A.x, A.y, A.z, None, A.w), None, True, ob, G)
            rx, _ = O.weight_grad(0, bs, (A.x, A.y, A.z, None, A.w), (B.x, B.y, B.z, None, None),
                                  False, ob, G)
            for g, r in ((eng.count_backward(A, None, b, G, isauto=True)[0], ra),
                         (eng.count_backward(A, B, b, G, need_b=False)[0], rx),
                         (eng.count_backward(B, A, b, G)[1], rx)):
                np.testing.assert_allclose(g, r, rtol=1e-6, atol=1e-9 * np.abs(r).max())
    finally:
        eng.set_tuning(0, 0, 0, 0)


def _large_survey_config(seed):
    """A larger survey configuration: binning type, integer or hybrid edges,
    catalogue sizes and a random grid, tile and workgroup shape."""
    rng = np.random.default_rng(7000 + seed)
    bt = (F.BIN_ISO, F.BIN_SMU, F.BIN_SPI)[seed % 3]
    if seed % 2 == 0:                 # integer-friendly edges; odd seeds: hybrid tables
        ds = float(rng.choice([2.0, 4.0, 5.0]))
        smin = float(rng.choice([0.0, ds]))
    else:
        ds = float(np.round(rng.uniform(1.3, 6.0), 2))
        smin = float(np.round(rng.uniform(0.0, 2.0 * ds), 2))
    ns = max(1, int((min(smin + int(rng.integers(5, 41)) * ds, 150.0) - smin) // ds))
    nmu = int(rng.choice([5, 40, 100])) if bt == F.BIN_SMU else 1
    dpi, pmax = (5.0, 5.0 * int(rng.integers(4, 21))) if bt == F.BIN_SPI else (1.0, 0.0)
    b = F.Bins.linear(F.GEOM_SURVEY, bt, smin, smin + ns * ds, ds, nmu=nmu, pmin=0.0,
                      pmax=pmax, dpi=dpi)
    x, y, z, w = F.gen_survey(int(rng.integers(1 << 30)), int(rng.integers(30000, 80000)))
    A = F.Catalog(b, x, y, z, w)
    x, y, z, w = F.gen_survey(int(rng.integers(1 << 30)), int(rng.integers(30000, 100000)))
    B = F.Catalog(b, x, y, z, w)
    shape = (int(rng.choice([2, 3, 4, 6])), int(rng.choice([256, 512, 1024, 2048])),
             int(rng.choice([1, 4, 16])), int(rng.choice([0, 7, 64, 300])))
    return bt, b, A, B, shape


@pytest.mark.parametrize("seed", range(6))
def test_random_large_survey(eng, seed):
    """Larger survey catalogues under random grid, tile (up to four objects
    per lane for weighted FP64 counts) and workgroup shapes, every form,
    auto and cross counts, against the oracle."""
    bt, b, A, B, (cd, tile, zd, nb) = _large_survey_config(seed)
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu, b.pbin if bt == F.BIN_SPI else None)
    print(f"case random large survey seed={seed} bt={bt} tab={b.tabtype} ns={b.ns} "
          f"nmu={b.nmu} nA={A.n} nB={B.n} shape={(cd, tile, zd, nb)}", flush=True)
    tA = (A.x, A.y, A.z, A.s, None)
    tB = (B.x, B.y, B.z, B.s, None)
    ref_a, cen_a = O.count(1, None, tA, None, True, ob, variant=0, census_mask=0b111111)
    ref_x, cen_x = O.count(1, None, tA, tB, False, ob, variant=0, census_mask=0b111111)
    wA = (A.x, A.y, A.z, A.s, A.w)
    wB = (B.x, B.y, B.z, B.s, B.w)
    ref_aw, _ = O.count(1, None, wA, None, True, ob, variant=0, weighted=True)
    ref_xw, _ = O.count(1, None, wA, wB, False, ob, variant=0, weighted=True)

    def close(g, r):
        scale = np.maximum(np.abs(r), 1e-300)
        assert np.max(np.abs(g - r) / scale) <= WT_RTOL

    try:
        eng.set_tuning(cd, tile, zd, nb)
        for prec in ("fp64", "fp32", "plain"):
            for g, r, cen in ((eng.count_pairs(A, None, b, isauto=True, precision=prec),
                               ref_a, cen_a),
                              (eng.count_pairs(A, B, b, precision=prec), ref_x, cen_x)):
                assert np.abs(g.astype(np.int64) - r.astype(np.int64)).sum() <= 2 * cen
                if cen == 0:
                    np.testing.assert_array_equal(g, r)
            close(eng.count_pairs(A, None, b, isauto=True, withwt=True, precision=prec), ref_aw)
            close(eng.count_pairs(A, B, b, withwt=True, precision=prec), ref_xw)
    finally:
        eng.set_tuning(0, 0, 0, 0)
