"""GPU parity: the HIP kernels against the CPU oracle on seeded inputs.

Every pair is evaluated on one rounding path of the reference (the
vectorised-B block, src/fcfc/2pt_box/metric_common.c:1154-1243, with the
larger catalogue of a cross count as "a" -- the other orientation is the
reference's a-side block).  Counts are compared against the oracle's
variant 1 together with its ambiguity census (the number of pairs whose bin
depends on the reference's rounding path): with census 0 they must be
identical.  Weighted counts: <= 1e-12 relative.
"""
import numpy as np
import pytest

import fcfc_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L = 1000.0
WT_RTOL = 1e-12


@pytest.fixture(scope="module")
def eng():
    with F.Engine() as e:
        yield e


def _bins(geom, bt, smax=200.0, ds=5.0, nmu=100, pmax=200.0, dpi=5.0, smin=0.0, bsize=L):
    return F.Bins.linear(geom, bt, smin, smax, ds, nmu=nmu if bt == F.BIN_SMU else 1,
                         pmin=0.0, pmax=pmax if bt == F.BIN_SPI else 0.0,
                         dpi=dpi, bsize=bsize if geom == F.GEOM_BOX else None)


def _obins(b):
    return O.oracle_bins(b.bintype, b.s2bin, b.nmu, b.pbin if b.bintype == F.BIN_SPI else None)


def _otuple(c):
    return (c.x, c.y, c.z, c.s, c.w)


def _cmp(gpu, ref, weighted, census=0):
    if weighted:
        scale = np.maximum(np.abs(ref), 1e-300)
        rel = np.max(np.abs(gpu - ref) / scale) if len(ref) else 0
        assert rel <= WT_RTOL, f"weighted counts differ: max rel {rel}"
    else:
        diff = np.abs(gpu.astype(np.int64) - ref.astype(np.int64)).sum()
        assert diff <= 2 * census, f"counts differ by {diff} (census {census})"
        if census == 0:
            np.testing.assert_array_equal(gpu, ref)


@pytest.fixture(scope="module")
def cats():
    x, y, z, w = F.gen_clustered(11, 3000, 30, 40, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(12, 6000, L)
    return (x, y, z, w), (xr, yr, zr)


@pytest.mark.parametrize("bt", [F.BIN_ISO, F.BIN_SMU, F.BIN_SPI])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "plain"])
def test_box_cross(eng, cats, bt, prec):
    (x, y, z, w), (xr, yr, zr) = cats
    b = _bins(F.GEOM_BOX, bt)
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr)
    print(f"case box cross bt={bt} prec={prec}", flush=True)
    g = eng.count_pairs(d, r, b, isauto=False, withwt=True, precision=prec)
    ref, _ = O.count(0, b.bsize_rescaled, _otuple(d), (r.x, r.y, r.z, None, None), False,
                     _obins(b), variant=1, weighted=True)
    _cmp(g, ref, True)
    gu = eng.count_pairs(d, r, b, isauto=False, withwt=False, precision=prec)
    refu, cen = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None),
                        (r.x, r.y, r.z, None, None), False, _obins(b), variant=1,
                        census_mask=0b11111)
    _cmp(gu, refu, False, cen)
    # cross counts do not depend on the order of the two catalogues
    np.testing.assert_array_equal(eng.count_pairs(r, d, b, isauto=False, withwt=False,
                                                  precision=prec), gu)


@pytest.mark.parametrize("bt", [F.BIN_ISO, F.BIN_SMU, F.BIN_SPI])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "plain"])
def test_box_auto(eng, cats, bt, prec):
    (x, y, z, w), _ = cats
    b = _bins(F.GEOM_BOX, bt)
    d = F.Catalog(b, x, y, z, w)
    print(f"case box auto bt={bt} prec={prec}", flush=True)
    ref, cen = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None), None, True, _obins(b),
                       variant=1, census_mask=0b11111)
    g = eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec)
    _cmp(g, ref, False, cen)
    refw, _ = O.count(0, b.bsize_rescaled, _otuple(d), None, True, _obins(b), variant=1,
                      weighted=True)
    gw = eng.count_pairs(d, None, b, isauto=True, withwt=True, precision=prec)
    _cmp(gw, refw, True)


@pytest.mark.parametrize("bt", [F.BIN_ISO, F.BIN_SMU, F.BIN_SPI])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "plain"])
def test_survey(eng, bt, prec):
    x, y, z, w = F.gen_survey(21, 4000)
    xr, yr, zr, wr = F.gen_survey(22, 6000)
    b = _bins(F.GEOM_SURVEY, bt)
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr, wr)
    print(f"case survey bt={bt} prec={prec}", flush=True)
    ob = _obins(b)
    refa, cen = O.count(1, None, (d.x, d.y, d.z, d.s, None), None, True, ob, variant=0,
                        census_mask=0b111111)
    ga = eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec)
    st = eng.stats()
    _cmp(ga, refa, False, cen)
    if prec == "fp32" and bt == F.BIN_SMU:
        # the FP32 kernel ran: pairs below its s^2 floor are re-evaluated
        assert st["rechecked"] > 0
    refx, _ = O.count(1, None, _otuple(d), _otuple(r), False, ob, variant=0, weighted=True)
    gx = eng.count_pairs(d, r, b, isauto=False, withwt=True, precision=prec)
    _cmp(gx, refx, True)


@pytest.mark.parametrize("bt", [F.BIN_ISO, F.BIN_SMU, F.BIN_SPI])
def test_survey_near_observer(eng, bt):
    """Objects around the observer: many pairs have both objects on opposite
    sides of it (|a + b| small), where the reference's pi and mu lose their
    accuracy and the FP32 form re-evaluates every pair in FP64."""
    rng = np.random.default_rng(31)
    n = 3000
    v = rng.normal(size=(n, 3))
    v *= (rng.uniform(0, 1, n) ** (1 / 3) * 120.0 / np.linalg.norm(v, axis=1))[:, None]
    x, y, z = v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()
    b = _bins(F.GEOM_SURVEY, bt)
    d = F.Catalog(b, x, y, z)
    ob = _obins(b)
    ref, cen = O.count(1, None, (d.x, d.y, d.z, d.s, None), None, True, ob, variant=0,
                       census_mask=0b111111)
    for prec in ("fp64", "fp32"):
        print(f"case survey near observer bt={bt} prec={prec} census={cen}", flush=True)
        _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec), ref, False, cen)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "plain"])
def test_survey_lattice_edges(eng, prec):
    """Survey (s_perp, pi) on an integer lattice (in rescaled units) placed
    symmetrically about the line of sight, far from the observer: mirror
    pairs (|a| = |b|) have pi^2 = 0 and an integer s_perp^2, pairs on the
    axis an integer pi^2, so about 2e4 pairs sit on integers, many of them
    bin edges (squares), where the fixed-point lookups of the FP64 form must
    defer to the division and the FP32 windows to FP64.  Every reference
    variant agrees here (census 0): each form equals the oracle exactly."""
    b = _bins(F.GEOM_SURVEY, F.BIN_SPI)
    f = b.rescale
    g = np.arange(-6, 7) / f
    zg = 1200.0 + np.arange(0, 13) / f
    x, y, z = [a.ravel().copy() for a in np.meshgrid(g, g, zg, indexing="ij")]
    d = F.Catalog(b, x, y, z)
    assert np.all(d.x == np.round(d.x)) and np.all(d.z == np.round(d.z))
    ref, cen = O.count(1, None, (d.x, d.y, d.z, d.s, None), None, True, _obins(b), variant=0,
                       census_mask=0b111111)
    assert cen == 0
    print(f"case survey lattice spi prec={prec}", flush=True)
    np.testing.assert_array_equal(eng.count_pairs(d, None, b, isauto=True, withwt=False,
                                                  precision=prec), ref)


@pytest.mark.parametrize("geom", [F.GEOM_BOX, F.GEOM_SURVEY])
def test_non_finite_coordinates_rejected(eng, cats, geom):
    """A catalogue with a NaN or infinite coordinate has no cell and no bin:
    the grid build flags it and the count fails with FCFC_ERR_DATA before
    any count kernel runs; the engine then counts valid catalogues as
    before."""
    (x, y, z, w), _ = cats
    b = _bins(geom, F.BIN_SMU)
    if geom == F.GEOM_SURVEY:
        x, y, z, w = F.gen_survey(65, 3000)
    good = F.Catalog(b, x, y, z, w)
    ref = eng.count_pairs(good, None, b, isauto=True, withwt=False)
    for bad_value in (np.nan, np.inf):
        xb = x.copy()
        xb[len(xb) // 2] = bad_value
        print(f"case non-finite geom={geom} value={bad_value}", flush=True)
        with pytest.raises(F.FcfcError, match="non-finite") as ei:
            eng.count_pairs(F.Catalog(b, xb, y, z, w), None, b, isauto=True, withwt=False)
        assert ei.value.code == -21
        np.testing.assert_array_equal(eng.count_pairs(good, None, b, isauto=True, withwt=False), ref)


@pytest.mark.parametrize("geom", [F.GEOM_BOX, F.GEOM_SURVEY])
def test_four_objects_per_lane(eng, geom):
    """Tiles of 2048 objects set by the caller: the weighted FP64 box
    counts with integer tables hold four A objects per lane, every other
    count (the survey ones included) cuts such tiles to 1024.  Dense halos (cells of thousands of objects,
    two cells per separation) fill the tiles; every count equals the
    oracle."""
    if geom == F.GEOM_BOX:
        x, y, z, w = F.gen_clustered(81, 1000, 4, 3000, 1.0, L, 0.3)
        xr, yr, zr = F.gen_uniform(82, 4000, L)
        wr = np.linspace(0.7, 1.3, len(xr))
        bts, g0, var = (F.BIN_SMU, F.BIN_SPI, F.BIN_ISO), 0, 1
    else:
        x, y, z, w = F.gen_survey(83, 6000)
        xr, yr, zr, wr = F.gen_survey(84, 8000)
        bts, g0, var = (F.BIN_SPI, F.BIN_SMU), 1, 0
    try:
        eng.set_tuning(cell_div=2, tile=2048, zdiv=1)
        for bt in bts:
            b = _bins(geom, bt)
            d = F.Catalog(b, x, y, z, w)
            r = F.Catalog(b, xr, yr, zr, wr)
            ob = _obins(b)
            bz = b.bsize_rescaled if geom == F.GEOM_BOX else None
            print(f"case four per lane geom={geom} bt={bt}", flush=True)
            refa, _ = O.count(g0, bz, _otuple(d), None, True, ob, variant=var, weighted=True)
            got = eng.count_pairs(d, None, b, isauto=True, withwt=True)
            assert eng.stats()["tile"] == (2048 if geom == F.GEOM_BOX else 1024)
            _cmp(got, refa, True)
            refx, _ = O.count(g0, bz, _otuple(d), _otuple(r), False, ob, variant=var, weighted=True)
            _cmp(eng.count_pairs(d, r, b, isauto=False, withwt=True), refx, True)
            for prec in ("fp32", "plain"):
                got = eng.count_pairs(d, r, b, isauto=False, withwt=True, precision=prec)
                assert eng.stats()["tile"] == 1024, prec
                _cmp(got, refx, True)
            refu, cen = O.count(g0, bz, (d.x, d.y, d.z, d.s, None), None, True, ob, variant=var,
                                census_mask=0b111111)
            got = eng.count_pairs(d, None, b, isauto=True, withwt=False)
            assert eng.stats()["tile"] == 1024
            _cmp(got, refu, False, cen)
    finally:
        eng.set_tuning()


@pytest.mark.parametrize("bt", [F.BIN_ISO, F.BIN_SMU, F.BIN_SPI])
def test_survey_hybrid(eng, bt):
    """Survey geometry with hybrid (non-integer) tables, every form."""
    x, y, z, w = F.gen_survey(23, 3000)
    xr, yr, zr, wr = F.gen_survey(24, 4000)
    b = F.Bins.linear(F.GEOM_SURVEY, bt, 0.5, 146.5, 7.3, nmu=7 if bt == F.BIN_SMU else 1,
                      pmin=1.5, pmax=96.5, dpi=3.7)
    assert b.tabtype == 1
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr, wr)
    ob = _obins(b)
    refa, cen = O.count(1, None, (d.x, d.y, d.z, d.s, None), None, True, ob, variant=0,
                        census_mask=0b111111)
    refx, _ = O.count(1, None, _otuple(d), _otuple(r), False, ob, variant=0, weighted=True)
    for prec in ("fp64", "fp32", "plain"):
        print(f"case survey hybrid bt={bt} prec={prec}", flush=True)
        _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec), refa, False, cen)
        _cmp(eng.count_pairs(d, r, b, isauto=False, withwt=True, precision=prec), refx, True)


@pytest.mark.parametrize("nmu", [100, 7, 12])
def test_lattice_mu_edges(eng, nmu):
    """Objects on an integer lattice (in rescaled units): many pairs sit
    exactly on mu and s edges (e.g. (3, 0, 4): mu = 4/5), where the fast mu
    bin must fall back to its exact test and the FP32 windows to FP64."""
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=nmu, bsize=L)
    f = b.rescale
    g = np.arange(14) / f + 100.0
    x, y, z = [a.ravel().copy() for a in np.meshgrid(g, g, g, indexing="ij")]
    d = F.Catalog(b, x, y, z)
    assert np.all(d.x == np.round(d.x))
    # exact coordinates: every path of the GPU is the oracle's variant 1 bit
    # for bit (the reference's own unfused scalar path, which divides before
    # multiplying by nmu^2, disagrees on some of these pairs: census > 0)
    ref, _ = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None), None, True, _obins(b),
                     variant=1)
    for prec in ("fp64", "fp32", "plain"):
        print(f"case lattice nmu={nmu} prec={prec}", flush=True)
        gl = eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec)
        np.testing.assert_array_equal(gl, ref)


@pytest.mark.parametrize("bt", [F.BIN_ISO, F.BIN_SMU, F.BIN_SPI])
def test_hybrid_bins(eng, cats, bt):
    """Non-integer bin edges take FCFC's hybrid lookup table route."""
    (x, y, z, w), (xr, yr, zr) = cats
    b = F.Bins.linear(F.GEOM_BOX, bt, 0.5, 146.5, 7.3, nmu=7 if bt == F.BIN_SMU else 1,
                      pmin=1.5, pmax=96.5, dpi=3.7, bsize=L)
    assert b.tabtype == 1
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr)
    ref, cen = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None),
                       (r.x, r.y, r.z, None, None), False, _obins(b), variant=1,
                       census_mask=0b11111)
    refw, _ = O.count(0, b.bsize_rescaled, _otuple(d), (r.x, r.y, r.z, None, None), False,
                      _obins(b), variant=1, weighted=True)
    for prec in ("fp64", "fp32", "plain"):
        print(f"case hybrid bt={bt} prec={prec}", flush=True)
        g = eng.count_pairs(d, r, b, isauto=False, withwt=False, precision=prec)
        _cmp(g, ref, False, cen)
        _cmp(eng.count_pairs(d, r, b, isauto=False, withwt=True, precision=prec), refw, True)


@pytest.mark.parametrize("geom", [F.GEOM_BOX, F.GEOM_SURVEY])
def test_coincident_points(eng, geom):
    """Duplicated objects (s = 0: the reference puts such pairs in the first
    mu bin, metric_common.c:184) and near-duplicates 1e-9 apart."""
    if geom == F.GEOM_BOX:
        x, y, z, w = F.gen_clustered(71, 1500, 10, 20, 2.0, L, 0.3)
    else:
        x, y, z, w = F.gen_survey(72, 2000)
    n = len(x)
    k = np.arange(0, n, 7)
    x = np.concatenate([x, x[k], x[k[::2]] + 1e-9])
    y = np.concatenate([y, y[k], y[k[::2]]])
    z = np.concatenate([z, z[k], z[k[::2]]])
    w = np.concatenate([w, w[k], w[k[::2]]])
    for bt in (F.BIN_SMU, F.BIN_SPI):
        b = _bins(geom, bt)
        d = F.Catalog(b, x, y, z, w)
        ob = _obins(b)
        g0 = 0 if geom == F.GEOM_BOX else 1
        var = 1 if geom == F.GEOM_BOX else 0
        ref, cen = O.count(g0, b.bsize_rescaled if geom == F.GEOM_BOX else None,
                           (d.x, d.y, d.z, d.s, None), None, True, ob, variant=var,
                           census_mask=0b111111)
        for prec in ("fp64", "fp32", "plain"):
            print(f"case coincident geom={geom} bt={bt} prec={prec}", flush=True)
            _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec), ref, False, cen)


@pytest.mark.parametrize("bt", [F.BIN_ISO, F.BIN_SMU, F.BIN_SPI])
def test_non_cubic_box(eng, bt):
    """A periodic box with three different side lengths (per-axis wrap)."""
    Lb = np.array([1000.0, 700.0, 450.0])
    rng = np.random.default_rng(61)
    x, y, z = (rng.random(6000) * Lb[k] for k in range(3))
    w = rng.lognormal(0, 0.3, 6000)
    b = F.Bins.linear(F.GEOM_BOX, bt, 0, 150, 5, nmu=20 if bt == F.BIN_SMU else 1,
                      pmin=0, pmax=150 if bt == F.BIN_SPI else 0, dpi=5, bsize=Lb)
    d = F.Catalog(b, x, y, z, w)
    ref, cen = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None), None, True, _obins(b),
                       variant=1, census_mask=0b11111)
    refw, _ = O.count(0, b.bsize_rescaled, _otuple(d), None, True, _obins(b), variant=1,
                      weighted=True)
    for prec in ("fp64", "fp32", "plain"):
        print(f"case non-cubic box bt={bt} prec={prec}", flush=True)
        _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec), ref, False, cen)
        _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=True, precision=prec), refw, True)


@pytest.mark.parametrize("bt", [F.BIN_SMU, F.BIN_SPI])
def test_large_box_small_bins(eng, bt):
    """A box of 20000 Mpc/h with separations up to 20 Mpc/h (coordinates
    three orders of magnitude above the bins): FP32 windows are wide
    relative to the bins there, more pairs are re-evaluated, counts stay
    exact."""
    Lb = 20000.0
    x, y, z, w = F.gen_clustered(51, 3000, 300, 10, 1.0, Lb, 0.3)
    b = F.Bins.linear(F.GEOM_BOX, bt, 0, 20, 0.5, nmu=50 if bt == F.BIN_SMU else 1,
                      pmin=0, pmax=20 if bt == F.BIN_SPI else 0, dpi=0.5, bsize=Lb)
    d = F.Catalog(b, x, y, z, w)
    ref, cen = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None), None, True, _obins(b),
                       variant=1, census_mask=0b11111)
    refw, _ = O.count(0, b.bsize_rescaled, _otuple(d), None, True, _obins(b), variant=1,
                      weighted=True)
    for prec in ("fp64", "fp32", "plain"):
        print(f"case large box bt={bt} prec={prec}", flush=True)
        _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec), ref, False, cen)
        _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=True, precision=prec), refw, True)


@pytest.mark.parametrize("nmu", [255, 127, 126, 3])
def test_mu_bins_many_and_few(eng, cats, nmu):
    """The finest mu binning the reference allows (MAX_NMU = 255,
    define_comm.h:148) and a coarse one: the mu windows scale with nmu.
    126 is the most mu bins the FP32 form's fixed-point mu value holds (it
    re-evaluates some pairs in FP64); with more the FP32 request runs the
    FP64 form (nothing re-evaluated), with the same counts."""
    (x, y, z, w), (xr, yr, zr) = cats
    b = _bins(F.GEOM_BOX, F.BIN_SMU, nmu=nmu)
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr)
    ref, cen = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None),
                       (r.x, r.y, r.z, None, None), False, _obins(b), variant=1,
                       census_mask=0b11111)
    refa, cena = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None), None, True,
                         _obins(b), variant=1, census_mask=0b11111)
    for prec in ("fp64", "fp32", "plain"):
        print(f"case nmu={nmu} prec={prec}", flush=True)
        _cmp(eng.count_pairs(d, r, b, isauto=False, withwt=False, precision=prec), ref, False, cen)
        if prec == "fp32":
            assert (eng.stats()["rechecked"] > 0) == (nmu <= 126)
        _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec), refa, False, cena)


@pytest.mark.parametrize("nmu", [255, 3])
def test_survey_mu_bins_many_and_few(eng, nmu):
    """Survey (s, mu) with the finest and a coarse mu binning: the capped
    fixed-point mu value of the survey form (K at most nmu + 1) at
    nmu = 255, where the FP32 request runs the FP64 form, and at nmu = 3."""
    x, y, z, w = F.gen_survey(63, 4000)
    xr, yr, zr, wr = F.gen_survey(64, 6000)
    b = _bins(F.GEOM_SURVEY, F.BIN_SMU, nmu=nmu)
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr, wr)
    ob = _obins(b)
    refa, cena = O.count(1, None, (d.x, d.y, d.z, d.s, None), None, True, ob, variant=0,
                         census_mask=0b111111)
    refx, _ = O.count(1, None, _otuple(d), _otuple(r), False, ob, variant=0, weighted=True)
    for prec in ("fp64", "fp32", "plain"):
        print(f"case survey nmu={nmu} prec={prec}", flush=True)
        _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec), refa, False, cena)
        if prec == "fp32":
            assert (eng.stats()["rechecked"] > 0) == (nmu <= 126)
        _cmp(eng.count_pairs(d, r, b, isauto=False, withwt=True, precision=prec), refx, True)


@pytest.mark.parametrize("bt,smin,smax,ds,pmin,pmax,dpi,soff", [
    (F.BIN_SMU, 20.0, 200.0, 5.0, 0.0, 0.0, 1.0, 16),       # zero-based table, 16 cells below
    (F.BIN_SPI, 20.0, 200.0, 5.0, 30.0, 150.0, 10.0, 16),
    (F.BIN_SMU, 70.0, 100.0, 1.0, 0.0, 0.0, 1.0, 4900),     # far from zero: hybrid lookups
    (F.BIN_SPI, 70.0, 100.0, 1.0, 80.0, 90.0, 1.0, 4900),
])
def test_offset_tables(eng, cats, bt, smin, smax, ds, pmin, pmax, dpi, soff):
    """Integer tables whose range starts above zero: near zero they are
    staged zero-based (cells below s2min are out of range), far from it the
    hybrid lookups with explicit range checks take over; both give the
    integer table's bins."""
    (x, y, z, w), (xr, yr, zr) = cats
    b = F.Bins.linear(F.GEOM_BOX, bt, smin, smax, ds, nmu=20 if bt == F.BIN_SMU else 1,
                      pmin=pmin, pmax=pmax, dpi=dpi, bsize=L)
    assert b.tabtype == 0 and int(b.s2bin[0]) == soff
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr)
    ref, cen = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None),
                       (r.x, r.y, r.z, None, None), False, _obins(b), variant=1,
                       census_mask=0b11111)
    refa, cena = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None), None, True,
                         _obins(b), variant=1, census_mask=0b11111)
    for prec in ("fp64", "fp32", "plain"):
        print(f"case offset table bt={bt} smin={smin} prec={prec}", flush=True)
        g = eng.count_pairs(d, r, b, isauto=False, withwt=False, precision=prec)
        _cmp(g, ref, False, cen)
        ga = eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec)
        _cmp(ga, refa, False, cena)


@pytest.mark.parametrize("bt,smin,pmin,soff,poff", [
    (F.BIN_SMU, 20.0, 0.0, 16, 0),
    (F.BIN_SPI, 20.0, 30.0, 16, 36),
])
def test_offset_tables_survey(eng, bt, smin, pmin, soff, poff):
    """Survey geometry with integer tables whose range starts above zero
    (s_perp from 20, pi from 30 Mpc/h): the cells below the first edge are
    out of range in the zero-based tables of every form, and in the
    fixed-point (s_perp, pi) lookups of the FP64 form, whose tables start
    at floor cell -2."""
    x, y, z, w = F.gen_survey(61, 4000)
    xr, yr, zr, wr = F.gen_survey(62, 6000)
    b = F.Bins.linear(F.GEOM_SURVEY, bt, smin, 200.0, 5.0, nmu=20 if bt == F.BIN_SMU else 1,
                      pmin=pmin, pmax=150.0 if bt == F.BIN_SPI else 0.0, dpi=10.0)
    assert b.tabtype == 0 and int(b.s2bin[0]) == soff
    if bt == F.BIN_SPI:
        assert int(b.pbin[0]) == poff
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr, wr)
    ob = _obins(b)
    refa, cena = O.count(1, None, (d.x, d.y, d.z, d.s, None), None, True, ob, variant=0,
                         census_mask=0b111111)
    refx, _ = O.count(1, None, _otuple(d), _otuple(r), False, ob, variant=0, weighted=True)
    for prec in ("fp64", "fp32", "plain"):
        print(f"case survey offset table bt={bt} prec={prec}", flush=True)
        _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec), refa, False, cena)
        _cmp(eng.count_pairs(d, r, b, isauto=False, withwt=True, precision=prec), refx, True)


def test_launch_shapes_identical(eng, cats):
    """Integer counts do not depend on the grid / tile / launch shape."""
    (x, y, z, w), (xr, yr, zr) = cats
    b = _bins(F.GEOM_BOX, F.BIN_SMU)
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr)
    res = []
    for cd, tile, zd, nb in [(0, 0, 0, 0), (1, 64, 1, 3), (3, 200, 7, 17), (6, 700, 2, 0), (4, 1024, 4, 0)]:
        eng.set_tuning(cd, tile, zd, nb)
        res.append(eng.count_pairs(d, r, b, isauto=False, withwt=False))
        res.append(eng.count_pairs(d, None, b, isauto=True, withwt=False))
    eng.set_tuning(0, 0, 0, 0)
    for k in range(2, len(res)):
        np.testing.assert_array_equal(res[k], res[k % 2])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_work_item_split(eng, cats, prec):
    """Few persistent workgroups and small tiles, so that the tile list is cut
    into coarse work items and a finer tail at many different boundaries
    (with one workgroup the finer tail is the last two tiles): weighted
    cross, unweighted auto and the weight gradient against the oracle."""
    (x, y, z, w), (xr, yr, zr) = cats
    b = _bins(F.GEOM_BOX, F.BIN_SMU)
    d = F.Catalog(b, x, y, z, w)
    r = F.Catalog(b, xr, yr, zr)
    ob = _obins(b)
    refw, _ = O.count(0, b.bsize_rescaled, _otuple(d), (r.x, r.y, r.z, None, None), False,
                      ob, variant=1, weighted=True)
    refu, cen = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None), None, True, ob,
                        variant=1, census_mask=0b11111)
    G = np.random.default_rng(7).normal(size=b.ntot)
    refg, _ = O.weight_grad(0, b.bsize_rescaled, _otuple(d), None, True, ob, G)
    try:
        for nb in (1, 2, 5, 33, 160):
            eng.set_tuning(3, 64, 4, nb)
            print(f"case work items prec={prec} nblock={nb}", flush=True)
            _cmp(eng.count_pairs(d, r, b, isauto=False, withwt=True, precision=prec), refw, True)
            _cmp(eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec),
                 refu, False, cen)
            if prec == "fp64":
                ga, _ = eng.count_backward(d, None, b, G, isauto=True)
                np.testing.assert_allclose(ga, refg, rtol=1e-6, atol=1e-9 * np.abs(refg).max())
    finally:
        eng.set_tuning(0, 0, 0, 0)


def test_empty_and_tiny(eng):
    b = _bins(F.GEOM_BOX, F.BIN_ISO)
    e = F.Catalog(b, np.zeros(0), np.zeros(0), np.zeros(0))
    one = F.Catalog(b, np.array([1.0]), np.array([2.0]), np.array([3.0]))
    assert eng.count_pairs(e, None, b, isauto=True).sum() == 0
    assert eng.count_pairs(one, None, b, isauto=True).sum() == 0
    assert eng.count_pairs(one, e, b).sum() == 0
    # two identical points: s = 0 lands in bin 0 (metric_common.c:184)
    two = F.Catalog(b, np.array([5.0, 5.0]), np.array([5.0, 5.0]), np.array([5.0, 5.0]))
    c = eng.count_pairs(two, None, b, isauto=True)
    assert c[0] == 1 and c.sum() == 1
    # a pair across the periodic boundary
    pb = F.Catalog(b, np.array([1.0, 999.0]), np.array([0.5, 0.5]), np.array([0.5, 0.5]))
    c = eng.count_pairs(pb, None, b, isauto=True)
    assert c[0] == 1
