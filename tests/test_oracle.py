"""Pins the CPU oracle (oracle/fcfc_oracle.c) against the upstream FCFC
binaries' outputs stored in tests/golden/, and against independent checks:
a numpy brute force, closed-form totals and counting invariants.

Reference flavours: "avx512" = OpenMP + AVX-512 build, whose blocks follow
the vectorised paths (oracle variants 0/1/2, FMA, RZ mu division); "scalar" =
serial build without SIMD (oracle variant 3/4, unfused).  Unweighted counts
must be identical whenever the census (pairs whose bin depends on the
rounding path) is zero; weighted counts agree to 1e-12 relative.
"""
import glob
import json
import os

import numpy as np
import pytest

import fcfc_amd as F
from oracle import oracle as O
from tests.golden.make_golden import catalogs, checksum

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# cfg_*: the cases at the configurations' own sizes (tests/test_gpu_sizes.py
# and the size tests below); cli_*: output files of the executables
# (tests/test_gpu_cli.py)
CASES = sorted(c for c in glob.glob(os.path.join(GOLD, "*.json"))
               if not os.path.basename(c).startswith(("cfg_", "cli_")))


def _bins(code, bk, L=1000.0):
    geom = F.GEOM_BOX if code == "FCFC_2PT_BOX" else F.GEOM_SURVEY
    return F.Bins.linear(geom, bk["bintype"], bk["smin"], bk["smax"], bk["ds"],
                         nmu=bk.get("nmu", 1), pmin=bk.get("pmin", 0.0),
                         pmax=bk.get("pmax", 0.0), dpi=bk.get("dpi", 1.0),
                         bsize=L if geom == F.GEOM_BOX else None)


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(c)[:-5] for c in CASES])
def test_oracle_matches_reference(path):
    _check_case(path)


@pytest.mark.parametrize("name", ["cfg_c1_full", "cfg_c2_twentieth"])
def test_oracle_at_config_sizes(name):
    """The oracle against the reference at the configurations' sizes: C1 as
    specified (AVX-512 and serial scalar builds) and the DD count of C2 at
    1/20 size (its DR and RR are left to the GPU tests, which compare the
    GPU with the reference's files directly; the oracle would need minutes
    for them here)."""
    g = json.load(open(os.path.join(GOLD, name + ".json")))
    assert g["census"]["DD"] == 0
    _check_case(os.path.join(GOLD, name + ".json"), only=("DD",))


def _check_case(path, only=None):
    g = json.load(open(path))
    cats, wts = catalogs(g["catalogs"])
    assert [checksum(*c) for c in cats] == g["checksums"], "generator output changed"
    b = _bins(g["code"], g["binning"])
    geom = 0 if g["code"] == "FCFC_2PT_BOX" else 1
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu, b.pbin if b.bintype == 2 else None)
    prepped = []
    for (c, wt) in zip(cats, wts):
        k = F.Catalog(b, c[0], c[1], c[2], c[3] if (wt and len(c) > 3) else None)
        prepped.append(k)
    for key, ref in g["results"].items():
        flav, pair = key.split(":")
        if only and pair not in only:
            continue
        A = prepped["DR".index(pair[0])]
        B = prepped["DR".index(pair[1])]
        isauto = pair[0] == pair[1]
        weighted = ref["spec"][4] == 1
        var = 1 if flav == "avx512" else 3
        ta = (A.x, A.y, A.z, A.s, A.w if weighted else None)
        tb = (B.x, B.y, B.z, B.s, B.w if weighted else None)
        cnt, cen = O.count(geom, b.bsize_rescaled, ta, None if isauto else tb, isauto, ob,
                           variant=var, census_mask=0b11111, weighted=weighted)
        rc = np.array(ref["cnt"])
        mine = cnt * 2 if isauto else cnt
        if weighted:
            rel = np.max(np.abs(mine - rc) / np.maximum(np.abs(rc), 1e-300))
            assert rel < 1e-12, f"{key}: weighted rel diff {rel}"
        else:
            if cen == 0:
                np.testing.assert_array_equal(mine, rc.astype(np.int64), err_msg=key)
            else:
                assert np.abs(mine - rc).sum() <= 2 * (2 if isauto else 1) * cen, key
        # the reference writes the RESCALED box size (save_res.c:213-217)
        if "bsize" in ref:
            np.testing.assert_allclose(ref["bsize"], b.bsize_rescaled, rtol=0, atol=0)
        np.testing.assert_array_equal(ref["sbin"], b.sbin_raw)
        # normalisation of eval_pairs (eval_cf.c:128-149)
        if isauto:
            n = A.sumw if weighted else A.n
            assert ref["norm"] == pytest.approx(n * (n - 1), rel=1e-14)
        else:
            n1, n2 = (A.sumw, B.sumw) if weighted else (A.n, B.n)
            assert ref["norm"] == pytest.approx(n1 * n2, rel=1e-14)


def test_reference_small_catalogue_quirk():
    """The serial reference is exact for tiny inputs; the OpenMP build drops
    shallow leaves (kdtree.c:405-413).  The oracle agrees with the serial
    build and with a brute force."""
    g = json.load(open(os.path.join(GOLD, "box_iso_tiny300.json")))
    (x, y, z), = catalogs("box_tiny300")[0]
    b = _bins("FCFC_2PT_BOX", g["binning"])
    c = F.Catalog(b, x, y, z)
    ob = O.oracle_bins(0, b.s2bin)
    cnt, _ = O.count(0, b.bsize_rescaled, (c.x, c.y, c.z, None, None), None, True, ob, variant=3)
    np.testing.assert_array_equal(2 * cnt, np.array(g["results"]["scalar:DD"]["cnt"]))
    assert 2 * cnt.sum() == _brute_iso(c.x, c.y, c.z, b)[1]


def _brute_iso(x, y, z, b):
    """numpy brute force, minimum image, unfused FP64 (variant 3 rounding)."""
    Lp = b.bsize_rescaled
    n = len(x)
    tot = np.zeros(b.ns, dtype=np.int64)
    for i in range(n - 1):
        d = []
        for k, v in enumerate((x, y, z)):
            dd = v[i] - v[i + 1:]
            dd = np.where(dd > Lp[k] / 2, dd - Lp[k], np.where(dd < -Lp[k] / 2, dd + Lp[k], dd))
            d.append(dd)
        s2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        ok = s2 < b.s2bin[-1]
        idx = np.searchsorted(b.s2bin, s2[ok], side="right") - 1
        tot += np.bincount(idx, minlength=b.ns)
    return tot, 2 * int(tot.sum())


def test_brute_force_smu_spi():
    """(s,mu) and (s_perp,pi) against a numpy brute force that bins with
    exact rational comparisons of the FP64 values (fractions)."""
    x, y, z = F.gen_uniform(77, 600, 1000.0)
    for bt in (1, 2):
        b = F.Bins.linear(F.GEOM_BOX, bt, 0, 200, 5, nmu=10, pmin=0, pmax=200, dpi=5,
                          bsize=1000.0)
        c = F.Catalog(b, x, y, z)
        ob = O.oracle_bins(bt, b.s2bin, b.nmu, b.pbin if bt == 2 else None)
        cnt, cen = O.count(0, b.bsize_rescaled, (c.x, c.y, c.z, None, None), None, True, ob,
                           variant=3, census_mask=0b11111)
        assert cen == 0
        Lp = b.bsize_rescaled
        ref = np.zeros(b.ntot, dtype=np.int64)
        n = len(c.x)
        for i in range(n - 1):
            d = []
            for k, v in enumerate((c.x, c.y, c.z)):
                dd = v[i] - v[i + 1:]
                dd = np.where(dd > Lp[k] / 2, dd - Lp[k], np.where(dd < -Lp[k] / 2, dd + Lp[k], dd))
                d.append(dd)
            if bt == 1:
                dz2 = d[2] * d[2]
                s2 = d[0] * d[0] + d[1] * d[1] + dz2
                ok = s2 < b.s2bin[-1]
                sb = np.searchsorted(b.s2bin, s2[ok], side="right") - 1
                mu = np.sqrt(dz2[ok] / np.maximum(s2[ok], 1e-300))
                mb = np.floor(mu * b.nmu).astype(int)
                keep = mb < b.nmu
                ref += np.bincount(sb[keep] + b.ns * mb[keep], minlength=b.ntot)
            else:
                pi = np.abs(d[2])
                s2 = d[0] * d[0] + d[1] * d[1]
                ok = (s2 < b.s2bin[-1]) & (pi < b.pbin[-1])
                sb = np.searchsorted(b.s2bin, s2[ok], side="right") - 1
                pb = np.searchsorted(b.pbin, pi[ok], side="right") - 1
                ref += np.bincount(sb + b.ns * pb, minlength=b.ntot)
        np.testing.assert_array_equal(cnt, ref)


def test_invariants():
    """Symmetry and additivity of the oracle's counts."""
    L = 1000.0
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=20, bsize=L)
    ob = O.oracle_bins(1, b.s2bin, b.nmu)
    xa, ya, za = F.gen_uniform(1, 1500, L)
    xb, yb, zb = F.gen_uniform(2, 1700, L)
    A = F.Catalog(b, xa, ya, za)
    B = F.Catalog(b, xb, yb, zb)
    U = F.Catalog(b, np.r_[xa, xb], np.r_[ya, yb], np.r_[za, zb])
    t = lambda c: (c.x, c.y, c.z, None, None)
    aa, _ = O.count(0, b.bsize_rescaled, t(A), None, True, ob)
    bb, _ = O.count(0, b.bsize_rescaled, t(B), None, True, ob)
    ab, _ = O.count(0, b.bsize_rescaled, t(A), t(B), False, ob)
    ba, _ = O.count(0, b.bsize_rescaled, t(B), t(A), False, ob)
    uu, _ = O.count(0, b.bsize_rescaled, t(U), None, True, ob)
    np.testing.assert_array_equal(ab, ba)                 # cross symmetry
    np.testing.assert_array_equal(uu, aa + bb + ab)       # DD(A u B) = DD(A)+DD(B)+DR
    # unit weights reproduce the unweighted counts
    w1 = np.ones(A.n)
    aw, _ = O.count(0, b.bsize_rescaled, (A.x, A.y, A.z, None, w1), None, True, ob, weighted=True)
    np.testing.assert_array_equal(aw, aa.astype(np.float64))
    # every in-range pair is binned once: total = pairs with s < smax
    iso = F.Bins.linear(F.GEOM_BOX, F.BIN_ISO, 0, 200, 5, bsize=L)
    oi = O.oracle_bins(0, iso.s2bin)
    ai, _ = O.count(0, iso.bsize_rescaled, t(A), None, True, oi)
    # iso = sum over mu (no mu = 1 pairs in random data)
    np.testing.assert_array_equal(ai, aa.reshape(b.nmu, b.ns).sum(axis=0))


def test_weight_grad_finite_difference():
    """The oracle's FP64 weight gradient equals central differences of the
    weighted counts (counts are bilinear in the weights: exact up to
    rounding)."""
    L = 1000.0
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=5, bsize=L)
    ob = O.oracle_bins(1, b.s2bin, b.nmu)
    x, y, z, w = F.gen_clustered(3, 300, 5, 20, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(4, 500, L)
    wr = np.linspace(0.5, 1.5, 500)
    D = F.Catalog(b, x, y, z, w)
    R = F.Catalog(b, xr, yr, zr, wr)
    rng = np.random.default_rng(0)
    G = rng.normal(size=b.ntot)
    for isauto in (True, False):
        ga, gb = O.weight_grad(0, b.bsize_rescaled, (D.x, D.y, D.z, None, D.w),
                               None if isauto else (R.x, R.y, R.z, None, R.w), isauto, ob, G)
        for i in (0, 7, 123, 321):
            h = 1e-3
            vals = []
            for sgn in (1, -1):
                w2 = D.w.copy()
                w2[i] += sgn * h
                c, _ = O.count(0, b.bsize_rescaled, (D.x, D.y, D.z, None, w2),
                               None if isauto else (R.x, R.y, R.z, None, R.w), isauto, ob,
                               weighted=True)
                vals.append(G @ c)
            fd = (vals[0] - vals[1]) / (2 * h)
            assert ga[i] == pytest.approx(fd, rel=1e-8, abs=1e-9)
        if not isauto:
            j = 42
            vals = []
            for sgn in (1, -1):
                w2 = R.w.copy()
                w2[j] += sgn * 1e-3
                c, _ = O.count(0, b.bsize_rescaled, (D.x, D.y, D.z, None, D.w),
                               (R.x, R.y, R.z, None, w2), False, ob, weighted=True)
                vals.append(G @ c)
            assert gb[j] == pytest.approx((vals[0] - vals[1]) / 2e-3, rel=1e-8, abs=1e-9)


def test_census_detects_path_dependence():
    """Planted case: two objects whose FP64 separation lands on a bin edge in
    one rounding path and not in another.  The census must flag it."""
    L = 1000.0
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_ISO, 0, 200, 5, bsize=L)
    ob = O.oracle_bins(0, b.s2bin)
    # pairs across the periodic boundary at s = 2 (an edge, s^2 = 4) within
    # ~1e-13: (a+S)-b, (b-S)-a and the wrapped a-b round differently
    rng = np.random.default_rng(5)
    found = 0
    for _ in range(4000):
        a = rng.uniform(199.0, 200.0)
        ay, by = 3.0, 3.0 + rng.uniform(0.0, 0.5)
        dx = np.sqrt(4.0 - (ay - by) ** 2)
        bx = a + dx - 200.0 + rng.uniform(-1e-13, 1e-13)
        if not (0 <= bx < 200):
            continue
        A = (np.array([a]), np.array([ay]), np.array([5.0]), None, None)
        B = (np.array([bx]), np.array([by]), np.array([5.0]), None, None)
        c, cen = O.count(0, b.bsize_rescaled, A, B, False, ob, variant=1, census_mask=0b11111)
        found += cen
    # a census that ignored the rounding paths would report 0 here
    found = found > 0
    assert found, "no path-dependent pair found"
