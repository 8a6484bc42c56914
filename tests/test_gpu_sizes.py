"""Parity at the configurations' own sizes, and the GPU rounding census.

cfg_c1_full: the C1 configuration as specified (box L = 1000, N = 1e5
uniform, DD auto, isotropic 40 bins in [0, 200)), counted by the upstream
FCFC_2PT_BOX in its AVX-512 and serial scalar builds.
cfg_c2_twentieth: the C2 recipe at 1/20 of its size (D = 1e5 clustered with
lognormal weights, R = 1e6 uniform; DD, DR weighted and RR unweighted in
(s, mu) 40 x 100), counted by the AVX-512 build.  Each fixture stores the
oracle's ambiguity census (in-range pairs whose bin depends on the
reference's rounding path), computed when the fixture was written.

Every form of the GPU count (fp64, fp32 + FP64 recheck, plain FP64) must
give the reference's counts: identical where the census is zero (the GPU
doubles nothing: the file's auto counts are twice the unordered ones),
within 1e-12 relative for weighted counts.
"""
import json
import os

import numpy as np
import pytest

import fcfc_amd as F
from oracle import oracle as O
from tests.golden.make_golden import catalogs, checksum

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
L = 1000.0
FORMS = ("fp64", "fp32", "plain")


@pytest.fixture(scope="module")
def eng():
    with F.Engine() as e:
        yield e


def _fixture(name):
    g = json.load(open(os.path.join(GOLD, name + ".json")))
    cats, wts = catalogs(g["catalogs"])
    assert [checksum(*c) for c in cats] == g["checksums"], "generator output changed"
    bk = g["binning"]
    b = F.Bins.linear(F.GEOM_BOX, bk["bintype"], bk["smin"], bk["smax"], bk["ds"],
                      nmu=bk.get("nmu", 1), bsize=L)
    return g, cats, wts, b


def _exact(gpu2, ref, census, what):
    diff = int(np.abs(gpu2 - ref).sum())
    assert diff <= 2 * 2 * census, f"{what}: counts differ by {diff} (census {census})"
    if census == 0:
        np.testing.assert_array_equal(gpu2, ref, err_msg=what)


@pytest.mark.parametrize("prec", FORMS)
def test_c1_full_size(eng, prec):
    g, cats, _, b = _fixture("cfg_c1_full")
    d = F.Catalog(b, *cats[0])
    print(f"case C1 full size prec={prec} n={d.n}", flush=True)
    cnt = eng.count_pairs(d, None, b, isauto=True, withwt=False, precision=prec)
    for flav in ("avx512", "scalar"):
        ref = np.array(g["results"][f"{flav}:DD"]["cnt"], dtype=np.int64)
        _exact(2 * cnt, ref, g["census"]["DD"], f"C1 {prec} vs {flav}")


@pytest.fixture(scope="module")
def c2t():
    g, cats, wts, b = _fixture("cfg_c2_twentieth")
    (x, y, z, w), (xr, yr, zr) = cats
    return g, b, F.Catalog(b, x, y, z, w), F.Catalog(b, xr, yr, zr)


@pytest.mark.parametrize("prec", FORMS)
def test_c2_twentieth_vs_reference(eng, c2t, prec):
    g, b, D, R = c2t
    res = g["results"]
    print(f"case C2/20 prec={prec} D={D.n} R={R.n}", flush=True)
    dd = eng.count_pairs(D, None, b, isauto=True, withwt=True, precision=prec)
    ref = np.array(res["avx512:DD"]["cnt"])
    assert np.max(np.abs(2 * dd - ref) / np.maximum(np.abs(ref), 1e-300)) <= 1e-12
    dr = eng.count_pairs(D, R, b, isauto=False, withwt=True, precision=prec)
    ref = np.array(res["avx512:DR"]["cnt"])
    assert np.max(np.abs(dr - ref) / np.maximum(np.abs(ref), 1e-300)) <= 1e-12
    rr = eng.count_pairs(R, None, b, isauto=True, withwt=False, precision=prec)
    _exact(2 * rr, np.array(res["avx512:RR"]["cnt"], dtype=np.int64), g["census"]["RR"],
           f"C2/20 RR {prec}")


def test_c2_twentieth_unweighted_vs_oracle(eng, c2t):
    """The unweighted DD and DR of the same catalogues (the reference fixture
    holds the weighted ones) against the oracle with its census."""
    g, b, D, R = c2t
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu)
    refdd, cdd = O.count(0, b.bsize_rescaled, (D.x, D.y, D.z, None, None), None, True, ob,
                         variant=1, census_mask=0b11111, nthread=16)
    refdr, cdr = O.count(0, b.bsize_rescaled, (D.x, D.y, D.z, None, None),
                         (R.x, R.y, R.z, None, None), False, ob, variant=1,
                         census_mask=0b11111, nthread=16)
    for prec in FORMS:
        print(f"case C2/20 unweighted prec={prec}", flush=True)
        dd = eng.count_pairs(D, None, b, isauto=True, withwt=False, precision=prec)
        _exact(2 * dd, 2 * refdd, cdd, f"DD {prec}")
        dr = eng.count_pairs(D, R, b, isauto=False, withwt=False, precision=prec)
        _exact(2 * dr, 2 * refdr, cdr, f"DR {prec}")


def _lattice(nmu):
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=nmu, bsize=L)
    g = np.arange(14) / b.rescale + 100.0
    x, y, z = [a.ravel().copy() for a in np.meshgrid(g, g, g, indexing="ij")]
    return b, F.Catalog(b, x, y, z)


@pytest.mark.parametrize("case", ["lattice7", "lattice12", "clustered"])
def test_gpu_census_matches_oracle(eng, case):
    """FCFC_GPU_CENSUS: the FP32 kernel evaluates the reference's other
    rounding paths (a-side shift, per-pair wrap, unfused scalar) for every
    pair it re-evaluates in FP64.  Its census must equal the oracle's over
    the same paths (variants 0-3): on an integer lattice many pairs sit on
    mu edges, and with nmu = 7 the scalar path disagrees on 9728 of them;
    with nmu = 12, and on clustered random catalogues, no pair's bin depends
    on the path (census 0)."""
    if case.startswith("lattice"):
        b, d = _lattice(int(case[7:]))
        r = None
    else:
        b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=100, bsize=L)
        x, y, z, _ = F.gen_clustered(81, 4000, 40, 50, 2.0, L)
        xr, yr, zr = F.gen_uniform(82, 8000, L)
        d, r = F.Catalog(b, x, y, z), F.Catalog(b, xr, yr, zr)
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu)
    for isauto in ((True,) if r is None else (True, False)):
        other = None if isauto else (r.x, r.y, r.z, None, None)
        ref, cen = O.count(0, b.bsize_rescaled, (d.x, d.y, d.z, None, None), other, isauto, ob,
                           variant=1, census_mask=0b01111)
        print(f"case census {case} isauto={isauto} oracle census {cen}", flush=True)
        got = eng.count_pairs(d, None if isauto else r, b, isauto=isauto, precision="fp32",
                              census=True)
        st = eng.stats()
        np.testing.assert_array_equal(got, ref)
        assert int(st["census"]) == cen, f"GPU census {st['census']} != oracle {cen}"
        # the oracle's census of these catalogues (computed on the CPU)
        assert cen == (9728 if case == "lattice7" else 0)


@pytest.mark.parametrize("bt", [F.BIN_SMU, F.BIN_SPI])
def test_survey_fp32_equals_fp64_at_scale(eng, bt):
    """A size-independent property of the FP32 survey form: its windows
    send every pair whose bin could differ to the FP64 re-evaluation, so its
    integer counts equal the FP64 form's exactly -- here for 3e5 objects of
    the survey shell (about 6e8 in-range pairs per count), auto and cross."""
    x, y, z, _ = F.gen_survey(41, 300000)
    xr, yr, zr, _ = F.gen_survey(42, 300000)
    b = F.Bins.linear(F.GEOM_SURVEY, bt, 0, 200, 5, nmu=100 if bt == F.BIN_SMU else 1,
                      pmin=0, pmax=200 if bt == F.BIN_SPI else 0, dpi=5)
    d, r = F.Catalog(b, x, y, z), F.Catalog(b, xr, yr, zr)
    for other in (None, r):
        print(f"case survey fp32 vs fp64 bt={bt} cross={other is not None}", flush=True)
        c64 = eng.count_pairs(d, other, b, isauto=other is None, precision="fp64")
        c32 = eng.count_pairs(d, other, b, isauto=other is None, precision="fp32")
        assert eng.stats()["rechecked"] > 0
        assert c64.sum() > 1e8
        np.testing.assert_array_equal(c32, c64)
