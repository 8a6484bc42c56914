"""Host-side logic of the engine (no GPU needed): the C ABI exports, the
binning setup (FCFC's cf_setup rules), catalogue preparation, the pair-count
file writer (byte-identical to FCFC's save_res) and the generator."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import fcfc_amd as F
from fcfc_amd import _lib as L
from oracle import oracle as O
from tests.golden.make_golden import catalogs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "fcfc_gpu.h")).read()
    names = set(re.findall(r"^\w[\w\s\*]*?\b(fcfc_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 15
    lib = C.CDLL(L.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} declared in fcfc_gpu.h but not exported"
    assert names == set(L.SIGNATURES), "ctypes signatures out of sync with the header"


def test_no_gpu_is_an_error_not_a_fallback():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    with pytest.raises(F.FcfcError) as e:
        F.Engine()
    assert e.value.code == L.FCFC_ERR_GPU


@pytest.mark.parametrize("smin,ds,fac", [(0, 5, 0.2), (0, 2.5, 0.4), (20, 10, 0.1), (0.5, 7.3, None)])
def test_rescale_factor(smin, ds, fac):
    """setup_cf.c:308-329 (least_fac2) and the hybrid route (:413-419)."""
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_ISO, smin, 146.5 if fac is None else 200, ds,
                      bsize=1000.0)
    if fac is None:
        assert b.tabtype == 1
        # hybrid: power-of-two factor with 8192 < fac * smax^2 <= 16384... not
        # above 2 * 8192 (setup_cf.c:416)
        f = b.rescale
        assert f == 2.0 ** np.floor(np.log2(f))
        assert f * 146.5 ** 2 <= 2 * 8192
    else:
        assert b.tabtype == 0
        assert b.rescale == pytest.approx(fac, rel=1e-15)
        np.testing.assert_array_equal(b.s2bin, np.round(b.sbin_raw * b.rescale) ** 2)


@pytest.mark.parametrize("kw", [dict(bt=0, smin=0, smax=200, ds=5),
                                dict(bt=1, smin=20, smax=150, ds=10, nmu=10),
                                dict(bt=2, smin=0, smax=200, ds=5, pmin=0, pmax=200, dpi=5),
                                dict(bt=1, smin=0.5, smax=146.5, ds=7.3, nmu=7)])
@pytest.mark.parametrize("geom", [F.GEOM_BOX, F.GEOM_SURVEY])
def test_lookup_table_is_edge_search(kw, geom):
    """Integer and hybrid tables (create_lut.c:58-140) + the downward
    refinement reproduce a plain search over the rescaled edges for every
    value in range (closed form of the lookup)."""
    b = F.Bins.linear(geom, kw["bt"], kw["smin"], kw["smax"], kw["ds"], nmu=kw.get("nmu", 1),
                      pmin=kw.get("pmin", 0), pmax=kw.get("pmax", 0), dpi=kw.get("dpi", 1),
                      bsize=1000.0 if geom == F.GEOM_BOX else None)
    s2 = b.s2bin
    tab = b.stab
    off = int(s2[0])
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.uniform(s2[0], s2[-1], 20000), s2[:-1], np.nextafter(s2[1:], 0)])
    for v in vals:
        k = int(v) - off
        e = int(tab[k])
        if e >= b.ns:
            e -= b.ns
            while e != 0 and v < s2[e]:
                e -= 1
        assert e == np.searchsorted(s2, v, side="right") - 1


def test_linear_bin_count_and_limits():
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_ISO, 0, 200, 5, bsize=1000.0)
    assert b.ns == 40 and b.ntot == 40
    np.testing.assert_array_equal(b.sbin_raw, np.arange(41) * 5.0)
    with pytest.raises(F.FcfcError):           # smax >= L/2 (load_conf.c:1000)
        F.Bins.linear(F.GEOM_BOX, F.BIN_ISO, 0, 500, 5, bsize=1000.0)
    with pytest.raises(F.FcfcError):           # nmu <= 1
        F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=1, bsize=1000.0)
    with pytest.raises(F.FcfcError):           # negative width
        F.Bins.linear(F.GEOM_BOX, F.BIN_ISO, 0, 200, -5, bsize=1000.0)
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SPI, 0, 200, 5, pmin=0, pmax=100, dpi=10, bsize=1000.0)
    assert (b.ns, b.np, b.ntot) == (40, 10, 400)


def test_survey_norm_lane_rule():
    """|r|^2 of FCFC_2PT: FMA for the first floor(n/8)*8 objects, unfused for
    the tail (2pt/build_tree.c:59-92)."""
    x, y, z, w = F.gen_survey(3, 1003)
    b = F.Bins.linear(F.GEOM_SURVEY, F.BIN_SMU, 0, 200, 5, nmu=10)
    c = F.Catalog(b, x, y, z, w)
    ref = O.survey_norm2(x * b.rescale, y * b.rescale, z * b.rescale, 8)
    np.testing.assert_array_equal(c.s, ref)
    np.testing.assert_array_equal(c.x, x * b.rescale)


def test_pair_count_file_byte_identical():
    """fcfc_save_pair_count writes exactly the bytes of FCFC's save_res for
    the same counts (binary format 0 and ASCII format 1)."""
    ref_bin = os.path.join(GOLD, "ref_box_smu_DR.bin")
    r = O.read_pair_count(ref_bin)
    (d, rr), wts = catalogs("box_clustered")
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=100, bsize=1000.0)
    D = F.Catalog(b, *d)
    R = F.Catalog(b, *rr)
    # the file stores the weighted totals of both catalogues (R unit weights)
    assert r["num"][0] == pytest.approx(D.sumw, rel=1e-15)
    for fmt, ref in ((0, ref_bin), (1, os.path.join(GOLD, "ref_box_smu_DR.txt"))):
        out = os.path.join("/tmp", f"pc_test.{fmt}")
        F.save_pair_count(out, b, np.array(r["cnt"]), D, R, isauto=False, withwt=True, ofmt=fmt)
        mine = open(out, "rb").read()
        theirs = open(ref, "rb").read()
        if fmt == 1:
            # the reference also prints the weighted totals with %.10lg
            assert mine == theirs
        else:
            assert mine == theirs


def test_generator_is_deterministic():
    a = F.gen_clustered(11, 100, 3, 10, 2.0, 1000.0, 0.3)
    b = F.gen_clustered(11, 100, 3, 10, 2.0, 1000.0, 0.3)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    x, y, z, w = a
    assert np.all((x >= 0) & (x < 1000)) and np.all(w == w.astype(np.float32))
    sx, sy, sz, sw = F.gen_survey(1, 2000)
    r = np.sqrt(sx ** 2 + sy ** 2 + sz ** 2)
    assert r.min() >= 1000 and r.max() < 2000
    assert np.all(np.abs(np.degrees(np.arctan2(sy, sx))) <= 30 + 1e-9)
    assert np.all(np.abs(np.degrees(np.arcsin(sz / r))) <= 30 + 1e-9)
    n = F.gen_sizes(3, 256, 5e4, 5e5)
    assert n.min() >= 5e4 and n.max() <= 5e5


def test_pair_counts_conventions():
    """eval_pairs doubling and normalisation (eval_cf.c:125-149)."""
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_ISO, 0, 200, 5, bsize=1000.0)
    A = F.Catalog(b, *F.gen_uniform(1, 10, 1000.0))
    c, norm, nc = F.pair_counts(np.arange(40), A, None, True, False)
    assert norm == 90 and c[3] == 6 and nc[3] == pytest.approx(6 / 90)
