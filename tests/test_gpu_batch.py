"""Batched counting (many catalogue pairs in one persistent launch) and the
autotuner: every job's histogram must equal the single-job count exactly."""
import numpy as np
import pytest

import fcfc_amd as F

pytestmark = pytest.mark.gpu
L = 1000.0


@pytest.fixture(scope="module")
def eng():
    with F.Engine() as e:
        yield e


def _mocks(b, n):
    sizes = F.gen_sizes(7, n, 300, 3000)
    out = []
    for k, nk in enumerate(sizes):
        nh = max(1, int(nk) // 100)
        x, y, z, w = F.gen_clustered(100 + k, int(nk) // 2, nh, 50, 2.0, L, 0.3)
        xr, yr, zr = F.gen_uniform(500 + k, 3 * int(nk), L)
        out.append((F.Catalog(b, x, y, z, w), F.Catalog(b, xr, yr, zr)))
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_equals_single(eng, prec):
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=20, bsize=L)
    mocks = _mocks(b, 12)
    jobs = []
    for D, R in mocks:
        jobs += [(D, None, True, True), (D, R, False, True), (R, None, True, False)]
    print(f"case batch prec={prec} jobs={len(jobs)}", flush=True)
    got = eng.count_batch(jobs, b, precision=prec)
    for k, (a, c, au, wt) in enumerate(jobs):
        ref = eng.count_pairs(a, c, b, isauto=au, withwt=wt, precision=prec)
        if wt:
            np.testing.assert_allclose(got[k], ref, rtol=1e-12, atol=0)
        else:
            np.testing.assert_array_equal(got[k], ref)


def test_autotune_keeps_counts(eng):
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_ISO, 0, 200, 5, bsize=L)
    x, y, z = F.gen_uniform(9, 20000, L)
    D = F.Catalog(b, x, y, z)
    ref = eng.count_pairs(D, None, b, isauto=True)
    cd, zd, tile, est = eng.autotune(D, None, b, isauto=True)
    assert cd in (2, 3, 4, 6) and zd in (1, 2, 4, 8, 16) and tile in (512, 1024) and est > 0
    np.testing.assert_array_equal(eng.count_pairs(D, None, b, isauto=True), ref)


def test_autotune_cross_and_gradient(eng):
    """Tuned shapes of a cross count whose first catalogue is the smaller one
    (the count swaps them) and of the weight gradient leave the results as
    they were: counts identical, gradients to rounding."""
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=10, bsize=L)
    x, y, z, w = F.gen_clustered(21, 3000, 20, 50, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(22, 12000, L)
    D = F.Catalog(b, x, y, z, w)
    R = F.Catalog(b, xr, yr, zr, np.linspace(0.8, 1.2, len(xr)))
    G = np.random.default_rng(3).normal(size=b.ntot)
    ref_c = eng.count_pairs(D, R, b, withwt=True)
    ref_g = [eng.count_backward(D, R, b, G), eng.count_backward(D, None, b, G, isauto=True)[0]]
    for g in (False, True):
        print(f"case autotune cross gradient={g}", flush=True)
        cd, zd, tile, est = eng.autotune(D, R, b, withwt=True, gradient=g)
        assert cd in (2, 3, 4, 6) and zd in (1, 2, 4, 8, 16) and tile in (512, 1024) and est > 0
        cd, zd, tile, est = eng.autotune(D, None, b, isauto=True, withwt=True, gradient=g)
        assert est > 0
    np.testing.assert_allclose(eng.count_pairs(D, R, b, withwt=True), ref_c, rtol=1e-12, atol=0)
    ga, gb = eng.count_backward(D, R, b, G)
    np.testing.assert_allclose(ga, ref_g[0][0], rtol=1e-10, atol=1e-12 * np.abs(ref_g[0][0]).max())
    np.testing.assert_allclose(gb, ref_g[0][1], rtol=1e-10, atol=1e-12 * np.abs(ref_g[0][1]).max())
    gd, _ = eng.count_backward(D, None, b, G, isauto=True)
    np.testing.assert_allclose(gd, ref_g[1], rtol=1e-10, atol=1e-12 * np.abs(ref_g[1]).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_plan_graph_replays(eng, prec):
    """A prepared batch replays its HIP graph: every replay must equal an
    eager batched count of the same jobs, also after the weights changed in
    place between replays (the graph re-reads them)."""
    import torch
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=20, bsize=L)
    f = b.rescale
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a * f)).cuda()
    jobs = []
    for k in range(3):
        x, y, z, w = F.gen_clustered(300 + k, 2000, 20, 50, 2.0, L, 0.3)
        xr, yr, zr = F.gen_uniform(400 + k, 5000, L)
        D = F.Catalog(b, dev(x), dev(y), dev(z), torch.from_numpy(w).cuda())
        R = F.Catalog(b, dev(xr), dev(yr), dev(zr))
        jobs += [(D, None, True, True), (D, R, False, True), (R, None, True, False)]
    plan = eng.plan(jobs, b, precision=prec)
    try:
        for rep in range(4):
            print(f"case plan prec={prec} replay={rep}", flush=True)
            if rep:
                for D, _, _, wt in jobs[::3]:
                    D.w.mul_(1.0 + 0.1 * rep)     # eager device work between replays
            got = plan.run()
            torch.cuda.synchronize()
            ref = eng.count_batch(jobs, b, precision=prec)
            for k, (_, _, _, wt) in enumerate(jobs):
                if wt:
                    np.testing.assert_allclose(got[k], ref[k], rtol=1e-12, atol=0)
                else:
                    np.testing.assert_array_equal(got[k], ref[k])
    finally:
        plan.close()


def test_batch_edge_catalogues(eng):
    """One batched call sorts all its catalogues together (one launch per
    stage over the concatenated objects): empty and one-object catalogues
    between others, host and device catalogues, unit weights filled on the
    device for a weighted count of an unweighted catalogue; every job is
    checked against the oracle."""
    import torch
    from oracle import oracle as O
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=20, bsize=L)
    f = b.rescale
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu)
    x, y, z, w = F.gen_clustered(31, 1500, 10, 50, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(32, 2500, L)
    D = F.Catalog(b, x, y, z, w)
    R = F.Catalog(b, xr, yr, zr)
    E = F.Catalog(b, np.zeros(0), np.zeros(0), np.zeros(0))
    one = F.Catalog(b, np.array([500.0]), np.array([500.0]), np.array([500.0]))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a * f)).cuda()
    Rd = F.Catalog(b, dev(xr), dev(yr), dev(zr), rescaled=True)      # device, no weights
    jobs = [(E, None, True, False), (D, None, True, True), (one, R, False, False),
            (E, R, False, True), (Rd, D, False, True), (R, None, True, False),
            (one, None, True, False), (Rd, None, True, True)]
    print(f"case batch edge catalogues jobs={len(jobs)}", flush=True)
    got = eng.count_batch(jobs, b)
    host = {id(Rd): R}
    for k, (a, c, au, wt) in enumerate(jobs):
        a = host.get(id(a), a)
        c = host.get(id(c), c) if c is not None else None
        # weighted counts give unweighted catalogues unit weights
        # (build_tree.c:143-156); the oracle takes them explicitly
        uw = lambda k: (k.w if k.w is not None else np.ones(k.n)) if wt else None
        ta = (a.x, a.y, a.z, None, uw(a))
        tc = None if au else (c.x, c.y, c.z, None, uw(c))
        ref, cen = O.count(0, b.bsize_rescaled, ta, tc, au, ob, variant=1, weighted=bool(wt),
                           census_mask=0 if wt else 0b11111)
        if wt:
            np.testing.assert_allclose(got[k], ref, rtol=1e-12, atol=0, err_msg=f"job {k}")
        else:
            diff = np.abs(got[k].astype(np.int64) - ref.astype(np.int64)).sum()
            assert diff <= 2 * cen, f"job {k}: counts differ by {diff} (census {cen})"
            if cen == 0:
                np.testing.assert_array_equal(got[k], ref)


def test_batch_survey_norms_on_device(eng):
    """Survey (s, mu): |r|^2 computed on the GPU inside the batched sort
    (device catalogues without s, with FCFC's lane rule) gives the counts of
    the host-prepared catalogues."""
    import torch
    b = F.Bins.linear(F.GEOM_SURVEY, F.BIN_SMU, 0, 200, 5, nmu=10)
    f = b.rescale
    x, y, z, w = F.gen_survey(41, 3001)
    xr, yr, zr, wr = F.gen_survey(42, 5003)
    D = F.Catalog(b, x, y, z, w)
    R = F.Catalog(b, xr, yr, zr, wr)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a * f)).cuda()
    Dd = F.Catalog(b, dev(x), dev(y), dev(z), torch.from_numpy(w).cuda(), rescaled=True)
    Rd = F.Catalog(b, dev(xr), dev(yr), dev(zr), torch.from_numpy(wr).cuda(), rescaled=True)
    assert Dd.s is None and Rd.s is None
    print("case batch survey norms", flush=True)
    host = eng.count_batch([(D, None, True, False), (D, R, False, True)], b)
    devc = eng.count_batch([(Dd, None, True, False), (Dd, Rd, False, True)], b)
    np.testing.assert_array_equal(devc[0], host[0])
    np.testing.assert_allclose(devc[1], host[1], rtol=1e-12, atol=0)


def test_batch_shared_positions_different_weights(eng):
    """Jobs on the same device position arrays with different weight arrays
    (jackknife or bootstrap weights) get sorted copies of their own: every
    job must carry its own weights, in a batch and in a prepared plan."""
    import torch
    from oracle import oracle as O
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=20, bsize=L)
    f = b.rescale
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu)
    x, y, z, w = F.gen_clustered(61, 2000, 15, 50, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(62, 4000, L)
    w2 = np.ascontiguousarray(w[::-1])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    X, Y, Z = dev(x * f), dev(y * f), dev(z * f)
    D1 = F.Catalog(b, X, Y, Z, dev(w), rescaled=True)
    D2 = F.Catalog(b, X, Y, Z, dev(w2), rescaled=True)
    R = F.Catalog(b, xr, yr, zr)
    jobs = [(D1, None, True, True), (D2, None, True, True), (D1, R, False, True),
            (D2, R, False, True)]
    Dh = F.Catalog(b, x, y, z)
    refs = []
    for ww in (w, w2):
        ta = (Dh.x, Dh.y, Dh.z, None, ww)
        refs.append(O.count(0, b.bsize_rescaled, ta, None, True, ob, variant=1, weighted=True)[0])
    for ww in (w, w2):
        ta = (Dh.x, Dh.y, Dh.z, None, ww)
        refs.append(O.count(0, b.bsize_rescaled, ta, (R.x, R.y, R.z, None, np.ones(R.n)), False,
                            ob, variant=1, weighted=True)[0])
    print("case batch shared positions, different weights", flush=True)
    got = eng.count_batch(jobs, b)
    for k in range(4):
        np.testing.assert_allclose(got[k], refs[k], rtol=1e-12, atol=0, err_msg=f"job {k}")
    plan = eng.plan(jobs, b)
    try:
        got = plan.run()
        for k in range(4):
            np.testing.assert_allclose(got[k], refs[k], rtol=1e-12, atol=0, err_msg=f"plan job {k}")
    finally:
        plan.close()
