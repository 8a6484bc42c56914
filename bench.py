"""Benchmark of the FCFC pair-count step on one MI355X.

Workload (the headline configuration): periodic box L = 1000 Mpc/h,
D = 2e6 clustered objects (1e6 uniform + 2000 Gaussian halos of 500 members,
sigma = 2 Mpc/h) with lognormal(0, 0.3) weights, R = 2e7 uniform randoms;
DD (weighted), DR (weighted, unit random weights) and RR (unweighted) in
(s, mu) bins: 40 linear s bins in [0, 200) x 100 mu bins.  Synthetic data
from the project's seeded generator, resident in HBM before timing.

One step = the three counts (grid build + pair counting).  Three forms of
the count are measured on the same data:
  fp64   FP64 distances; the mu bin from the hardware reciprocal square root,
         settled exactly near its edges (the headline path, `value`);
  plain  FP64 with the reference's per-pair arithmetic transcribed directly
         (an IEEE division per (s,mu) pair) -- the "plain FP64 GPU kernel";
  fp32   FP32 distances, pairs near a bin edge re-evaluated in FP64.
Integer counts of all forms must be identical, weighted ones within 1e-12.

Prints ONE JSON line.  `value` = candidate pair-distance evaluations per
second of the timed steps (whole job: grid build, tiling and counting, data
resident in HBM).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair distance evals/s and mocks/s vs plain FP64 GPU, counts bit-exact"
# FLOPs per evaluated pair (+ - x count 1, FMA 2, division 1): box (s,mu) 10,
# +2 for the weight product and its accumulation (weighted counts)
FLOPS_PAIR = {"smu": 10, "smu_w": 12}
PEAK_FP32_VALU = 157.3e12
PEAK_FP64_VALU = 78.6e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(scale):
    import fcfc_amd as F
    L = 1000.0
    nh = max(1, int(round(2000 * scale)))
    x, y, z, w = F.gen_clustered(20261, int(1e6 * scale), nh, 500, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(20262, int(2e7 * scale), L)
    return L, (x, y, z, w), (xr, yr, zr)


def to_device(b, cat, F, torch):
    f = b.rescale
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a * f)).to("cuda")
    w = torch.from_numpy(np.ascontiguousarray(cat[3])).to("cuda") if len(cat) > 3 and cat[3] is not None else None
    return F.Catalog(b, dev(cat[0]), dev(cat[1]), dev(cat[2]), w)


def run_step(eng, b, D, R, prec):
    """DD, DR, RR; returns (counts, evals, kernel ms per count, rechecked, inrange)."""
    out, ev, kms, rc = {}, 0.0, {}, 0.0
    for name, a, c, au, wt in (("DD", D, None, True, True), ("DR", D, R, False, True),
                               ("RR", R, None, True, False)):
        out[name] = eng.count_pairs(a, c, b, isauto=au, withwt=wt, precision=prec)
        s = eng.stats()
        ev += s["evals"]
        kms[name] = s["kernel_ms"]
        rc += s["rechecked"]
        out[name + "_evals"] = s["evals"]
    return out, ev, kms, rc


def cpu_baseline(target_s=6.0, runs=3):
    """The upstream FCFC_2PT_BOX (OpenMP + SIMD build from the unmodified
    sources, oracle/_ref) on the host cores, OMP_PROC_BIND=spread
    OMP_PLACES=cores, count-phase time from its own progress lines:
      * a bounded sample of the headline workload: RR of a uniform subsample
        in the L = 1000 box, (s,mu) 40x100, sized for ~target_s per run,
        median of `runs` runs;
      * the C1 configuration as specified (N = 1e5 uniform, DD, iso 40 bins).
    Returns the baseline record plus the catalogues and the reference's
    counts, which the caller compares with the GPU's (the reference is the
    checker there, never the thing timed for `value`)."""
    from oracle import oracle as O
    import fcfc_amd as F
    flav = O.host_flavour()
    exe = O.ref_binary("FCFC_2PT_BOX", flav)
    if not os.path.exists(exe):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env = {"OMP_PROC_BIND": "spread", "OMP_PLACES": "cores"}
    L = 1000.0

    def rr(n, wd=None):
        x, y, z = F.gen_uniform(777, n, L)
        res, phase, wall, _ = O.run_reference("FCFC_2PT_BOX", flav, [(x, y, z)], ["R"], ["RR"],
                                              1, 0, 200, 5, nmu=100, bsize=L, threads=threads,
                                              extra_env=env, workdir=wd, reuse_files=wd is not None)
        return (x, y, z), res["RR"]["cnt"].astype(np.int64), phase["RR"], wall

    _, c0, t0, _ = rr(400000)
    n = int(min(8e6, max(400000, 400000 * np.sqrt(target_s / max(t0, 1e-3)))))
    times, cnts = [], []
    import shutil
    import tempfile
    wd = tempfile.mkdtemp(prefix="fcfc_base_")
    try:
        for _ in range(runs):
            cat, cnt, t, _ = rr(n, wd)
            times.append(t)
            cnts.append(cnt)
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    tmed = float(np.median(times))
    pairs = float(cnts[0].sum()) / 2          # the file holds doubled auto counts
    xc, yc, zc = F.gen_uniform(1001, 100000, L)
    res, phase, wall, _ = O.run_reference("FCFC_2PT_BOX", flav, [(xc, yc, zc)], ["D"], ["DD"],
                                          0, 0, 200, 5, bsize=L, threads=threads, extra_env=env)
    c1 = res["DD"]["cnt"].astype(np.int64)
    rec = {"value": pairs / tmed, "unit": "unordered in-range pairs/s", "cores": threads,
           "kind": "reference",
           "sample": f"upstream FCFC_2PT_BOX ({flav} build, OpenMP {threads} threads, "
                     f"OMP_PROC_BIND=spread OMP_PLACES=cores), RR auto count of {n} uniform "
                     f"points in the L=1000 box, (s,mu) 40x100; count phase median of {runs} "
                     f"runs {tmed:.2f} s ({', '.join(f'{t:.2f}' for t in times)}), "
                     f"{pairs:.4g} unordered in-range pairs",
           "runs_identical": all(np.array_equal(c, cnts[0]) for c in cnts),
           "c1": {"value": float(c1.sum()) / 2 / phase["DD"], "unit": "unordered in-range pairs/s",
                  "count_phase_s": phase["DD"], "wall_s": wall,
                  "sample": "C1 as specified: N=1e5 uniform, DD auto, iso 40 bins in [0,200)"}}
    return {"record": rec, "rr": (cat, cnts[0]), "c1": ((xc, yc, zc), c1)}


def parity_vs_reference(eng, base, forms=("fp64", "fp32", "plain")):
    """The GPU counts of the baseline's catalogues in every form against the
    reference's own counts (auto counts doubled as in its files)."""
    import fcfc_amd as F
    L = 1000.0
    out = {}
    for key, src, bt in (("rr_sample", "rr", F.BIN_SMU), ("c1", "c1", F.BIN_ISO)):
        (x, y, z), ref = base[src]
        b = F.Bins.linear(F.GEOM_BOX, bt, 0, 200, 5, nmu=100 if bt == F.BIN_SMU else 1, bsize=L)
        c = F.Catalog(b, x, y, z)
        rec = {"n": int(c.n), "reference_pairs": int(ref.sum()) // 2}
        for prec in forms:
            got = 2 * eng.count_pairs(c, None, b, isauto=True, withwt=False, precision=prec)
            st = eng.stats()
            rec[prec] = {"identical": bool(np.array_equal(got, ref)),
                         "abs_diff": int(np.abs(got - ref).sum()),
                         "kernel_ms": st["kernel_ms"],
                         "inrange_pairs_per_s": float(ref.sum()) / 2 / (st["kernel_ms"] * 1e-3)}
        out[key] = rec
    return out


PROFILE = os.path.join(ROOT, "profiles", "k_count_smu_box_RR_counters.txt")
PROFILE32 = os.path.join(ROOT, "profiles", "k_count_smu_box_RR_fp32_counters.txt")


def traffic_from_profile():
    """HBM bytes per launch of the RR count kernel, from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    kernel calibration recorded there); None when absent."""
    try:
        for l in open(PROFILE):
            if l.startswith("hbm_bytes_per_launch"):
                return float(l.split()[1])
    except OSError:
        pass
    return None


def issue_from_profile(fn=PROFILE):
    """VALU busy and VALU issue-slot fractions of the RR count kernel from the
    same summary (rocprofv3 SQ counters); for the FP32 form's summary the
    instruction rate against the FP32 vector peak.  None when absent."""
    out = {}
    try:
        for l in open(fn):
            t = l.split()
            if t[:1] == ["VALUBusy"]:
                out["valu_busy"] = float(t[-2]) / 100
            elif t[:2] == ["issue", "slots"]:
                out["valu_issue_slots"] = float(t[-2]) / 100
            elif t[:1] == ["measured"] and "%" in t:
                out["valu_issue_rate_vs_fp32_peak"] = float(t[t.index("%") - 1]) / 100
    except (OSError, ValueError, IndexError):
        return None
    return out or None


def gpu_mocks(b, nm, mock_scale, seed=4040):
    """256 periodic-box mocks (L = 1000) generated on the GPU from a seeded
    torch generator: sizes log-uniform in [5e4, 5e5] (times mock_scale),
    half uniform and half in halos of 500 members (sigma = 2, wrapped), with
    lognormal(0, 0.3) weights rounded to float32; 5x uniform randoms.
    Returns (sizes, [(D, R) device catalogues, rescaled])."""
    import torch
    import fcfc_amd as F
    L = 1000.0
    f = b.rescale
    sizes = F.gen_sizes(seed, nm, 5e4 * mock_scale, 5e5 * mock_scale)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    kw = dict(generator=g, device="cuda", dtype=torch.float64)

    def wrap(p):
        p = torch.remainder(p, L)
        return torch.where(p >= L, p - L, p)

    def cat(p, w=None):
        p = p * f
        return F.Catalog(b, p[:, 0].contiguous(), p[:, 1].contiguous(), p[:, 2].contiguous(), w)

    mocks = []
    for n in sizes:
        n = int(n)
        nh = max(1, n // 1000)
        nu = max(n - nh * 500, 1)
        cen = torch.rand((nh, 1, 3), **kw) * L
        halo = (cen + 2.0 * torch.randn((nh, 500, 3), **kw)).reshape(-1, 3)
        p = wrap(torch.cat([torch.rand((nu, 3), **kw) * L, halo]))
        w = torch.exp(0.3 * torch.randn(p.shape[0], **kw)).to(torch.float32).to(torch.float64)
        r = torch.rand((5 * n, 3), **kw) * L
        mocks.append((cat(p, w), cat(torch.where(r >= L, r - L, r))))
    torch.cuda.synchronize()
    return sizes, mocks


def mocks_run(eng, b, sizes, mocks, prec, steps=1, warmup=0, tune=True):
    """All 768 counts (DD, DR, RR of every mock) in one batched call (one
    persistent launch per kernel variant) against one call per count, with
    xi_0,2,4 of every mock; shapes from the autotuner for both (run on mocks
    of four sizes)."""
    import torch
    from fcfc_amd.xi import multipoles
    nm = len(mocks)
    jobs = []
    for D, R in mocks:
        jobs += [(D, None, True, True), (D, R, False, True), (R, None, True, False)]
    if tune:
        order = np.argsort(sizes)
        for k in sorted(set(int(order[int(q * (nm - 1))]) for q in (0.0, 0.33, 0.67, 1.0))):
            for a, c, au, wt in jobs[3 * k:3 * k + 3]:
                eng.autotune(a, c, b, isauto=au, withwt=wt, precision=prec)
        log("mocks: autotuned on mocks of four sizes")

    def xi_all(res):
        """Landy-Szalay xi(s, mu) and xi_0,2,4 of every mock (the same code
        for the batched and the per-count results)."""
        out = []
        for k, (D, R) in enumerate(mocks):
            dd, dr, rr = res[3 * k], res[3 * k + 1], res[3 * k + 2]
            ddn = dd * (2 / (D.sumw * (D.sumw - 1)))
            drn = dr * (1 / (D.sumw * R.n))
            rrn = rr * (2 / (float(R.n) * (R.n - 1)))
            xi = (ddn - 2 * drn + rrn) / np.where(rrn > 0, rrn, np.inf)
            out.append(multipoles(xi.reshape(b.nmu, b.ns), b.nmu, (0, 2, 4)))
        return out

    for i in range(warmup):
        eng.count_batch(jobs, b, precision=prec)
    torch.cuda.synchronize()
    t0 = time.time()
    for i in range(steps):
        res = eng.count_batch(jobs, b, precision=prec)
        xi_b = xi_all(res)
        log(f"mocks: batched pass {i} done")
    torch.cuda.synchronize()
    tb = (time.time() - t0) / steps
    stb = eng.stats()
    t1 = time.time()
    res1 = [eng.count_pairs(a, c, b, isauto=au, withwt=wt, precision=prec) for a, c, au, wt in jobs]
    xi_s = xi_all(res1)
    torch.cuda.synchronize()
    ts = time.time() - t1
    log("mocks: per-count pass done")
    exact = all(np.array_equal(res[3 * k + 2], res1[3 * k + 2]) for k in range(nm))
    wrel = max(float(np.max(np.abs(res[j] - res1[j]) / np.maximum(np.abs(res1[j]), 1e-300)))
               for j in range(len(jobs)) if jobs[j][3])
    xrel = max(float(np.nanmax(np.abs(a - c) / np.maximum(np.abs(c), 1e-300)))
               for a, c in zip(xi_b, xi_s))
    return {"mocks_per_s": nm / tb, "per_catalogue_mocks_per_s": nm / ts,
            "speedup_vs_per_catalogue": ts / tb, "batched_ms": tb * 1e3, "per_count_ms": ts * 1e3,
            "evals_per_s": stb["evals"] / tb, "kernel_ms": stb["kernel_ms"],
            "n_mocks": nm, "n_data_total": int(sum(int(v) for v in sizes)),
            "counts_identical_batched_vs_single": {"RR": exact, "weighted_max_rel": wrel,
                                                   "xi_max_rel": xrel}}


def mocks_workload(args):
    """--workload mocks: the batched-mocks configuration on its own line."""
    import fcfc_amd as F
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=100, bsize=1000.0)
    nm = max(1, int(round(256 * args.scale)))
    sizes, mocks = gpu_mocks(b, nm, args.mock_scale)
    log(f"{nm} mocks ready, {sum(int(v) for v in sizes)} data objects")
    eng = F.Engine(0)
    prec = args.precision
    m = mocks_run(eng, b, sizes, mocks, prec, steps=args.steps, warmup=args.warmup,
                  tune=not args.no_autotune)
    line = {
        "metric": METRIC, "value": m["mocks_per_s"], "unit": "mocks/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["batched_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"{nm} box mocks L=1000, N log-uniform in "
                               f"[{5e4 * args.mock_scale:.3g},{5e5 * args.mock_scale:.3g}], 5x randoms, "
                               "DD+DR+RR (s,mu) 40x100 -> xi_0,2,4, batched",
                   "n_data_total": m["n_data_total"], "precision": prec},
    }
    line.update({k: v for k, v in m.items() if k not in ("mocks_per_s",)})
    eng.close()
    print(json.dumps(line), flush=True)


def survey_workload(args):
    """Survey-like geometry: data 1e6 and randoms 2e7 in the comoving shell
    r in [1000, 2000] Mpc/h over a 60 x 60 deg cap, observer at the origin,
    FKP weights; weighted DD, DR, RR in (s_perp, pi) 40 x 40 bins up to 200
    Mpc/h, and w_p(s_perp).  One step = the three counts."""
    import torch
    import fcfc_amd as F
    from fcfc_amd.xi import wp
    nd, nr = int(1e6 * args.scale), int(2e7 * args.scale)
    b = F.Bins.linear(F.GEOM_SURVEY, F.BIN_SPI, 0, 200, 5, pmin=0, pmax=200, dpi=5)
    t0 = time.time()
    x, y, z, w = F.gen_survey(31, nd)
    xr, yr, zr, wr = F.gen_survey(32, nr)
    D = to_device(b, (x, y, z, w), F, torch)
    R = to_device(b, (xr, yr, zr, wr), F, torch)
    torch.cuda.synchronize()
    log(f"survey workload ready in {time.time() - t0:.1f} s: D={D.n} R={R.n}")
    eng = F.Engine(0)
    if args.shape:
        cd, zd, tl = (int(v) for v in args.shape.split(","))
        eng.set_tuning(cell_div=cd, tile=tl, zdiv=zd)
    prec = args.precision
    others = [f for f in ("fp64", "fp32", "plain") if f != prec and f not in args.skip]
    if not args.no_autotune and not args.shape:
        for p in [prec] + others:
            for name, a, c, au in (("DD", D, None, True), ("DR", D, R, False), ("RR", R, None, True)):
                log(f"autotune {p} {name}: {eng.autotune(a, c, b, isauto=au, withwt=True, precision=p)}")

    rechk = [0.0]

    def step(p):
        out, ev, km = {}, 0.0, {}
        for name, a, c, au in (("DD", D, None, True), ("DR", D, R, False), ("RR", R, None, True)):
            out[name] = eng.count_pairs(a, c, b, isauto=au, withwt=True, precision=p)
            st = eng.stats()
            ev += st["evals"]
            km[name] = st["kernel_ms"]
            if p == prec:
                rechk[0] += st["rechecked"]
        return out, ev, km

    for i in range(args.warmup):
        step(prec)
    torch.cuda.synchronize()
    t1 = time.time()
    ev_tot = 0.0
    for i in range(args.steps):
        res, ev, km = step(prec)
        ev_tot += ev
        log(f"survey step {i} done")
    torch.cuda.synchronize()
    dt = time.time() - t1
    sw, swr = D.sumw, R.sumw
    dd = 2 * res["DD"] / (sw * sw)
    dr = res["DR"] / (sw * swr)
    rr = 2 * res["RR"] / (swr * swr)
    xi = np.where(rr > 0, (dd - 2 * dr + rr) / np.where(rr > 0, rr, 1), 0.0).reshape(40, 40)
    line = {
        "metric": METRIC, "value": ev_tot / dt, "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": "survey shell r in [1000,2000] Mpc/h, 60x60 deg cap, FKP weights; "
                               "D=1e6, R=2e7; weighted DD+DR+RR (s_perp,pi) 40x40 to 200 + w_p",
                   "scale": args.scale, "n_data": D.n, "n_random": R.n, "precision": prec},
        "kernel_ms": km, "evals_per_step": ev_tot / args.steps,
        "wp_first_bins": [float(v) for v in wp(xi, b.pbin_raw)[:5]],
    }
    if prec == "fp32":
        line["rechecked_fraction"] = rechk[0] / max(ev_tot * (1 + args.warmup / max(args.steps, 1)), 1)
    line["forms"] = {}
    for other in others:
        t2 = time.time()
        rp, evp, kmp = step(other)
        torch.cuda.synchronize()
        dtp = time.time() - t2
        line["forms"][other] = {"evals_per_s": evp / dtp, "ms_per_step": dtp * 1e3, "kernel_ms": kmp,
                                "weighted_max_rel": max(float(np.max(np.abs(res[k] - rp[k]) /
                                                        np.maximum(np.abs(rp[k]), 1e-300)))
                                                        for k in res)}
        log(f"survey {other} step done")
    eng.close()
    print(json.dumps(line), flush=True)


def grad_workload(args):
    """Differentiable weights: box L = 1000, D = 1e6 clustered (C2 recipe at
    half size), R = 5e6 uniform; float32 per-object weights with gradient;
    weighted (s,mu) 40x100 counts -> xi_0,2 (Landy-Szalay); loss = sum of
    squared differences to a seeded target.  One step = forward (DD and DR
    recounted through a prepared plan, HIP graph replay) + backward (the
    gradient kernel for DD and DR) + a weight update."""
    import torch
    import fcfc_amd as F
    from fcfc_amd.xi import WeightFitCounts, WeightedPairCount, xi_multipoles_torch
    L = 1000.0
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=100, bsize=L)
    nh = max(1, int(round(1000 * args.scale)))
    x, y, z, w = F.gen_clustered(20271, int(5e5 * args.scale), nh, 500, 2.0, L, 0.3)
    xr, yr, zr = F.gen_uniform(20272, int(5e6 * args.scale), L)
    D = to_device(b, (x, y, z), F, torch)
    R = to_device(b, (xr, yr, zr), F, torch)
    wr = torch.ones(R.n, dtype=torch.float64, device="cuda")
    wd = torch.tensor(w, dtype=torch.float32, device="cuda", requires_grad=True)
    eng = F.Engine(0)
    t0 = time.time()
    rr = WeightedPairCount.apply(wr, None, eng, R, None, b, True).detach()
    log(f"RR once: {time.time() - t0:.2f} s")
    if not args.no_autotune:
        Dw = F.Catalog(b, D.x, D.y, D.z, wd.detach().to(torch.float64).contiguous())
        Rw = F.Catalog(b, R.x, R.y, R.z, wr)
        for c2, au in ((None, True), (Rw, False)):
            for g in (False, True):
                log(f"autotune {'gradient' if g else 'count'} {'DD' if au else 'DR'}: "
                    f"{eng.autotune(Dw, c2, b, isauto=au, withwt=True, gradient=g)}")
    fit = WeightFitCounts(eng, D, R, b)
    rng = np.random.default_rng(20273)
    target = torch.tensor(rng.normal(0, 0.01, (2, b.ns)), dtype=torch.float64, device="cuda")
    opt = torch.optim.SGD([wd], lr=1e-3)

    def step():
        opt.zero_grad()
        t1 = time.time()
        xi = xi_multipoles_torch(wd, wr, eng, D, R, b, rr_cnt=rr, poles=(0, 2), fit=fit)
        loss = ((xi - target) ** 2).sum()
        torch.cuda.synchronize()
        tf = time.time() - t1
        ev_f = eng.stats()["evals"]
        t2 = time.time()
        loss.backward()
        torch.cuda.synchronize()
        tb = time.time() - t2
        opt.step()
        return float(loss.item()), tf, tb, ev_f

    for i in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t3 = time.time()
    tfs, tbs = 0.0, 0.0
    for i in range(args.steps):
        loss, tf, tb, _ = step()
        tfs += tf
        tbs += tb
        log(f"fit step {i}: loss {loss:.6g} forward {tf:.2f} s backward {tb:.2f} s")
    torch.cuda.synchronize()
    dt = (time.time() - t3) / args.steps
    line = {
        "metric": METRIC, "value": 1.0 / dt, "unit": "fit steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "box L=1000; D=1e6 clustered, float32 weights with gradient; "
                               "R=5e6; weighted (s,mu) 40x100 -> xi_0,2; loss to a seeded target; "
                               "forward (plan replay) + backward (gradient kernel) + SGD step",
                   "scale": args.scale, "n_data": D.n, "n_random": R.n},
        "forward_ms": tfs / args.steps * 1e3, "backward_ms": tbs / args.steps * 1e3,
    }
    fit.close()
    eng.close()
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0,
                    help="fraction of the workload sizes (1 = the headline configuration)")
    ap.add_argument("--precision", choices=["fp64", "fp32", "plain"], default="fp64",
                    help="form of the timed steps (the others run one step each)")
    ap.add_argument("--skip", nargs="*", default=[],
                    choices=["fp64", "fp32", "plain", "census", "mocks"],
                    help="forms not to run for comparison; census: no full-size census "
                         "count; mocks: no batched-mocks leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mock-scale", type=float, default=1.0,
                    help="mocks workload: factor on the catalogue sizes")
    ap.add_argument("--workload", choices=["c2", "mocks", "survey", "grad"], default="c2",
                    help="c2: the headline clustered box DD/DR/RR; mocks: 256 batched mocks; "
                         "survey: the survey-shell (s_perp, pi) counts; grad: weight fitting "
                         "with the backward kernel")
    ap.add_argument("--copy-calib", action="store_true",
                    help="start with a 1 GiB device-to-device copy (calibrates HBM byte "
                         "counters of a profiler run)")
    ap.add_argument("--shape", default=None, metavar="CD,ZD,T",
                    help="c2 and survey workloads: fix the grid/tile shape of every count "
                         "(cells per separation, z-cells per cell, A objects per tile) instead "
                         "of autotuning, so that separate runs and forms see one shape")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default grid/tile shapes (by default the autotuner picks "
                         "them for each count and form before timing; its results are cached "
                         "in the engine like a production run's)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this "
              "build (the pair count would shard by ranges of the tile list, each GPU "
              "holding the sorted catalogues, with one histogram all-reduce at the end)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import fcfc_amd as F
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calib:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del src, dst
    if args.workload == "mocks":
        mocks_workload(args)
        return
    if args.workload == "survey":
        survey_workload(args)
        return
    if args.workload == "grad":
        grad_workload(args)
        return

    t0 = time.time()
    L, dcat, rcat = make_workload(args.scale)
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=100, bsize=L)
    D = to_device(b, dcat, F, torch)
    R = to_device(b, rcat, F, torch)
    torch.cuda.synchronize()
    log(f"workload ready in {time.time() - t0:.1f} s: D={D.n} R={R.n}")

    eng = F.Engine(0)
    if args.shape:
        cd, zd, tl = (int(v) for v in args.shape.split(","))
        eng.set_tuning(cell_div=cd, tile=tl, zdiv=zd)
        log(f"shape fixed: cells per separation {cd}, z-cells per cell {zd}, tile {tl}")

    def tune(p):
        if args.no_autotune or args.shape:
            return
        for name, a, c, au, wt in (("DD", D, None, True, True), ("DR", D, R, False, True),
                                   ("RR", R, None, True, False)):
            log(f"autotune {p} {name}: {eng.autotune(a, c, b, isauto=au, withwt=wt, precision=p)}")

    prec = args.precision
    tune(prec)
    for i in range(args.warmup):
        run_step(eng, b, D, R, prec)
        log(f"warmup step {i} done")
    torch.cuda.synchronize()
    t1 = time.time()
    evals = 0.0
    kms = {"DD": 0.0, "DR": 0.0, "RR": 0.0}
    rechk = 0.0
    res = None
    for i in range(args.steps):
        res, ev, km, rc = run_step(eng, b, D, R, prec)
        evals += ev
        rechk += rc
        for k in kms:
            kms[k] += km[k]
        log(f"step {i} done")
    torch.cuda.synchronize()
    dt = time.time() - t1
    value = evals / dt
    line = {
        "metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": "box L=1000 Mpc/h; D=2e6 clustered (1e6 uniform + 2000 halos x 500, "
                               "sigma=2) lognormal(0,0.3) weights; R=2e7 uniform; DD+DR+RR "
                               "(s,mu) 40x100 in [0,200)", "scale": args.scale,
                   "n_data": D.n, "n_random": R.n, "precision": prec},
    }
    step_evals = evals / args.steps
    line["evals_per_step"] = step_evals
    if prec == "fp32":
        line["rechecked_fraction"] = rechk / max(evals, 1)
    # dominant kernel: the RR count (unweighted)
    rr_ms = kms["RR"] / args.steps
    rr_evals = res["RR_evals"]
    peak = PEAK_FP32_VALU if prec == "fp32" else PEAK_FP64_VALU
    ach = rr_evals * FLOPS_PAIR["smu"] / (rr_ms * 1e-3)
    line["roofline"] = {"bound": "valu", "achieved": ach / 1e12, "peak": peak / 1e12,
                        "unit": "TFLOP/s", "frac": ach / peak, "traffic": traffic_from_profile(),
                        "kernel": f"k_count (s,mu) box {prec}, RR auto count", "kernel_ms": rr_ms,
                        "flops_per_pair": FLOPS_PAIR["smu"], "evals": rr_evals,
                        "evals_per_s": rr_evals / (rr_ms * 1e-3),
                        "issue_profile": issue_from_profile()}
    line["kernel_ms"] = {k: v / args.steps for k, v in kms.items()}
    # in-range pairs of RR (unordered), the algorithm-independent rate
    line["inrange_pairs_per_s_RR"] = float(res["RR"].sum()) / (rr_ms * 1e-3)
    # the same over the whole step: the unordered in-range pairs of DD and
    # DR come from one unweighted count of each, outside the timed region
    npair = float(res["RR"].sum())
    for name, a, c, au in (("DD", D, None, True), ("DR", D, R, False)):
        npair += float(eng.count_pairs(a, c, b, isauto=au, withwt=False, precision=prec).sum())
    line["inrange_pairs_per_step"] = npair
    line["inrange_pairs_per_s"] = npair * args.steps / dt

    # the other forms, one step each, on the same data
    forms = {}
    for other in [f for f in ("plain", "fp64", "fp32") if f != prec and f not in args.skip]:
        tune(other)
        t2 = time.time()
        ro, evo, kmo, rco = run_step(eng, b, D, R, other)
        torch.cuda.synchronize()
        dto = time.time() - t2
        eq = {k: bool(np.array_equal(res[k], ro[k])) for k in ("RR",)}
        wrel = max(float(np.max(np.abs(res[k] - ro[k]) / np.maximum(np.abs(ro[k]), 1e-300)))
                   for k in ("DD", "DR"))
        forms[other] = {"evals_per_s": evo / dto, "ms_per_step": dto * 1e3, "kernel_ms": kmo,
                        "RR_evals_per_s_kernel": ro["RR_evals"] / (kmo["RR"] * 1e-3),
                        "RR_identical": eq["RR"], "weighted_max_rel": wrel}
        if other == "fp32":
            forms[other]["rechecked_fraction"] = rco / max(evo, 1)
            ach32 = ro["RR_evals"] * FLOPS_PAIR["smu"] / (kmo["RR"] * 1e-3)
            forms[other]["roofline"] = {
                "bound": "valu", "achieved": ach32 / 1e12, "peak": PEAK_FP32_VALU / 1e12,
                "unit": "TFLOP/s", "frac": ach32 / PEAK_FP32_VALU,
                "kernel": "k_count (s,mu) box fp32, RR auto count",
                "issue_profile": issue_from_profile(PROFILE32)}
        log(f"{other} step done: {evo / dto:.4g} evals/s")
    line["forms"] = forms
    mine = {"evals_per_s": value, "ms_per_step": dt / args.steps * 1e3, "kernel_ms": line["kernel_ms"]}
    allf = dict(forms, **{prec: mine})
    if "plain" in allf and prec != "plain":
        line["speedup_vs_plain_fp64"] = value / allf["plain"]["evals_per_s"]
    if "fp32" in allf and "fp64" in allf:
        f32, f64 = allf["fp32"], allf["fp64"]
        line["speedup_fp32_vs_fp64_form"] = {
            "step": f64["ms_per_step"] / f32["ms_per_step"],
            "RR_kernel": f64["kernel_ms"]["RR"] / f32["kernel_ms"]["RR"],
            "DD_kernel": f64["kernel_ms"]["DD"] / f32["kernel_ms"]["DD"],
            "DR_kernel": f64["kernel_ms"]["DR"] / f32["kernel_ms"]["DR"]}
    if "plain" in allf and "fp32" in allf:
        line["speedup_fp32_vs_plain_fp64"] = {
            "step": allf["plain"]["ms_per_step"] / allf["fp32"]["ms_per_step"],
            "RR_kernel": allf["plain"]["kernel_ms"]["RR"] / allf["fp32"]["kernel_ms"]["RR"]}
    line["counts_bit_exact"] = all(f["RR_identical"] and f["weighted_max_rel"] <= 1e-12
                                   for f in forms.values())
    # the rounding census of the full-size RR count: the FP32 kernel evaluates
    # the reference's other rounding paths for every pair it re-evaluates in
    # FP64; census 0 means every path of the reference bins every pair as the
    # GPU does
    if "census" not in args.skip:
        rc = eng.count_pairs(R, None, b, isauto=True, withwt=False, precision="fp32", census=True)
        st = eng.stats()
        line["census_RR_full_size"] = {"census": st["census"], "rechecked": st["rechecked"],
                                       "evals": st["evals"], "kernel_ms": st["kernel_ms"],
                                       "counts_identical": bool(np.array_equal(rc, res["RR"]))}
        log(f"census RR: {st['census']} of {st['rechecked']} re-evaluated pairs")
    # batched mocks at the configured sizes (256 mocks, N in [5e4, 5e5])
    if "mocks" not in args.skip:
        nm = max(1, int(round(256 * args.scale)))
        del D, R
        torch.cuda.empty_cache()
        sizes, mocks = gpu_mocks(b, nm, args.mock_scale)
        log(f"mocks: {nm} ready, {int(sum(int(v) for v in sizes))} data objects")
        line["mocks"] = mocks_run(eng, b, sizes, mocks, prec, tune=not args.no_autotune)
        line["mocks"]["config"] = (f"{nm} box mocks L=1000, N log-uniform in [{5e4 * args.mock_scale:.3g},"
                                   f"{5e5 * args.mock_scale:.3g}], 5x randoms, DD+DR+RR (s,mu) "
                                   "40x100 -> xi_0,2,4, one batched call vs one call per count")
        del mocks
        torch.cuda.empty_cache()
    base = None if args.no_cpu else cpu_baseline()
    line["cpu_baseline"] = base["record"] if base else None
    if base:
        line["parity_vs_reference"] = parity_vs_reference(
            eng, base, forms=tuple(f for f in ("fp64", "fp32", "plain") if f not in args.skip))
        log("parity vs reference: " + json.dumps(
            {k: {f: v[f]["identical"] for f in v if isinstance(v[f], dict)}
             for k, v in line["parity_vs_reference"].items()}))
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
