"""Writes the golden pair-count fixtures under tests/golden/.

Runs the upstream FCFC executables (built from the unmodified sources into
oracle/_ref/ by `make -C oracle ref`) on catalogues from the project's seeded
generator and stores, per case, the input checksum and the reference's
outputs (counts, normalisation, bin edges, box size as written in its binary
pair-count file).  The inputs themselves are not stored: the tests
regenerate them from the seeds and check the checksum first.

Each PAIR_COUNT runs in its own invocation: requesting DD, DR and RR in one
run with a weighted D and an unweighted R makes the upstream binary crash.

Usage:  python tests/golden/make_golden.py          (needs oracle/_ref built)
        python tests/golden/make_golden.py --sizes  (the cfg_* cases at the
                                                     configurations' sizes)
        python tests/golden/make_golden.py --cli    (the cli_* cases: outputs
                                                     of the executables)
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import fcfc_amd as F                     # noqa: E402  (generator only)
from oracle import oracle as O           # noqa: E402

L = 1000.0


def checksum(*arrs):
    h = hashlib.sha256()
    for a in arrs:
        if a is not None:
            h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def catalogs(kind):
    if kind == "box_clustered":
        x, y, z, w = F.gen_clustered(11, 3000, 30, 40, 2.0, L, 0.3)
        xr, yr, zr = F.gen_uniform(12, 6000, L)
        return [(x, y, z, w), (xr, yr, zr)], [True, False]
    if kind == "box_uniform20k":
        return [F.gen_uniform(31, 20000, L)], [False]
    if kind == "box_tiny300":
        return [F.gen_uniform(41, 300, L)], [False]
    if kind == "c1_full":
        # the C1 configuration as specified: N = 1e5 uniform in the L = 1000 box
        return [F.gen_uniform(1001, 100000, L)], [False]
    if kind == "c2_twentieth":
        # the C2 recipe at 1/20 of its size: D = 5e4 uniform + 100 halos of
        # 500 members (sigma = 2), lognormal(0, 0.3) weights; R = 1e6 uniform
        x, y, z, w = F.gen_clustered(2001, 50000, 100, 500, 2.0, L, 0.3)
        xr, yr, zr = F.gen_uniform(2002, 1000000, L)
        return [(x, y, z, w), (xr, yr, zr)], [True, False]
    if kind == "survey":
        x, y, z, w = F.gen_survey(21, 4000)
        xr, yr, zr, wr = F.gen_survey(22, 6000)
        return [(x, y, z, w), (xr, yr, zr, wr)], [True, True]
    raise KeyError(kind)


CASES = [
    # name, code, catalogue kind, pairs, binning kwargs, flavours
    ("box_iso_uniform20k", "FCFC_2PT_BOX", "box_uniform20k", ["DD"],
     dict(bintype=0, smin=0, smax=200, ds=5), ["avx512", "scalar"]),
    ("box_iso_tiny300", "FCFC_2PT_BOX", "box_tiny300", ["DD"],
     dict(bintype=0, smin=0, smax=200, ds=5), ["scalar"]),
    ("box_iso_clustered", "FCFC_2PT_BOX", "box_clustered", ["DD", "DR", "RR"],
     dict(bintype=0, smin=0, smax=200, ds=5), ["avx512", "scalar"]),
    ("box_smu_clustered", "FCFC_2PT_BOX", "box_clustered", ["DD", "DR", "RR"],
     dict(bintype=1, smin=0, smax=200, ds=5, nmu=100), ["avx512", "scalar"]),
    ("box_spi_clustered", "FCFC_2PT_BOX", "box_clustered", ["DD", "DR", "RR"],
     dict(bintype=2, smin=0, smax=200, ds=5, pmin=0, pmax=200, dpi=5), ["avx512", "scalar"]),
    ("box_smu_hybrid", "FCFC_2PT_BOX", "box_clustered", ["DD", "DR"],
     dict(bintype=1, smin=0.5, smax=146.5, ds=7.3, nmu=7), ["avx512"]),
    ("box_smu_smin", "FCFC_2PT_BOX", "box_clustered", ["DR", "RR"],
     dict(bintype=1, smin=20, smax=150, ds=10, nmu=10), ["avx512"]),
    ("box_spi_pmin", "FCFC_2PT_BOX", "box_clustered", ["DD", "DR"],
     dict(bintype=2, smin=10, smax=100, ds=10, pmin=20, pmax=120, dpi=20), ["avx512"]),
    ("survey_iso", "FCFC_2PT", "survey", ["DD", "DR"],
     dict(bintype=0, smin=0, smax=200, ds=5), ["avx512", "scalar"]),
    ("survey_smu", "FCFC_2PT", "survey", ["DD", "DR"],
     dict(bintype=1, smin=0, smax=200, ds=5, nmu=100), ["avx512", "scalar"]),
    ("survey_spi", "FCFC_2PT", "survey", ["DD", "DR"],
     dict(bintype=2, smin=0, smax=200, ds=5, pmin=0, pmax=200, dpi=5), ["avx512", "scalar"]),
]


# Cases at the configurations' own sizes (kept apart from CASES: the oracle
# census over their 1e8 - 1e10 pairs is computed once, here, and stored)
SIZE_CASES = [
    ("cfg_c1_full", "FCFC_2PT_BOX", "c1_full", ["DD"],
     dict(bintype=0, smin=0, smax=200, ds=5), ["avx512", "scalar"]),
    ("cfg_c2_twentieth", "FCFC_2PT_BOX", "c2_twentieth", ["DD", "DR", "RR"],
     dict(bintype=1, smin=0, smax=200, ds=5, nmu=100), ["avx512"]),
]


def oracle_census(kind, pairs, bk):
    """Unweighted oracle counts (variant 1, the GPU's rounding path) and the
    census of each pair: in-range pairs whose bin differs between the
    reference's paths (variants 0-4)."""
    cats, wts = catalogs(kind)
    b = F.Bins.linear(F.GEOM_BOX, bk["bintype"], bk["smin"], bk["smax"], bk["ds"],
                      nmu=bk.get("nmu", 1), bsize=L)
    ob = O.oracle_bins(b.bintype, b.s2bin, b.nmu)
    prepped = [F.Catalog(b, c[0], c[1], c[2]) for c in cats]
    out = {}
    for p in pairs:
        A, B = prepped["DR".index(p[0])], prepped["DR".index(p[1])]
        isauto = p[0] == p[1]
        _, cen = O.count(0, b.bsize_rescaled, (A.x, A.y, A.z, None, None),
                         None if isauto else (B.x, B.y, B.z, None, None), isauto, ob,
                         variant=1, census_mask=0b11111)
        out[p] = int(cen)
        print(f"  census {p}: {cen}", flush=True)
    return out


def run_case(name, code, kind, pairs, bk, flavours):
    cats, wts = catalogs(kind)
    out = {"name": name, "code": code, "catalogs": kind, "binning": bk,
           "checksums": [checksum(*c) for c in cats], "results": {}}
    for fl in flavours:
        for p in pairs:
            idx = sorted(set("DR".index(ch) for ch in p))
            sub = [cats[i] for i in idx]
            labels = ["DR"[i] for i in idx]
            res, phase, wall, _ = O.run_reference(
                code, fl, sub, labels, [p], bk["bintype"], bk["smin"], bk["smax"], bk["ds"],
                nmu=bk.get("nmu"), pmin=bk.get("pmin"), pmax=bk.get("pmax"), dpi=bk.get("dpi"),
                bsize=L if code == "FCFC_2PT_BOX" else None,
                weights=[wts[i] for i in idx], threads=1 if fl == "scalar" else 8)
            r = res[p]
            ent = {"header": r["header"], "spec": r["spec"].tolist(),
                   "sbin": r["sbin"].tolist(), "norm": float(r["norm"]),
                   "num": r["num"].tolist(),
                   "cnt": r["cnt"].tolist()}
            if "bsize" in r:
                ent["bsize"] = r["bsize"].tolist()
            if "pbin" in r:
                ent["pbin"] = r["pbin"].tolist()
            out["results"][f"{fl}:{p}"] = ent
    return out


# Cases of the drop-in executables: a configuration template ({d} = the
# working directory holding the catalogues cat0.txt, cat1.txt, ...) run by
# the upstream executable; the fixture keeps its output files, parsed.
def cli_catalogs(kind):
    """Text columns of the catalogues of a CLI case (lists of arrays)."""
    if kind == "box_plain":
        (x, y, z, w), (xr, yr, zr) = catalogs("box_clustered")[0]
        return [[x, y, z], [xr, yr, zr]]
    if kind == "box_expr":
        # x, y, z, w plus a uniform flag column: POSITION / WEIGHT /
        # SELECTION expressions over five columns
        (x, y, z, w), _ = catalogs("box_clustered")[0]
        flag = np.random.default_rng(51).random(len(x))
        return [[x, y, z, w, flag]]
    if kind == "box_types":
        # float, double, float, int, a string column, float, a suppressed
        # column: typed ASCII_FORMATTER conversions and integer SELECTION
        (x, y, z, w), _ = catalogs("box_clustered")[0]
        g = np.random.default_rng(52)
        ids = g.integers(0, 10, len(x)).astype(float)
        junk = g.random(len(x))
        return [[x, y, z, ids, junk, w, junk]]
    if kind == "survey_radec":
        # RA, Dec in degrees over a 10 x 10 deg patch, redshift in [0.4, 0.45]
        g = np.random.default_rng(61)
        out = []
        for n in (3000, 6000):
            out.append([g.uniform(-5, 5, n), g.uniform(-5, 5, n), g.uniform(0.4, 0.45, n)])
        return out
    raise KeyError(kind)


def eds_table():
    """(z, comoving distance) samples of the Einstein-de Sitter closed form
    D(z) = 2 (c / H0) (1 - 1 / sqrt(1 + z)), c / H0 = 2997.92458 Mpc/h, for
    the Z_CMVDST_FILE case."""
    z = np.linspace(0.3, 0.6, 61)
    return z, 2 * 2997.92458 * (1 - 1 / np.sqrt(1 + z))


SMU_CF = ("CATALOG = [{d}/cat0.txt, {d}/cat1.txt]\nCATALOG_LABEL = [D, R]\n"
          'ASCII_FORMATTER = ["%lf %lf %lf", "%lf %lf %lf"]\n'
          "POSITION = [$1, $2, $3, $1, $2, $3]\nBOX_SIZE = 1000\nBINNING_SCHEME = 1\n"
          "PAIR_COUNT = [DD, DR, RR]\nPAIR_COUNT_FILE = [{d}/DD.txt, {d}/DR.txt, {d}/RR.txt]\n"
          "CF_ESTIMATOR = [(DD - 2*DR + RR)/RR, DD/@@ - 1]\n"
          "CF_OUTPUT_FILE = [{d}/xi_ls.txt, {d}/xi_nat.txt]\nMULTIPOLE = [4, 0, 2, 2]\n"
          "MULTIPOLE_FILE = [{d}/mp_ls.txt, {d}/mp_nat.txt]\n"
          "SEP_BIN_MIN = 0\nSEP_BIN_MAX = 150\nSEP_BIN_SIZE = 5\nMU_BIN_NUM = 20\n"
          "OUTPUT_FORMAT = 1\nOVERWRITE = 2\n")
SPI_WP = ("CATALOG = [{d}/cat0.txt, {d}/cat1.txt]\nCATALOG_LABEL = [D, R]\n"
          'ASCII_FORMATTER = ["%lf %lf %lf", "%lf %lf %lf"]\n'
          "POSITION = [$1, $2, $3, $1, $2, $3]\nBOX_SIZE = 1000\nBINNING_SCHEME = 2\n"
          "PAIR_COUNT = [DD, DR, RR]\nPAIR_COUNT_FILE = [{d}/DD.bin, {d}/DR.bin, {d}/RR.bin]\n"
          "CF_ESTIMATOR = (DD - 2*DR + RR)/RR\nCF_OUTPUT_FILE = {d}/xi.txt\n"
          "PROJECTED_CF = T\nPROJECTED_FILE = {d}/wp.txt\n"
          "SEP_BIN_MIN = 0\nSEP_BIN_MAX = 100\nSEP_BIN_SIZE = 5\n"
          "PI_BIN_MIN = 0\nPI_BIN_MAX = 60\nPI_BIN_SIZE = 4\nOUTPUT_FORMAT = 0\nOVERWRITE = 2\n")
EXPR = ("CATALOG = {d}/cat0.txt\nCATALOG_LABEL = D\n"
        'ASCII_FORMATTER = "%lf %lf %lf %lf %lf"\n'
        "POSITION = [$1*0.5 + 250, $2, ($3 + 100) % 1000]\nWEIGHT = $4 * 2 - 0.5*$5\n"
        "SELECTION = $5 < 0.7 && ($4 > 0.8 || $1 < 300)\n"
        "BOX_SIZE = 1000\nBINNING_SCHEME = 1\nPAIR_COUNT = DD\nPAIR_COUNT_FILE = {d}/DD.bin\n"
        "SEP_BIN_MIN = 0\nSEP_BIN_MAX = 200\nSEP_BIN_SIZE = 5\nMU_BIN_NUM = 10\n"
        "OUTPUT_FORMAT = 0\nOVERWRITE = 2\n")
TYPES = ("CATALOG = {d}/cat0.txt\nCATALOG_LABEL = D\n"
         'ASCII_FORMATTER = "%f %lf %f %d %s %f %*lf"\n'
         "POSITION = [$1, $2, ($3 + $4 / 3) % 1000]\nWEIGHT = $6 * 2\n"
         "SELECTION = ($4 / 3 == 1 || $4 % 4 == 0) && 2 ** $4 < 200 && $6 > 0.4\n"
         "BOX_SIZE = 1000\nBINNING_SCHEME = 1\nPAIR_COUNT = DD\nPAIR_COUNT_FILE = {d}/DD.bin\n"
         "SEP_BIN_MIN = 0\nSEP_BIN_MAX = 200\nSEP_BIN_SIZE = 5\nMU_BIN_NUM = 10\n"
         "OUTPUT_FORMAT = 0\nOVERWRITE = 2\n")
CONV = ("CATALOG = [{d}/cat0.txt, {d}/cat1.txt]\nCATALOG_LABEL = [D, R]\n"
        'ASCII_FORMATTER = ["%lf %lf %lf", "%lf %lf %lf"]\n'
        "POSITION = [$1, $2, $3, $1, $2, $3]\nCOORD_CONVERT = T\n{cosmo}"
        "BINNING_SCHEME = 1\nPAIR_COUNT = [DD, DR]\nPAIR_COUNT_FILE = [{d}/DD.bin, {d}/DR.bin]\n"
        "SEP_BIN_MIN = 0\nSEP_BIN_MAX = 100\nSEP_BIN_SIZE = 5\nMU_BIN_NUM = 10\n"
        "OUTPUT_FORMAT = 0\nOVERWRITE = 2\n")
CLI_CASES = [
    # name, code, catalogue kind, template, outputs
    ("box_smu_cf", "FCFC_2PT_BOX", "box_plain", SMU_CF,
     ["DD.txt", "DR.txt", "RR.txt", "xi_ls.txt", "xi_nat.txt", "mp_ls.txt", "mp_nat.txt"]),
    ("box_spi_wp", "FCFC_2PT_BOX", "box_plain", SPI_WP, ["DD.bin", "DR.bin", "RR.bin", "xi.txt", "wp.txt"]),
    ("box_expr", "FCFC_2PT_BOX", "box_expr", EXPR, ["DD.bin"]),
    ("box_types", "FCFC_2PT_BOX", "box_types", TYPES, ["DD.bin"]),
    ("survey_convert_integral", "FCFC_2PT", "survey_radec",
     CONV.replace("{cosmo}", "OMEGA_M = 0.31\nDE_EOS_W = -0.9\nCMVDST_ERR = 1e-10\n"), ["DD.bin", "DR.bin"]),
    ("survey_convert_table", "FCFC_2PT", "survey_radec",
     CONV.replace("{cosmo}", "Z_CMVDST_FILE = {d}/zdist.txt\n"), ["DD.bin", "DR.bin"]),
]


def write_cli_inputs(kind, d):
    """Catalogue files (and the distance table) of a CLI case into d."""
    for k, cols in enumerate(cli_catalogs(kind)):
        O.write_ascii(os.path.join(d, f"cat{k}.txt"), cols)
    z, dist = eds_table()
    O.write_ascii(os.path.join(d, "zdist.txt"), [z, dist])


def read_output(path):
    """A pair-count file (binary or ASCII) or a whitespace table."""
    with open(path, "rb") as f:
        head = f.readline().decode(errors="replace")
    if "format = 0" in head:
        r = O.read_pair_count(path)
        return {"kind": "pc", "spec": r["spec"].tolist(), "norm": float(r["norm"]),
                "num": r["num"].tolist(), "ncnt": r["ncnt"].tolist(), "cnt": r["cnt"].tolist()}
    lines = open(path).read().splitlines()
    body = [l for l in lines if l and not l.startswith("#")]
    return {"kind": "table", "header": [l for l in lines if l.startswith("#")], "text": body}


def run_cli_case(name, code, kind, template, outs):
    import subprocess
    import tempfile
    d = tempfile.mkdtemp(prefix="fcfc_cli_")
    write_cli_inputs(kind, d)
    conf = os.path.join(d, "fcfc.conf")
    with open(conf, "w") as f:
        f.write(template.replace("{d}", d))
    exe = O.ref_binary(code, "avx512")
    r = subprocess.run([exe, "-c", conf], cwd=d, capture_output=True, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="8"))
    if r.returncode:
        raise RuntimeError(f"{name}: reference failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    res = {o: read_output(os.path.join(d, o)) for o in outs}
    shutil.rmtree(d)
    return {"name": name, "code": code, "catalogs": kind, "template": template, "outputs": res,
            "checksums": [checksum(*c) for c in cli_catalogs(kind)]}


def main():
    if "--cli" in sys.argv:
        only = [a for a in sys.argv[1:] if not a.startswith("--")]
        for case in CLI_CASES:
            if only and case[0] not in only:
                continue
            print("cli case", case[0], flush=True)
            data = run_cli_case(*case)
            with open(os.path.join(HERE, "cli_" + case[0] + ".json"), "w") as fp:
                json.dump(data, fp, separators=(",", ":"))
        return
    if "--sizes" in sys.argv:
        for f in SIZE_CASES:
            print("case", f[0], flush=True)
            data = run_case(*f)
            data["census"] = oracle_census(f[2], f[3], f[4])
            with open(os.path.join(HERE, f[0] + ".json"), "w") as fp:
                json.dump(data, fp, separators=(",", ":"))
        return
    for f in [c for c in CASES]:
        name = f[0]
        print("case", name, flush=True)
        data = run_case(*f)
        with open(os.path.join(HERE, name + ".json"), "w") as fp:
            json.dump(data, fp, separators=(",", ":"))
    # one reference binary and one ASCII pair-count file, for the writer test
    cats, wts = catalogs("box_clustered")
    for ofmt, ext in ((0, "bin"), (1, "txt")):
        tmp = os.path.join(HERE, "_tmp")
        os.makedirs(tmp, exist_ok=True)
        res, _, _, _ = O.run_reference("FCFC_2PT_BOX", "scalar", [cats[0], cats[1]], ["D", "R"],
                                       ["DR"], 1, 0, 200, 5, nmu=100, bsize=L,
                                       weights=wts, threads=1, workdir=tmp)
        src = os.path.join(tmp, "pc_DR.bin")
        if ofmt == 1:
            # rerun with the ASCII output format
            exe = O.ref_binary("FCFC_2PT_BOX", "scalar")
            import subprocess
            args = [exe, "-i", f"[{tmp}/cat0.txt,{tmp}/cat1.txt]", "-l", "[D,R]",
                    "-f", '["%lf %lf %lf %lf","%lf %lf %lf"]',
                    "-x", "[${1},${2},${3},${1},${2},${3}]", "-w", "[${4},1]",
                    "-B", "1", "-p", "[DR]", "-P", f"[{tmp}/pc_DR.txt]",
                    "--s-min", "0", "--s-max", "200", "--s-step", "5", "--mu-num", "100",
                    "-b", "1000", "-F", "1", "-O", "2"]
            subprocess.run(args, cwd=tmp, check=True, capture_output=True)
            src = os.path.join(tmp, "pc_DR.txt")
        shutil.copy(src, os.path.join(HERE, f"ref_box_smu_DR.{ext}"))
        shutil.rmtree(tmp)
    print("done")


if __name__ == "__main__":
    main()
