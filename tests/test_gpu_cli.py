"""The drop-in executables (FCFC_2PT_BOX_GPU / FCFC_2PT_GPU) end to end:
same configuration keys and command line as FCFC, same pair-count files.
Expected counts come from the upstream binaries (tests/golden/)."""
import json
import os
import subprocess

import numpy as np
import pytest

import fcfc_amd as F
from oracle import oracle as O
from tests.golden.make_golden import catalogs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fcfc_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")


def _write(tmp, cats, wts):
    files = []
    for k, (c, wt) in enumerate(zip(cats, wts)):
        fn = os.path.join(tmp, f"cat{k}.txt")
        O.write_ascii(fn, c if wt else c[:3])
        files.append(fn)
    return files


def test_box_cli_matches_reference(tmp_path):
    g = json.load(open(os.path.join(GOLD, "box_smu_clustered.json")))
    cats, wts = catalogs("box_clustered")
    files = _write(str(tmp_path), cats, wts)
    conf = tmp_path / "fcfc_2pt_box.conf"
    conf.write_text(
        f"CATALOG = [{files[0]}, {files[1]}]\nCATALOG_LABEL = [D, R]\n"
        'ASCII_FORMATTER = ["%lf %lf %lf %lf", "%lf %lf %lf"]\n'
        "POSITION = [${1}, ${2}, ${3}, ${1}, ${2}, ${3}]\nWEIGHT = [${4}, 1]\n"
        "BOX_SIZE = 1000\nBINNING_SCHEME = 1\nPAIR_COUNT = [DD, DR, RR]\n"
        f"PAIR_COUNT_FILE = [{tmp_path}/DD.bin, {tmp_path}/DR.bin, {tmp_path}/RR.bin]\n"
        "CF_ESTIMATOR = (DD - 2*DR + RR)/RR\n"
        f"CF_OUTPUT_FILE = {tmp_path}/xi.txt\nMULTIPOLE = [0, 2, 4]\n"
        f"MULTIPOLE_FILE = {tmp_path}/xil.txt\n"
        "SEP_BIN_MIN = 0\nSEP_BIN_MAX = 200\nSEP_BIN_SIZE = 5\nMU_BIN_NUM = 100\n"
        "OUTPUT_FORMAT = 0\nOVERWRITE = 2\n")
    for prec in ("0", "1"):
        r = subprocess.run([os.path.join(BIN, "FCFC_2PT_BOX_GPU"), "-c", str(conf), "--gpu-prec", prec],
                           capture_output=True, text=True, cwd=str(tmp_path), timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        for p in ("DD", "DR", "RR"):
            ref = g["results"][f"avx512:{p}"]
            mine = O.read_pair_count(str(tmp_path / f"{p}.bin"))
            assert list(mine["spec"]) == ref["spec"]
            assert mine["header"] == ref["header"]
            np.testing.assert_array_equal(mine["sbin"], ref["sbin"])
            np.testing.assert_array_equal(mine["bsize"], ref["bsize"])
            assert mine["norm"] == pytest.approx(ref["norm"], rel=1e-15)
            rc = np.array(ref["cnt"])
            if ref["spec"][4]:
                np.testing.assert_allclose(mine["cnt"], rc, rtol=1e-12)
            else:
                np.testing.assert_array_equal(mine["cnt"], rc.astype(np.int64))
        xi = np.loadtxt(tmp_path / "xi.txt")
        assert xi.shape == (4000, 5)
        xil = np.loadtxt(tmp_path / "xil.txt")
        assert xil.shape == (40, 6)
        # multipoles from xi(s,mu) with the midpoint rule (eval_cf.c:326-354)
        from fcfc_amd.xi import multipoles
        ml = multipoles(xi[:, 4].reshape(100, 40), 100, (0, 2, 4))
        # xi.txt holds 10 significant digits (%.10lg)
        np.testing.assert_allclose(xil[:, 3:].T, ml, rtol=1e-7,
                                   atol=1e-9 * np.abs(xi[:, 4]).max())


def test_survey_cli_matches_reference(tmp_path):
    g = json.load(open(os.path.join(GOLD, "survey_spi.json")))
    cats, wts = catalogs("survey")
    files = _write(str(tmp_path), cats, wts)
    args = [os.path.join(BIN, "FCFC_2PT_GPU"), "-i", f"[{files[0]},{files[1]}]", "-l", "[D,R]",
            "-f", '["%lf %lf %lf %lf","%lf %lf %lf %lf"]',
            "-x", "[${1},${2},${3},${1},${2},${3}]", "-w", "[${4},${4}]",
            "-B", "2", "-p", "[DD,DR]", "-P", f"[{tmp_path}/DD.bin,{tmp_path}/DR.bin]",
            "--s-min", "0", "--s-max", "200", "--s-step", "5",
            "--pi-min", "0", "--pi-max", "200", "--pi-step", "5", "-F", "0", "-O", "2"]
    r = subprocess.run(args, capture_output=True, text=True, cwd=str(tmp_path), timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    for p in ("DD", "DR"):
        ref = g["results"][f"avx512:{p}"]
        mine = O.read_pair_count(str(tmp_path / f"{p}.bin"))
        assert list(mine["spec"]) == ref["spec"]
        np.testing.assert_allclose(mine["cnt"], np.array(ref["cnt"]), rtol=1e-12)


def test_cli_rejects_unsupported(tmp_path):
    r = subprocess.run([os.path.join(BIN, "FCFC_2PT_BOX_GPU"), "-i", "x.txt", "-f", "%lf %lf %lf",
                        "-x", "[${1}*2,${2},${3}]", "-b", "1000", "-p", "DD", "-P", "o.bin",
                        "--s-min", "0", "--s-max", "200", "--s-step", "5"],
                       capture_output=True, text=True, cwd=str(tmp_path), timeout=60)
    assert r.returncode != 0
