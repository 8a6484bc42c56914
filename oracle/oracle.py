"""Python side of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() (as the checker) and bench.py's
cpu_baseline leg use this module.  It offers
  * count()/weight_grad(): the C restatement in fcfc_oracle.c (ctypes);
  * run_reference(): runs an upstream FCFC executable built from the
    unmodified sources into oracle/_ref/ and parses its binary pair-count
    file (layout of src/fcfc/2pt_box/save_res.c:177-296);
  * read_pair_count(): reader for FCFC binary pair-count files.
"""
import ctypes as C
import os
import re
import struct
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
REF = os.path.join(HERE, "_ref")

_dp = C.POINTER(C.c_double)


class orc_bins(C.Structure):
    _fields_ = [("bintype", C.c_int), ("ns", C.c_int), ("nmu", C.c_int), ("np", C.c_int),
                ("s2bin", _dp), ("pbin", _dp)]


class orc_cat(C.Structure):
    _fields_ = [("n", C.c_int64), ("x", _dp), ("y", _dp), ("z", _dp), ("s", _dp), ("w", _dp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", HERE], check=True, capture_output=True)
        _lib = C.CDLL(LIB)
        _lib.orc_count.restype = C.c_int
        _lib.orc_count.argtypes = [C.c_int, _dp, C.POINTER(orc_cat), C.POINTER(orc_cat), C.c_int,
                                   C.POINTER(orc_bins), C.c_int, C.c_int,
                                   C.POINTER(C.c_int64), _dp, C.POINTER(C.c_int64), C.c_int]
        _lib.orc_weight_grad.restype = C.c_int
        _lib.orc_weight_grad.argtypes = [C.c_int, _dp, C.POINTER(orc_cat), C.POINTER(orc_cat),
                                         C.c_int, C.POINTER(orc_bins), _dp, _dp, _dp, C.c_int]
        _lib.orc_survey_norm2.restype = None
        _lib.orc_survey_norm2.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_int, _dp]
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class _Cat:
    def __init__(self, x, y, z, s=None, w=None):
        self.arrs = [np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
                     for v in (x, y, z, s, w)]
        x, y, z, s, w = self.arrs
        self.c = orc_cat(len(x), _p(x), _p(y), _p(z), _p(s), _p(w))


def oracle_bins(bintype, s2bin, nmu=1, pbin=None):
    s2 = np.ascontiguousarray(s2bin, dtype=np.float64)
    pb = np.ascontiguousarray(pbin, dtype=np.float64) if pbin is not None else None
    b = orc_bins(bintype, len(s2) - 1, nmu, (len(pb) - 1) if pb is not None else 0, _p(s2), _p(pb))
    b._keep = (s2, pb)
    return b


def count(geom, bsize, A, B, isauto, bins, variant=0, census_mask=0, weighted=False,
          nthread=8):
    """Counts with the reference arithmetic variant (see fcfc_oracle.c).
    A, B: tuples (x, y, z, s, w) of RESCALED coordinates.  Returns
    (counts, census)."""
    ntot = bins.ns * (bins.nmu if bins.bintype == 1 else (bins.np if bins.bintype == 2 else 1))
    ca = _Cat(*A)
    cb = _Cat(*B) if B is not None else ca
    icnt = np.zeros(ntot, dtype=np.int64)
    dcnt = np.zeros(ntot, dtype=np.float64)
    cens = C.c_int64(0)
    bs = np.ascontiguousarray(bsize if bsize is not None else [0, 0, 0], dtype=np.float64)
    e = lib().orc_count(geom, _p(bs), C.byref(ca.c), C.byref(cb.c), int(isauto), C.byref(bins),
                        variant, census_mask, icnt.ctypes.data_as(C.POINTER(C.c_int64)),
                        _p(dcnt), C.byref(cens), nthread)
    if e:
        raise RuntimeError("oracle failed")
    return (dcnt if weighted else icnt), cens.value


def weight_grad(geom, bsize, A, B, isauto, bins, G, nthread=8):
    ca = _Cat(*A)
    cb = _Cat(*B) if B is not None else ca
    ga = np.zeros(len(ca.arrs[0]))
    gb = np.zeros(len(cb.arrs[0]))
    G = np.ascontiguousarray(G, dtype=np.float64)
    bs = np.ascontiguousarray(bsize if bsize is not None else [0, 0, 0], dtype=np.float64)
    e = lib().orc_weight_grad(geom, _p(bs), C.byref(ca.c), C.byref(cb.c), int(isauto),
                              C.byref(bins), _p(G), _p(ga), _p(gb), nthread)
    if e:
        raise RuntimeError("oracle failed")
    return ga, (None if isauto else gb)


def survey_norm2(x, y, z, lanes=8):
    x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
    s = np.empty(len(x))
    lib().orc_survey_norm2(len(x), _p(x), _p(y), _p(z), lanes, _p(s))
    return s


# ---------------------------------------------------------------------------
def read_pair_count(fname):
    """Parse an FCFC binary pair-count file (save_res.c:177-296)."""
    with open(fname, "rb") as f:
        raw = f.read()
    nl = raw.index(b"\n")
    header = raw[:nl].decode()
    pos = nl + 1
    blocks = []
    while pos < len(raw):
        (sz,) = struct.unpack_from("<Q", raw, pos)
        pos += 8
        blocks.append(raw[pos:pos + sz])
        pos += sz
        (sz2,) = struct.unpack_from("<Q", raw, pos)
        assert sz2 == sz, "corrupt block"
        pos += 8
    spec = np.frombuffer(blocks[0], dtype=np.int32)
    out = {"header": header, "spec": spec, "blocks": blocks}
    i = 1
    periodic = spec[2] == 1
    if periodic:
        out["bsize"] = np.frombuffer(blocks[i], dtype=np.float64); i += 1
    out["sbin"] = np.frombuffer(blocks[i], dtype=np.float64); i += 1
    if spec[5] == 2:
        out["pbin"] = np.frombuffer(blocks[i], dtype=np.float64); i += 1
    wt = spec[4] == 1
    out["num"] = np.frombuffer(blocks[i], dtype=np.float64 if wt else np.uint64); i += 1
    out["norm"] = np.frombuffer(blocks[i], dtype=np.float64)[0]; i += 1
    out["ncnt"] = np.frombuffer(blocks[i], dtype=np.float64); i += 1
    out["cnt"] = np.frombuffer(blocks[i], dtype=np.float64 if wt else np.int64); i += 1
    return out


def ref_binary(name, flavour="scalar"):
    """Path of an upstream executable built by `make -C oracle ref`:
    name in {FCFC_2PT_BOX, FCFC_2PT}, flavour in {scalar, avx2, avx512}."""
    return os.path.join(REF, f"{name}.{flavour}")


def host_flavour():
    """The SIMD flavour of the reference this host can run."""
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return "scalar"
    if " avx512f " in flags and " avx512dq " in flags:
        return "avx512"
    if " avx2 " in flags and " fma " in flags:
        return "avx2"
    return "scalar"


def write_ascii(fname, cols):
    arr = np.column_stack(cols)
    np.savetxt(fname, arr, fmt="%.17g")


def run_reference(name, flavour, catalogs, labels, pairs, bintype, smin, smax, ds,
                  nmu=None, pmin=None, pmax=None, dpi=None, bsize=None, weights=None,
                  threads=None, workdir=None, timeout=3600, extra_env=None,
                  reuse_files=False):
    """Run the upstream executable on raw (unrescaled) catalogues and return
    ({pair: parsed pair-count file}, {pair: count-phase seconds}, wall).

    catalogs: list of (x, y, z[, w]) arrays; weights: list of bools.
    reuse_files: keep catalogue files already written into workdir (repeated
    timing runs of the same catalogues)."""
    exe = ref_binary(name, flavour)
    tmp = workdir or tempfile.mkdtemp(prefix="fcfc_ref_")
    files = []
    for k, cat in enumerate(catalogs):
        fn = os.path.join(tmp, f"cat{k}.txt")
        if not (reuse_files and os.path.exists(fn)):
            write_ascii(fn, cat)
        files.append(fn)
    outs = [os.path.join(tmp, f"pc_{p}.bin") for p in pairs]
    for o in outs:
        if os.path.exists(o):
            os.remove(o)
    args = [exe, "-i", "[" + ",".join(files) + "]", "-l", "[" + ",".join(labels) + "]",
            "-f", "[" + ",".join(("\"%lf %lf %lf %lf\"" if (weights and weights[k]) else "\"%lf %lf %lf\"") for k in range(len(files))) + "]",
            "-x", "[" + ",".join("${1},${2},${3}" for _ in files) + "]",
            "-B", str(bintype), "-p", "[" + ",".join(pairs) + "]",
            "-P", "[" + ",".join(outs) + "]", "--s-min", repr(float(smin)),
            "--s-max", repr(float(smax)), "--s-step", repr(float(ds)), "-F", "0", "-O", "2"]
    if weights and any(weights):
        args += ["-w", "[" + ",".join("${4}" if w else "1" for w in weights) + "]"]
    if nmu is not None:
        args += ["--mu-num", str(nmu)]
    if pmin is not None:
        args += ["--pi-min", repr(float(pmin)), "--pi-max", repr(float(pmax)),
                 "--pi-step", repr(float(dpi))]
    if bsize is not None:
        args += ["-b", repr(float(bsize))]
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    if extra_env:
        env.update(extra_env)
    # the reference reads fcfc_2pt_box.conf from the cwd when present: use the
    # empty temporary directory
    t0 = time.time()
    p = subprocess.Popen(["stdbuf", "-oL"] + args, cwd=tmp, env=env, stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True)
    stamps = []
    for line in p.stdout:
        stamps.append((time.time(), line))
    rc = p.wait(timeout=timeout)
    wall = time.time() - t0
    text = "".join(l for _, l in stamps)
    if rc != 0:
        raise RuntimeError(f"reference failed ({rc}):\n{text[-3000:]}")
    # count-phase time: from the "Counting XY pairs" line to the next line
    phase = {}
    for i, (t, line) in enumerate(stamps):
        m = re.search(r"Counting (\w\w) pairs", line)
        if m:
            nxt = stamps[i + 1][0] if i + 1 < len(stamps) else time.time()
            phase[m.group(1)] = nxt - t
    res = {p_: read_pair_count(o) for p_, o in zip(pairs, outs)}
    return res, phase, wall, text
