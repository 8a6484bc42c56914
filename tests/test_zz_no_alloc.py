"""The no-allocation entry point (fcfc_gpu_count_noalloc) through the C ABI:
its counts must equal fcfc_gpu_count's on the same inputs -- integer counts
identical, weighted counts within 1e-12 relative -- on consecutive calls
whose second input is smaller (the scratch array keeps the previous call's
contents), for host and device catalogues, and inputs too large for its
scratch must be refused with FCFC_ERR_MEMORY before any work is queued."""
import ctypes as C

import numpy as np
import pytest

import fcfc_amd as F
from fcfc_amd import _lib as L

pytestmark = pytest.mark.gpu
BOX = 1000.0


def _count(eng, fn, a, b, bins, flags, out):
    ca = a._c()
    cb = (b if b is not None else a)._c()
    bs = bins.bsize_rescaled
    e = fn(eng._h, C.byref(ca), C.byref(cb), C.byref(bins._b),
           bs.ctypes.data_as(C.POINTER(C.c_double)) if bs is not None else None,
           flags, out.ctypes.data)
    return e


@pytest.fixture(scope="module")
def eng():
    with F.Engine() as e:
        yield e


@pytest.mark.parametrize("bt", [F.BIN_ISO, F.BIN_SMU, F.BIN_SPI])
def test_noalloc_equals_count(eng, bt):
    lib = F.lib()
    b = F.Bins.linear(F.GEOM_BOX, bt, 0, 200, 5, nmu=100 if bt == F.BIN_SMU else 1,
                      pmin=0, pmax=200 if bt == F.BIN_SPI else 0, dpi=5, bsize=BOX)
    big = F.gen_clustered(91, 30000, 60, 200, 2.0, BOX, 0.3)
    small = F.gen_clustered(92, 4000, 10, 100, 2.0, BOX, 0.3)
    rnd = F.gen_uniform(93, 60000, BOX)
    for tag, (x, y, z, w) in (("first", big), ("second, smaller", small)):
        D = F.Catalog(b, x, y, z, w)
        R = F.Catalog(b, *rnd) if tag == "first" else F.Catalog(b, *F.gen_uniform(94, 9000, BOX))
        for isauto, wt, prec in ((True, False, "fp64"), (True, True, "fp32"),
                                 (False, True, "fp64"), (False, False, "plain")):
            flags = ((L.FCFC_GPU_AUTO if isauto else 0) | (L.FCFC_GPU_WEIGHTED if wt else 0)
                     | F._prec_flag(prec))
            print(f"case noalloc bt={bt} {tag} auto={isauto} wt={wt} prec={prec}", flush=True)
            ref = np.zeros(b.ntot, dtype=np.float64 if wt else np.int64)
            got = np.zeros_like(ref)
            assert _count(eng, lib.fcfc_gpu_count, D, None if isauto else R, b, flags, ref) == 0
            e = _count(eng, lib.fcfc_gpu_count_noalloc, D, None if isauto else R, b, flags, got)
            assert e == 0, lib.fcfc_gpu_last_error(eng._h)
            if wt:
                np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
            else:
                np.testing.assert_array_equal(got, ref)


def test_noalloc_device_catalogues_and_counts(eng):
    """Device catalogues and a device histogram (FCFC_GPU_CNT_DEVICE), the
    histogram accumulated into as fcfc_gpu_count does."""
    import torch
    lib = F.lib()
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_SMU, 0, 200, 5, nmu=100, bsize=BOX)
    f = b.rescale
    for seed, n in ((95, 50000), (96, 7000)):
        x, y, z = F.gen_uniform(seed, n, BOX)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a * f)).cuda()
        D = F.Catalog(b, dev(x), dev(y), dev(z))
        ref = torch.full((b.ntot,), 3, dtype=torch.int64, device="cuda")
        got = torch.full((b.ntot,), 3, dtype=torch.int64, device="cuda")
        flags = L.FCFC_GPU_AUTO | L.FCFC_GPU_CNT_DEVICE
        torch.cuda.synchronize()
        print(f"case noalloc device n={n}", flush=True)
        for fn, out in ((lib.fcfc_gpu_count, ref), (lib.fcfc_gpu_count_noalloc, got)):
            ca = D._c()
            e = fn(eng._h, C.byref(ca), C.byref(ca), C.byref(b._b),
                   b.bsize_rescaled.ctypes.data_as(C.POINTER(C.c_double)), flags,
                   C.c_void_p(out.data_ptr()))
            assert e == 0, lib.fcfc_gpu_last_error(eng._h)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        assert int(ref.sum()) > 3 * b.ntot


def test_noalloc_rejects_oversize(eng):
    """Catalogues whose sorted copies cannot fit the 128 MiB scratch are
    refused up front (a host catalogue of 4e6 objects needs at least 60
    bytes of sorted copy, upload and sort scratch per object: 240 MB)."""
    lib = F.lib()
    b = F.Bins.linear(F.GEOM_BOX, F.BIN_ISO, 0, 20, 5, bsize=BOX)
    n = 4_000_000
    c = L.fcfc_cat()
    xs = np.zeros(n)
    c.n, c.x, c.y, c.z, c.s, c.w, c.on_device = n, xs.ctypes.data, xs.ctypes.data, xs.ctypes.data, None, None, 0
    out = np.zeros(b.ntot, dtype=np.int64)
    print("case noalloc oversize", flush=True)
    e = lib.fcfc_gpu_count_noalloc(eng._h, C.byref(c), C.byref(c), C.byref(b._b),
                                   b.bsize_rescaled.ctypes.data_as(C.POINTER(C.c_double)),
                                   L.FCFC_GPU_AUTO, out.ctypes.data)
    assert e == L.FCFC_ERR_MEMORY
    assert b"scratch" in lib.fcfc_gpu_last_error(eng._h)
    assert not out.any()
