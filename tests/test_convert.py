"""Survey coordinate conversion (src/fcfc/2pt/cnvt_coord.c) on the host:
the comoving distance integral against the Einstein-de Sitter closed form,
the spline path against its own samples and the closed form, the cosmology
checks of load_conf.c, and the Gauss-Legendre table against its generator.
Counts of converted catalogues against the upstream FCFC_2PT are GPU tests
(tests/test_gpu_cli.py)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import fcfc_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C_H0 = 299792.458 * 0.01          # c / H0 in Mpc/h


def eds(z):
    """Einstein-de Sitter (Omega_m = 1): D = 2 c/H0 (1 - 1/sqrt(1 + z))."""
    return 2 * C_H0 * (1 - 1 / np.sqrt(1 + z))


def test_integral_matches_closed_form():
    """Omega_m = 1 with Omega_Lambda = 1e-12 (FCFC's H^2 > 0 check refuses
    Omega_Lambda = 0 exactly): within 1e-12 of Einstein-de Sitter."""
    g = np.random.default_rng(1)
    n = 2000
    ra, dec, z = g.uniform(0, 360, n), g.uniform(-90, 90, n), g.uniform(0.01, 3.0, n)
    x, y, zz, order = F.convert_coords(ra, dec, z, omega_m=1.0, omega_l=1e-12, err=1e-12)
    assert 4 <= order <= 32
    d = np.sqrt(x * x + y * y + zz * zz)
    np.testing.assert_allclose(d, eds(z), rtol=1e-11)
    # direction: RA and Dec survive the conversion
    np.testing.assert_allclose(np.degrees(np.arcsin(zz / d)), dec, atol=1e-9)
    np.testing.assert_allclose(np.degrees(np.arctan2(y, x)) % 360, ra, atol=1e-9)


def test_order_grows_with_accuracy_and_range():
    z = np.linspace(0.1, 0.5, 100)
    orders = [F.convert_coords(z * 0, z * 0, z, 0.31, err=e)[3] for e in (1e-4, 1e-8, 1e-13)]
    assert orders == sorted(orders) and orders[0] < orders[-1]
    # flat LCDM (Omega_m = 0.31) against adaptive quadrature
    from scipy.integrate import quad
    zq = np.array([0.2, 1.0, 2.5])
    x, _, _, _ = F.convert_coords(zq * 0, zq * 0, zq, 0.31, err=1e-12)
    ref = [quad(lambda t: C_H0 / np.sqrt(0.31 * (1 + t) ** 3 + 0.69), 0, v, epsabs=0,
                epsrel=1e-13)[0] for v in zq]
    np.testing.assert_allclose(x, ref, rtol=1e-11)


def test_dark_energy_w_and_curvature():
    """w = -1 (cosmological constant) equals the limit of w -> -1; an open
    model (Omega_Lambda unset is flat) gives a different distance."""
    z = np.array([0.5, 1.0, 2.0])
    zero = np.zeros(3)
    d1 = F.convert_coords(zero, zero, z, 0.3, err=1e-12)[0]
    d2 = F.convert_coords(zero, zero, z, 0.3, w=-1.0 + 1e-9, err=1e-12)[0]
    np.testing.assert_allclose(d1, d2, rtol=1e-8)
    d3 = F.convert_coords(zero, zero, z, 0.3, omega_l=0.5, err=1e-12)[0]
    assert np.all(d3 < d1)


@pytest.mark.parametrize("bad", [dict(omega_m=0.0), dict(omega_m=1.5),
                                 dict(omega_m=0.3, omega_l=-0.1),
                                 dict(omega_m=0.3, w=-0.2), dict(omega_m=0.3, err=1e-20)])
def test_cosmology_checks(bad):
    with pytest.raises(F.FcfcError):
        F.convert_coords([0.0], [0.0], [0.5], **bad)


def test_negative_redshift_is_refused():
    with pytest.raises(F.FcfcError):
        F.convert_coords([0.0, 0.0], [0.0, 0.0], [0.5, -0.1], 0.3)


def test_spline_table():
    zs = np.linspace(0.3, 0.6, 61)
    ds = eds(zs)
    perm = np.random.default_rng(2).permutation(len(zs))       # any order
    # at the samples the spline returns the sample (to rounding)
    x, _, _ = F.convert_coords_table(zs[perm], ds[perm], np.zeros(60), np.zeros(60), zs[:-1])
    np.testing.assert_allclose(x, ds[:-1], rtol=4e-16)
    # between them it follows the smooth closed form (the natural end
    # conditions, zero second derivative, cost O(h^2) near the ends)
    zq = np.random.default_rng(3).uniform(0.3, 0.6, 500)
    x, _, _ = F.convert_coords_table(zs, ds, np.zeros(500), np.zeros(500), zq)
    np.testing.assert_allclose(x, eds(zq), rtol=1e-5)
    inner = (zq > 0.35) & (zq < 0.55)
    np.testing.assert_allclose(x[inner], eds(zq[inner]), rtol=1e-8)
    with pytest.raises(F.FcfcError):              # beyond the samples
        F.convert_coords_table(zs, ds, [0.0], [0.0], [0.6])


def test_legauss_table_is_the_generator_output(tmp_path):
    """fcfc_legauss.h is what gen_legauss.py writes (rules computed from the
    Legendre polynomials at 60 digits, correctly rounded)."""
    src = os.path.join(ROOT, "fcfc_amd", "csrc")
    code = open(os.path.join(src, "gen_legauss.py")).read().replace(
        'HERE = os.path.dirname(os.path.abspath(__file__))', f'HERE = {str(tmp_path)!r}')
    (tmp_path / "gen.py").write_text(code)
    subprocess.run([sys.executable, str(tmp_path / "gen.py")], check=True)
    assert (tmp_path / "fcfc_legauss.h").read_text() == open(os.path.join(src, "fcfc_legauss.h")).read()
    # and the rules integrate polynomials of degree 2n - 1 exactly
    import re
    txt = open(os.path.join(src, "fcfc_legauss.h")).read()
    for n, xs, ws, mid in re.findall(r"\{(\d+), \{([^}]*)\}, \{([^}]*)\}, ([^}]*)\}", txt):
        n = int(n)
        x = np.array([float.fromhex(t) for t in xs.split(",")])
        w = np.array([float.fromhex(t) for t in ws.split(",")])
        m = float.fromhex(mid.strip()) if n % 2 else 0.0
        for k in range(0, 2 * n, 2):                 # even moments over [-1, 1]
            got = 2 * np.sum(w * x ** k) + (m if k == 0 else 0.0)
            assert got == pytest.approx(2 / (k + 1), rel=1e-13, abs=1e-15), (n, k)
