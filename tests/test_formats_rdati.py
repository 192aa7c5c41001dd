"""formats.read_data (measured data files) against the reference's rdati
compiled from source (golden vectors tests/golden/ref_rdati.npz, made from the
data files in tests/golden/rdati/ by make_golden.py): the data vector and the
three weight arrays bit for bit, and the reference's error number for each
rejected file."""
from pathlib import Path

import numpy as np
import pytest

from conftest import golden
from titatest_amd import formats as F

HERE = Path(__file__).resolve().parent / "golden"
MODEL = dict(stabw0=3.0, stabm0=1e-3, stabpA1=0.5, stabpB=-0.8, stabpA2=2.0, stabp0=0.1)
CASES = [("crt_model", 8, False), ("abmn_model", 8, False), ("crt_indiv", 8, False), ("dc_indiv", 6, True),
         ("dc_model", 6, True), ("spanning", 8, False), ("err_dup", 8, False), ("err_range", 4, False),
         ("err_phase_error", 8, False), ("err_resistance", 8, False), ("err_phase", 8, False),
         ("err_eof", 8, False)]


@pytest.mark.parametrize("name,eanz,ldc", CASES)
def test_read_data_matches_rdati(name, eanz, ldc):
    g = golden("rdati")
    path = HERE / "rdati" / f"{name}.dat"
    errnr = int(g[f"{name}__errnr"])
    if errnr:
        with pytest.raises(F.FormatError) as e:
            F.read_data(path, eanz, ldc, **MODEL)
        assert e.value.errnr == errnr
        return
    d = F.read_data(path, eanz, ldc, **MODEL)
    assert d.lindiv == bool(g[f"{name}__lindiv"])
    for k in ("strnr", "vnr", "dat", "wmatdr", "wmatdp", "wmatd_cri", "wdfak"):
        np.testing.assert_array_equal(getattr(d, k), g[f"{name}__{k}"], err_msg=k)


def test_read_data_error_model_closed_form(tmp_path):
    # sigma_{ln R} = stabw0/100 + stabm0/R, sigma_phi = (A1 R^B + A2/100 |phi| + p0) mrad
    f = tmp_path / "v.dat"
    f.write_text("2\n10002 40003 2.0 -10.0\n20003 50004 0.5 4.0\n")
    d = F.read_data(f, 8, False, **MODEL)
    R = np.array([2.0, 0.5])
    phi = np.array([-10.0, 4.0])
    sr = 0.03 + 1e-3 / R
    sp = (0.5 * R ** -0.8 + 0.02 * np.abs(phi) + 0.1) * 1e-3
    np.testing.assert_allclose(d.wmatdr, 1 / sr ** 2, rtol=1e-15)
    np.testing.assert_allclose(d.wmatdp, 1 / sp ** 2, rtol=1e-15)
    np.testing.assert_allclose(d.wmatd_cri, 1 / (sr ** 2 + sp ** 2), rtol=1e-15)
    np.testing.assert_allclose(d.dat, -np.log(R) - 1j * phi / 1e3, rtol=1e-15)


NOISE = [(1, 2.0, 1e-3, 0.3, 0.5, 1.5, 0.2), (12345, 5.0, 0.0, 0.0, 0.0, 0.0, 0.0)]


@pytest.mark.parametrize("name,eanz,ldc", [("crt_model", 8, False), ("abmn_model", 8, False), ("dc_model", 6, True)])
@pytest.mark.parametrize("j", [0, 1])
def test_read_data_synthetic_noise_matches_rdati(name, eanz, ldc, j):
    """rdati's synthetic noise: the reference's generator (make_noise) seeded
    with iseed and -iseed, bit for bit."""
    g = golden("rdati")
    nz = NOISE[j]
    key = f"{name}__noise{j}"
    assert int(g[f"{key}__errnr"]) == 0
    d = F.read_data(HERE / "rdati" / f"{name}.dat", eanz, ldc, **MODEL, noise=F.NoiseModel(int(nz[0]), *nz[1:]))
    for k in ("dat", "wmatdr", "wmatdp", "wmatd_cri"):
        np.testing.assert_array_equal(getattr(d, k), g[f"{key}__{k}"], err_msg=k)
    assert not np.array_equal(d.dat, F.read_data(HERE / "rdati" / f"{name}.dat", eanz, ldc, **MODEL).dat)


def test_ref_random_gauss_moments():
    # the generator's deviates: mean 0, variance 1 (50,000 draws)
    g = F.RefRandom(7)
    x = np.array([g.gauss() for _ in range(50000)])
    assert abs(x.mean()) < 0.02 and abs(x.var() - 1.0) < 0.03
