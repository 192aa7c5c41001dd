"""Inversion-loop pieces of the oracle (chkpo2, dmisft, bpar, bsigma, the
model update step) against the reference's own routines' golden vectors
(tests/golden/ref_inversion_steps.npz, written by make_golden.py from
src/dmisft.f90, src/bpar.f90 and src/bsigma.f90 compiled from source) and
against what the mathematics fixes."""
import numpy as np
import pytest

from conftest import golden
from oracle import cpu

G = None


def g():
    global G
    if G is None:
        G = golden("inversion_steps")
    return G


def test_chkpo2_matches_reference_bitwise():
    d = g()
    for i in range(len(d["chk_in_d"])):
        cd, cs, wd = cpu.chkpo2(complex(d["chk_in_d"][i]), complex(d["chk_in_s"][i]), int(d["chk_in_ld"][i]))
        assert cd == complex(d["chk_cdat"][i]) and cs == complex(d["chk_csig"][i]), i
        assert wd == int(d["chk_wdlok"][i]), i


def test_chkpo2_folds_into_the_principal_range():
    # datum and response in the same class: both folded by pi, the datum is kept
    pi = np.pi
    cd, cs, wd = cpu.chkpo2(complex(1.0, 2.0), complex(2.0, 3.0), 0)
    assert wd == 1 and cd == complex(1.0, 2.0 - pi) and cs == complex(2.0, 3.0 - pi)
    cd, cs, wd = cpu.chkpo2(complex(1.0, -2.0), complex(2.0, -3.0), 0)
    assert wd == 1 and cd.imag == -2.0 + pi and cs.imag == -3.0 + pi
    # mismatch: the datum is excluded; with ldum its assumed polarity is corrected
    _, _, wd = cpu.chkpo2(complex(1.0, 0.1), complex(2.0, 3.0), 1)
    assert wd == 0
    cd, _, wd = cpu.chkpo2(complex(1.0, 3.0), complex(2.0, 0.1), 1)
    assert wd == 0 and cd.imag == 3.0 - pi
    # the boundaries: +pi/2 is in range, -pi/2 is not
    assert cpu.chkpo2(complex(0, pi / 2), complex(0, pi / 2), 0)[2] == 1
    cd, cs, _ = cpu.chkpo2(complex(0, -pi / 2), complex(0, -pi / 2), 0)
    assert cd.imag == -pi / 2 + pi and cs.imag == -pi / 2 + pi


@pytest.mark.parametrize("flags", list(range(16)))
def test_dmisft_matches_reference_bitwise(flags):
    d = g()
    st, w, _ = cpu.dmisft(d["dat"], d["sigmaa"], d["wmatd"], d["wmatdr"], d["wmatdp"], flags,
                          float(d["l1min"]), d["d0"], d["fm0"])
    np.testing.assert_array_equal(st, d["stats"][flags])
    np.testing.assert_array_equal(w, d["wmatd_out"][flags])


def test_dmisft_closed_form():
    # all data in range: nrmsd is the weighted RMS of |d - f|, betrms/pharms
    # the RMS of the weighted real/imaginary residuals, rmssum the chi^2 sum
    rng = np.random.default_rng(7)
    N = 64
    dat = rng.normal(size=N) + 1j * rng.uniform(-1, 1, N)
    sig = dat + 0.1 * (rng.normal(size=N) + 1j * rng.normal(size=N))
    w = rng.uniform(1, 4, N)
    wr = rng.uniform(1, 4, N)
    wp = rng.uniform(1, 4, N)
    st, wout, wd = cpu.dmisft(dat, sig, w, wr, wp, cpu.DMISFT_LPOL)
    r = dat - sig
    assert np.all(wd == 1) and st[5] == 0
    np.testing.assert_allclose(st[0], np.sqrt(np.mean(w * np.abs(r) ** 2)), rtol=1e-14)
    np.testing.assert_allclose(st[1], np.sqrt(np.mean(wr * r.real ** 2)), rtol=1e-14)
    np.testing.assert_allclose(st[2], np.sqrt(np.mean(wp * r.imag ** 2)), rtol=1e-14)
    np.testing.assert_allclose(st[3], np.sum(w * np.abs(r) ** 2), rtol=1e-14)
    np.testing.assert_array_equal(wout, w)
    # FPI: only the imaginary part of the residual counts
    st, _, _ = cpu.dmisft(dat, sig, w, wr, wp, cpu.DMISFT_LFPI)
    np.testing.assert_allclose(st[0], np.sqrt(np.mean(w * r.imag ** 2)), rtol=1e-14)
    # difference inversion subtracts d0 - f(m0)
    d0 = rng.normal(size=N) + 0j
    fm0 = d0 + 0.05
    st, _, _ = cpu.dmisft(dat, sig, w, wr, wp, cpu.DMISFT_LDIFF, d0=d0, fm0=fm0)
    np.testing.assert_allclose(st[0], np.sqrt(np.mean(w * np.abs(r - (d0 - fm0)) ** 2)), rtol=1e-14)


def test_dmisft_all_excluded_is_error_99():
    pi = np.pi
    dat = np.full(4, complex(1.0, 0.1))
    sig = np.full(4, complex(1.0, pi - 0.1))
    with pytest.raises(cpu.OracleError) as e:
        cpu.dmisft(dat, sig, np.ones(4), np.ones(4), np.ones(4), 0)
    assert e.value.code == 99


def test_robust_reweighting_closed_form():
    # with psi_i = c (all residuals equal in weighted size) the estimated
    # standard deviations all equal eps_i = 1/sqrt(w_i), l1rat = 1
    N = 10
    w = np.full(N, 4.0)
    dat = np.full(N, 1.0 + 0.0j)
    sig = dat - 0.25
    st, wout, _ = cpu.dmisft(dat, sig, w, w, w, cpu.DMISFT_LROBUST, l1min=0.5)
    assert st[4] == pytest.approx(1.0, rel=1e-15)
    np.testing.assert_allclose(wout, w, rtol=1e-14)
    # one outlier: its standard deviation grows, the others keep theirs
    sig2 = sig.copy()
    sig2[3] = dat[3] - 25.0
    st, wout, _ = cpu.dmisft(dat, sig2, w, w, w, cpu.DMISFT_LROBUST, l1min=0.5)
    assert st[4] > 1.0 and wout[3] < 0.5 * w[3]
    np.testing.assert_allclose(np.delete(wout, 3), np.delete(w, 3), rtol=1e-14)


def test_bpar_bsigma_match_reference_bitwise():
    from titatest_amd import meshgen
    d = g()
    w = meshgen.small_block()
    el = w.mesh.elanz
    np.testing.assert_array_equal(cpu.bpar(w.sigma, el), d["par_id"])
    mz = d["mnr_z"]
    np.testing.assert_array_equal(cpu.bpar(w.sigma, len(d["par_z"]), mz), d["par_z"])
    np.testing.assert_array_equal(cpu.bsigma(d["par_step"], el, mz), d["sigma_z"])


def test_bpar_bsigma_inverse_and_zoned_last_writer():
    rng = np.random.default_rng(5)
    s = np.exp(rng.normal(size=50)) * np.exp(-1j * rng.uniform(0, 0.03, 50))
    p = cpu.bpar(s)
    np.testing.assert_allclose(cpu.bsigma(p), s, rtol=1e-15)
    np.testing.assert_allclose(p, np.log(s), rtol=0, atol=4e-16)
    # several elements per parameter: the last element in order wins (bpar's loop)
    mnr = np.array([1, 1, 2, 2, 2], np.int32)
    p = cpu.bpar(s[:5], 2, mnr)
    assert p[0] == np.log(s[1]) or abs(p[0] - np.log(s[1])) < 1e-15
    assert abs(p[1] - np.log(s[4])) < 1e-15


def test_update_step_exact_cases():
    # dyadic values: every product and sum is exact
    par = np.array([1.0 + 0.5j, -2.0 + 0.25j, 0.5 - 0.125j, 0.0 + 0.0j])
    dpar = np.array([0.5 - 0.25j, 1.0 + 0.0j, -0.25 - 0.5j, 2.0 - 1.0j])
    p, bd, n = cpu.update_step(par, dpar, 0.5)
    np.testing.assert_array_equal(p, par + 0.5 * dpar)
    assert bd == 0.5 * np.sum(dpar.real ** 2 + dpar.imag ** 2) and n == 0
    # lphi0: a negative imaginary part (conductivity phase) becomes 0, counted
    p, bd, n = cpu.update_step(par, dpar, 0.5, lphi0=True)
    expect = par + 0.5 * dpar
    neg = expect.imag < 0
    expect[neg] = expect[neg].real
    np.testing.assert_array_equal(p, expect)
    assert n == int(neg.sum()) == 2


def _small_block_j():
    from titatest_amd import meshgen
    w = meshgen.small_block()
    o = cpu.Oracle(w.mesh)
    o.rwaven(w.strnr, w.vnr)
    o.precal()
    o.forward(w.sigma)
    o.volt(w.strnr, w.vnr, 1)
    return o.sens(1)


def test_blam0_matches_reference_bitwise():
    d = golden("blam0")
    J = _small_block_j()
    for mode, ref_mode in ((0, 0), (1, 1), (1, 2)):  # complex, lfpi, ldc (real parts)
        assert cpu.blam0(J, d["w"], mode, float(d["alfx"]), float(d["alfz"])) == d["lammax"][ref_mode]


def test_blam0_is_the_mean_of_jhwj1():
    # the O(N M^2) loop of the reference equals |J^H W (J 1)| averaged over
    # the columns, times 2/(alfx+alfz)*5
    d = golden("blam0")
    J = _small_block_j()
    w = d["w"]
    for mode in (0, 1):
        jj = J.real if mode else J
        want = np.mean(np.abs((jj.conj().T * w) @ jj.sum(axis=1))) * 2.0 / 1.5 * 5.0
        assert cpu.blam0(J, w, mode, 1.0, 0.5) == pytest.approx(want, rel=1e-13)


def test_update_and_cjggra_match_reference_bitwise():
    """The oracle's chain for one model update -- update_rhs (cgfac and the
    right-hand side with R m summed in update's order), cjggra and the scaled
    step -- equals the reference's update (src/update.f90 with cg_mod's cjg ->
    cjggra, compiled from source) bit for bit, for 10, 20 and 30 iterations."""
    from titatest_amd import meshgen
    d = golden("update")
    w = meshgen.small_block()
    o = cpu.Oracle(w.mesh)
    o.rwaven(w.strnr, w.vnr)
    o.precal()
    o.forward(w.sigma)
    v, sa = o.volt(w.strnr, w.vnr, 1)
    J = o.sens(1)
    N = J.shape[0]
    nx = int(d["nx"])
    sm = cpu.bsmatmreg(o, nx, int(d["nz"]), 1.0, 1.0)
    par = np.log(w.sigma)
    ones = np.ones(N, np.int32)
    cgfac, bvec = cpu.update_rhs(J, d["w"], ones, d["dat"], sa, float(d["lam"]), 0, sm[:, 0], cpu.rm_reg(sm, nx, par))
    for name in ("a", "b", "c"):
        ncg, eps = int(d[f"{name}_ncg"]), float(d[f"{name}_eps"])
        dpar, cgres, n = cpu.cjggra(J, d["w"], ones, cgfac, bvec, float(d["lam"]), 0, nx, sm, ncg, eps)
        np.testing.assert_array_equal(cgfac, d[f"{name}_cgfac"])
        np.testing.assert_array_equal(dpar * cgfac, d[f"{name}_dpar"])
        assert n == int(d[f"{name}_cgres"][0])
        if eps > 0:  # with eps = 0 the reference's residual entries are 0/0
            np.testing.assert_array_equal(cgres[:ncg + 1], d[f"{name}_cgres"])
        p, bd, _ = cpu.update_step(par, dpar * cgfac, 1.0)
        np.testing.assert_array_equal(p, d[f"{name}_par"])
        assert bd == d[f"{name}_bdpar"]
