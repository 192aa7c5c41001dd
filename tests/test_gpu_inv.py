"""The inversion-loop pieces on the device (crt_dmisft, crt_bpar, crt_bsigma,
crt_update_step and their device-pointer variants, crt_blam0) against the
oracle and the reference's golden vectors (tests/golden/ref_inversion_steps.npz,
ref_blam0.npz).

Tolerances: dmisft's sums run in a fixed tree order instead of the
reference's sequential one, and the device's sqrt/hypot/log/exp may differ
from the host library's by an ulp, so the statistics and weights are held to
1e-13 relative; the polarity flags and npol are exact; the update step's
arithmetic (no fused multiply-add) is bit-exact, its bdpar sum 1e-14."""
import numpy as np
import pytest

from conftest import golden, rel_err
from titatest_amd import meshgen

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def crtmod(hip_lib):
    from titatest_amd import crt
    return crt


@pytest.fixture(scope="module")
def handle(crtmod):
    w = meshgen.small_block()
    return w, crtmod.CRT(w.mesh)


@pytest.mark.parametrize("flags", list(range(16)))
def test_dmisft_matches_oracle_and_reference(handle, oracle_lib, flags):
    from oracle import cpu
    _, g = handle
    d = golden("inversion_steps")
    say("dmisft flags", flags)
    args = (d["dat"], d["sigmaa"], d["wmatd"], d["wmatdr"], d["wmatdp"], flags, float(d["l1min"]))
    st, w, wd = g.dmisft(*args, d0=d["d0"], fm0=d["fm0"])
    so, wo, wdo = cpu.dmisft(*args, d0=d["d0"], fm0=d["fm0"])
    np.testing.assert_array_equal(wd, wdo)
    assert st[5] == so[5] == d["stats"][flags][5]
    np.testing.assert_allclose(st[:5], d["stats"][flags][:5], rtol=1e-13, atol=0)
    np.testing.assert_allclose(w, d["wmatd_out"][flags], rtol=1e-13, atol=0)


def test_dmisft_every_datum_excluded_is_error_99(handle, crtmod):
    _, g = handle
    say("dmisft all excluded")
    dat = np.full(8, complex(1.0, 0.1))
    sig = np.full(8, complex(1.0, np.pi - 0.1))
    with pytest.raises(crtmod.CRTError) as e:
        g.dmisft(dat, sig, np.ones(8), np.ones(8), np.ones(8), 0)
    assert e.value.code == 99


def test_bpar_bsigma_match_reference(handle):
    w, g = handle
    d = golden("inversion_steps")
    say("bpar/bsigma identity and zoned maps")
    g.set_model_map()
    assert rel_err(g.bpar(w.sigma), d["par_id"]) < 1e-15
    mz = d["mnr_z"]
    g.set_model_map(len(d["par_z"]), mz)
    assert rel_err(g.bpar(w.sigma), d["par_z"]) < 1e-15
    assert rel_err(g.bsigma(d["par_step"]), d["sigma_z"]) < 1e-15
    g.set_model_map()


def test_update_step_matches_oracle(handle, oracle_lib):
    from oracle import cpu
    _, g = handle
    say("update step")
    rng = np.random.default_rng(9)
    M = 3000
    par = rng.normal(size=M) + 1j * rng.normal(0, 0.01, M)
    dpar = rng.normal(size=M) + 1j * rng.normal(0, 0.01, M)
    for lphi0 in (False, True):
        p, bd, n = g.update_step(par, dpar, 0.37, lphi0)
        po, bdo, no = cpu.update_step(par, dpar, 0.37, lphi0)
        np.testing.assert_array_equal(p, po)
        assert n == no and (n > 0) == lphi0
        assert abs(bd - bdo) <= 1e-14 * bdo


def test_device_variants_equal_host_variants(handle, crtmod):
    import torch
    w, g = handle
    d = golden("inversion_steps")
    say("device-pointer variants")
    dev = torch.device("cuda:0")
    N = len(d["dat"])
    t = lambda a, dt: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)  # noqa: E731
    flags = crtmod.MISFIT_LPOL | crtmod.MISFIT_LROBUST
    st_h, w_h, wd_h = g.dmisft(d["dat"], d["sigmaa"], d["wmatd"], d["wmatdr"], d["wmatdp"], flags, 1.0)
    dat, sig = t(d["dat"], torch.complex128), t(d["sigmaa"], torch.complex128)
    wm, wr, wp = t(d["wmatd"], torch.float64), t(d["wmatdr"], torch.float64), t(d["wmatdp"], torch.float64)
    st = torch.zeros(6, dtype=torch.float64, device=dev)
    wd = torch.zeros(N, dtype=torch.int32, device=dev)
    # the inputs come from torch's stream; the entry points queue on the
    # handle's own stream without waiting for it
    torch.cuda.synchronize()
    rc = g.lib.crt_dmisft_dev(g.h, N, dat.data_ptr(), sig.data_ptr(), wm.data_ptr(), wr.data_ptr(), wp.data_ptr(),
                              None, None, flags, 1.0, st.data_ptr(), wd.data_ptr())
    assert rc == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(st.cpu().numpy(), st_h)
    np.testing.assert_array_equal(wm.cpu().numpy(), w_h)
    np.testing.assert_array_equal(wd.cpu().numpy(), wd_h)
    # bpar -> update -> bsigma without leaving the device
    g.set_model_map()
    M = w.mesh.elanz
    sg = t(w.sigma, torch.complex128)
    par = torch.zeros(M, dtype=torch.complex128, device=dev)
    rng = np.random.default_rng(4)
    dp_h = 0.01 * (rng.normal(size=M) + 1j * rng.normal(size=M))
    dpar = t(dp_h, torch.complex128)
    out = torch.zeros(2, dtype=torch.float64, device=dev)
    sg2 = torch.zeros_like(sg)
    torch.cuda.synchronize()
    assert g.lib.crt_bpar_dev(g.h, sg.data_ptr(), par.data_ptr()) == 0
    assert g.lib.crt_update_step_dev(g.h, M, par.data_ptr(), dpar.data_ptr(), 0.5, 1, out.data_ptr()) == 0
    assert g.lib.crt_bsigma_dev(g.h, par.data_ptr(), sg2.data_ptr()) == 0
    torch.cuda.synchronize()
    p_h = g.bpar(w.sigma)
    p_h, bd_h, n_h = g.update_step(p_h, dp_h, 0.5, True)
    np.testing.assert_array_equal(par.cpu().numpy(), p_h)
    assert out[0].item() == bd_h and int(out[1].item()) == n_h
    np.testing.assert_array_equal(sg2.cpu().numpy(), g.bsigma(p_h))


def test_blam0_matches_reference(crtmod):
    """blam0 from the device Jacobian, J^H (w o (J 1)) in O(N M), against the
    reference's O(N M^2) loop (golden vectors).  Held to the sensitivity
    tolerance: the device J itself agrees with the reference's to 1e-8."""
    w = meshgen.small_block()
    d = golden("blam0")
    say("blam0 small_block")
    g = crtmod.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    g.forward(w.sigma)
    g.bvolti(w.strnr, w.vnr)
    g.sensitivities(1, to_host=False)
    for mode, ref_mode in ((0, 0), (1, 1)):
        lam = g.blam0(d["w"], mode, float(d["alfx"]), float(d["alfz"]))
        assert abs(lam - d["lammax"][ref_mode]) <= 1e-8 * d["lammax"][ref_mode], (mode, lam)
    with pytest.raises(crtmod.CRTError):
        g.blam0(d["w"], 2)


def test_cg_update_matches_reference_update(crtmod, oracle_lib):
    """One model update on the device -- crt_cg_solve (cgfac, right-hand side,
    10 CG iterations) and crt_update_step -- against the reference's update
    (src/update.f90 -> cjg -> cjggra, golden vectors ref_update.npz) at the
    sensitivity tolerance: the device's J and modelled data agree with the
    reference's to ~1e-12, which 10 iterations amplify only mildly."""
    from oracle import cpu
    w = meshgen.small_block()
    d = golden("update")
    say("cg update vs reference update, small_block")
    g = crtmod.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    g.forward(w.sigma)
    v, sa = g.bvolti(w.strnr, w.vnr)
    g.sensitivities(1, to_host=False)
    nx, nz = int(d["nx"]), int(d["nz"])
    sm = g.smoothness(0, 1.0, 1.0, nx=nx, nz=nz)
    par = np.log(w.sigma)
    rm = cpu.rm_reg(sm, nx, par)
    dpar, cgfac, cgres = g.cg_solve(d["w"], d["dat"], sa, float(d["lam"]), 0, sm, sm[:, 0], int(d["a_ncg"]),
                                    eps=float(d["a_eps"]), nx=nx, rm=rm)
    assert rel_err(cgfac, d["a_cgfac"]) < 1e-12
    assert int(cgres[0]) == int(d["a_cgres"][0])
    assert rel_err(dpar, d["a_dpar"]) < 1e-8
    p, bd, _ = g.update_step(par, dpar, 1.0)
    assert rel_err(p, d["a_par"]) < 1e-8 and abs(bd - d["a_bdpar"]) < 1e-8 * d["a_bdpar"]
