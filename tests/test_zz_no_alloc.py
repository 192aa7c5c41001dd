"""crt_dmisft_dev_noalloc (the misfit entry point that allocates nothing: its
per-datum scratch is a static device array) against crt_dmisft_dev through
the C-ABI: two calls in a row on device buffers, the second with fewer data,
results bit for bit equal (the same kernel on the same inputs), with the
robust reweighting on so the scratch is read back; sizes above the array's
limit refused with 1004."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIMIT = 1 << 18


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _data(N, seed):
    rng = np.random.default_rng(seed)
    dat = rng.normal(0, 1, N) + 1j * rng.uniform(-2.5, 2.5, N)
    sig = dat + 0.05 * (rng.normal(size=N) + 1j * rng.normal(size=N))
    sig[::17] += 2.0j
    wmatd = 1.0 / (0.02 + 1e-3 * rng.uniform(size=N)) ** 2
    wr = rng.uniform(0.5, 2, N) * wmatd
    wp = rng.uniform(0.5, 2, N) * wmatd
    return dat, sig, wmatd, wr, wp


def _run(lib, g, fn, N, arrays, flags, l1min):
    import torch
    dev = torch.device("cuda:0")
    dat, sig, wmatd, wr, wp = arrays
    td = torch.tensor(dat, dtype=torch.complex128, device=dev)
    ts = torch.tensor(sig, dtype=torch.complex128, device=dev)
    tw = torch.tensor(wmatd, dtype=torch.float64, device=dev)
    twr = torch.tensor(wr, dtype=torch.float64, device=dev)
    twp = torch.tensor(wp, dtype=torch.float64, device=dev)
    st = torch.zeros(6, dtype=torch.float64, device=dev)
    wd = torch.zeros(N, dtype=torch.int32, device=dev)
    # the inputs come from torch's stream; the entry point queues on the
    # handle's own stream without waiting for it
    torch.cuda.synchronize()
    rc = fn(g.h, N, td.data_ptr(), ts.data_ptr(), tw.data_ptr(), twr.data_ptr(), twp.data_ptr(), None, None,
            flags, l1min, st.data_ptr(), wd.data_ptr())
    assert rc == 0, rc
    torch.cuda.synchronize()
    return st.cpu().numpy(), tw.cpu().numpy(), wd.cpu().numpy()


def test_noalloc_misfit_equals_dmisft_dev(hip_lib):
    from titatest_amd import crt, meshgen
    w = meshgen.small_block()
    g = crt.CRT(w.mesh)
    lib = g.lib
    flags = crt.MISFIT_LPOL | crt.MISFIT_LROBUST
    for N, seed in ((20000, 1), (3000, 2)):
        say("noalloc misfit N=", N)
        arrays = _data(N, seed)
        a = _run(lib, g, lib.crt_dmisft_dev, N, arrays, flags, 0.5)
        b = _run(lib, g, lib.crt_dmisft_dev_noalloc, N, arrays, flags, 0.5)
        assert a[0][4] > 0.5  # the robust reweighting ran (l1rat > l1min)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    # a fresh handle whose first misfit call is the no-allocation one
    g2 = crt.CRT(w.mesh)
    arrays = _data(777, 3)
    b = _run(lib, g2, lib.crt_dmisft_dev_noalloc, 777, arrays, flags, 0.5)
    a = _run(lib, g2, lib.crt_dmisft_dev, 777, arrays, flags, 0.5)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    rc = lib.crt_dmisft_dev_noalloc(g.h, LIMIT + 1, 1, 1, 1, 1, 1, None, None, 0, 1.0, 1, 1)
    assert rc == 1004
