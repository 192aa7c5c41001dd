"""The C-ABI library loads and exports every entry point include/crt_hip.h
declares (no compute calls: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from titatest_amd import crt

HEADER = Path(__file__).resolve().parents[1] / "include" / "crt_hip.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^int (crt_\w+)\(", text, re.M)))


def test_header_declares_the_binding_set():
    assert declared() == sorted(crt.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not crt.LIB_PATH.exists():
        pytest.fail(f"{crt.LIB_PATH} not built; run python -m titatest_amd.build")
    lib = ctypes.CDLL(str(crt.LIB_PATH))
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_configures_all_symbols():
    lib = crt.load_library()
    for s in crt.EXPORTS:
        assert getattr(lib, s).argtypes is not None


def test_null_arguments_are_rejected_without_a_device():
    lib = crt.load_library()
    # argument validation happens before any HIP call
    assert lib.crt_create(None, None) == 1000
    assert lib.crt_destroy(None) == 0
    assert lib.crt_rwaven(None, 0, None, None, None, None, None) == 1000
    assert lib.crt_forward(None, None, None, None) == 1000
    assert lib.crt_forward_dc(None, None, None, None) == 1000
    assert lib.crt_sensitivities_dc(None, None) == 1000
    assert lib.crt_set_boundary_values(None, 0, None, None) == 1000
    assert lib.crt_smoothness(None, 0, 1.0, 1.0, 1, 1, None, None) == 1000
    assert lib.crt_set_decouplings(None, 0, None, None) == 1000
    assert lib.crt_set_graph(None, 1) == 1000
    assert lib.crt_set_device_budget(None, 1 << 30) == 1000
    assert lib.crt_set_tuning(None, 0, 0) == 1000
    assert lib.crt_forward_many(0, None, None) == 1000
    assert lib.crt_dmisft(None, 4, None, None, None, None, None, None, None, 0, 1.0, None, None) == 1000
    assert lib.crt_dmisft_dev(None, 4, None, None, None, None, None, None, None, 0, 1.0, None, None) == 1000
    assert lib.crt_set_model_map(None, 0, None) == 1000
    assert lib.crt_bpar(None, None, None) == 1000
    assert lib.crt_bsigma_dev(None, None, None) == 1000
    assert lib.crt_update_step(None, 4, None, None, 1.0, 0, None) == 1000
    assert lib.crt_blam0(None, None, 0, 1.0, 1.0, None) == 1000


def test_fortran_host_built_and_linked():
    import subprocess
    exe = crt.LIB_PATH.parent / "crmod_hip"
    if not exe.exists():
        pytest.skip("no Fortran compiler in this environment")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stdout
