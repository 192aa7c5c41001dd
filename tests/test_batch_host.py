"""Host pieces of the batch driver (titatest_amd/batch.py) that need no
device: R m of the rectangular smoothness operator against the oracle's
restatement of update's bvec (src/update.f90:63-84)."""
import numpy as np

from conftest import rel_err
from titatest_amd import meshgen


def test_rm_rect_matches_oracle(oracle_lib):
    from titatest_amd.batch import rm_rect
    w = meshgen.small_block()
    o = oracle_lib.Oracle(w.mesh)
    sm = oracle_lib.bsmatmreg(o, 40, 20, 1.5, 0.5)
    rng = np.random.default_rng(1)
    m = rng.standard_normal(o.elanz) + 1j * rng.standard_normal(o.elanz)
    assert rel_err(rm_rect(sm, 40, m), oracle_lib.rm_reg(sm, 40, m)) < 1e-14
    # constants are in the operator's null space
    assert np.max(np.abs(rm_rect(sm, 40, np.ones(o.elanz, complex)))) < 1e-12
