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


def test_rm_tri_is_the_dense_operator(oracle_lib):
    """R m of the triangular operator (src/update.f90:99-117) on the
    reference-pinned bsmatmtri/bnachbar of a terrain mesh: equal to the dense
    R assembled from smatm and nachbar (diagonal + neighbour couplings), and
    constants in its null space."""
    from titatest_amd.batch import rm_tri
    w = meshgen.topo_triangles(nx_nodes=25, nz_nodes=10, eanz=12)
    o = oracle_lib.Oracle(w.mesh)
    nb = oracle_lib.bnachbar(o)
    sm = oracle_lib.bsmatmtri(o, 1.0, 0.5, nb)
    M = o.elanz
    R = np.diag(sm[:, 3])
    for i in range(M):
        for k in range(3):
            if nb[i, k] > 0:
                R[i, nb[i, k] - 1] += sm[i, k]
    rng = np.random.default_rng(2)
    m = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    assert rel_err(rm_tri(sm, nb[:, :3], m), R @ m) < 1e-14
    assert np.max(np.abs(rm_tri(sm, nb[:, :3], np.ones(M, complex)))) < 1e-12
