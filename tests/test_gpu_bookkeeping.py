"""Device bookkeeping (rbhip_body_bookkeeping) against the oracle's
cleanBodyIndices and atom order (RigidBodySystem.cpp:28-105), bit for bit:
labels with gaps, negative and zero labels, labels up to 2^31 - 1,
interleaved atoms, massless particles labelled free, and a 3M-atom system
(the 1M-water labels, shuffled).  The oracle itself is pinned by a
brute-force statement of the ordering (test_oracle_system)."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems

pytestmark = pytest.mark.gpu


def run_device(hip, labels, masses):
    from openmm_rigidbody_plugin_amd.hipmem import DeviceArray
    n = labels.size
    dl, dm = DeviceArray(labels.astype(np.int32)), DeviceArray(masses.astype(np.float64))
    bi = DeviceArray(np.zeros(n, np.int32))
    ai = DeviceArray(np.zeros(n, np.int32))
    bs = DeviceArray(np.zeros(n + 1, np.int32))
    c = hip.body_bookkeeping(n, dl.ptr.value, dm.ptr.value, bi.ptr.value, ai.ptr.value, bs.ptr.value)
    num_free, num_bodies, num_massless, num_body_atoms = c
    return c, bi.download(), ai.download()[:num_free + num_body_atoms], bs.download()[:num_bodies + 1]


def expected(oracle, labels, masses):
    """The oracle's bookkeeping; body starts from its body sizes."""
    bi = oracle.clean_body_indices(labels)
    massive = masses > 0
    free = np.nonzero((bi == 0) & massive)[0]
    nb = int(bi.max(initial=0))
    sizes = np.bincount(bi[bi > 0], minlength=nb + 1)[1:]
    atoms = np.concatenate([free] + [np.nonzero(bi == b)[0] for b in range(1, nb + 1)]).astype(np.int32)
    return bi, atoms, np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32), free.size


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bookkeeping_matches_oracle(hip, oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2000, 20000))
    pool = np.unique(rng.integers(1, 2**31 - 1, size=n // 3, dtype=np.int64))
    labels = rng.choice(np.concatenate([pool, [0, 0, -1, -5]]), size=n).astype(np.int32)
    masses = rng.uniform(1.0, 20.0, n)
    masses[(labels <= 0) & (rng.uniform(size=n) < 0.1)] = 0.0      # massless particles, labelled free
    c, bi, ai, bs = run_device(hip, labels, masses)
    ebi, eai, ebs, nfree = expected(oracle, labels, masses)
    np.testing.assert_array_equal(bi, ebi)
    np.testing.assert_array_equal(ai, eai)
    np.testing.assert_array_equal(bs, ebs)
    assert c == (nfree, int(ebi.max()), int(np.count_nonzero(masses == 0)), int(np.count_nonzero(ebi)))


def test_bookkeeping_of_the_headline_system(hip, oracle):
    spec = systems.water_c2()
    rng = np.random.default_rng(9)
    perm = rng.permutation(spec.num_atoms)
    labels = spec.body_indices[perm] * 7 + 3          # sparse labels, atoms interleaved
    c, bi, ai, bs = run_device(hip, labels, spec.masses[perm])
    np.testing.assert_array_equal(bi, oracle.clean_body_indices(labels))
    order = np.argsort(np.where(labels > 0, labels, 0), kind="stable")
    np.testing.assert_array_equal(ai, order)
    assert c == (0, 1_000_000, 0, 3_000_000)
    np.testing.assert_array_equal(bs, np.arange(0, 3_000_001, 3))


def test_bookkeeping_refusals(hip):
    labels = np.array([1, 1, 0, 2, 2], np.int32)
    with pytest.raises(hip.RbhipError, match="massless"):
        run_device(hip, labels, np.array([1.0, 0.0, 1.0, 2.0, 2.0]))
    with pytest.raises(hip.RbhipError, match="mass"):
        run_device(hip, labels, np.array([1.0, 1.0, -1.0, 2.0, 2.0]))
