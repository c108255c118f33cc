"""The force slot (ReferenceRigidBodyKernels.cpp:98-104: part 1, the
context's force evaluation, part 2) with forces computed outside the library,
through the C ABI:

- host forces: numpy tethers evaluated at the GPU's positions after each
  part 1 and handed over with rbhip_set_forces drive the same trajectory as
  the FP64 oracle's tether (C1 tolerances), and forces equal to the
  library's own give the library's trajectory bit for bit;
- adopted arrays in OpenMM's GPU layout (posq + posqCorrection as float4,
  velm as double4, fixed-point forces at 2^32, a permuted atom order that
  changes mid-run as OpenMM's reordering does): the test owns the arrays as
  an OpenMM context would, evaluates the tether on the host into the
  fixed-point array, and the trajectory matches the oracle; with the same
  forces it equals the library-layout trajectory bit for bit.

Tolerances after 10 steps (as test_gpu_parity): mixed 1e-6 nm positions,
velocities 1e-5 of their RMS, KE 1e-5 relative; double 1e-10, 1e-9, 1e-11.
"""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems

pytestmark = pytest.mark.gpu

TOL = {"mixed": dict(x=1e-6, v=1e-5, ke=1e-5), "double": dict(x=1e-10, v=1e-9, ke=1e-11)}
FIXED = 4294967296.0


def tether(spec, x):
    """The benchmark force field, on the host in FP64."""
    return -spec.tether_k[:, None] * (x - spec.tether_r0)


def quantized(f):
    """OpenMM's fixed-point force (truncation of f 2^32) and its value."""
    q = np.trunc(f * FIXED).astype(np.int64)
    return q, q.astype(np.float64) / FIXED


def check_against_oracle(spec, o, x, v, ke, precision):
    tol = TOL[precision]
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
    dx = np.abs(o.positions() - x).max()
    assert dx <= tol["x"], dx
    vo = o.velocities()
    vscale = np.sqrt((vo ** 2).sum(1).mean())
    assert np.abs(vo - v).max() <= tol["v"] * vscale, np.abs(vo - v).max() / vscale
    np.testing.assert_allclose(ke, o.kinetic_energies(), rtol=tol["ke"], atol=0)


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_host_forces_match_oracle(hip, oracle, precision, mode):
    spec = systems.water_c1(rotation_mode=mode)
    o = oracle.OracleSim.from_spec(spec)
    # no force field in the library: every force comes through the slot
    g = hip.Context(spec.masses, spec.body_indices, spec.dt, mode, precision)
    g.set_positions(spec.positions)
    g.set_velocities(spec.velocities)
    g.set_forces(tether(spec, spec.positions))   # the state change's force evaluation
    for _ in range(10):
        g.part1()
        g.set_forces(tether(spec, g.positions()))
        g.part2()
    o.step(10)
    check_against_oracle(spec, o, g.positions(), g.velocities(), g.kinetic_energies(), precision)
    t, n = g.time()
    assert n == 10 and t == pytest.approx(o.time(), rel=1e-15)


@pytest.mark.parametrize("mode", [10, 0])
def test_slot_forces_equal_library_forces_bitwise(hip, mode):
    """Forces handed over through the slot that equal the library's own give
    its trajectory bit for bit, mid-step and when set again after part 2
    (the folded half-kick is redone from an exactly zero difference)."""
    spec = systems.heterogeneous(600, 400, seed=21, rotation_mode=mode)
    a = hip.Context.from_spec(spec)
    b = hip.Context.from_spec(spec)
    for _ in range(6):
        a.part1()
        b.part1()
        a.compute_forces()
        b.set_forces(a.forces())
        a.part2()
        b.part2()
    b.set_forces(a.forces())
    a.step(5)
    b.step(5)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())
    assert np.array_equal(a.kinetic_energies(), b.kinetic_energies())


def test_set_forces_after_a_state_change_redoes_the_half_kick(hip, oracle):
    """Forces set after rbhip_set_positions replace those the first half-kick
    was made with: the trajectory is the one the forces would have given at
    the state change itself (here compared with the library's own tether)."""
    spec = systems.water_c1(rotation_mode=10)
    a = hip.Context.from_spec(spec, precision="double")
    b = hip.Context(spec.masses, spec.body_indices, spec.dt, 10, "double")
    b.set_positions(spec.positions)
    b.set_velocities(spec.velocities)
    b.set_forces(tether(spec, spec.positions))
    for _ in range(3):
        a.part1()
        b.part1()
        a.compute_forces()
        b.set_forces(tether(spec, b.positions()))
        a.part2()
        b.part2()
    assert np.abs(a.positions() - b.positions()).max() < 1e-11
    assert np.abs(a.velocities() - b.velocities()).max() < 1e-9


# ---------------------------------------------------------------------------
# adopted arrays in OpenMM's GPU layout
# ---------------------------------------------------------------------------
class OpenMMArrays:
    """The atom arrays of an OpenMM context on the GPU, owned by the test:
    slot s holds original atom order[s]; padded slots beyond the atoms."""

    def __init__(self, spec, precision, order, padded):
        from openmm_rigidbody_plugin_amd.hipmem import DeviceArray
        n = spec.num_atoms
        self.n, self.padded, self.mixed = n, padded, precision == "mixed"
        self.order = np.asarray(order, dtype=np.int32)
        rt = np.float32 if self.mixed else np.float64
        posq = np.zeros((padded, 4), dtype=rt)
        corr = np.zeros((padded, 4), dtype=np.float32)
        velm = np.zeros((padded, 4), dtype=np.float64)
        self.charges = np.linspace(-1.0, 1.0, padded).astype(rt)
        self.inv_mass = np.zeros(padded)
        posq[:, 3] = self.charges
        m = spec.masses[self.order]
        self.inv_mass[:len(self.order)] = np.where(m > 0, 1.0 / np.where(m > 0, m, 1.0), 0.0)
        velm[:, 3] = self.inv_mass
        self.posq = DeviceArray(posq)
        self.corr = DeviceArray(corr) if self.mixed else None
        self.velm = DeviceArray(velm)
        self.force = DeviceArray(np.zeros((3, padded), dtype=np.int64))

    def pointers(self):
        return (self.posq.ptr.value, self.corr.ptr.value if self.corr else None, self.velm.ptr.value,
                self.force.ptr.value, self.padded, self.order)

    # what OpenMM's setPositions / setVelocities / getState do with the arrays
    def write_positions(self, x):
        posq = self.posq.download()
        xs = x[self.order]
        hi = xs.astype(posq.dtype)
        posq[:len(self.order), :3] = hi
        self.posq.upload(posq)
        if self.corr:
            corr = self.corr.download()
            corr[:len(self.order), :3] = (xs - hi.astype(np.float64)).astype(np.float32)
            self.corr.upload(corr)

    def write_velocities(self, v):
        velm = self.velm.download()
        velm[:len(self.order), :3] = v[self.order]
        self.velm.upload(velm)

    def positions(self):
        posq = self.posq.download()
        x = posq[:len(self.order), :3].astype(np.float64)
        if self.corr:
            x = x + self.corr.download()[:len(self.order), :3].astype(np.float64)
        out = np.empty((self.n, 3))
        out[self.order] = x
        return out

    def velocities(self):
        out = np.empty((self.n, 3))
        out[self.order] = self.velm.download()[:len(self.order), :3]
        return out

    def write_forces(self, f):
        """A force evaluation: fixed-point forces per slot."""
        q, _ = quantized(f)
        arr = np.zeros((3, self.padded), dtype=np.int64)
        arr[:, :len(self.order)] = q[self.order].T
        self.force.upload(arr)

    def reorder(self, new_order):
        """OpenMM's reorderAtoms: positions and velocities move to the new
        slots (the force array is not touched)."""
        posq, velm = self.posq.download(), self.velm.download()
        corr = self.corr.download() if self.corr else None
        inv_new = np.empty_like(new_order)
        inv_new[new_order] = np.arange(len(new_order))
        src = np.empty(len(new_order), dtype=np.int64)   # new slot s takes old slot of atom new_order[s]
        inv_old = np.empty_like(self.order)
        inv_old[self.order] = np.arange(len(self.order))
        src[:] = inv_old[new_order]
        n = len(new_order)
        posq[:n], velm[:n] = posq[src], velm[src]
        self.posq.upload(posq)
        self.velm.upload(velm)
        if corr is not None:
            corr[:n] = corr[src]
            self.corr.upload(corr)
        self.charges[:n] = self.charges[src]
        self.inv_mass[:n] = self.inv_mass[src]
        self.order = np.asarray(new_order, dtype=np.int32)

    def check_w_untouched(self):
        np.testing.assert_array_equal(self.posq.download()[:, 3], self.charges)
        np.testing.assert_array_equal(self.velm.download()[:, 3], self.inv_mass)


def adopted_context(hip, spec, precision, mode, seed=5, tethered=False):
    rng = np.random.default_rng(seed)
    n = spec.num_atoms
    order = rng.permutation(n).astype(np.int32)
    padded = (n + 31) // 32 * 32 + 32
    arrays = OpenMMArrays(spec, precision, order, padded)
    g = hip.Context(spec.masses, spec.body_indices, spec.dt, mode, precision)
    if tethered:
        g.set_tether(spec.tether_k, spec.tether_r0)
    g.adopt_arrays(*arrays.pointers())
    return g, arrays, rng


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_adopted_arrays_match_oracle(hip, oracle, precision, mode):
    spec = systems.water_c1(rotation_mode=mode)
    o = oracle.OracleSim.from_spec(spec)
    g, arr, rng = adopted_context(hip, spec, precision, mode)
    # Context::setPositions / setVelocities, then the integrator's stateChanged
    # after the force evaluation at the new positions
    arr.write_positions(spec.positions)
    arr.write_velocities(spec.velocities)
    arr.write_forces(tether(spec, spec.positions))
    g.state_changed(g.STATE_POSITIONS)
    for step in range(10):
        g.part1()
        arr.write_forces(tether(spec, arr.positions()))
        g.part2()
        if step == 4:
            # OpenMM reorders its atoms between steps; the listener call
            # moves the forces along
            arr.reorder(rng.permutation(spec.num_atoms).astype(np.int32))
            g.set_atom_order(arr.order)
    o.step(10)
    check_against_oracle(spec, o, arr.positions(), arr.velocities(), g.kinetic_energies(), precision)
    # the library's own readers see the same arrays
    np.testing.assert_array_equal(g.positions(), arr.positions())
    np.testing.assert_array_equal(g.velocities(), arr.velocities())
    arr.check_w_untouched()


@pytest.mark.parametrize("refined", [False, True])
@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_adopted_layout_equals_library_layout(hip, precision, mode, refined):
    """The same forces through the fixed-point array and through
    rbhip_set_forces: the two atom layouts give one trajectory, and with
    refined energies on (each part1/part2 pair brackets its step) the same
    refined kinetic energies and potential-energy refinement."""
    spec = systems.heterogeneous(500, 300, seed=23, rotation_mode=mode)
    g, arr, _ = adopted_context(hip, spec, precision, mode, seed=7)
    b = hip.Context(spec.masses, spec.body_indices, spec.dt, mode, precision)
    # the same state changes on both: positions (velocities still zero),
    # then velocities (a dynamics rebuild on the stored body frame)
    arr.write_positions(spec.positions)
    f0 = quantized(tether(spec, spec.positions))[1]
    arr.write_forces(f0)
    g.state_changed(g.STATE_POSITIONS)
    arr.write_velocities(spec.velocities)
    g.state_changed(g.STATE_VELOCITIES)
    b.set_positions(arr.positions())
    b.set_velocities(arr.velocities())
    b.set_forces(f0)
    for c in (g, b):
        c.set_refined(refined)
    for _ in range(8):
        g.part1()
        b.part1()
        f = quantized(tether(spec, arr.positions()))[1]
        arr.write_forces(f)
        b.set_forces(f)
        g.part2()
        b.part2()
    np.testing.assert_array_equal(g.velocities(), b.velocities())
    np.testing.assert_array_equal(g.kinetic_energies(), b.kinetic_energies())
    if refined:
        np.testing.assert_array_equal(g.refined_kinetic_energies(), b.refined_kinetic_energies())
        assert g.potential_energy_refinement() == b.potential_energy_refinement() != 0.0
    free = spec.body_indices <= 0
    np.testing.assert_array_equal(g.positions()[free], b.positions()[free])
    # body atoms: the adopted arrays hold the step kernel's r_cm + A(q)^T d at
    # full precision; the library's own layout rebuilds it for the download
    # in a second kernel, whose fp32 rotation of d may round differently
    # (different contraction), by ~ulp(|d|) in fp32
    tol = 2e-7 if precision == "mixed" else 1e-13
    assert np.abs(g.positions() - b.positions()).max() <= tol


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_adopted_arrays_with_the_library_force_field(hip, oracle, precision, mode):
    """rbhip_execute on adopted arrays: the library's tether writes the
    fixed-point force array, and the fused, graph-replayed steps match the
    oracle; host-set positions and velocities are written through."""
    spec = systems.water_c1(rotation_mode=mode)
    o = oracle.OracleSim.from_spec(spec)
    g, arr, _ = adopted_context(hip, spec, precision, mode, tethered=True)
    g.set_positions(spec.positions)
    g.set_velocities(spec.velocities)
    x = arr.positions()
    assert np.abs(x - spec.positions).max() <= (1e-13 if precision == "mixed" else 0.0)
    g.step(10)
    o.step(10)
    check_against_oracle(spec, o, arr.positions(), arr.velocities(), g.kinetic_energies(), precision)
    # a velocity change written by the owner, announced by state_changed
    v = spec.velocities * 0.5
    arr.write_velocities(v)
    o.set_velocities(v)
    g.state_changed(g.STATE_VELOCITIES)
    np.testing.assert_allclose(g.kinetic_energies(), o.kinetic_energies(), rtol=TOL[precision]["ke"] * 10)
    g.step(4)
    o.step(4)
    check_against_oracle(spec, o, arr.positions(), arr.velocities(), g.kinetic_energies(), precision)
    arr.check_w_untouched()


def test_adopted_arrays_refusals(hip):
    spec = systems.water_box(8, 0.3125, num_free=4, seed=3)
    g, arr, _ = adopted_context(hip, spec, "mixed", 10)
    with pytest.raises(hip.RbhipError):
        g.set_forces(np.zeros((spec.num_atoms, 3)))      # the forces live in the array
    with pytest.raises(hip.RbhipError):
        g.adopt_arrays(*arr.pointers())                   # adopted once
    with pytest.raises(hip.RbhipError):
        g.set_atom_order(np.arange(spec.num_atoms - 1))   # too short
    with pytest.raises(hip.RbhipError):
        g.set_atom_order(np.zeros(spec.num_atoms, dtype=np.int32))   # not a permutation
    with pytest.raises(hip.RbhipError):
        g.step(1)                                         # no state read yet
    b = hip.Context(spec.masses, spec.body_indices, spec.dt, 10, "mixed")
    with pytest.raises(hip.RbhipError):
        b.state_changed(b.STATE_POSITIONS)                # no adopted arrays
    ptrs = list(arr.pointers())
    ptrs[1] = None                                        # mixed precision needs posqCorrection
    with pytest.raises(hip.RbhipError):
        b.adopt_arrays(*ptrs)
    c = hip.Context(spec.masses, spec.body_indices, spec.dt, 10, "double")
    with pytest.raises(hip.RbhipError):
        c.adopt_arrays(*arr.pointers())                   # double precision has no correction array
