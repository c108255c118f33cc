"""Constraints between free atoms applied by the caller to adopted arrays, in
the CUDA platform's order (CudaRigidBodyKernels.cpp:405-432): the library
writes the free atoms' drift displacements into the caller's posDelta
(rbhip_free_delta), the caller's constraint solver adjusts them, part 1 moves
the free atoms by them (rbhip_part1_constrained), the caller evaluates the
forces, part 2 adds the correction over dt to the velocities (the
reference's (R - saved) / dt), and the caller applies its velocity
constraints.  The test plays OpenMM: it owns the arrays and solves the
constraints on the host with SHAKE and RATTLE's velocity stage in the
oracle's form (Gauss-Seidel in the order given, OpenMM's convergence tests)
at a tight tolerance, so the trajectory must match the FP64 oracle's
constrained one at the tolerances of test_gpu_constraints."""
import numpy as np
import pytest

from test_gpu_constraints import TOL, constrained_spec
from test_gpu_force_slot import adopted_context, tether

pytestmark = pytest.mark.gpu

CTOL = 1e-10


def shake(r, r_old, cons, inv_mass, tol=CTOL):
    """Positions r (modified) to the constraint lengths, moving along the
    pre-step bond directions r_old."""
    lo, hi = (1.0 - tol) ** 2, (1.0 + tol) ** 2
    for _ in range(1000):
        done = True
        for a, b, d in cons:
            s = r[a] - r[b]
            s2, d2 = s @ s, d * d
            if lo * d2 <= s2 <= hi * d2:
                continue
            done = False
            ro = r_old[a] - r_old[b]
            g = (d2 - s2) / (2.0 * (inv_mass[a] + inv_mass[b]) * (s @ ro))
            r[a] += ro * (g * inv_mass[a])
            r[b] -= ro * (g * inv_mass[b])
        if done:
            return
    raise AssertionError("SHAKE did not converge")


def rattle(r, v, cons, inv_mass, tol=CTOL):
    """Velocities v (modified) without relative motion along each bond."""
    for _ in range(1000):
        done = True
        for a, b, _d in cons:
            s = r[a] - r[b]
            dv = v[a] - v[b]
            rv, s2 = s @ dv, s @ s
            if rv * rv <= tol * tol * s2 * (dv @ dv):
                continue
            done = False
            g = -rv / ((inv_mass[a] + inv_mass[b]) * s2)
            v[a] += s * (g * inv_mass[a])
            v[b] -= s * (g * inv_mass[b])
        if done:
            return
    raise AssertionError("RATTLE did not converge")


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_caller_constraints_on_adopted_arrays_match_oracle(hip, oracle, precision, mode):
    from openmm_rigidbody_plugin_amd.hipmem import DeviceArray
    spec, cons = constrained_spec(mode)
    inv_mass = 1.0 / spec.masses
    o = oracle.OracleSim.from_spec(spec)
    for a, b, d in cons:
        o.add_constraint(a, b, d)
    o.set_constraint_tolerance(CTOL)
    o.set_positions(spec.positions)
    o.set_velocities(spec.velocities)

    g, arr, rng = adopted_context(hip, spec, precision, mode)
    pos_delta = DeviceArray(np.zeros((arr.padded, 4)))
    arr.write_positions(spec.positions)
    arr.write_velocities(spec.velocities)
    arr.write_forces(tether(spec, spec.positions))
    g.state_changed(g.STATE_POSITIONS)
    free = np.flatnonzero(spec.body_indices <= 0)
    for step in range(10):
        # freeAtomsDelta, applyConstraints, part 1 (CudaRigidBodyKernels.cpp:414-421)
        r_old = arr.positions()
        pos_delta.upload(np.zeros((arr.padded, 4)))
        g.free_delta(pos_delta.ptr.value)
        pd = pos_delta.download()
        d = np.zeros((spec.num_atoms, 3))
        d[arr.order] = pd[:len(arr.order), :3]
        assert np.all(d[spec.body_indices > 0] == 0.0)   # only the free atoms' slots are written
        r = r_old + d
        shake(r, r_old, cons, inv_mass)
        pd[:len(arr.order), :3] = (r - r_old)[arr.order]
        pos_delta.upload(pd)
        g.part1_constrained(pos_delta.ptr.value)
        # the free atoms sit at the constrained positions
        np.testing.assert_allclose(arr.positions()[free], r[free], rtol=0, atol=1e-6 if precision == "mixed" else 1e-12)
        # forces, part 2, applyVelocityConstraints (:424-429)
        arr.write_forces(tether(spec, arr.positions()))
        g.part2()
        v = arr.velocities()
        rattle(arr.positions(), v, cons, inv_mass)
        arr.write_velocities(v)
        if step == 4:
            arr.reorder(rng.permutation(spec.num_atoms).astype(np.int32))
            g.set_atom_order(arr.order)
    o.step(10)
    t = TOL[precision]
    np.testing.assert_allclose(arr.positions(), o.positions(), rtol=0, atol=t["pos"])
    vo = o.velocities()
    np.testing.assert_allclose(arr.velocities(), vo, rtol=0, atol=t["vel"] * np.abs(vo).max())
    # the kinetic energies re-read the free atoms' constrained velocities
    np.testing.assert_allclose(g.kinetic_energies(), o.kinetic_energies(), rtol=t["vel"])
    arr.check_w_untouched()


def test_constrained_part1_refusals(hip):
    from openmm_rigidbody_plugin_amd.hipmem import DeviceArray
    spec, _ = constrained_spec(10)
    lib = hip.Context.from_spec(spec, precision="mixed")
    pos_delta = DeviceArray(np.zeros((spec.num_atoms + 64, 4)))
    with pytest.raises(hip.RbhipError):
        lib.free_delta(pos_delta.ptr.value)        # not adopted arrays
    with pytest.raises(hip.RbhipError):
        lib.part1_constrained(pos_delta.ptr.value)
    g, arr, _ = adopted_context(hip, spec, "mixed", 10)
    arr.write_positions(spec.positions)
    arr.write_velocities(spec.velocities)
    arr.write_forces(tether(spec, spec.positions))
    g.state_changed(g.STATE_POSITIONS)
    pd = DeviceArray(np.zeros((arr.padded, 4)))
    # refined energies do not combine with the caller's constraints
    g.set_refined(True)
    with pytest.raises(hip.RbhipError):
        g.free_delta(pd.ptr.value)
    with pytest.raises(hip.RbhipError):
        g.part1_constrained(pd.ptr.value)
    g.set_refined(False)
    g.free_delta(pd.ptr.value)
    g.part1_constrained(pd.ptr.value)
    # the constrained part 1 must be completed by part 2
    for call in (lambda: g.part1(), lambda: g.step(1), lambda: g.free_delta(pd.ptr.value),
                 lambda: g.part1_constrained(pd.ptr.value), lambda: g.time_kernels(1), lambda: g.autotune()):
        with pytest.raises(hip.RbhipError):
            call()
    g.part2()
    g.free_delta(pd.ptr.value)
    # a state change abandons a pending constrained step
    g.part1_constrained(pd.ptr.value)
    g.state_changed(g.STATE_POSITIONS)
    g.part1()
    g.part2()
    with pytest.raises(hip.RbhipError):
        g.set_refined(True)           # after constrained steps
