"""Body-atom velocities in the oracle (RigidBody::updateAtomicVelocities,
RigidBody.cpp:159-168: V_j = v_cm + w_s x delta_j), pinned by what rigid-body
kinematics fixes independently of the restatement:

- a dimer spun about an axis perpendicular to its bond keeps spinning: every
  atom moves with v_cm + w e_z x (r_j - c), with w and v_cm set by the test;
- every body atom's velocity is the time derivative of its position: a central
  difference of the trajectory's positions over +-dt agrees with the atom
  velocity to O(dt^2) (the error falls 4x per halving);
- (test_oracle_constraints) a triangle of constrained free atoms has the rigid
  body's atom velocities to O(dt^2).

The kinetic-energy identity 1/2 sum m |V|^2 = KE_t + KE_r cannot tell
w x delta from delta x w (sum m delta = 0), so each check here is also run
with that sign slip planted in the oracle, and must fail there.
"""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems


def spinning_dimer(oracle, mode, flags, steps=7, dt=0.002):
    """Atoms of mass 3 and 1, 0.2 nm apart along x, centre of mass moving at
    v_cm and the bond turning about z at w rad/ps; no forces."""
    m = np.array([3.0, 1.0])
    x = np.array([[1.00, 2.00, 3.00], [1.20, 2.00, 3.00]])
    c = (m[:, None] * x).sum(0) / m.sum()
    vcm = np.array([0.1, -0.2, 0.3])
    w = 5.0
    ez = np.array([0.0, 0.0, 1.0])
    v = vcm + w * np.cross(ez, x - c)
    o = oracle.OracleSim(m, [1, 1], dt, mode, flags=flags)
    o.set_positions(x)
    o.set_velocities(v)
    o.step(steps)
    return o, m, vcm, w, ez, (c, x - c), steps * dt


@pytest.mark.parametrize("mode", [0, 10])
def test_spinning_dimer_atom_velocities(oracle, mode):
    o, m, vcm, w, ez, (c0, d0), t = spinning_dimer(oracle, mode, oracle.FIX_ALL)
    x, v = o.positions(), o.velocities()
    c = (m[:, None] * x).sum(0) / m.sum()
    np.testing.assert_allclose(c, c0 + vcm * t, rtol=0, atol=1e-12)
    # the bond has turned by w t about z
    ct, st = np.cos(w * t), np.sin(w * t)
    rot = np.array([[ct, -st, 0.0], [st, ct, 0.0], [0.0, 0.0, 1.0]])
    np.testing.assert_allclose(x - c, d0 @ rot.T, rtol=0, atol=1e-9)
    np.testing.assert_allclose(v, vcm + w * np.cross(ez, x - c), rtol=0, atol=1e-9)


def test_spinning_dimer_catches_a_planted_velocity_sign(oracle):
    o, m, vcm, w, ez, _, _ = spinning_dimer(oracle, 0, oracle.FIX_ALL | oracle.PLANT_VELOCITY_SIGN)
    x, v = o.positions(), o.velocities()
    c = (m[:, None] * x).sum(0) / m.sum()
    err = np.abs(v - (vcm + w * np.cross(ez, x - c))).max()
    assert err > 0.5   # 2 w |r_j - c| = 1.5 nm/ps for the light atom


def central_difference_error(oracle, mode, dt, flags, steps=0.01):
    """max_j |V_j(t) - (r_j(t + dt) - r_j(t - dt)) / (2 dt)| over the body atoms
    of a small heterogeneous system in the tether field, and the largest
    |V_j - v_cm| (the rotational part that a sign slip reverses)."""
    spec = systems.heterogeneous(num_bodies=40, num_free=10, seed=11, rotation_mode=mode, dt=dt)
    o = oracle.OracleSim.from_spec(spec, flags=flags)
    n = int(round(steps / dt))
    o.step(n - 1)
    xm = o.positions()
    o.step(1)
    x0, v0 = o.positions(), o.velocities()
    o.step(1)
    xp = o.positions()
    body = spec.body_indices > 0
    err = np.abs(v0 - (xp - xm) / (2.0 * dt))[body].max()
    # rotational velocity scale: body atoms relative to their body's centre of mass
    rot = 0.0
    for b in np.unique(spec.body_indices[body]):
        sel = spec.body_indices == b
        mm = spec.masses[sel][:, None]
        vcm = (mm * v0[sel]).sum(0) / mm.sum()
        rot = max(rot, np.abs(v0[sel] - vcm).max())
    return err, rot


@pytest.mark.parametrize("mode", [0, 10])
def test_body_atom_velocities_are_position_derivatives(oracle, mode):
    e1, rot = central_difference_error(oracle, mode, 0.001, oracle.FIX_ALL)
    e2, _ = central_difference_error(oracle, mode, 0.0005, oracle.FIX_ALL)
    assert rot > 0.5
    assert e1 < 1e-3 * rot
    assert 3.0 < e1 / e2 < 5.0, (e1, e2)


def test_position_derivatives_catch_a_planted_velocity_sign(oracle):
    e, rot = central_difference_error(oracle, 0, 0.0005, oracle.FIX_ALL | oracle.PLANT_VELOCITY_SIGN)
    assert e > 0.5 * rot
