"""The reference's Python interface to the integrator, by its own names.

The reference wraps ``RigidBodyIntegrator`` and ``RigidBodySystem`` with SWIG
(python/rigidbodyplugin.i:51-72).  Here the same methods sit over the C++
plugin layer on the HIP platform (``plugin.Simulation``), and a ``Context``
binds an integrator to a system's masses and forces as OpenMM's Context does.
As in the reference (openmmapi/src/RigidBodyIntegrator.cpp:26-38), the
rotation mode and the refined-energy switch can be set only before the
integrator is bound, and a negative rotation mode is refused.

    integrator = RigidBodyIntegrator(0.002, bodyIndices)
    integrator.setRotationMode(10)
    context = Context(masses, integrator, tether_k=k, tether_r0=r0)
    context.setPositions(x)
    context.setVelocities(v)
    integrator.step(1000)
    ke_trans, ke_rot = integrator.getKineticEnergies()
"""
from __future__ import annotations

import numpy as np

from . import plugin


class RigidBodySystem:
    """The body bookkeeping and kinetic energies of a bound integrator at the
    time it was asked for (RigidBodySystem's getters).  The reference hands out
    the copy it built when the state was last set, which goes stale as the
    integrator steps (its TODO.md:5 lists downloading it first); here the
    values are read from the device when asked for."""

    def __init__(self, num_dof, kinetic):
        self._dof = int(num_dof)
        self._kt, self._kr = float(kinetic[0]), float(kinetic[1])

    def getNumDOF(self):
        return self._dof

    def getTranslationalEnergy(self):
        return self._kt

    def getRotationalEnergy(self):
        return self._kr

    def getKineticEnergy(self):
        return self._kt + self._kr


class RigidBodyIntegrator:
    """RigidBodyIntegrator(stepSize, bodyIndices): body index 0 marks a free
    atom, equal positive indices one rigid body."""

    def __init__(self, stepSize, bodyIndices):
        self._dt = float(stepSize)
        self._bodies = np.ascontiguousarray(bodyIndices, dtype=np.int32)
        self._mode = 0
        self._refined = False
        self._sim = None

    def _bound(self):
        if self._sim is None:
            raise RuntimeError("the integrator is not bound to a context")
        return self._sim

    def setRotationMode(self, mode):
        if mode < 0:
            raise RuntimeError("Rotation mode cannot be negative")
        if self._sim is not None:
            raise RuntimeError("Cannot set rotation mode: integrator already bound to a context")
        self._mode = int(mode)

    def getRotationMode(self):
        return self._mode

    def setComputeRefinedEnergies(self, compute):
        if self._sim is not None:
            raise RuntimeError("Cannot set refined energy computation: integrator already bound to a context")
        self._refined = bool(compute)

    def getComputeRefinedEnergies(self):
        return self._refined

    def getStepSize(self):
        return self._dt

    def setStepSize(self, size):
        if self._sim is not None:
            self._sim.set_step_size(float(size))
        self._dt = float(size)

    def step(self, steps):
        self._bound().step(int(steps))

    def getKineticEnergies(self):
        return [float(v) for v in self._bound().energies()["kinetic"]]

    def getRefinedKineticEnergies(self):
        return [float(v) for v in self._bound().energies()["refined_kinetic"]]

    def getPotentialEnergyRefinement(self):
        return float(self._bound().energies()["potential_refinement"])

    def getRigidBodySystem(self):
        sim = self._bound()
        return RigidBodySystem(sim.counts()["numDOF"], sim.energies()["kinetic"])


class Context:
    """Binds a RigidBodyIntegrator to a system on the HIP platform: per-atom
    masses and the force field: harmonic bonds as (i, j, length, k) and
    harmonic tethers of constant k to points r0, evaluated on the device, and
    `force`, any force field written in Python, f(positions) -> (forces,
    energy) with (n, 3) arrays, evaluated between part 1 and part 2 of every
    step as the reference's execute() evaluates the context's forces."""

    def __init__(self, masses, integrator, precision="mixed", device=0, bonds=None, tether_k=None,
                 tether_r0=None, force=None):
        if integrator._sim is not None:
            raise RuntimeError("the integrator is already bound to a context")
        self._sim = plugin.Simulation(masses, integrator._bodies, integrator._dt, integrator._mode, precision,
                                      device, bonds=bonds, tether_k=tether_k, tether_r0=tether_r0,
                                      refined_energies=integrator._refined, host_force=force)
        integrator._sim = self._sim
        self._integrator = integrator

    def setPositions(self, positions):
        self._sim.set_positions(positions)

    def setVelocities(self, velocities):
        self._sim.set_velocities(velocities)

    def getState(self, getPositions=False, getVelocities=False, getForces=False, getEnergy=False):
        """A dict with the arrays asked for and, with getEnergy, the kinetic
        and potential energies and the time."""
        return self._sim.state(positions=getPositions, velocities=getVelocities, forces=getForces,
                               energy=getEnergy)
