"""The reference's application layer, standalone.

The reference extends OpenMM's ``app.ForceField`` with rigid-body templates
(python/forcefield.py:3-178) and ``app.StateDataReporter`` with the
translational / rotational and refined energies (python/statedatareporter.py:
3-88).  OpenMM's ``app`` package is not part of this build, so the same
behaviour sits here on small stand-ins that follow OpenMM's duck types:

  * ``ForceField`` holds residue templates (atom names and which of them are
    virtual sites) registered directly, since the force-field XML reader is
    OpenMM's; ``registerBodyTemplate``, ``resolveBodies``,
    ``removeConstraints`` and ``removeForces`` keep the reference's names,
    arguments, numbering and errors, and accept OpenMM's own ``Topology`` and
    ``System`` objects as well as the stand-ins below.
  * ``Simulation`` steps an ``integrator.Context`` and calls its reporters on
    the steps they ask for, as ``app.Simulation`` does.
  * ``StateDataReporter`` writes OpenMM's columns (step, time, potential,
    kinetic and total energy, temperature) plus the reference's columns, with
    the temperature over the rigid system's degrees of freedom.

    ff = ForceField()
    ff.registerResidueTemplate('HOH', ['O', 'H1', 'H2'])
    ff.registerBodyTemplate('water', 'HOH')
    bodies = ff.resolveBodies(topology)
"""
from __future__ import annotations

import re
import sys

# kJ/(mol K), OpenMM's MOLAR_GAS_CONSTANT_R (CODATA 2018)
MOLAR_GAS_CONSTANT_R = 8.31446261815324e-3


# ---------------------------------------------------------------------------
# topology stand-in (OpenMM's app.Topology duck type: residues(), atoms(),
# getNumAtoms(); residue.name, residue.atoms(); atom.name, atom.index)
# ---------------------------------------------------------------------------
class Atom:
    def __init__(self, name, index, residue):
        self.name = name
        self.index = index
        self.residue = residue


class Residue:
    def __init__(self, name, index):
        self.name = name
        self.index = index
        self._atoms = []

    def atoms(self):
        return iter(self._atoms)


class Topology:
    def __init__(self):
        self._residues = []
        self._atoms = []

    def addResidue(self, name):
        r = Residue(name, len(self._residues))
        self._residues.append(r)
        return r

    def addAtom(self, name, residue):
        a = Atom(name, len(self._atoms), residue)
        residue._atoms.append(a)
        self._atoms.append(a)
        return a

    def residues(self):
        return iter(self._residues)

    def atoms(self):
        return iter(self._atoms)

    def getNumAtoms(self):
        return len(self._atoms)


# ---------------------------------------------------------------------------
# rigid-body templates (python/forcefield.py)
# ---------------------------------------------------------------------------
class ForceField:
    """Rigid-body templates over residue templates.

    Residue templates come from ``registerResidueTemplate`` (OpenMM reads them
    from force-field files).  A body template selects, in every residue of
    its residue template's name, the actual atoms (not virtual sites) whose
    names match a regular expression (``re.match``)."""

    class _BodyTemplate:
        def __init__(self, residue, pattern, virtualSites):
            self.residue = residue
            self.pattern = pattern
            self.virtualSites = virtualSites

    def __init__(self):
        self._templates = {}
        self._bodyTemplates = {}

    def registerResidueTemplate(self, name, atomNames, virtualSites=()):
        """A residue template: its atom names, and the names among them that
        are virtual sites."""
        atomNames = list(atomNames)
        extra = set(virtualSites) - set(atomNames)
        if extra:
            raise ValueError('virtual sites %s are not atoms of residue %s' % (sorted(extra), name))
        self._templates[name] = (atomNames, set(virtualSites))

    def getBodyTemplate(self, name):
        return self._bodyTemplates[name]

    def registerBodyTemplate(self, name, residue, pattern='.+'):
        """Register a rigid-body template (forcefield.py:21-45): `pattern`
        selects the body's atoms of `residue` by name; virtual sites never
        belong to a body."""
        if name in self._bodyTemplates:
            raise ValueError('A rigid body template named %s has already been registered.' % name)
        if residue not in self._templates:
            raise ValueError('Unknown residue %s in rigid body registration.' % residue)
        self._bodyTemplates[name] = self._BodyTemplate(residue, pattern, set(self._templates[residue][1]))

    @staticmethod
    def _mergeSets(sets):
        """Groups that share an index joined until all groups are disjoint
        (the connected components of the merge list), in order of first
        appearance (forcefield.py:47-60)."""
        groups = []
        for s in sets:
            s = set(s)
            joined = [g for g in groups if not g.isdisjoint(s)]
            if not joined:
                groups.append(s)
                continue
            first = joined[0]
            for g in joined[1:]:
                first |= g
                groups.remove(g)
            first |= s
        return groups

    def resolveBodies(self, topology, merge=None):
        """The rigid-body index of every atom of `topology` (0: not in a body),
        forcefield.py:63-108.  Every (residue, body template of that residue's
        name) pair, in topology order and in template registration order,
        opens the next body number; `merge` (a list of body numbers, or a
        list of such lists) joins bodies, each group taking its smallest
        number."""
        index = [0] * topology.getNumAtoms()
        n = 0
        for res in topology.residues():
            for body in (b for b in self._bodyTemplates.values() if b.residue == res.name):
                n += 1
                atoms = [a for a in res.atoms() if a.name not in body.virtualSites and re.match(body.pattern, a.name)]
                if not atoms:
                    raise ValueError('no atom in residue %s matches pattern %s' % (res.name, body.pattern))
                for a in atoms:
                    index[a.index] = n
        if merge is not None:
            if not hasattr(merge, '__iter__'):
                raise ValueError('merge parameter is not a sequence')
            merge = list(merge)
            if all(hasattr(m, '__iter__') for m in merge):
                groups = self._mergeSets(merge)
            else:
                groups = [set(merge)]
            target = {}
            for g in groups:
                for b in g:
                    target[b] = min(g)
            index = [target.get(b, b) for b in index]
        return index

    @staticmethod
    def intraBodyPairs(bodyIndices):
        """Every pair (i, j) of atoms of one body (forcefield.py:110-126): per
        body, each atom with each later-listed partner, atoms taken from the
        highest index down, bodies by ascending number."""
        members = {}
        for i, b in enumerate(bodyIndices):
            if b != 0:
                members.setdefault(b, []).append(i)
        pairs = []
        for b in sorted(members):
            atoms = members[b][::-1]
            for k, i in enumerate(atoms):
                for j in atoms[k + 1:]:
                    pairs.append((i, j))
        return pairs

    @staticmethod
    def removeConstraints(system, bodyIndices):
        """Removes every constraint that involves a body atom
        (forcefield.py:128-132); `system` is an OpenMM System or anything with
        getNumConstraints / getConstraintParameters / removeConstraint."""
        for i in reversed(range(system.getNumConstraints())):
            a, b, _ = system.getConstraintParameters(i)
            if bodyIndices[a] != 0 or bodyIndices[b] != 0:
                system.removeConstraint(i)

    @staticmethod
    def removeForces(system, bodyIndices):
        """Zeroes the nonbonded interaction of every pair of atoms of one body
        by an exception (charge product 0, sigma 1, epsilon 0, replacing any
        exception there), forcefield.py:134-141.  The nonbonded forces are the
        system's forces with an addException(i, j, qq, sigma, epsilon,
        replace=True) method (OpenMM's NonbondedForce)."""
        forces = [system.getForce(k) for k in range(system.getNumForces())]
        nonbonded = [f for f in forces if hasattr(f, 'addException')]
        for i, j in ForceField.intraBodyPairs(bodyIndices):
            for f in nonbonded:
                f.addException(i, j, 0, 1, 0, replace=True)


# ---------------------------------------------------------------------------
# simulation and reporter (OpenMM's app.Simulation reporter protocol)
# ---------------------------------------------------------------------------
class Simulation:
    """Steps an ``integrator.Context`` and calls each reporter at the steps it
    asks for (``describeNextReport``), as OpenMM's app.Simulation does.
    `removesCOMMotion` stands for a CMMotionRemover in the system: the
    reporters' temperatures then count three degrees of freedom fewer."""

    def __init__(self, integrator, context, removesCOMMotion=False):
        self.integrator = integrator
        self.context = context
        self.removesCOMMotion = bool(removesCOMMotion)
        self.reporters = []
        self.currentStep = 0

    def step(self, steps):
        end = self.currentStep + int(steps)
        while self.currentStep < end:
            nextStep = end
            due = []
            for r in self.reporters:
                d = r.describeNextReport(self)
                at = self.currentStep + d[0]
                if at <= nextStep:
                    if at < nextStep:
                        due = []
                    nextStep = at
                    due.append((r, d))
            if nextStep > self.currentStep:
                self.integrator.step(nextStep - self.currentStep)
                self.currentStep = nextStep
            if due:
                want = [any(d[k] for _, d in due) for k in range(1, 5)]
                state = self.context.getState(getPositions=want[0], getVelocities=want[1], getForces=want[2],
                                              getEnergy=True)
                for r, _ in due:
                    r.report(self, state)


class StateDataReporter:
    """OpenMM's StateDataReporter columns plus the reference's
    (statedatareporter.py:8-88): translational and rotational kinetic energy,
    and the refined potential, kinetic, total, translational and rotational
    energies and temperature.  Temperatures use the rigid system's degrees of
    freedom (RigidBodySystem::getNumDOF, three fewer with COM-motion
    removal).  Energies in kJ/mol, time in ps."""

    def __init__(self, file, reportInterval, step=False, time=False, potentialEnergy=False, kineticEnergy=False,
                 totalEnergy=False, temperature=False, separator=',', translationalEnergy=False,
                 rotationalEnergy=False, refinedPotentialEnergy=False, refinedKineticEnergy=False,
                 refinedTotalEnergy=False, refinedTemperature=False, refinedTranslationalEnergy=False,
                 refinedRotationalEnergy=False):
        self._own = isinstance(file, str)
        self._out = open(file, 'w') if self._own else file
        self._interval = int(reportInterval)
        self._step, self._time = step, time
        self._pe, self._ke, self._total, self._temperature = potentialEnergy, kineticEnergy, totalEnergy, temperature
        self._separator = separator
        self._trans, self._rot = translationalEnergy, rotationalEnergy
        self._rPe, self._rKe, self._rTotal = refinedPotentialEnergy, refinedKineticEnergy, refinedTotalEnergy
        self._rTemperature = refinedTemperature
        self._rTrans, self._rRot = refinedTranslationalEnergy, refinedRotationalEnergy
        self._dof = None
        self._hasInitialized = False

    def describeNextReport(self, simulation):
        steps = self._interval - simulation.currentStep % self._interval
        return (steps, False, False, False, True)

    def _initializeConstants(self, simulation):
        # statedatareporter.py:20-29
        if self._temperature or self._rTemperature:
            self._dof = simulation.integrator.getRigidBodySystem().getNumDOF()
            if simulation.removesCOMMotion:
                self._dof -= 3

    def _constructHeaders(self):
        h = []
        if self._step:
            h.append('Step')
        if self._time:
            h.append('Time (ps)')
        if self._pe:
            h.append('Potential Energy (kJ/mole)')
        if self._ke:
            h.append('Kinetic Energy (kJ/mole)')
        if self._total:
            h.append('Total Energy (kJ/mole)')
        if self._temperature:
            h.append('Temperature (K)')
        for on, name in ((self._trans, 'Translational Energy (kJ/mole)'),
                         (self._rot, 'Rotational Energy (kJ/mole)'),
                         (self._rPe, 'Refined Potential Energy (kJ/mole)'),
                         (self._rKe, 'Refined Kinetic Energy (kJ/mole)'),
                         (self._rTotal, 'Refined Total Energy (kJ/mole)'),
                         (self._rTemperature, 'Refined Temperature (K)'),
                         (self._rTrans, 'Refined Translational Energy (kJ/mole)'),
                         (self._rRot, 'Refined Rotational Energy (kJ/mole)')):
            if on:
                h.append(name)
        return h

    def _constructReportValues(self, simulation, state):
        integ = simulation.integrator
        v = []
        pe, ke = state['potential'], state['kinetic']
        if self._step:
            v.append(simulation.currentStep)
        if self._time:
            v.append(state['time'])
        if self._pe:
            v.append(pe)
        if self._ke:
            v.append(ke)
        if self._total:
            v.append(pe + ke)
        if self._temperature:
            v.append(2.0 * ke / (self._dof * MOLAR_GAS_CONSTANT_R))
        if self._trans or self._rot:
            kt, kr = integ.getKineticEnergies()
            if self._trans:
                v.append(kt)
            if self._rot:
                v.append(kr)
        u = pe
        if self._rPe or self._rTotal:
            u = pe + integ.getPotentialEnergyRefinement()
            if self._rPe:
                v.append(u)
        if self._rKe or self._rTotal or self._rTemperature or self._rTrans or self._rRot:
            kt, kr = integ.getRefinedKineticEnergies()
            if self._rKe:
                v.append(kt + kr)
            if self._rTotal:
                v.append(kt + kr + u)
            if self._rTemperature:
                v.append(2.0 * (kt + kr) / (self._dof * MOLAR_GAS_CONSTANT_R))
            if self._rTrans:
                v.append(kt)
            if self._rRot:
                v.append(kr)
        return v

    def report(self, simulation, state):
        if not self._hasInitialized:
            self._initializeConstants(simulation)
            print('#"%s"' % ('"' + self._separator + '"').join(self._constructHeaders()), file=self._out)
            self._hasInitialized = True
        values = self._constructReportValues(simulation, state)
        print(self._separator.join(str(x) for x in values), file=self._out)
        if hasattr(self._out, 'flush'):
            self._out.flush()

    def __del__(self):
        if getattr(self, '_own', False):
            self._out.close()


__all__ = ['Topology', 'ForceField', 'Simulation', 'StateDataReporter', 'MOLAR_GAS_CONSTANT_R']
