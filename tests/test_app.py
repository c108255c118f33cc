"""The application layer on the CPU: rigid-body templates (python/forcefield.py)
and the reporter's columns (python/statedatareporter.py) on stand-in objects.
Expected body numbers, pairs and temperatures are written out by hand for
small topologies, not recomputed by the code under test."""
import io

import pytest

from openmm_rigidbody_plugin_amd.app import (MOLAR_GAS_CONSTANT_R, ForceField, Simulation, StateDataReporter,
                                             Topology)


def waters(n, names=("O", "H1", "H2"), extra=()):
    top = Topology()
    for _ in range(n):
        r = top.addResidue("HOH")
        for a in names:
            top.addAtom(a, r)
    for resname, atoms in extra:
        r = top.addResidue(resname)
        for a in atoms:
            top.addAtom(a, r)
    return top


def forcefield():
    ff = ForceField()
    ff.registerResidueTemplate("HOH", ["O", "H1", "H2"])
    ff.registerResidueTemplate("HO4", ["O", "H1", "H2", "M"], virtualSites=["M"])
    ff.registerResidueTemplate("NA", ["Na"])
    return ff


def test_whole_residues_are_bodies_in_topology_order():
    ff = forcefield()
    ff.registerBodyTemplate("water", "HOH")
    top = waters(3, extra=[("NA", ["Na"])])
    assert ff.resolveBodies(top) == [1, 1, 1, 2, 2, 2, 3, 3, 3, 0]


def test_virtual_sites_and_patterns_select_atoms():
    ff = forcefield()
    ff.registerBodyTemplate("tip4p", "HO4")
    top = waters(2, names=("O", "H1", "H2", "M"))
    for r in top.residues():
        r.name = "HO4"
    # the virtual site M stays out of its body
    assert ff.resolveBodies(top) == [1, 1, 1, 0, 2, 2, 2, 0]
    ff2 = forcefield()
    ff2.registerBodyTemplate("OH", "HOH", pattern="(O|H1)$")
    assert ff2.resolveBodies(waters(2)) == [1, 1, 0, 2, 2, 0]


def test_two_templates_of_one_residue_number_in_registration_order():
    ff = forcefield()
    ff.registerBodyTemplate("hydrogens", "HOH", pattern="H")
    ff.registerBodyTemplate("oxygen", "HOH", pattern="O")
    # per residue: the hydrogens body first, then the oxygen one
    assert ff.resolveBodies(waters(2)) == [2, 1, 1, 4, 3, 3]


def test_merge_lists():
    ff = forcefield()
    ff.registerBodyTemplate("water", "HOH")
    top = waters(5)
    base = [1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5]
    assert ff.resolveBodies(top, merge=None) == base
    # a flat list merges its bodies into the smallest
    assert ff.resolveBodies(top, merge=[4, 2]) == [1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 5, 5, 5]
    # overlapping groups are joined: {2, 3} and {3, 5} are one group
    assert ff.resolveBodies(top, merge=[[3, 5], [1], [2, 3]]) == [1, 1, 1, 2, 2, 2, 2, 2, 2, 4, 4, 4, 2, 2, 2]
    # a chain joined out of order: {4, 5}, {1, 2}, {2, 4}
    assert ff.resolveBodies(top, merge=[[4, 5], [1, 2], [2, 4]]) == [1] * 6 + [3, 3, 3] + [1] * 6
    with pytest.raises(ValueError, match="not a sequence"):
        ff.resolveBodies(top, merge=3)


def test_template_errors():
    ff = forcefield()
    ff.registerBodyTemplate("water", "HOH")
    with pytest.raises(ValueError, match="already been registered"):
        ff.registerBodyTemplate("water", "HOH")
    with pytest.raises(ValueError, match="Unknown residue"):
        ff.registerBodyTemplate("x", "XYZ")
    ff.registerBodyTemplate("none", "NA", pattern="Cl")
    with pytest.raises(ValueError, match="no atom in residue NA matches pattern Cl"):
        ff.resolveBodies(waters(1, extra=[("NA", ["Na"])]))
    with pytest.raises(ValueError, match="virtual sites"):
        ff.registerResidueTemplate("BAD", ["A"], virtualSites=["B"])


def test_intra_body_pairs():
    # body 2 = atoms 0, 3; body 1 = atoms 1, 2, 5; atom 4 free
    pairs = ForceField.intraBodyPairs([2, 1, 1, 2, 0, 1])
    assert pairs == [(5, 2), (5, 1), (2, 1), (3, 0)]
    assert ForceField.intraBodyPairs([0, 0]) == []


class FakeSystem:
    def __init__(self, constraints, forces=()):
        self.c = list(constraints)
        self.forces = list(forces)

    def getNumConstraints(self):
        return len(self.c)

    def getConstraintParameters(self, i):
        return self.c[i]

    def removeConstraint(self, i):
        del self.c[i]

    def getNumForces(self):
        return len(self.forces)

    def getForce(self, i):
        return self.forces[i]


class FakeNonbonded:
    def __init__(self):
        self.exceptions = []

    def addException(self, i, j, qq, sigma, eps, replace=False):
        self.exceptions.append((i, j, qq, sigma, eps, replace))


def test_remove_constraints_and_forces():
    bodies = [1, 1, 1, 0, 0, 2, 2]
    s = FakeSystem([(0, 1, 0.1), (3, 4, 0.2), (4, 5, 0.3), (5, 6, 0.4)],
                   forces=[object(), FakeNonbonded(), FakeNonbonded()])
    ForceField.removeConstraints(s, bodies)
    assert s.c == [(3, 4, 0.2)]
    ForceField.removeForces(s, bodies)
    want = [(2, 1, 0, 1, 0, True), (2, 0, 0, 1, 0, True), (1, 0, 0, 1, 0, True), (6, 5, 0, 1, 0, True)]
    assert s.forces[1].exceptions == want and s.forces[2].exceptions == want


class FakeIntegrator:
    def __init__(self):
        self.steps = []

    def step(self, n):
        self.steps.append(n)

    def getKineticEnergies(self):
        return [3.0, 1.5]

    def getRefinedKineticEnergies(self):
        return [2.0, 1.0]

    def getPotentialEnergyRefinement(self):
        return -0.25

    def getRigidBodySystem(self):
        class R:
            def getNumDOF(self):
                return 12
        return R()


class FakeContext:
    def __init__(self, integ):
        self.integ = integ

    def getState(self, getPositions=False, getVelocities=False, getForces=False, getEnergy=False):
        t = sum(self.integ.steps) * 0.002
        return {"kinetic": 4.5, "potential": -10.0, "time": t}


def test_reporter_columns_and_schedule():
    integ = FakeIntegrator()
    sim = Simulation(integ, FakeContext(integ))
    out = io.StringIO()
    sim.reporters.append(StateDataReporter(out, 4, step=True, time=True, potentialEnergy=True, kineticEnergy=True,
                                           totalEnergy=True, temperature=True, translationalEnergy=True,
                                           rotationalEnergy=True, refinedPotentialEnergy=True,
                                           refinedKineticEnergy=True, refinedTotalEnergy=True,
                                           refinedTemperature=True, refinedTranslationalEnergy=True,
                                           refinedRotationalEnergy=True))
    out2 = io.StringIO()
    sim.reporters.append(StateDataReporter(out2, 6, step=True, separator="\t"))
    sim.step(10)
    # steps end at every report of either reporter: 4, 6, 8, then 10
    assert integ.steps == [4, 2, 2, 2] and sim.currentStep == 10
    lines = out.getvalue().splitlines()
    assert lines[0] == ('#"Step","Time (ps)","Potential Energy (kJ/mole)","Kinetic Energy (kJ/mole)",'
                        '"Total Energy (kJ/mole)","Temperature (K)","Translational Energy (kJ/mole)",'
                        '"Rotational Energy (kJ/mole)","Refined Potential Energy (kJ/mole)",'
                        '"Refined Kinetic Energy (kJ/mole)","Refined Total Energy (kJ/mole)",'
                        '"Refined Temperature (K)","Refined Translational Energy (kJ/mole)",'
                        '"Refined Rotational Energy (kJ/mole)"')
    assert len(lines) == 3
    v = [float(x) for x in lines[2].split(",")]
    T = 2 * 4.5 / (12 * 0.00831446261815324)
    Tr = 2 * 3.0 / (12 * 0.00831446261815324)
    want = [8, 0.016, -10.0, 4.5, -5.5, T, 3.0, 1.5, -10.25, 3.0, -7.25, Tr, 2.0, 1.0]
    assert v == pytest.approx(want, rel=1e-12)
    assert out2.getvalue().splitlines() == ['#"Step"', "6"]
    assert MOLAR_GAS_CONSTANT_R == 8.31446261815324e-3


def test_reporter_temperature_with_com_removal():
    integ = FakeIntegrator()
    sim = Simulation(integ, FakeContext(integ), removesCOMMotion=True)
    out = io.StringIO()
    sim.reporters.append(StateDataReporter(out, 1, temperature=True))
    sim.step(1)
    assert float(out.getvalue().splitlines()[1]) == pytest.approx(2 * 4.5 / (9 * 0.00831446261815324), rel=1e-12)
