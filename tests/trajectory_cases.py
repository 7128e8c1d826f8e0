"""Trajectory cases shared by the host-logic test (CPU, the oracle standing in
for the device) and the GPU test: a model written to a file, the hook plan
(lattice, site-type and rate changes at given steps) and the records the
reference's generateTrajectory would produce, computed by stepping the oracle
to the same steps (MpdRdmeSolver.cu:485-668, 824-985)."""
import dataclasses

import numpy as np

from lattice_microbes_amd import models
import oracle

SEED = 11


def small_spec():
    return models.random_model(shape=(16, 12, 10), P=8, nS=4, nST=3, R=6, density=0.3, seed=7)


def type_counts(a, n):
    c = np.bincount(np.asarray(a).reshape(-1), minlength=256)[:n].astype(np.uint32)
    c[0] = 0
    return c


def expected_plain(spec, replicate, steps, wi, lwi):
    """Counts at 0, wi, 2wi, ...; frames at 0, lwi, ...; the max particle
    counts over the initial lattice and the frames."""
    orc = oracle.Oracle(spec, seed_top=SEED, replicate=replicate)
    counts, frames = {0: orc.species_counts()}, {0: orc.row_major()}
    maxp = type_counts(orc.lattice.view(np.uint8), spec.nS + 1)
    for t in range(1, steps + 1):
        assert orc.run(1) == 0
        if t % wi == 0:
            counts[t] = orc.species_counts()
        if t % lwi == 0:
            frames[t] = orc.row_major()
            maxp = np.maximum(maxp, type_counts(orc.lattice.view(np.uint8), spec.nS + 1))
    return counts, frames, maxp


class Edit:
    """At (x, y, z): replace the particles by one of species s and, if given,
    set the site type."""

    def __init__(self, x, y, z, species, site_type=None):
        self.x, self.y, self.z, self.s, self.st = x, y, z, species, site_type

    def apply(self, view, sites, off=0):
        view[:, self.z, self.y, self.x, :] = 0
        view[0, self.z, self.y, self.x, 0] = self.s + off
        if self.st is not None:
            sites[self.z, self.y, self.x] = self.st


BEGIN_EDIT = Edit(0, 0, 0, 4)
# hook step -> (edit, return code, new rate constant of reaction 0)
PLAN = {
    0: (Edit(1, 2, 3, 2), 1, None),
    4: (Edit(5, 6, 7, 3), 1, 2.5e6),
    8: (Edit(9, 1, 2, 1, 2), 2, None),
    12: (None, 0, None),
    16: (None, 3, None),
}
HOOK_STEPS, HOOK_WI, HOOK_LWI, HOOK_HI = 18, 2, 4, 4


class EditingHooks:
    """Mixin overriding the solver's hooks with BEGIN_EDIT and PLAN (species
    ids raised by ``species_offset``)."""

    species_offset = 0

    def onBeginTrajectory(self, lattice):
        self.calls = []
        BEGIN_EDIT.apply(lattice.getParticleLatticeView(), lattice.getSiteLatticeView(), self.species_offset)
        return 0

    def hookSimulation(self, time, lattice):
        step = int(round(time / self.tau))
        self.calls.append(step)
        edit, code, rate = PLAN[step]
        if edit is not None:
            edit.apply(lattice.getParticleLatticeView(), lattice.getSiteLatticeView(), self.species_offset)
        if rate is not None:
            self.setReactionRate(0, rate)
        assert lattice.getXSize() == 16 and lattice.getMaxOccupancy() == 8
        assert lattice.getParticleCounts() == {int(k): int(v) for k, v in
                                               enumerate(np.bincount(lattice.getParticleLatticeView().reshape(-1)))
                                               if k and v}
        return code

    def onEndTrajectory(self, lattice):
        self.ended = lattice.getParticleLatticeView().copy()
        return 0


def oracle_with_rate(spec, orc, rxid, rate, replicate):
    """A fresh oracle for the model with reaction rxid's constant replaced
    (as the float setReactionRate takes), carrying over the state."""
    K = spec.K.copy()
    K[rxid] = float(np.float32(rate))
    spec = dataclasses.replace(spec, K=K)
    nxt = oracle.Oracle(spec, seed_top=SEED, replicate=replicate)
    nxt.lattice[:] = orc.lattice
    nxt.sites[:] = orc.sites
    nxt.timestep = orc.timestep
    return spec, nxt


def expected_hooked(spec, replicate, stop=HOOK_STEPS):
    """Frames, counts, site-frame steps and the final lattice bytes of the
    hooked trajectory.  At time 0 the counts and the frame are recorded
    before the first hook call (MpdRdmeSolver.cu:539-557), at later events
    after it (:620-660).  stop < HOOK_STEPS: the trajectory ends after the
    events of step stop (a stop request made by its hook)."""
    orc = oracle.Oracle(spec, seed_top=SEED, replicate=replicate)
    W, (X, Y, Z) = spec.P // 4, spec.shape

    def views(o):
        return o.lattice.view(np.uint8).reshape(W, Z, Y, X, 4), o.sites.reshape(Z, Y, X)

    BEGIN_EDIT.apply(*views(orc))
    frames, counts, site_steps = [orc.row_major()], [orc.species_counts()], []
    for t in range(0, stop + 1):
        if t > 0:
            assert orc.run(1) == 0
        if t % HOOK_HI == 0:
            edit, code, rate = PLAN[t]
            if edit is not None:
                edit.apply(*views(orc))
            if rate is not None:
                spec, orc = oracle_with_rate(spec, orc, 0, rate, replicate)
            if code in (2, 3):
                site_steps.append(t)
        if t > 0 and t % HOOK_LWI == 0:
            frames.append(orc.row_major())
        if t > 0 and t % HOOK_WI == 0:
            counts.append(orc.species_counts())
    return frames, np.array(counts), site_steps, orc.lattice.view(np.uint8).copy()
