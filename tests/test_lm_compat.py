"""The ``lm`` compatibility module: a model file built the way jLM's
``LmWriteMixin.finalize`` and pyLM's ``RDMESimulation.save`` build it
(src/jlm/jLM/LmInteract.py:47-100,231-330; src/pylm/pyLM/RDME.py:480-600),
through ``lm.ReactionModel``/``lm.DiffusionModel``/``lm.ByteLattice``/
``lm.SimulationFile``, holds the same model, lattice and sites as the file
written from the same arrays by ModelSpec.to_file; solver composition as
jLM's ``makeSolver`` does it (src/jlm/jLM/Solvers.py:45-65)."""
import numpy as np
import pytest

from lattice_microbes_amd import InvalidArgException, compat
from lattice_microbes_amd.simfile import SimulationFile as RawFile
from trajectory_cases import small_spec

compat.install()
import lm  # noqa: E402


def build_with_lm(spec, path):
    """The jLM sequence of calls, from a ModelSpec's arrays."""
    nS, R, nST = spec.nS, spec.R, spec.nST
    X, Y, Z = spec.shape
    rm = lm.ReactionModel()
    rm.set_number_reactions(R)
    rm.set_number_species(nS)
    for c in spec.initial_counts():
        rm.add_initial_species_count(int(c))
    for r in range(R):
        rm.add_reaction()
        rm.mutable_reaction(r).set_type(int(spec.types[r]))
        rm.mutable_reaction(r).add_rate_constant(float(spec.K[r]))
    for s in range(nS):
        for r in range(R):
            rm.add_stoichiometric_matrix(int(spec.S[s, r]))
    for s in range(nS):
        for r in range(R):
            rm.add_dependency_matrix(int(spec.D[s, r]))

    dm = lm.DiffusionModel()
    dm.set_number_reactions(R)
    dm.set_number_species(nS)
    dm.set_number_site_types(nST)
    dm.set_lattice_x_size(X)
    dm.set_lattice_y_size(Y)
    dm.set_lattice_z_size(Z)
    dm.set_lattice_spacing(spec.spacing)
    dm.set_particles_per_site(spec.P)
    dm.set_bytes_per_particle(1)
    for a in range(nST):
        for b in range(nST):
            for p in range(nS):
                dm.add_diffusion_matrix(float(spec.DF[a, b, p]))
    for r in range(R):
        for s in range(nST):
            dm.add_reaction_location_matrix(int(spec.RL[r, s]))

    lattice = lm.ByteLattice(X, Y, Z, spec.spacing, spec.P)
    # particles through the view (jLM fills getParticleLatticeView) for half
    # of the sites and through addParticle for the rest
    view = lattice.getParticleLatticeView()
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                lattice.setSiteType(x, y, z, int(spec.sites[x, y, z]))
                parts = [int(p) for p in spec.lattice[x, y, z] if p]
                if (x + y + z) % 2:
                    for k, p in enumerate(parts):
                        view[k // 4, z, y, x, k % 4] = p
                else:
                    for p in parts:
                        lattice.addParticle(x, y, z, p)

    lm.SimulationFile.create(path)
    f = lm.SimulationFile(path)
    f.setReactionModel(rm)
    f.setDiffusionModel(dm)
    f.setSpatialModel(lm.SpatialModel())
    f.setDiffusionModelLattice(dm, lattice)
    f.setParameter("writeInterval", "5")
    f.setParameter("latticeWriteInterval", "10")
    f.setParameter("hookInterval", "")
    f.setParameter("maxTime", "{:.10e}".format(20 * spec.tau))
    f.setParameter("timestep", "{:.10e}".format(spec.tau))
    f.close()
    return rm, dm, lattice


def test_lm_built_file_matches_model_spec_file(tmp_path):
    spec = small_spec()
    a, b = str(tmp_path / "lm.h5"), str(tmp_path / "spec.h5")
    build_with_lm(spec, a)
    spec.to_file(b)
    with RawFile(a, writable=False) as fa, RawFile(b, writable=False) as fb:
        ra, rb = fa.getReactionModel(), fb.getReactionModel()
        for k in ("number_species", "number_reactions"):
            assert getattr(ra, k) == getattr(rb, k)
        for k in ("initial_species_count", "types", "S", "D"):
            np.testing.assert_array_equal(getattr(ra, k), getattr(rb, k))
        assert ra.rate_constants == rb.rate_constants
        da, db = fa.getDiffusionModel(), fb.getDiffusionModel()
        assert da.shape == db.shape and da.particles_per_site == db.particles_per_site
        np.testing.assert_array_equal(da.diffusion_matrix, db.diffusion_matrix)
        np.testing.assert_array_equal(da.reaction_location_matrix, db.reaction_location_matrix)
        la, sa = fa.getDiffusionModelLattice()
        lb, sb = fb.getDiffusionModelLattice()
        np.testing.assert_array_equal(la, lb)
        np.testing.assert_array_equal(sa, sb)
        assert fa.getParameter("timestep") == "{:.10e}".format(spec.tau)


def test_messages_read_back(tmp_path):
    spec = small_spec()
    path = str(tmp_path / "lm.h5")
    rm0, dm0, lat0 = build_with_lm(spec, path)
    f = lm.SimulationFile(path)
    rm = lm.ReactionModel()
    f.getReactionModel(rm)
    assert rm.number_species() == spec.nS and rm.number_reactions() == spec.R
    assert rm.reaction_size() == spec.R and rm.reaction(2).type() == int(spec.types[2])
    assert rm.reaction(2).rate_constant(0) == float(spec.K[2]) and rm.reaction(2).rate_constant_size() == 1
    assert rm.stoichiometric_matrix() == rm0.stoichiometric_matrix()
    assert rm.initial_species_count_size() == spec.nS
    dm = lm.DiffusionModel()
    f.getDiffusionModel(dm)
    assert dm.lattice_x_size() == spec.shape[0] and dm.particles_per_site() == spec.P
    assert dm.diffusion_matrix() == dm0.diffusion_matrix()
    lat = lm.ByteLattice(*spec.shape, spec.spacing, spec.P)
    f.getDiffusionModelLattice(dm, lat)
    np.testing.assert_array_equal(lat.getParticleLatticeView(), lat0.getParticleLatticeView())
    np.testing.assert_array_equal(lat.getSiteLatticeView(), lat0.getSiteLatticeView())
    assert lat.getParticleCounts() == lat0.getParticleCounts()
    f.close()


def test_byte_lattice_accessors():
    lat = lm.ByteLattice(4, 3, 2, 1e-8, 8)
    lat.addParticle(1, 2, 1, 5)
    lat.addParticle(1, 2, 1, 7)
    assert lat.getOccupancy(1, 2, 1) == 2 and lat.getParticle(1, 2, 1, 1) == 7
    assert lat.getParticleCounts() == {5: 1, 7: 1}
    assert lat.findParticles(6, 9) == [(1, 2, 1, 1, 7)]
    assert lat.row_major()[1, 2, 1, :2].tolist() == [5, 7]
    for _ in range(6):
        lat.addParticle(0, 0, 0, 1)
    lat.addParticle(0, 0, 0, 2)
    lat.addParticle(0, 0, 0, 3)
    assert lat.getParticleLatticeView()[1, 0, 0, 0].tolist() == [1, 1, 2, 3]
    with pytest.raises(InvalidArgException, match="full"):
        lat.addParticle(0, 0, 0, 4)
    with pytest.raises(InvalidArgException, match="InvalidSite"):
        lat.getSiteType(4, 0, 0)
    lat.removeParticles(1, 2, 1)
    assert lat.getOccupancy(1, 2, 1) == 0
    assert lm.getCompiledLatticeMaxOccupancy() == 8


def test_solver_composition_and_spatial_limits(tmp_path):
    class Custom:
        def __init__(self, device=0):
            super().__init__()
            self.seen = []

        def hookSimulation(self, t, lattice):
            self.seen.append(t)
            return 0

    composed = type("Custom_MpdRdmeSolver", (Custom, lm.MpdRdmeSolver), {})
    assert composed.hookSimulation is Custom.hookSimulation
    assert issubclass(composed, lm.RDMESolver)
    sm = lm.SpatialModel()
    sm.add_region().set_site_type(1)
    path = str(tmp_path / "s.h5")
    lm.SimulationFile.create(path)
    with lm.SimulationFile(path) as f:
        with pytest.raises(NotImplementedError):
            f.setSpatialModel(sm)


def test_int_lattice_accessors():
    lat = lm.IntLattice(4, 3, 2, 1e-8, 8)
    assert lat.getParticleLatticeView().shape == (8, 2, 3, 4, 1)
    assert lat.getParticleLatticeView().dtype == np.uint32
    lat.addParticle(1, 2, 1, 70000)
    lat.addParticle(1, 2, 1, 300)
    assert lat.getOccupancy(1, 2, 1) == 2 and lat.getParticle(1, 2, 1, 1) == 300
    assert lat.getParticleCounts() == {300: 1, 70000: 1}
    assert lat.findParticles(256, 1000) == [(1, 2, 1, 1, 300)]
    rm = lat.row_major()
    assert rm.shape == (4, 3, 2, 8) and rm.dtype == np.uint32 and rm[1, 2, 1, :2].tolist() == [70000, 300]
    other = lm.IntLattice(4, 3, 2, 1e-8, 8)
    scattered = np.zeros_like(rm)
    scattered[1, 2, 1, 3], scattered[1, 2, 1, 6] = 70000, 300   # compacted to the first slots in order
    other.setFromRowMajorData(scattered)
    np.testing.assert_array_equal(other.getParticleLatticeView(), lat.getParticleLatticeView())
    for _ in range(6):
        lat.addParticle(1, 2, 1, 5)
    with pytest.raises(InvalidArgException, match="full"):
        lat.addParticle(1, 2, 1, 5)
    with pytest.raises(InvalidArgException, match="out of range"):
        lat.addParticle(0, 0, 0, 2 ** 32)


def test_int_lattice_file_round_trip(tmp_path):
    """A uint32 lattice written through lm.SimulationFile.setDiffusionModelLattice
    with bytes_per_particle 4 (SimulationFile.cpp:860-900) reads back as the
    same IntLattice."""
    spec = small_spec()
    nS = 400
    X, Y, Z = spec.shape
    dm = lm.DiffusionModel()
    dm.set_number_species(nS)
    dm.set_number_reactions(0)
    dm.set_number_site_types(spec.nST)
    for f, v in (("lattice_x_size", X), ("lattice_y_size", Y), ("lattice_z_size", Z),
                 ("lattice_spacing", spec.spacing), ("particles_per_site", spec.P), ("bytes_per_particle", 4)):
        getattr(dm, "set_" + f)(v)
    for _ in range(spec.nST * spec.nST * nS):
        dm.add_diffusion_matrix(0.0)
    lat = lm.IntLattice(X, Y, Z, spec.spacing, spec.P)
    rng = np.random.default_rng(0)
    for _ in range(50):
        lat.addParticle(int(rng.integers(X)), int(rng.integers(Y)), int(rng.integers(Z)), int(rng.integers(1, nS + 1)))
    path = str(tmp_path / "int.h5")
    lm.SimulationFile.create(path)
    with lm.SimulationFile(path) as f:
        f.setDiffusionModel(dm)
        f.setDiffusionModelLattice(dm, lat)
    with lm.SimulationFile(path) as f:
        back = lm.IntLattice(X, Y, Z, spec.spacing, spec.P)
        f.getDiffusionModelLattice(dm, back)
        d = f.getDiffusionModel()
    assert d.bytes_per_particle == 4
    np.testing.assert_array_equal(back.getParticleLatticeView(), lat.getParticleLatticeView())
    assert back.getParticleCounts() == lat.getParticleCounts()
