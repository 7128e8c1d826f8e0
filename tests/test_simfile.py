"""SimulationFile (HDF5) reader and writer, on the CPU.

Mirrors the reference's own file tests (testing/c/lm/io/hdf5/simulation_file.cpp)
on the fixture files it ships (copied as data into tests/golden/hdf5), with the
expected values that test asserts, and checks the layout of written files
(group and dataset names, types, chunking, gzip level) with h5dump, a reader
independent of this library.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from lattice_microbes_amd import simfile as sf
from lattice_microbes_amd.simfile import DiffusionModel, IOException, ReactionModel, SimulationFile

GOLD = os.path.join(os.path.dirname(__file__), "golden", "hdf5")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def gold(name):
    return os.path.join(GOLD, name)


def h5dump(path, *args):
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    return subprocess.run([H5DUMP, *args, path], capture_output=True, text=True, check=True).stdout


# ---------------------------------------------------------------- open / create
def test_open_rejects_non_hdf5_and_directories():
    # OpenNonHDF5File (simulation_file.cpp:82-89)
    with pytest.raises(IOException):
        SimulationFile(gold("SimulationFile_1.h5"))
    with pytest.raises(IOException):
        SimulationFile(GOLD + "/")
    with pytest.raises(IOException):
        SimulationFile(gold("does_not_exist.h5"))


def test_open_good_file_version_2():
    # OpenGoodFile (:91-100)
    with SimulationFile(gold("SimulationFile_2.h5"), writable=False) as f:
        assert f.version == 2


def test_create_file(tmp_path):
    # CreateFile (:102-117) plus the magic and the existing-file check of
    # SimulationFile_create.cpp:79-80,114-121
    p = str(tmp_path / "create.h5")
    SimulationFile.create(p)
    with open(p, "rb") as fh:
        assert fh.read(4) == b"LMH5"
    with SimulationFile(p) as f:
        assert f.version == 4
        assert f.getParameters() == {}
        assert f.getReactionModel().number_species == 0
    with pytest.raises(IOException):
        SimulationFile.create(p)
    dump = h5dump(p, "-H")
    for g in ("Parameters", "Model", "Simulations"):
        assert f'GROUP "{g}"' in dump
    assert 'ATTRIBUTE "formatVersion"' in dump and "H5T_STD_U32LE" in dump
    assert "USER_BLOCK" in h5dump(p, "-B", "-H") and "USERBLOCK_SIZE 512" in h5dump(p, "-B", "-H")


def test_files_without_magic_are_rejected(tmp_path):
    # isValidFile (SimulationFile.cpp:225-252): an HDF5 file without the magic
    p = str(tmp_path / "nomagic.h5")
    SimulationFile.create(p)
    with open(p, "r+b") as fh:
        fh.write(b"XXXX")
    with pytest.raises(IOException):
        SimulationFile(p)


# ---------------------------------------------------------------- parameters
def test_get_parameter_v2():
    # GetParameterV2 (:119-128): version-2 doubles formatted with %g
    with SimulationFile(gold("SimulationFile_2.h5"), writable=False) as f:
        assert f.version == 2
        assert float(f.getParameter("unused1", "11.0")) == 11.0
        assert float(f.getParameter("test1")) == pytest.approx(2.3, rel=1e-4)
        assert float(f.getParameter("test2")) == pytest.approx(5.7e-16, rel=1e-4)
        assert f.getParameter("test1") == "2.3" and f.getParameter("test2") == "5.7e-16"


def test_get_parameter_v3():
    # GetParameterV3 (:130-141): strings; the integer attribute test4 is skipped
    with SimulationFile(gold("SimulationFile_3.h5"), writable=False) as f:
        assert f.version == 3
        assert float(f.getParameter("unused1", "11.0")) == 11.0
        assert float(f.getParameter("test1")) == pytest.approx(2.3, rel=1e-4)
        assert float(f.getParameter("test2")) == pytest.approx(5.7e-16, rel=1e-4)
        assert f.getParameter("test3") == "testing"
        assert f.getParameter("test4") == ""


def test_set_parameter(tmp_path):
    # SetParameter (:143-181)
    p = str(tmp_path / "params.h5")
    SimulationFile.create(p)
    with SimulationFile(p) as f:
        f.setParameter("test1", "23.0")
        f.setParameter("test2", "24.0")
        f.setParameter("test3", "1.7e7")
        f.setParameter("test4", "newtest")
    with SimulationFile(p) as f:
        assert float(f.getParameter("test1")) == 23.0
        assert float(f.getParameter("test2")) == 24.0
        assert float(f.getParameter("test3")) == 1.7e7
        assert f.getParameter("test4") == "newtest"
        f.setParameter("test2", "-25.0")
        f.setParameter("test3", "1.8e-7")
        f.setParameter("test4", "newtest2")
    with SimulationFile(p) as f:
        assert f.getParameters() == {"test1": "23.0", "test2": "-25.0", "test3": "1.8e-7", "test4": "newtest2"}
    with SimulationFile(p) as f:
        with pytest.raises(sf.InvalidArgException):
            f.setParameter("bad=key", "1")
    with SimulationFile(p, writable=False) as f:
        with pytest.raises(sf.InvalidArgException):
            f.setParameter("test1", "1")


def test_set_parameter_version_2_writes_doubles(tmp_path):
    # setParameter on a version-2 file stores atof(value) (SimulationFile.cpp:327-331)
    p = str(tmp_path / "v2.h5")
    shutil.copyfile(gold("SimulationFile_2.h5"), p)
    os.chmod(p, 0o644)
    with SimulationFile(p) as f:
        f.setParameter("test5", "4.25")
    with SimulationFile(p) as f:
        assert f.getParameter("test5") == "4.25"
    assert "H5T_IEEE_F64LE" in h5dump(p, "-a", "/Parameters/test5")


def test_checkpoint(tmp_path):
    # Checkpoint (:183-222)
    p = str(tmp_path / "chk.h5")
    SimulationFile.create(p)
    f = SimulationFile(p)
    f.setParameter("test1", "23.0")
    f.setParameter("test2", "24.0")
    f.setParameter("test3", "1.7e7")
    chk = f.checkpoint()
    assert chk == p + ".chk"
    f.setParameter("test1", "23.0")
    f.setParameter("test2", "-25.0")
    f.setParameter("test3", "1.8e-7")
    f.close()
    with SimulationFile(p) as g:
        assert [float(g.getParameter(k)) for k in ("test1", "test2", "test3")] == [23.0, -25.0, 1.8e-7]
    with SimulationFile(chk) as g:
        assert [float(g.getParameter(k)) for k in ("test1", "test2", "test3")] == [23.0, 24.0, 1.7e7]


# ---------------------------------------------------------------- reaction model
S7 = [[1, -1, 0, -1, 0, 1, -1], [-1, 1, 0, 0, 1, 1, -1], [0, 0, 1, 1, 1, 1, -1], [0, 0, -1, 0, 0, 1, -1],
      [0, 0, 0, -1, 0, 1, -1]]
D7 = [[1, 0, 0, 1, 0, 0, 1], [0, 1, 0, 1, 0, 0, 1], [0, 0, 1, 0, 1, 0, 1], [0, 1, 0, 0, 1, 1, 1],
      [1, 0, 0, 0, 0, 1, 0]]
K7 = [[1.5], [7.0], [16.7, 17.3], [1678.0, 1.0e12], [2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0, 11.0], [5.6], [6.7]]


def check_model7(m):
    assert m.number_species == 5 and m.number_reactions == 7
    assert m.initial_species_count.tolist() == [11, 13, 17, 19, 23]
    assert m.types.tolist() == [1, 1, 2, 2, 91, 98, 99]
    assert [len(k) for k in m.rate_constants] == [1, 1, 2, 2, 10, 1, 1]
    for got, want in zip(m.rate_constants, K7):
        np.testing.assert_allclose(got, want, rtol=1e-11)
    assert m.S.tolist() == S7
    assert m.D.tolist() == D7


def test_get_reaction_model_fixtures():
    # GetReactionModel (:501-668)
    with SimulationFile(gold("SimulationFile_GetReactionModel_NoReactionCount.h5"), writable=False) as f:
        m = f.getReactionModel()
        assert m.number_species == 5 and m.number_reactions == 0 and m.initial_species_count.size == 0
    with SimulationFile(gold("SimulationFile_GetReactionModel.h5"), writable=False) as f:
        m = f.getReactionModel()
        check_model7(m)
        assert not any(m.rate_has_noise(r) for r in range(7))
    with SimulationFile(gold("SimulationFile_GetReactionModel_Noise.h5"), writable=False) as f:
        m = f.getReactionModel()
        check_model7(m)
        assert [m.rate_has_noise(r) for r in range(7)] == [False, True, False, False, True, True, False]
        np.testing.assert_allclose(m.rate_noise[[1, 4, 5], 0], [11.11, 4.4, 5.555], rtol=1e-11)
        np.testing.assert_allclose(m.rate_noise[[1, 4, 5], 1], [1e5, 1e14, 1e15], rtol=1e-11)


S11 = [[0, 1, 1], [1, 0, 1], [0, 1, 0], [0, 1, 0], [-1, 0, 1], [0, 1, 1], [0, 1, 0], [0, 1, 0], [1, 0, 1],
       [1, 0, 1], [1, 0, -1]]
D11 = [[0, 1, 1], [1, 0, 1]] * 5 + [[0, 1, 1]]


@pytest.mark.parametrize("noise", [False, True])
def test_set_reaction_model_round_trip(tmp_path, noise):
    # SetReactionModel (:670-1056)
    p = str(tmp_path / "rx.h5")
    SimulationFile.create(p)
    m = ReactionModel(number_species=11, number_reactions=3, initial_species_count=np.arange(1, 12, dtype=np.uint32),
                      types=np.array([17, 37, 47], np.uint32),
                      rate_constants=[[1.7], [11.7, 111.7], [21.7, 211.7, 2111.7]],
                      S=np.array(S11, np.int32), D=np.array(D11, np.uint32))
    if noise:
        m.rate_noise = np.array([[17.178, 1e5], [123.456789, 1e9], [0.0, 0.0]])
    with SimulationFile(p) as f:
        f.setReactionModel(m)
    with SimulationFile(p) as f:
        g = f.getReactionModel()
    assert g.number_species == 11 and g.number_reactions == 3
    assert g.initial_species_count.tolist() == list(range(1, 12))
    assert g.types.tolist() == [17, 37, 47]
    assert g.rate_constants == [[1.7], [11.7, 111.7], [21.7, 211.7, 2111.7]]
    assert g.S.tolist() == S11 and g.D.tolist() == D11
    assert [g.rate_has_noise(r) for r in range(3)] == ([True, True, False] if noise else [False] * 3)
    if noise:
        assert g.rate_noise[:2].tolist() == [[17.178, 1e5], [123.456789, 1e9]]
    dump = h5dump(p, "-H", "-g", "/Model/Reaction")
    assert ("ReactionRateNoise" in dump) == noise
    assert "( 3, 10 )" in dump                       # rate constants [R][10]
    assert "( 11, 3 )" in dump                       # S and D species-major
    # replacing the model replaces every dataset
    m2 = ReactionModel(number_species=2, number_reactions=0, initial_species_count=np.array([4, 5], np.uint32))
    with SimulationFile(p) as f:
        f.setReactionModel(m2)
        g = f.getReactionModel()
    assert g.number_species == 2 and g.number_reactions == 0 and g.initial_species_count.tolist() == [4, 5]


def test_set_reaction_model_rejects_inconsistent_sizes(tmp_path):
    p = str(tmp_path / "bad.h5")
    SimulationFile.create(p)
    with SimulationFile(p) as f:
        with pytest.raises(sf.InvalidArgException):
            f.setReactionModel(ReactionModel(number_species=0))
        with pytest.raises(sf.InvalidArgException):
            f.setReactionModel(ReactionModel(number_species=2, number_reactions=1,
                                             initial_species_count=np.array([1, 2], np.uint32),
                                             types=np.array([1], np.uint32), rate_constants=[[1.0]],
                                             S=np.zeros((2, 2), np.int32), D=np.zeros((2, 1), np.uint32)))


# ---------------------------------------------------------------- diffusion model
def test_get_diffusion_model_fixture():
    # GetDiffusionModel (:1058-1073)
    with SimulationFile(gold("SimulationFile_GetDiffusionModel.h5"), writable=False) as f:
        d = f.getDiffusionModel()
        assert d.lattice_spacing == 16.0
        assert (d.lattice_x_size, d.lattice_y_size, d.lattice_z_size, d.particles_per_site) == (3, 4, 5, 2)
        assert d.bytes_per_particle == 1
        # GetDiffusionModelLattice (:1099-1148) expects the column-major
        # (Matlab h5write) layout of an older reader: the stored dataset is
        # (P, Z, Y, X) = (2, 5, 4, 3).  The current reader requires (X, Y, Z, P)
        # (SimulationFile.cpp:627-629) and rejects it; so does this one.
        with pytest.raises(IOException, match="invalid dataset dimensions"):
            f.getDiffusionModelLattice()
    dump = h5dump(gold("SimulationFile_GetDiffusionModel.h5"), "-H", "-d", "/Model/Diffusion/Lattice")
    assert "( 2, 5, 4, 3 )" in dump


def diffusion_case(X=20, Y=18, Z=17, P=8, nS=3, nST=2, R=2, bpp=1, seed=0):
    rng = np.random.default_rng(seed)
    d = DiffusionModel(number_species=nS, number_reactions=R, number_site_types=nST, lattice_spacing=3.2e-8,
                       lattice_x_size=X, lattice_y_size=Y, lattice_z_size=Z, particles_per_site=P,
                       bytes_per_particle=bpp, diffusion_matrix=rng.random((nST, nST, nS)),
                       reaction_location_matrix=rng.integers(0, 2, (R, nST)).astype(np.uint32))
    hi = 256 if bpp == 1 else 70000
    lat = rng.integers(0, hi, (X, Y, Z, P)).astype(np.uint8 if bpp == 1 else np.uint32)
    sites = rng.integers(0, nST, (X, Y, Z)).astype(np.uint8)
    return d, lat, sites


@pytest.mark.parametrize("bpp", [1, 4])
def test_set_diffusion_model_round_trip(tmp_path, bpp):
    p = str(tmp_path / "diff.h5")
    SimulationFile.create(p)
    d, lat, sites = diffusion_case(bpp=bpp)
    with SimulationFile(p) as f:
        f.setDiffusionModel(d, lat, sites)
    with SimulationFile(p) as f:
        g = f.getDiffusionModel()
        lat2, sites2 = f.getDiffusionModelLattice(g)
    assert (g.number_species, g.number_reactions, g.number_site_types) == (3, 2, 2)
    assert g.lattice_spacing == 3.2e-8 and g.shape == (20, 18, 17) and g.particles_per_site == 8
    assert g.bytes_per_particle == bpp
    np.testing.assert_array_equal(g.diffusion_matrix, d.diffusion_matrix)
    np.testing.assert_array_equal(g.reaction_location_matrix, d.reaction_location_matrix)
    assert lat2.dtype == lat.dtype
    np.testing.assert_array_equal(lat2, lat)
    np.testing.assert_array_equal(sites2, sites)
    dump = h5dump(p, "-p", "-H", "-g", "/Model/Diffusion")
    assert "( 20, 18, 17, 8 )" in dump and "( 20, 18, 17 )" in dump
    assert "CHUNKED ( 16, 16, 16, 8 )" in dump and "CHUNKED ( 16, 16, 16 )" in dump
    assert "DEFLATE { LEVEL 1 }" in dump
    assert ("H5T_STD_U32LE" if bpp == 4 else "H5T_STD_U8LE") in dump


def test_set_diffusion_model_rejects_bad_shapes(tmp_path):
    p = str(tmp_path / "bad.h5")
    SimulationFile.create(p)
    d, lat, sites = diffusion_case()
    with SimulationFile(p) as f:
        with pytest.raises(sf.InvalidArgException):
            f.setDiffusionModel(d, lat[:-1], sites)
        d.bytes_per_particle = 2
        with pytest.raises(sf.InvalidArgException):
            f.setDiffusionModel(d, None, None)


# ---------------------------------------------------------------- trajectories
def test_replicates_and_species_counts(tmp_path):
    # CreateReplicateHandles / OpenReplicateHandles (:224-325): a replicate is
    # created with empty SpeciesCounts [0][numberSpecies] and SpeciesCountTimes
    p = str(tmp_path / "traj.h5")
    SimulationFile.create(p, 10)
    with SimulationFile(p) as f:
        assert not f.replicateExists(0)
        rng = np.random.default_rng(1)
        c1 = rng.integers(0, 1000, (150, 10)).astype(np.uint32)
        t1 = np.arange(150) * 0.5
        f.appendSpeciesCounts(0, t1, c1)
        c2 = rng.integers(0, 1000, (3, 10)).astype(np.uint32)
        f.appendSpeciesCounts(0, [75.0, 75.5, 76.0], c2)
        f.appendSpeciesCounts(10, [0.0], c2[:1])
        assert f.replicateExists(0) and f.replicateExists(10) and not f.replicateExists(1)
        with pytest.raises(IOException):
            f.appendSpeciesCounts(0, [1.0], np.zeros((1, 9), np.uint32))   # column count is fixed
    with SimulationFile(p, writable=False) as f:
        np.testing.assert_array_equal(f.getSpeciesCounts(0), np.vstack([c1, c2]))
        np.testing.assert_array_equal(f.getSpeciesCountTimes(0), np.r_[t1, 75.0, 75.5, 76.0])
        np.testing.assert_array_equal(f.getSpeciesCounts(10), c2[:1])
        with pytest.raises(IOException):
            f.getSpeciesCounts(1)
    dump = h5dump(p, "-p", "-H", "-g", "/Simulations/0000000")
    assert 'DATASET "SpeciesCounts"' in dump and "H5T_STD_I32LE" in dump
    assert "( 153, 10 ) / ( H5S_UNLIMITED, 10 )" in dump and "CHUNKED ( 100, 10 )" in dump
    assert 'DATASET "SpeciesCountTimes"' in dump and "CHUNKED ( 100 )" in dump
    assert 'GROUP "0000010"' in h5dump(p, "-H", "-g", "/Simulations")


def test_replicate_needs_a_model(tmp_path):
    # createReplicateHandles -> loadModel: "No model has been defined."
    p = str(tmp_path / "nomodel.h5")
    SimulationFile.create(p)
    with SimulationFile(p) as f:
        with pytest.raises(IOException, match="no model"):
            f.appendSpeciesCounts(0, [0.0], np.zeros((1, 3), np.uint32))


def test_lattice_and_site_frames(tmp_path):
    p = str(tmp_path / "frames.h5")
    SimulationFile.create(p, 3)
    d, lat0, sites = diffusion_case(X=9, Y=33, Z=5, P=16)
    rng = np.random.default_rng(5)
    frames = [lat0] + [rng.integers(0, 4, lat0.shape).astype(np.uint8) for _ in range(2)]
    with SimulationFile(p) as f:
        f.setDiffusionModel(d, lat0, sites)
        for k, fr in enumerate(frames):
            f.appendLattice(2, 0.25 * k, fr)
        f.appendSites(2, 0.5, sites)
        f.appendSites(2, 0.75, sites[::-1].copy())
        f.writeReplicateDataset(2, "MaxParticleCounts", [0, 5, 6, 7])
        with pytest.raises(sf.InvalidArgException):
            f.writeReplicateDataset(2, "MaxParticleCounts", [1])
    with SimulationFile(p, writable=False) as f:
        np.testing.assert_array_equal(f.getLatticeTimes(2), [0.0, 0.25, 0.5])
        for k, fr in enumerate(frames):
            np.testing.assert_array_equal(f.getLattice(2, k), fr)
        with pytest.raises(sf.InvalidArgException):
            f.getLattice(2, 3)
        with pytest.raises(sf.InvalidArgException):
            f.getLattice(2, 0, shape=(9, 33, 5, 8))
        assert f.getReplicateDataset(2, "MaxParticleCounts").tolist() == [0, 5, 6, 7]
        assert f.getSpeciesCounts(2).shape == (0, 3)
    dump = h5dump(p, "-p", "-H", "-g", "/Simulations/0000002")
    for name in ("Lattice", "Sites"):
        assert f'GROUP "{name}"' in dump
    for name in ("0000000000", "0000000001", "0000000002"):
        assert f'DATASET "{name}"' in dump
    assert "CHUNKED ( 9, 16, 5, 16 )" in dump and "CHUNKED ( 9, 16, 5 )" in dump
    assert 'DATASET "LatticeTimes"' in dump and 'DATASET "SiteTimes"' in dump
    assert "DEFLATE { LEVEL 1 }" in dump
    assert 'DATASET "MaxParticleCounts"' in dump
