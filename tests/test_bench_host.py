"""Host-side pieces of bench.py, on the CPU: the HBM-traffic file is built
from rocprofv3 counter CSVs scaled by a plain copy's counters and stamped
with the kernel sources, and counters stamped with other sources are not
used; the CPU description reads the host."""
import csv
import json
import os

import pytest

import bench


def write_csv(path, rows):
    keys = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=keys)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(keys, r)))


@pytest.fixture
def profiles_dir(tmp_path, monkeypatch):
    root = tmp_path / "repo"
    (root / "profiles").mkdir(parents=True)
    csrc = root / "lattice_microbes_amd" / "csrc"
    csrc.mkdir(parents=True)
    (csrc / "k.hip").write_text("kernel source one\n")
    (csrc / "lmx_a.h").write_text("header\n")
    monkeypatch.setattr(bench, "ROOT", str(root))
    return root


def test_traffic_file_scaled_and_stamped(profiles_dir, tmp_path):
    x = bench.TRAFFIC_KERNELS["sweep_x"] + "(lmx::Params, unsigned int const*, unsigned int*)"
    z = bench.TRAFFIC_KERNELS["sweep_z"] + "(lmx::Params, ...)"
    # two dispatches each: means 100 and 300 KiB read, 50 KiB written
    write_csv(tmp_path / "f.csv", [(1, x, "FETCH_SIZE", 90), (2, x, "FETCH_SIZE", 110), (3, z, "FETCH_SIZE", 300),
                                   (4, z, "FETCH_SIZE", 300)])
    write_csv(tmp_path / "w.csv", [(1, x, "WRITE_SIZE", 50), (2, x, "WRITE_SIZE", 50), (3, z, "WRITE_SIZE", 40),
                                   (4, z, "WRITE_SIZE", 60)])
    # a 1 GiB copy reads as 512 MiB of FETCH_SIZE and 1 GiB of WRITE_SIZE
    write_csv(tmp_path / "cf.csv", [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 524288)])
    write_csv(tmp_path / "cw.csv", [(1, "__amd_rocclr_copyBuffer", "WRITE_SIZE", 1048576)])
    data = bench.write_traffic(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), str(tmp_path / "cf.csv"),
                               str(tmp_path / "cw.csv"), "wc")
    assert data["read_scale"] == pytest.approx(2.0) and data["write_scale"] == pytest.approx(1.0)
    assert data["bytes_per_launch"]["wc:sweep_x"] == (100 * 2 + 50) * 1024
    assert data["bytes_per_launch"]["wc:sweep_z"] == (300 * 2 + 50) * 1024
    saved = json.load(open(profiles_dir / "profiles" / "traffic.json"))
    assert saved["kernel_sources_sha256"] == bench.kernel_source_digest()
    value, note = bench.measured_traffic("wc:sweep_x")
    assert value == (100 * 2 + 50) * 1024 and "FETCH_SIZE" in note
    # a kernel source changes: the counters no longer describe the kernels
    (profiles_dir / "lattice_microbes_amd" / "csrc" / "k.hip").write_text("kernel source two\n")
    value, note = bench.measured_traffic("wc:sweep_x")
    assert value is None and "other kernel sources" in note


def test_missing_traffic_file(profiles_dir):
    value, note = bench.measured_traffic("wc:sweep_x")
    assert value is None and "no counters" in note


def test_host_cpu():
    info = bench.host_cpu()
    assert info["nproc"] == os.cpu_count() and 1 <= info["affinity"] <= os.cpu_count()
