"""Register budget of the default-path kernels, from the compiler's report of
the last build (hipcc -Rpass-analysis=kernel-resource-usage, written by
lattice_microbes_amd/build.py).  A kernel of the default step that spills
registers to scratch runs several times slower (the P = 4 tile sweeps once
took 207 us against 96 us when a select change pushed them over their
budget), so a spill here is a regression even when every result stays
bit-exact."""
import os

import pytest

from lattice_microbes_amd import build

pytestmark = pytest.mark.skipif(not os.path.exists(build.RESOURCES),
                                reason="no kernel resource report (build the library first)")

# Kernels of the default step and of batches, for 4 and 8 slots per site
# (W = 1, 2): (mangled-name fragment, least occupancy in waves per SIMD,
# most scratch bytes per lane).  The batched tile sweeps loop over several
# tiles per block within a budget of eight waves per SIMD; computing the
# step hash inside that loop once cost them an 8-byte spill (ensemble step
# 1.44 against 1.38 ms without it).
DEFAULT_KERNELS = [
    ("sweep_x4_kernelILi1E", 8, 0), ("sweep_x4_kernelILi2E", 8, 0),
    ("sweep_tile_kernelILi1ELb0E", 8, 0), ("sweep_tile_kernelILi1ELb1E", 8, 0),
    ("sweep_tile_kernelILi2ELb0E", 8, 0), ("sweep_tile_kernelILi2ELb1E", 8, 0),
    ("reaction_list_kernelILi1E", 4, 0), ("reaction_list_kernelILi2E", 4, 0),
    ("batch_x4_kernelILi1E", 8, 0), ("batch_x4_kernelILi2E", 8, 0),
    ("batch_tile_kernelILi1ELi256ELi1E", 8, 0), ("batch_tile_kernelILi1ELi256ELi2E", 8, 0),
    ("batch_tile_kernelILi2ELi256ELi1E", 8, 0), ("batch_tile_kernelILi2ELi256ELi2E", 8, 0),
    ("batch_reaction_kernelILi1E", 4, 0), ("batch_reaction_kernelILi2E", 4, 0),
    ("3lmx14resolve_kernel", 1, 0), ("batch_resolve_kernel", 1, 0),
]


def _find(res, frag):
    names = [n for n in res if frag in n]
    assert len(names) == 1, f"{frag}: {names}"
    return res[names[0]]


@pytest.mark.parametrize("frag,occupancy,scratch", DEFAULT_KERNELS, ids=[k[0] for k in DEFAULT_KERNELS])
def test_default_kernels_do_not_spill(frag, occupancy, scratch):
    r = _find(build.kernel_resources(), frag)
    assert r["ScratchSize"] <= scratch, f"{frag} spills {r['ScratchSize']} bytes per lane ({r['VGPRs']} VGPRs)"
    assert r["Occupancy"] >= occupancy, f"{frag}: occupancy {r['Occupancy']} < {occupancy}"


def test_report_parses():
    res = build.kernel_resources()
    assert len(res) > 20
    assert all("VGPRs" in v and "ScratchSize" in v for v in res.values())
