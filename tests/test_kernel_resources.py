"""The compiler's resource report of the built library (written by the build
next to libgevpm.so).  No kernel may use scratch memory: a register spill or
a stack object costs memory traffic in every launch and, for the particle
kernels, more than half their speed.  The hot kernels of the 512^3 cycle keep
the waves per SIMD they were tuned for.  CPU only."""
import csv
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TSV = os.path.join(ROOT, "gevolution-1.2_amd", "lib", "kernel_resources.tsv")


@pytest.fixture(scope="module")
def rows():
    if not os.path.exists(TSV):
        pytest.skip("library not built with its resource report")
    with open(TSV) as f:
        return list(csv.DictReader(f, delimiter="\t"))


def test_every_kernel_is_reported(rows):
    names = [r["kernel"] for r in rows]
    for k in ("k_deposit_tiled", "k_update_tiled", "k_zsolve", "k_r2c_x_prep", "k_c2c_cols", "k_c2r_x",
              "k_scan_final", "k_realize", "k_spectrum_rows"):
        assert any(k in n for n in names), k


def test_no_kernel_uses_scratch(rows):
    spilled = [(r["kernel"], r["scratch"]) for r in rows if r["scratch"] not in ("0", "")]
    assert not spilled


@pytest.mark.parametrize("mangled,waves", [
    ("k_update_tiledILb1ELi0ELb1ELi0ELb1E", 3),        # kick + drift, GR, power-of-two N (49 KB LDS)
    ("k_deposit_tiledILi3ELb1ELb0ELb1ELb1E", 3),       # T00 + Tij, full tiles
    ("k_r2c_x_prepILi512ELi6E", 3),                    # tensor prepare + x pass
    ("k_zsolveILi512ELi1E", 2),                        # tensor z solve
    ("k_c2c_colsIdLi512ELb0E", 2),                     # y pass
])
def test_hot_kernel_occupancy(rows, mangled, waves):
    occ = [int(r["occupancy"]) for r in rows if mangled in r["kernel"]]
    assert occ and min(occ) >= waves, (mangled, occ)
