"""bench.py on the CPU: the algorithmic byte counts it divides by kernel
times, the measured-traffic lookup, and the one-GPU contract."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

N, NP = 512, 512 ** 3


def test_dominant_kernels_algorithmic_bytes():
    # fused T00 + Tij deposit: pos and q read (48 B per particle); phi read,
    # T00 and the six Tij written (64 B per cell)
    assert bench.algorithmic_bytes("projection_T00_Tij_project", N, NP) == 48 * NP + 64 * N ** 3 == 15032385536
    # kick + drift: particles read and written (96 B), phi, chi, B read and
    # the cell index and count written (40 B per cell)
    assert bench.algorithmic_bytes("kick_drift", N, NP) == 96 * NP + 40 * N ** 3
    # the FP32 form halves the field bytes, not the particle bytes
    assert bench.algorithmic_bytes("projection_T00_Tij_project", N, NP, 4) == 48 * NP + 32 * N ** 3


def test_step_bytes_match_the_stage_table():
    """92.5 GB per GR-parabolic cycle at 512^3 (688.8 B per particle-step),
    the FP64 column of the stage table; the data-dependency minimum is
    smaller."""
    b = bench.stage_bytes(N, NP)
    assert b == 92496986112
    assert abs(b / NP - 688.8) < 0.5
    assert bench.cycle_bytes(N, NP) < b


def test_measured_traffic_reads_the_committed_profile():
    path = os.path.join(ROOT, "profiles", "hbm_traffic_gr_run_512_fp64.txt")
    t = bench.measured_traffic(path, "projection_T00_Tij_project")
    assert t is not None and t > bench.algorithmic_bytes("projection_T00_Tij_project", N, NP)
    assert bench.measured_traffic(path, "no_such_kernel") is None
    assert bench.measured_traffic(os.path.join(ROOT, "no_such_file.txt"), "kick_drift") is None


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_profile_commands_parse():
    """Every bench.py command quoted in a committed profile summary is one the
    bench still accepts (the summaries name the command that made them)."""
    import re
    pdir = os.path.join(ROOT, "profiles")
    seen = 0
    for name in sorted(os.listdir(pdir)):
        for line in open(os.path.join(pdir, name)):
            m = re.search(r"bench\.py((?:\s+--[\w-]+(?:\s+[^\s(-][^\s]*)?)*)", line)
            if not m or "--" not in m.group(1):
                continue
            argv = m.group(1).split()
            bench.arg_parser().parse_args(argv)  # exits (SystemExit) on an unknown flag
            seen += 1
    assert seen >= 4, seen
