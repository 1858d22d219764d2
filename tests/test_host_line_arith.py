"""Host checks of the arithmetic shortcuts of the per-wave line kernel
(partia.jl_amd/csrc): the float32 table argument q / dq + 1 formed with an
fma-corrected product instead of the division (line_W_rdq), and the node test
decided from precomputed thresholds (line_node_prune).  Both must give the
per-ray kernel's results bit for bit."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_table_step_division_is_exact(tmp_path):
    """Exhaustively over every float32 q with 2^-40 <= |q| <= 16 and the three
    float32 table steps of line_integrated_kernel.jl:14: the division's
    result, bit for bit (accepted neighbours have q < 6; below 2^-40 both are
    exactly 1)."""
    exe = str(tmp_path / "table_step_division")
    src = os.path.join(HERE, "native", "table_step_division.c")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm", "-lpthread"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert ", 0 mismatches" in r.stdout


def test_line_node_thresholds_match_the_per_ray_test(tmp_path):
    """The line walk's node test decides most lanes from two precomputed
    thresholds on the squared centre distance (line_node / line_node_prune in
    partia.jl_amd/csrc/ph_interp_dev.h); randomised cases placed within
    2^-30 .. 2^-2 of the pruning threshold, and degenerate nodes, give the
    per-ray test's decision in float and double."""
    exe = str(tmp_path / "line_prune_thresholds")
    src = os.path.join(HERE, "native", "line_prune_thresholds.c")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
    r = subprocess.run([exe, "3000000"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count(", 0 mismatches") == 2
