"""The float32 per-wave line kernel forms the table argument q / dq + 1 with
an fma-corrected product instead of the division (line_W_rdq in
partia.jl_amd/csrc/ph_sph.h).  This checks, exhaustively over every float32 q
with 2^-40 <= |q| <= 16 and the three float32 table steps of
line_integrated_kernel.jl:14, that the result is the division's, bit for bit
(accepted neighbours have q < 6; below 2^-40 both are exactly 1)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_table_step_division_is_exact(tmp_path):
    exe = str(tmp_path / "table_step_division")
    src = os.path.join(HERE, "native", "table_step_division.c")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm", "-lpthread"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert ", 0 mismatches" in r.stdout
