"""The kernels' modulus (gt_fastmod, and gt_fastmod30 for tables below 2^30
bins) equals the % operator: tests/cpp/test_fastmod.cc built for the host
with ROCm's clang (the header uses clang builtins) on the header the kernels
include."""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fastmod_equals_remainder(tmp_path):
    cxx = shutil.which('clang++') or '/opt/rocm/llvm/bin/clang++'
    exe = str(tmp_path / 'fastmod')
    subprocess.run([cxx, '-std=c++17', '-O2', f'-I{os.path.join(ROOT, "goetia_amd", "csrc")}',
                    os.path.join(ROOT, 'tests', 'cpp', 'test_fastmod.cc'), '-o', exe], check=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert 'mismatches 0' in res.stdout
