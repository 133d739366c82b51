"""The C++ facade (include/goetia/*.hh): compiles with the host compiler
against the C ABI header only (no HIP headers), links against
libgoetia_amd.so, and on a GPU runs goetia's dBG assertions."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, 'tests', 'cpp', 'test_facade.cc')
LIBDIR = os.path.join(ROOT, 'goetia_amd')


def _compile(out):
    cxx = shutil.which('g++') or shutil.which('c++')
    cmd = [cxx, '-std=c++17', '-O1', '-Wall', '-Werror', f'-I{os.path.join(ROOT, "include")}', SRC,
           f'-L{LIBDIR}', '-lgoetia_amd', f'-Wl,-rpath,{LIBDIR}', '-o', out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def test_facade_compiles_without_hip_headers(tmp_path):
    _compile(str(tmp_path / 'facade'))
    # the public headers never include HIP
    for dirpath, _, files in os.walk(os.path.join(ROOT, 'include')):
        for f in files:
            text = open(os.path.join(dirpath, f)).read()
            assert 'hip_runtime' not in text and '<hip/' not in text, f


@pytest.mark.gpu
def test_facade_runs_on_gpu(tmp_path):
    exe = str(tmp_path / 'facade')
    _compile(exe)
    fq = os.path.join(ROOT, 'tests', 'golden', 'data', 'test-fastq-reads.fq')
    res = subprocess.run([exe, fq, str(tmp_path / 'filtered.fq')], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert 'FACADE OK' in res.stdout
