"""The C++ facade (include/goetia/*.hh): compiles with the host compiler
against the C ABI header only (no HIP headers), links against
libgoetia_amd.so, and on a GPU runs goetia's dBG assertions."""
import os
import shutil
import subprocess

import numpy as np
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


SEG_SRC = os.path.join(ROOT, 'tests', 'cpp', 'segments_facade.cc')


def _compile_segments(out):
    cxx = shutil.which('g++') or shutil.which('c++')
    cmd = [cxx, '-std=c++17', '-O1', '-Wall', '-Werror', f'-I{os.path.join(ROOT, "include")}', SEG_SRC,
           f'-L{LIBDIR}', '-lgoetia_amd', f'-Wl,-rpath,{LIBDIR}', '-o', out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def test_compactor_facade_compiles_and_refuses_canonical(tmp_path):
    """find_new_segments compiles for forward graphs; for a canonical one it
    does not, as in goetia (compactor.hh:405,428)."""
    _compile_segments(str(tmp_path / 'segments'))
    bad = tmp_path / 'canonical.cc'
    bad.write_text('#include "goetia/cdbg/compactor.hh"\n'
                   'using namespace goetia;\n'
                   'int main() {\n'
                   '    auto g = dBG<ByteStorage, CanLemireShifter>::build(ByteStorage::build(1000, 4), 21);\n'
                   '    StreamingCompactor<dBG<ByteStorage, CanLemireShifter>>::Compactor c(g);\n'
                   '    return (int)c.find_new_segments("ACGT").size();\n'
                   '}\n')
    cxx = shutil.which('g++') or shutil.which('c++')
    res = subprocess.run([cxx, '-std=c++17', '-fsyntax-only', f'-I{os.path.join(ROOT, "include")}', str(bad)],
                         capture_output=True, text=True)
    assert res.returncode != 0 and 'forward hashes only' in res.stderr


@pytest.mark.gpu
def test_compactor_facade_vs_reference(tmp_path):
    """The C++ StreamingCompactor<dBG<S, FwdLemireShifter>>::Compactor gives
    every golden case's records (one process for all cases)."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from segments_golden import CASES, case, parse
    exe = str(tmp_path / 'segments')
    _compile_segments(exe)
    cases = [case(n) for n in CASES]
    with open(tmp_path / 'cases.txt', 'w') as f:
        f.write(f'{len(cases)}\n')
        for c in cases:
            f.write(f"{c['kind']} {c['K']} {c['max_table']} {c['n_tables']} {int(c['persist'])} "
                    f"{int(c['insert_after'])}\n")
            for group in (c['bg'], c['reads']):
                f.write(f'{len(group)}\n')
                for r in group:
                    f.write(r + '\n')
    res = subprocess.run([exe, str(tmp_path / 'cases.txt'), str(tmp_path / 'words.txt'), str(tmp_path / 'kmers.txt')],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    assert f'SEGMENTS OK {len(cases)}' in res.stdout
    words = open(tmp_path / 'words.txt').read().split('\n')
    klines = open(tmp_path / 'kmers.txt').read().split('\n')
    wi = ki = 0
    for name, c in zip(CASES, cases):
        n = int(words[wi])
        w = np.array([int(x) for x in words[wi + 1:wi + 1 + n]], np.uint64)
        wi += 1 + n
        kb = int(klines[ki])
        km = np.frombuffer(klines[ki + 1].encode(), np.uint8)
        assert len(km) == kb
        ki += 2
        got = parse(w, km, c['K'], len(c['reads']), c['persist'])
        assert got == c['expected'], name
