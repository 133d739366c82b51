"""Builds goetia_amd's native library for gfx950 in-tree.

    python -m goetia_amd.build        # -> goetia_amd/libgoetia_amd.so

hipcc compiles the HIP kernels and the C ABI into one shared library; no
GPU is needed to build.  Packed-FP32 instructions are disabled for every HIP
translation unit (no kernel here uses floating point, and the flag keeps the
compiler from emitting them).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libgoetia_amd.so')
ARCH = os.environ.get('GOETIA_AMD_ARCH', 'gfx950')

SOURCES = ['gt_kernels.hip', 'gt_capi.cpp', 'gt_serving.cpp', 'gt_fastx.cpp', 'gt_membench.hip', 'gt_segments.cpp',
           'gt_oxli.cpp', 'gt_sorted_iq.hip', 'gt_bigcount.hip']
HIP_FLAGS = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-Wall', '-Wno-unused-function',
             f'-I{os.path.join(ROOT, "include")}', f'-I{CSRC}']


def hipcc():
    exe = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(exe):
        raise RuntimeError('hipcc not found; goetia_amd needs ROCm to build')
    return exe


def _compile(src, objdir, verbose):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + '.o')
    cmd = [hipcc(), *HIP_FLAGS, '-c', os.path.join(CSRC, src), '-o', obj]
    if src.endswith('.cpp'):
        cmd[1:1] = ['-x', 'hip']
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f'hipcc failed on {src}')
    return obj


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    jobs = max(1, min(len(SOURCES), int(os.environ.get('MAX_JOBS', os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, objdir, verbose), SOURCES))
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '-shared', f'--offload-arch={ARCH}', '-o', tmp, *objs, '-lz']
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('link of libgoetia_amd.so failed')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
