"""Build the gfx950 C-ABI library ``libn3ds.so`` in-tree with hipcc.

Usage: ``python -m neural3dstrokes_amd.build`` (or ``__graft_entry__.build()``).
Each source is compiled to an object in parallel, then linked into
``neural3dstrokes_amd/libn3ds.so``.  Packed FP32 instructions are disabled
(``-packed-fp32-ops``) for every device compilation.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'libn3ds.so')
BUILD_DIR = os.path.join(HERE, '_build')
SOURCES = ['n3ds_capi.cpp', 'n3ds_strokes.hip', 'n3ds_compose.hip', 'n3ds_grid.hip', 'n3ds_render.hip', 'n3ds_optim.hip',
           'n3ds_texture.hip', 'n3ds_stepfun.hip']
ARCH = os.environ.get('N3DS_OFFLOAD_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

INCLUDE = os.path.join(os.path.dirname(HERE), 'include')  # public C-ABI header n3ds.h

COMMON = ['-O3', '-std=c++17', '-fPIC', '-fvisibility=hidden', f'--offload-arch={ARCH}', f'-I{INCLUDE}',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-result']


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + COMMON + ['-c', os.path.join(CSRC, src), '-o', obj]
    if src.endswith('.cpp'):
        cmd = [HIPCC] + COMMON + ['-x', 'hip', '-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'compile failed: {" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    return obj


def build(verbose=True):
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', f'--offload-arch={ARCH}', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed: {" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, OUT)
    if verbose:
        print(f'built {OUT}')
    return OUT


if __name__ == '__main__':
    build()
    sys.exit(0)
