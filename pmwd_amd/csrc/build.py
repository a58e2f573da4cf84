"""Build libpmx.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m pmwd_amd.csrc.build  (or build() from Python).
The library lands next to the package (pmwd_amd/libpmx.so) so that it
travels with the source tree.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, 'libpmx.so')
SOURCES = ['pmx_ctx.hip', 'pmx_fft.hip', 'pmx_cic.hip', 'pmx_adjoint.hip', 'pmx_generic.hip',
           'pmx_sort.hip', 'pmx_pipeline.hip', 'pmx_random.hip', 'pmx_noalloc.hip', 'pmx_cosmo.hip', 'pmx_lpt.hip']

HIPCC = os.environ.get('HIPCC', shutil.which('hipcc') or '/opt/rocm/bin/hipcc')

FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
    # separate float32 roundings, as the reference's XLA ops
    '-ffp-contract=off',
    # no packed-FP32 instructions
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I' + os.path.join(ROOT, 'include'),
    '-Wno-unused-result', '-Wno-unused-value',
]


def _compile(src, objdir):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC, *FLAGS, '-c', os.path.join(HERE, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{r.stdout}\n{r.stderr}')
    return obj


def build(verbose=False):
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir), SOURCES))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', *objs, '-o', tmp,
           '-Wl,-rpath,/opt/rocm/lib']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
