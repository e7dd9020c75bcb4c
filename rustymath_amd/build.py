"""Builds the hip-sys shared library (libhip_sys.so) in-tree from rustymath_amd/csrc/*.hip for gfx950.

Run `python -m rustymath_amd.build`. Every source compiles in parallel to an object under
rustymath_amd/_build/, then one link makes rustymath_amd/libhip_sys.so.
"""
import os, subprocess, sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libhip_sys.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = ['--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17', '-ffp-contract=off', '-fno-gpu-flush-denormals-to-zero',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-I', os.path.join(ROOT, 'include'), '-Wno-unused-result']


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src)[:-4] + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'hipcc failed on {src}:\n{r.stderr[-6000:]}')
    return obj


def build(verbose=False):
    "Compile every HIP source and link libhip_sys.so; raises with the compiler's message on failure."
    os.makedirs(OUT, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, sources()))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'link failed:\n{r.stderr[-6000:]}')
    os.replace(tmp, LIB)
    if verbose: print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
