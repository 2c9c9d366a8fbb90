"""Build libvmbd_hip.so (all HIP sources under csrc/) for gfx950, in-tree.

    python -m vmbd_amd.build [-j N] [--verbose]

Every call recompiles.  Objects go to vmbd_amd/_build/, the shared library to
vmbd_amd/libvmbd_hip.so.  No GPU is needed (hipcc cross-compiles).
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libvmbd_hip.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = 'gfx950'

# Packed-FP32 VALU forms are disabled for every HIP translation unit.
CFLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
    '-mcode-object-version=5',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
    '-Wno-unused-command-line-argument',
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def _compile(src, verbose):
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + '.o')
    cmd = [HIPCC] + CFLAGS + ['-c', src, '-o', obj]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed on {src}:\n{r.stdout}\n{r.stderr}')
    # the host pass warns that it does not know the packed-fp32 feature
    noise = [ln for ln in r.stderr.splitlines() if ln.strip() and 'packed-fp32-ops' not in ln]
    if verbose and noise:
        print('\n'.join(noise), flush=True)
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', f'--offload-arch={ARCH}', '-o', tmp] + objs
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('-j', type=int, default=None)
    ap.add_argument('--verbose', action='store_true')
    a = ap.parse_args(argv)
    print(build(a.j, a.verbose))


if __name__ == '__main__':
    sys.exit(main())
