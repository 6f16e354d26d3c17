"""Build libbqcjk.so (HIP kernels for gfx950 + host C++) in-tree.

    python -m byteqc_amd.build [--incremental]

--incremental recompiles only sources older than their object or than the
newest header under csrc/ and include/.
Objects go to byteqc_amd/_build/, the library to byteqc_amd/lib/libbqcjk.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIBDIR = os.path.join(HERE, 'lib')
LIB = os.path.join(LIBDIR, 'libbqcjk.so')
ARCH = os.environ.get('BQC_OFFLOAD_ARCH', 'gfx950')

SOURCES = ['jk.hip', 'jk_bra_00.hip', 'jk_bra_10.hip', 'jk_bra_11.hip', 'jk_bra_20.hip',
           'jk_bra_21.hip', 'jk_bra_22.hip', 'jk_bra_30.hip', 'jk_bra_31.hip', 'jk_bra_32.hip',
           'jk_bra_33.hip', 'jk_general.hip', 'schwarz.hip', 'int1e.hip', 'dmcond.hip', 'fill.hip', 'fill_bra_a.hip',
           'fill_bra_b.hip', 'fill_bra_c.hip', 'fill_bra_d.hip', 'pair_data.cpp', 'planner.cpp']
HIP_FLAGS = ['-O3', '-fPIC', '-std=c++17', f'--offload-arch={ARCH}', '-munsafe-fp-atomics',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-Wno-unused-command-line-argument']


def _hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.exists(cand) or cand == 'hipcc'):
            return cand
    raise RuntimeError('hipcc not found')


def _newest_header():
    inc = os.path.join(os.path.dirname(HERE), 'include')
    paths = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(('.h', '.inc'))]
    paths += [os.path.join(inc, f) for f in os.listdir(inc)]
    return max(os.path.getmtime(p) for p in paths)


def _compile(src, incremental=False):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    if incremental and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if t > os.path.getmtime(os.path.join(CSRC, src)) and t > _newest_header():
            return obj
    cmd = [_hipcc(), *HIP_FLAGS, '-c', os.path.join(CSRC, src), '-o', obj]
    if src.endswith('.cpp'):
        cmd = [_hipcc(), '-O3', '-fPIC', '-std=c++17', '-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'compile {src} failed:\n{r.stdout}\n{r.stderr}')
    return obj


def build(verbose=False, incremental=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda f: _compile(f, incremental), SOURCES))
    tmp = LIB + '.tmp'
    cmd = [_hipcc(), '-shared', f'--offload-arch={ARCH}', '-o', tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True, incremental='--incremental' in sys.argv[1:])
    sys.exit(0)
