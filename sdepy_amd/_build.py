"""Build recipe of libsdegpu.so (hand-written HIP for gfx950).

Every ``*.hip`` file under ``sdepy_amd/csrc`` is compiled by ``hipcc`` into an
object file (in parallel), and the objects are linked into
``sdepy_amd/libsdegpu.so`` next to this file.  The library is built in-tree so
that it travels with the source tree; nothing is installed.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')
OBJDIR = os.path.join(HERE, '_objs')
LIB = os.path.join(HERE, 'libsdegpu.so')

ARCH = 'gfx950'

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in generated code.
HIPCC_FLAGS = [
    '--offload-arch=' + ARCH, '-O3', '-fPIC', '-std=c++17',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result',
]


def hipcc():
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(path):
        raise RuntimeError('hipcc not found: the HIP toolchain is required '
                           'to build libsdegpu.so')
    return path


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith('.hip'))


def _compile(src, verbose):
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + '.o')
    cmd = [hipcc(), *HIPCC_FLAGS, '-I', INCLUDE, '-c', src, '-o', obj]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature
    err = '\n'.join(line for line in res.stderr.splitlines()
                    if 'packed-fp32-ops' not in line)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed on {}:\n{}'.format(src, err))
    if err.strip() and verbose:
        print(err, file=sys.stderr)
    return obj


def build(verbose=False, jobs=None):
    """Compile every HIP source and link libsdegpu.so; returns its path."""
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1), 16)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '--offload-arch=' + ARCH, '-shared', '-fPIC',
           *objs, '-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('link of libsdegpu.so failed:\n' + res.stderr)
    os.replace(tmp, LIB)
    build_c_smoke(verbose)
    return LIB


C_SMOKE_SRC = os.path.join(os.path.dirname(HERE), 'tests', 'c', 'abi_smoke.c')
C_SMOKE_BIN = os.path.join(os.path.dirname(HERE), 'tests', 'c', 'abi_smoke')


def build_c_smoke(verbose=False):
    """Compile the plain-C caller of the library (tests/c/abi_smoke.c) with
    gcc against include/sdegpu.h and the HIP runtime's C API; returns the
    binary's path, or None when the test sources are not present."""
    if not os.path.exists(C_SMOKE_SRC):
        return None
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    cmd = ['gcc', '-std=c11', '-O2', '-D__HIP_PLATFORM_AMD__',
           '-I', INCLUDE, '-I', os.path.join(rocm, 'include'),
           C_SMOKE_SRC, '-o', C_SMOKE_BIN,
           '-L', HERE, '-lsdegpu', '-Wl,-rpath,' + HERE,
           '-L', os.path.join(rocm, 'lib'), '-lamdhip64',
           '-Wl,-rpath,' + os.path.join(rocm, 'lib'), '-lm']
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('gcc failed on abi_smoke.c:\n' + res.stderr)
    return C_SMOKE_BIN


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
