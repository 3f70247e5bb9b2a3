"""Build libgpawhip.so (the HIP kernels + C ABI) for gfx950 in-tree.

    python -m gpaw_amd.build

Compiles every gpaw_amd/csrc/*.hip with hipcc into gpaw_amd/lib/.  Packed
FP32 instructions are disabled (-packed-fp32-ops) for all device code.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIBDIR = os.path.join(HERE, 'lib')
LIB = os.path.join(LIBDIR, 'libgpawhip.so')
ARCH = os.environ.get('GPAW_AMD_ARCH', 'gfx950')

FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH,
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-munsafe-fp-atomics', '-Wall', '-Wno-unused-function',
         '-Wno-unused-variable']


def sources():
    return sorted(glob.glob(os.path.join(CSRC, '*.hip')))


def build(verbose=False, jobs=8):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, 'obj')
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src)[:-4] + '.o')
        objs.append(obj)
        cmd = ['hipcc', *FLAGS, '-I', os.path.join(HERE, '..', 'include'),
               '-c', src, '-o', obj]
        if verbose:
            print(' '.join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + '.tmp'
    cmd = ['hipcc', '-shared', '-fPIC', '--offload-arch=' + ARCH, *objs,
           '-o', tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = out.decode(errors='replace')
    if p.returncode != 0:
        sys.stderr.write(text)
        raise RuntimeError('hipcc failed for ' + src)
    if text.strip():
        sys.stderr.write(text)


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
