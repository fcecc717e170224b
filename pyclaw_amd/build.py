"""Builds libclawhip.so (all HIP kernels + the C ABI) in-tree for gfx950.

Flags: -ffp-contract=off keeps every multiply and add separately rounded, as
in the reference's CPU build, so the FP64 kernels can match it bit for bit;
packed-FP32 instructions are disabled.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('CLAWHIP_ARCH', 'gfx950')

FLAGS = ['-O3', '--offload-arch=' + ARCH, '-fPIC', '-shared', '-std=c++17', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']


def sources():
    return sorted(glob.glob(os.path.join(HERE, 'csrc', '*.hip')))


def build(verbose=False):
    out = os.path.join(HERE, 'libclawhip.so')
    tmp = out + '.tmp'
    cmd = [HIPCC] + FLAGS + ['-o', tmp] + sources()
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("hipcc failed building libclawhip.so")
    os.replace(tmp, out)
    return out


if __name__ == '__main__':
    print(build(verbose=True))
