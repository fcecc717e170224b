"""Builds libclawhip.so (all HIP kernels + the C ABI) in-tree for gfx950.

Flags: -ffp-contract=off keeps every multiply and add separately rounded, as
in the reference's CPU build, so the FP64 kernels can match it bit for bit;
packed-FP32 instructions are disabled.  Each source is compiled to an object
on its own (in parallel), then the objects are linked into one shared
library; no device code refers across files.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('CLAWHIP_ARCH', 'gfx950')

FLAGS = ['-O3', '--offload-arch=' + ARCH, '-fPIC', '-std=c++17', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']


def sources():
    return sorted(glob.glob(os.path.join(HERE, 'csrc', '*.hip')))


def _run(cmd, verbose):
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return r.returncode, r.stdout


def build(verbose=False, jobs=None):
    out = os.path.join(HERE, 'libclawhip.so')
    tmp = out + '.tmp'
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 1)))
    with tempfile.TemporaryDirectory(prefix='clawhip_build_') as d:
        objs = [os.path.join(d, os.path.splitext(os.path.basename(s))[0] + '.o') for s in srcs]
        cmds = [[HIPCC] + FLAGS + ['-c', '-o', o, s] for s, o in zip(srcs, objs)]
        with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
            results = list(ex.map(lambda c: _run(c, verbose), cmds))
        for (rc, text), s in zip(results, srcs):
            if rc != 0:
                sys.stderr.write(text)
                raise RuntimeError("hipcc failed compiling %s" % os.path.basename(s))
        rc, text = _run([HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs, verbose)
        if rc != 0:
            sys.stderr.write(text)
            raise RuntimeError("hipcc failed linking libclawhip.so")
    os.replace(tmp, out)
    return out


if __name__ == '__main__':
    print(build(verbose=True))
