"""Converts the reference's own regression data files into one compressed
numpy archive (tests/golden/reference_fixtures.npz).  Values are parsed from
the text files exactly (float64); nothing is recomputed.

Sources (all under the reference's apps/ directory):
  acoustics_2d_homogeneous/verify_classic.txt     p at t=0.12, 100x100, classic split
  acoustics_2d_homogeneous/verify_sharpclaw.txt   p at t=0.12, 100x100, SharpClaw
  acoustics_2d_variable/pressure_classic.txt      p at frame 20 (t=0.6), 200x200
  acoustics_2d_variable/io_test_verification/fort.q0020, fort.a0000 (ascii frames,
                                                  and the SHA-256 of fort.q0020, fort.t0020, fort.a0000)
  euler_2d/verify_shockbubble_classic.txt         density at t=0.2, 160x40
Run in the container that has the reference: python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

REF = os.environ.get('REF', '/root/reference')
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main():
    from pyclaw_amd import io
    from pyclaw_amd.solution import Solution
    apps = os.path.join(REF, 'apps')
    out = {}
    out['acoustics_homog_classic_p'] = np.loadtxt(os.path.join(apps, 'acoustics_2d_homogeneous', 'verify_classic.txt'))
    out['acoustics_homog_sharpclaw_p'] = np.loadtxt(os.path.join(apps, 'acoustics_2d_homogeneous', 'verify_sharpclaw.txt'))
    out['acoustics_vc_pressure_classic'] = np.loadtxt(os.path.join(apps, 'acoustics_2d_variable', 'pressure_classic.txt'))
    out['shockbubble_density'] = np.loadtxt(os.path.join(apps, 'euler_2d', 'verify_shockbubble_classic.txt'))
    d = os.path.join(apps, 'acoustics_2d_variable', 'io_test_verification')
    t, num_eqn, nstates, num_aux, num_dim = io.read_t(20, d)
    s = Solution()
    io.read(s, 20, d, read_aux=False)
    out['acoustics_vc_fortq0020'] = s.state.q
    out['acoustics_vc_fortq0020_t'] = np.array(t)
    s0 = Solution()
    io.read(s0, 0, d, read_aux=True)
    out['acoustics_vc_forta0000'] = s0.state.aux
    # SHA-256 of the reference's frame files, to pin the ascii writer byte for byte
    import hashlib
    for name in ('fort.q0020', 'fort.t0020', 'fort.a0000'):
        with open(os.path.join(d, name), 'rb') as f:
            out['sha256_' + name.replace('.', '_')] = np.array(hashlib.sha256(f.read()).hexdigest())
    np.savez_compressed(os.path.join(HERE, 'reference_fixtures.npz'), **out)
    for k, v in out.items():
        print(k, np.shape(v))


if __name__ == '__main__':
    main()
