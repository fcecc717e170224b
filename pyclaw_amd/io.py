"""Clawpack ascii frame files (fort.tNNNN / fort.qNNNN / fort.aNNNN).

Same file format as the reference's writer and reader
(src/pyclaw/io/ascii.py:14-152 write, :154-357 read): a fort.t header with
time, num_eqn, nstates, num_aux, num_dim; per patch a header (patch_number,
AMR_level, m<dim>, <dim>low, d<dim>) followed by one line per cell, x fastest,
with "%18.8e" per component and a blank line after each row.
"""
import os

import numpy as np


def _fname(prefix, kind, frame):
    return '%s.%s%s' % (prefix, kind, str(frame).zfill(4))


def _write_field(f, patch, data):
    f.write("%5i                  patch_number\n" % patch.patch_index)
    f.write("%5i                  AMR_level\n" % patch.level)
    for d in patch.dimensions:
        f.write("%5i                  m%s\n" % (d.num_cells, d.name))
    for d in patch.dimensions:
        f.write("%18.8e     %slow\n" % (d.lower, d.name))
    for d in patch.dimensions:
        f.write("%18.8e     d%s\n" % (d.delta, d.name))
    f.write("\n")
    ncomp, mx, my = data.shape
    for j in range(my):
        rows = data[:, :, j].T          # (mx, ncomp)
        f.write(''.join(''.join("%18.8e" % v for v in row) + '\n' for row in rows))
        f.write('\n')


def write(solution, frame, path='./', file_format='ascii', file_prefix=None, write_aux=False,
          options={}, write_p=False):
    if file_format not in ('ascii',):
        raise NotImplementedError("output format %s is not available" % file_format)
    prefix = file_prefix or 'fort'
    os.makedirs(path, exist_ok=True)
    state0 = solution.states[0]
    with open(os.path.join(path, _fname(prefix, 't', frame)), 'w') as f:
        f.write("%18.8e     time\n" % state0.t)
        f.write("%5i                  num_eqn\n" % state0.num_eqn)
        f.write("%5i                  nstates\n" % len(solution.states))
        f.write("%5i                  num_aux\n" % state0.num_aux)
        f.write("%5i                  num_dim\n" % solution.domain.num_dim)
    with open(os.path.join(path, 'fort.q' + str(frame).zfill(4)), 'w') as f:
        for s in solution.states:
            _write_field(f, s.patch, s.p if write_p else s.q)
    if state0.num_aux > 0 and write_aux:
        with open(os.path.join(path, 'fort.a' + str(frame).zfill(4)), 'w') as f:
            for s in solution.states:
                _write_field(f, s.patch, s.aux)


def read_t(frame, path='./_output', file_prefix='fort'):
    with open(os.path.join(path, _fname(file_prefix or 'fort', 't', frame))) as f:
        t = float(f.readline().split()[0])
        num_eqn = int(f.readline().split()[0])
        nstates = int(f.readline().split()[0])
        num_aux = int(f.readline().split()[0])
        num_dim = int(f.readline().split()[0])
    return t, num_eqn, nstates, num_aux, num_dim


def _read_field(lines, pos, num_dim, ncomp):
    hdr = {}
    for _ in range(2 + 3 * num_dim):
        val, key = lines[pos].split()[:2]
        hdr[key] = val
        pos += 1
    names = [k[1:] for k in hdr if k.startswith('m') and k not in ('patch_number',)][:num_dim]
    n = [int(hdr['m' + nm]) for nm in names]
    low = [float(hdr[nm + 'low']) for nm in names]
    dlt = [float(hdr['d' + nm]) for nm in names]
    ncell = int(np.prod(n))
    vals = []
    while len(vals) < ncell * ncomp:
        ln = lines[pos]
        pos += 1
        if ln.strip():
            vals.extend(float(x) for x in ln.split())
    data = np.array(vals).reshape(n[1], n[0], ncomp).transpose(2, 1, 0)
    return pos, names, n, low, dlt, int(hdr['patch_number']), int(hdr['AMR_level']), np.asfortranarray(data)


def read(solution, frame, path='./_output', file_format='ascii', file_prefix=None, read_aux=True,
         options={}):
    """Fill `solution` (possibly empty) from frame files (ascii.py:154-357)."""
    from .geometry import Dimension, Domain, Patch
    from .state import State
    prefix = file_prefix or 'fort'
    t, num_eqn, nstates, num_aux, num_dim = read_t(frame, path, prefix)
    with open(os.path.join(path, 'fort.q' + str(frame).zfill(4))) as f:
        lines = f.readlines()
    alines = None
    apath = os.path.join(path, 'fort.a' + str(frame).zfill(4))
    if read_aux and num_aux > 0 and os.path.exists(apath):
        with open(apath) as f:
            alines = f.readlines()
    pos = apos = 0
    states = []
    patches = []
    for _ in range(nstates):
        pos, names, n, low, dlt, pidx, lvl, q = _read_field(lines, pos, num_dim, num_eqn)
        dims = [Dimension(nm, lo, lo + nn * d, nn) for nm, lo, nn, d in zip(names, low, n, dlt)]
        patch = Patch(dims)
        patch.patch_index = pidx
        patch.level = lvl
        s = State(patch, num_eqn, num_aux)
        s.t = t
        s.q = q
        if alines is not None:
            apos, _, _, _, _, _, _, a = _read_field(alines, apos, num_dim, num_aux)
            s.aux = a
        states.append(s)
        patches.append(patch)
    solution.__dict__['states'] = states
    solution.__dict__['domain'] = Domain(patches)
