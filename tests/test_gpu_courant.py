"""Courant numbers of step launches combined through a context's group words
(each wavefront max-combines its value into its group's slot, passes it to
the output when it raised the slot, and counts itself in; the group's last
wavefront leaves slot and count at zero).  One context serves a sequence of
classic and SharpClaw launches of different grids, launch shapes and step
sizes, a large Courant number followed by a smaller one: every launch's
Courant number and state equal the oracle's bit for bit, which they would
not if a launch left a maximum or a count behind for the next, or if a
wavefront past the grid's bands did not count itself in."""
import numpy as np
import pytest

import oracle as O
from helpers import (RP_CASES, gpu_classic_step, gpu_sharp_dq, oracle_classic_step, padded_with_bc,
                     vc_aux)

pytestmark = pytest.mark.gpu


def _classic(ctx, name, mx, my, seed, dt_scale, shape):
    case = RP_CASES[name]
    rng = np.random.default_rng(seed)
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    lo, hi = (3, 3), (1, 1)
    qbc = padded_with_bc(q, 2, lo, hi, True, ghost_rng=rng)
    abc = padded_with_bc(aux, 2, (1, 3), (3, 1), False) if aux is not None else None
    dt = case['dt'] * dt_scale
    ref, cref = oracle_classic_step(qbc, abc, case, dt)
    got, cgot = gpu_classic_step(qbc, abc, case, dt, 2, 1, False, None, lo, hi, (1, 3), (3, 1),
                                 shape=shape, ctx=ctx)
    return ref, cref, got, cgot


def _sharp(ctx, name, mx, my, seed, dt_scale):
    case = RP_CASES[name]
    rng = np.random.default_rng(seed)
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    lo, hi = (1, 1), (1, 1)
    qbc = padded_with_bc(q, 3, lo, hi, True, ghost_rng=rng)
    abc = padded_with_bc(aux, 3, (3, 1), (1, 3), False) if aux is not None else None
    O.set_params(case['rp'], case['pd'])
    dt = case['dt'] * dt_scale
    ref, cref = O.sharp_dq(case['rp'], qbc, abc, 0.1, 0.07, dt, case['nw'])
    got, cgot = gpu_sharp_dq(qbc, abc, case, dt, lo, hi, (3, 1), (1, 3), ctx=ctx)
    return ref, cref, got, cgot


def test_courant_groups_over_a_sequence_of_launches_on_one_context():
    from pyclaw_amd import _lib
    ctx = _lib.Context(0)
    # (kind, solver, mx, my, seed, dt scale, launch shape): 5-row bands in
    # 4-wavefront blocks leave 3 of 12 wavefronts of each strip's blocks past
    # the bands; (6, 2, 1) gives 170 one-wavefront blocks (groups of 5-6);
    # the paired fill on 2000x301
    seq = [('classic', 'acoustics', 70, 45, 3, 1.0, (5, 1, 4)),
           ('classic', 'acoustics', 70, 45, 3, 0.25, (5, 1, 4)),
           ('classic', 'euler', 300, 200, 4, 1.0, (6, 2, 1)),
           ('sharp', 'shallow', 100, 80, 5, 1.0, None),
           ('classic', 'shallow', 2000, 301, 6, 1.0, (0, 2, 1)),
           ('sharp', 'shallow', 100, 80, 5, 0.5, None),
           ('classic', 'euler', 300, 200, 4, 0.5, (6, 2, 1)),
           ('classic', 'acoustics', 70, 45, 3, 1.0, (5, 1, 4))]
    seen = []
    for kind, name, mx, my, seed, scale, shape in seq:
        print("launch", kind, name, mx, my, "dt x", scale, "shape", shape, flush=True)
        if kind == 'classic':
            ref, cref, got, cgot = _classic(ctx, name, mx, my, seed, scale, shape)
        else:
            ref, cref, got, cgot = _sharp(ctx, name, mx, my, seed, scale)
        assert np.array_equal(ref, got), np.abs(ref - got).max()
        assert cgot == cref, (cgot, cref)
        seen.append(cref)
    # the sequence does step down in Courant number
    assert seen[1] < seen[0] and seen[6] < seen[2] and seen[5] < seen[3]
