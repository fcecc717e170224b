"""GPU parity: the fused classic kernel against the CPU oracle, one step from
identical seeded inputs.  FP64 is required to be bit-exact (the kernel keeps
the reference's operation order and is built without contraction); the CFL
number must be identical.  FP32 is checked against FP64 with a stated bound."""
import numpy as np
import pytest

import oracle as O
from helpers import RP_CASES, gpu_classic_step, oracle_classic_step, padded_with_bc, vc_aux

pytestmark = pytest.mark.gpu

BC_SETS = [
    ((1, 1), (1, 1)),      # extrap everywhere
    ((3, 3), (1, 1)),      # walls low, extrap high (config 1 / vc-acoustics app)
    ((2, 2), (2, 2)),      # periodic
    ((0, 3), (1, 0)),      # stored ("custom") x-low and y-high
    ((3, 1), (3, 3)),
]


def _run(name, mx, my, seed, order=2, trans=1, split=False, mthlim=None, bcs=BC_SETS[0]):
    case = RP_CASES[name]
    rng = np.random.default_rng(seed)
    q = case['make'](rng, mx, my)
    aux = vc_aux(rng, mx, my) if case['naux'] else None
    lo, hi = bcs
    qbc = padded_with_bc(q, 2, lo, hi, True, ghost_rng=rng)
    abc = padded_with_bc(aux, 2, (1, 3), (3, 1), False) if aux is not None else None
    ref, cref = oracle_classic_step(qbc, abc, case, case['dt'], order, trans, split, mthlim)
    got, cgot = gpu_classic_step(qbc, abc, case, case['dt'], order, trans, split, mthlim, lo, hi,
                                 (1, 3), (3, 1))
    return ref, cref, got, cgot


@pytest.mark.parametrize("name", list(RP_CASES))
@pytest.mark.parametrize("trans", [0, 1, 2])
def test_unsplit_bitexact(name, trans):
    print("case", name, "trans", trans, flush=True)
    ref, cref, got, cgot = _run(name, 70, 45, 1, trans=trans)
    assert np.array_equal(ref, got), np.abs(ref - got).max()
    assert cref == cgot


@pytest.mark.parametrize("bcs", BC_SETS)
@pytest.mark.parametrize("name", ['euler', 'vc_acoustics'])
def test_unsplit_bcs(name, bcs):
    print("case", name, bcs, flush=True)
    ref, cref, got, cgot = _run(name, 67, 38, 2, trans=2, bcs=bcs)
    assert np.array_equal(ref, got), np.abs(ref - got).max()
    assert cref == cgot


@pytest.mark.parametrize("lim", [0, 1, 2, 3, 4, 5])
def test_unsplit_limiters(lim):
    ref, cref, got, cgot = _run('euler', 130, 33, 3, trans=2, mthlim=[lim] * 5)
    assert np.array_equal(ref, got)
    assert cref == cgot


def test_unsplit_order1():
    ref, cref, got, cgot = _run('euler', 64, 64, 4, order=1, trans=1)
    assert np.array_equal(ref, got)
    assert cref == cgot


@pytest.mark.parametrize("name", list(RP_CASES))
def test_split_bitexact(name):
    ref, cref, got, cgot = _run(name, 75, 41, 5, split=True, mthlim=None, bcs=BC_SETS[1])
    assert np.array_equal(ref, got), np.abs(ref - got).max()
    assert cref == cgot


def test_split_shockbubble_limiters():
    ref, cref, got, cgot = _run('euler', 160, 40, 6, split=True, mthlim=[4, 4, 4, 4, 2], bcs=BC_SETS[3])
    assert np.array_equal(ref, got)
    assert cref == cgot


@pytest.mark.parametrize("shape", [(2, 2), (3, 200), (200, 3), (61, 61), (62, 1), (244, 123)])
def test_unsplit_ragged_shapes(shape):
    mx, my = shape
    if my < 2:
        pytest.skip("grid needs at least 2 cells per side")
    ref, cref, got, cgot = _run('euler', mx, my, 7, trans=2, bcs=BC_SETS[1])
    assert np.array_equal(ref, got)
    assert cref == cgot


def test_f32_error_vs_f64():
    """FP32 variant: relative L-inf error of one Euler step against the FP64 oracle."""
    case = RP_CASES['euler']
    rng = np.random.default_rng(8)
    q = case['make'](rng, 96, 64)
    qbc = padded_with_bc(q, 2, (1, 1), (1, 1), True)
    ref, cref = oracle_classic_step(qbc, None, case, case['dt'], 2, 2)
    got, cgot = gpu_classic_step(qbc, None, case, case['dt'], 2, 2, precision='f32')
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    print("f32 rel Linf", rel)
    assert rel < 1e-5
    assert abs(cgot - cref) / cref < 1e-5
