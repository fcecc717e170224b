"""Host-side logic (no GPU): the C ABI library, geometry, ghost fill, State,
solver options, frame IO."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import oracle as O
import pyclaw_amd as pc
from pyclaw_amd import _lib
from pyclaw_amd.solver import fill_ghosts

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, 'include', 'clawhip.h')).read()
    names = set(re.findall(r'^int\s+(claw\w+)\s*\(', hdr, flags=re.M))
    assert len(names) >= 8
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_abi_struct_layout():
    # the C structs as declared in include/clawhip.h (natural alignment)
    assert ctypes.sizeof(_lib.ClawGrid) == 4 * 4 + 8 * 2 + 8 * 2
    assert ctypes.sizeof(_lib.ClawMethod) == 4 * 5 + 4 * 8
    assert ctypes.sizeof(_lib.ClawRiemann) == 8 + 8 * 8
    assert ctypes.sizeof(_lib.ClawBC) == 4 * 8


def test_invalid_arguments_return_einval_without_gpu():
    L = _lib.lib()
    ctx = _lib.Context(0)
    g = _lib.ClawGrid(1, 1, 2, 0, 5, 25, 0.1, 0.1)         # too small a grid
    m = _lib.ClawMethod(); m.order = 2; m.num_waves = 5
    r = _lib.ClawRiemann(); r.kind = _lib.RP_EULER
    bc = _lib.ClawBC()
    rc = L.claw2d_classic_step(ctx.handle, g, m, r, bc, _lib.F64, ctypes.c_void_p(16), ctypes.c_void_p(32),
                               None, ctypes.c_void_p(48), ctypes.c_void_p(64), None)
    assert rc == _lib.CLAW_EINVAL
    # the allocation-free step validates before it looks up its device array
    rc = L.claw2d_classic_step_noalloc(ctx.handle, g, m, r, bc, _lib.F64, ctypes.c_void_p(16), ctypes.c_void_p(32),
                                       None, ctypes.c_void_p(48), ctypes.c_void_p(64), None)
    assert rc == _lib.CLAW_EINVAL
    assert L.claw2d_classic_step_noalloc(None, g, m, r, bc, _lib.F64, ctypes.c_void_p(16), ctypes.c_void_p(32),
                                         None, ctypes.c_void_p(48), ctypes.c_void_p(64), None) == _lib.CLAW_EINVAL
    g = _lib.ClawGrid(8, 8, 3, 0, 14, 196, 0.1, 0.1)
    bc.lower[0], bc.upper[0] = _lib.BC_PERIODIC, _lib.BC_EXTRAP   # periodic on one side only
    rc = L.claw2d_classic_step(ctx.handle, g, m, r, bc, _lib.F64, ctypes.c_void_p(16), ctypes.c_void_p(32),
                               None, ctypes.c_void_p(48), ctypes.c_void_p(64), None)
    assert rc == _lib.CLAW_EINVAL
    # the context switch and the source terms validate before any launch
    assert L.claw_ctx_set_wide_offsets(None, 1) == _lib.CLAW_EINVAL
    assert L.claw_ctx_set_wide_offsets(ctx.handle, 2) == _lib.CLAW_EINVAL
    assert L.claw_ctx_set_wide_offsets(ctx.handle, 0) == _lib.CLAW_OK
    q, a, d = ctypes.c_void_p(16), ctypes.c_void_p(32), ctypes.c_void_p(48)
    assert L.claw2d_euler_radial_step(g, _lib.F64, 5, q, None, 0.1, 0.4, None) == _lib.CLAW_EINVAL   # no aux
    assert L.claw2d_euler_radial_step(g, _lib.F64, 3, q, a, 0.1, 0.4, None) == _lib.CLAW_EINVAL      # num_eqn < 4
    assert L.claw2d_euler_radial_step(g, 7, 5, q, a, 0.1, 0.4, None) == _lib.CLAW_EINVAL             # precision
    assert L.claw2d_euler_radial_step(g, _lib.F64, 5, q, a, -0.1, 0.4, None) == _lib.CLAW_EINVAL     # dt < 0
    assert L.claw2d_euler_radial_dq(g, _lib.F64, 5, q, a, q, 0.1, 0.4, None) == _lib.CLAW_EINVAL     # dq aliases q
    # the device step loop's bookkeeping
    rule = _lib.ClawLoopRule(1.0, 0.9, 1e99, 1.0, 0.0, 10, 0)
    s0, s1, c0, c1 = ctypes.c_void_p(256), ctypes.c_void_p(512), ctypes.c_void_p(768), ctypes.c_void_p(776)
    qa, qb = ctypes.c_void_p(4096), ctypes.c_void_p(8192)
    adv = L.claw2d_loop_advance
    assert adv(None, s0, s1, c0, c1, qa, qb, 64, None) == _lib.CLAW_EINVAL
    assert adv(rule, s0, s0, c0, c1, qa, qb, 64, None) == _lib.CLAW_EINVAL        # cur == next
    assert adv(rule, s0, s1, c0, c0, qa, qb, 64, None) == _lib.CLAW_EINVAL        # cfl slots alias
    assert adv(rule, s0, s1, c0, c1, qa, qa, 64, None) == _lib.CLAW_EINVAL        # qold == qnew
    assert adv(rule, s0, s1, c0, c1, qa, qb, 62, None) == _lib.CLAW_EINVAL        # not whole 4-byte words
    assert adv(rule, s0, s1, c0, c1, qa, ctypes.c_void_p(8200), 64, None) == _lib.CLAW_EINVAL   # misaligned
    assert L.claw2d_classic_loop(ctx.handle, g, m, r, bc, _lib.F64, qa, qb, None, rule, s0, c0, 64, 0, -1,
                                 None, None) == _lib.CLAW_EINVAL
    assert L.claw2d_sharpclaw_loop(ctx.handle, g, r, bc, _lib.F64, qa, qb, qa, qa, qa, None, rule, s0, c0, 64, 0, 0,
                                   None, None) == _lib.CLAW_EINVAL                     # the SSP104 loop needs ncfl = 10
    rule10 = _lib.ClawLoopRule(1.0, 0.9, 1e99, 1.0, 0.0, 10, 10)
    assert adv(rule10, s0, s1, c0, c1, qa, qb, 64, None) == _lib.CLAW_EINVAL      # 10-value slots overlap
    assert adv(_lib.ClawLoopRule(1.0, 0.9, 1e99, 1.0, 0.0, 10, 17), s0, s1, c0, ctypes.c_void_p(1024), qa, qb, 64,
               None) == _lib.CLAW_EINVAL
    assert L.claw2d_classic_loop(ctx.handle, g, m, r, bc, _lib.F64, qa, qb, None, rule, s0, c0, 64, 0, 0,
                                 None, None) == _lib.CLAW_OK                     # nothing to launch
    G = _lib.ClawLoopGauges(None, None, None, 3, 0, 16, 0)
    assert L.claw2d_classic_loop(ctx.handle, g, m, r, bc, _lib.F32, qa, qb, None, rule, s0, c0, 64, 0, 0,
                                 ctypes.byref(G), None) == _lib.CLAW_EINVAL     # gauges are recorded in FP64
    bad = _lib.ClawGrid(8, 8, 3, 0, 13, 196, 0.1, 0.1)   # ldq < mx + 2 ng
    assert L.claw2d_euler_radial_dq(bad, _lib.F64, 5, q, a, d, 0.1, 0.4, None) == _lib.CLAW_EINVAL


def test_dimension_and_domain_match_reference_doctests():
    # geometry.py:397-421
    x = pc.Dimension('x', 0., 1., 100)
    assert x.num_cells == 100 and x.delta == 0.01
    assert x.edges[0] == 0.0 and x.edges[1] == 0.01 and x.edges[-1] == 1.0
    assert x.centers[-1] == 0.995 and len(x.centers) == 100 and len(x.edges) == 101
    d = pc.Domain((0., -1.), (1., 1.), (10, 25))
    assert d.grid.num_cells == [10, 25] and d.grid.lower == [0.0, -1.0]
    assert d.grid.delta == [0.1, 0.08]
    assert d.patch.x.name == 'x' and d.patch.y.name == 'y'


@pytest.mark.parametrize("lo,hi", [((1, 1), (1, 1)), ((3, 3), (1, 1)), ((2, 2), (2, 2)), ((3, 1), (1, 3)),
                                   ((2, 3), (2, 3))])
@pytest.mark.parametrize("ng", [2, 3])
def test_host_ghost_fill_matches_oracle(lo, hi, ng):
    rng = np.random.default_rng(5)
    q = rng.standard_normal((3, 9, 7))
    ref = np.zeros((3, 9 + 2 * ng, 7 + 2 * ng), order='F')
    ref[:, ng:-ng, ng:-ng] = q
    O.fill_bc(ref, ng, lo, hi, True)
    t = torch.zeros((3, 9 + 2 * ng, 7 + 2 * ng), dtype=torch.float64)
    t[:, ng:-ng, ng:-ng] = torch.from_numpy(q)
    fill_ghosts(t, ng, lo, hi, True)
    assert np.array_equal(t.numpy(), ref)


def test_custom_bc_callback_sees_reference_layout():
    seen = {}

    def cb(state, dim, t, qbc, ng):
        seen['shape'] = tuple(qbc.shape)
        for i in range(ng):
            qbc[0, i, ...] = 7.0
    t = torch.zeros((3, 6 + 4, 5 + 4), dtype=torch.float64)
    fill_ghosts(t, 2, (0, 1), (1, 1), True, user_lower=cb)
    assert seen['shape'] == (3, 10, 9)
    # x-lower ghosts hold the user value except the corners, refilled by the y rules
    assert np.all(t[0, :2, 2:-2].numpy() == 7.0)


def test_state_host_mirror_and_device_buffer():
    x = pc.Dimension('x', 0., 1., 6)
    y = pc.Dimension('y', 0., 2., 4)
    dom = pc.Domain([x, y])
    s = pc.State(dom, 3, 2, device='cpu')
    assert s.q.shape == (3, 6, 4) and s.aux.shape == (2, 6, 4)
    q = np.arange(72.).reshape(3, 6, 4)
    s.q[...] = q
    s.sync_to_device()
    ng = 3
    buf = s._q.buf
    assert np.array_equal(buf[:, ng:ng + 4, ng:ng + 6].permute(0, 2, 1).numpy(), q)
    # device-side change is seen by the next host read
    s.q_device[1, 2, 3] = -1.0
    assert s.q[1, 2, 3] == -1.0
    assert np.array_equal(s.get_q_global()[0], q[0])


def test_solver_rejects_unknown_attributes_and_checks_problem_data():
    s = pc.ClawSolver2D(pc.riemann.rp2_euler_5wave)
    with pytest.raises(TypeError):
        s.no_such_option = 1
    assert s.num_waves == 5 and s.dimensional_split and s.transverse_waves == 1
    assert s.cfl_max == 1.0 and s.cfl_desired == 0.9 and s.dt_initial == 0.1
    with pytest.raises(Exception):
        pc.riemann.rp2_euler_5wave.params({'gamma': 1.4})
    sh = pc.SharpClawSolver2D(pc.riemann.rp2_shallow_roe_with_efix)
    assert sh.cfl_max == 2.5 and sh.cfl_desired == 2.45 and sh.time_integrator == 'SSP104'


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_solver_fails_loudly_without_gpu():
    x = pc.Dimension('x', 0., 1., 8)
    y = pc.Dimension('y', 0., 1., 8)
    dom = pc.Domain([x, y])
    st = pc.State(dom, 3)
    st.problem_data.update(rho=1., bulk=1., cc=1., zz=1.)
    sol = pc.Solution(st, dom)
    s = pc.ClawSolver2D(pc.riemann.rp2_acoustics)
    s.bc_lower = [1, 1]; s.bc_upper = [1, 1]
    with pytest.raises(_lib.ClawError):
        s.setup(sol)


def test_ascii_frame_roundtrip(tmp_path):
    x = pc.Dimension('x', -1., 1., 5)
    y = pc.Dimension('y', 0., 1., 3)
    dom = pc.Domain([x, y])
    st = pc.State(dom, 3, 2, device='cpu')
    rng = np.random.default_rng(0)
    st.q[...] = rng.standard_normal((3, 5, 3))
    st.aux[...] = rng.random((2, 5, 3))
    st.t = 0.25
    sol = pc.Solution(st, dom)
    sol.write(4, str(tmp_path), write_aux=True)
    back = pc.Solution()
    back.read(4, str(tmp_path), read_aux=True)
    assert back.state.t == 0.25
    assert np.allclose(back.state.q, st.q, rtol=1e-7, atol=1e-8)
    assert np.allclose(back.state.aux, st.aux, rtol=1e-7, atol=1e-8)
    assert back.domain.grid.num_cells == [5, 3]


def test_ascii_writer_reproduces_the_reference_frame_files(tmp_path):
    """Frame 20 of the reference's variable-coefficient acoustics io test
    (apps/acoustics_2d_variable/io_test_verification), written by
    Solution.write from the values parsed out of those files: fort.q0020,
    fort.t0020 and the aux file are byte-identical to the reference's (SHA-256
    recorded by tests/golden/make_fixtures.py)."""
    import hashlib
    G = np.load(os.path.join(ROOT, 'tests', 'golden', 'reference_fixtures.npz'))
    x = pc.Dimension('x', -1., 1., 200)
    y = pc.Dimension('y', -1., 1., 200)
    dom = pc.Domain([x, y])
    st = pc.State(dom, 3, 2, device='cpu')
    st.q[...] = G['acoustics_vc_fortq0020']
    st.aux[...] = G['acoustics_vc_forta0000']
    st.t = float(G['acoustics_vc_fortq0020_t'])
    pc.Solution(st, dom).write(20, str(tmp_path), write_aux=True)
    for mine, ref in (('fort.q0020', 'fort_q0020'), ('fort.t0020', 'fort_t0020'), ('fort.a0020', 'fort_a0000')):
        with open(os.path.join(str(tmp_path), mine), 'rb') as f:
            digest = hashlib.sha256(f.read()).hexdigest()
        assert digest == str(G['sha256_' + ref]), mine


def test_gauge_indices_round_as_python2():
    """geometry.py:349-350 rounds with Python 2's round (halves away from zero);
    on power-of-two spacings gauges at exact cell edges hit the halves."""
    from pyclaw_amd.geometry import py2_round
    assert [py2_round(v) for v in (0.5, 1.5, 2.5, 3.5, 0.49999999999999994, 2.4999999999999996, 7.0)] == \
        [1, 2, 3, 4, 0, 2, 7]
    assert py2_round(-2.5) == -3
    x = pc.Dimension('x', 0.0, 1.0, 4)     # delta 0.25
    y = pc.Dimension('y', 0.0, 2.0, 1024)  # delta 2**-9
    dom = pc.Domain([x, y])
    st = pc.State(dom, 3)
    st.grid.add_gauges([(0.125, 2.0 ** -10), (0.375, 3 * 2.0 ** -10), (0.2, 0.0)])
    # Python 3's round would give [0, 0], [2, 2], [1, 0]
    assert st.grid.gauges == [[1, 1], [2, 2], [1, 0]]


def test_gauge_in_last_half_cell_raises_index_error():
    """A gauge in the last half cell gets index num_cells (geometry.py:349-350);
    the reference then raises IndexError reading q[:, ix, iy] (solver.py:722-723)."""
    from pyclaw_amd.solver import check_gauge_cells, gauge_cells_inside
    x = pc.Dimension('x', 0.0, 1.0, 10)
    y = pc.Dimension('y', 0.0, 1.0, 8)
    st = pc.State(pc.Domain([x, y]), 3)
    st.grid.add_gauges([(0.5, 0.5)])
    check_gauge_cells(st)
    assert gauge_cells_inside(st)
    st.grid.add_gauges([(0.97, 0.5)])
    assert st.grid.gauges[-1] == [10, 4]
    assert not gauge_cells_inside(st)
    with pytest.raises(IndexError, match="out of bounds for axis 1 with size 10"):
        check_gauge_cells(st)
    st2 = pc.State(pc.Domain([pc.Dimension('x', 0.0, 1.0, 10), pc.Dimension('y', 0.0, 1.0, 8)]), 3)
    st2.grid.add_gauges([(0.5, 0.99)])
    with pytest.raises(IndexError, match="axis 2 with size 8"):
        check_gauge_cells(st2)


def test_gauge_lines_use_python2_str():
    """The reference writes str(t) and str(value) (solver.py:736-737) under
    Python 2: 12 significant digits, '.0' after integral values."""
    from pyclaw_amd.solver import py2_str
    cases = {0.1 + 0.2: '0.3', 1.0 / 3.0: '0.333333333333', 100.0: '100.0', 1e-5: '1e-05',
             123456789012345.0: '1.23456789012e+14', -0.0: '-0.0', 2.0 ** 60: '1.15292150461e+18',
             float('inf'): 'inf', 0.6000000000000001: '0.6', 5e-324: '4.94065645841e-324'}
    for v, s in cases.items():
        assert py2_str(v) == s, (v, py2_str(v), s)
        assert py2_str(np.float64(v)) == s
    assert py2_str(float('nan')) == 'nan'
    assert py2_str(3) == '3'
    # a line as the solver writes it
    solver = pc.ClawSolver2D(pc.riemann.rp2_acoustics)

    class F(object):
        def __init__(self):
            self.lines = []

        def write(self, s):
            self.lines.append(s)
    st = pc.State(pc.Domain([pc.Dimension('x', 0.0, 1.0, 4), pc.Dimension('y', 0.0, 1.0, 4)]), 3)
    st.grid.gauges = [[0, 0]]
    f = F()
    st.grid.gauge_files = [f]
    solver._emit_gauge(st, 0, 0.1 + 0.2, np.array([1.0, 2.0 / 3.0, -1e-20]))
    assert f.lines == ['0.3 1.0 0.666666666667 -1e-20\n']
