"""End-to-end runs of the CPU oracle against the reference's stored regression
data (tests/golden/reference_fixtures.npz, converted by make_fixtures.py from
the reference's own fixture files), with the reference's own tolerances:
  verify_classic.txt / verify_sharpclaw.txt  L2 rel 1e-3  (apps/acoustics_2d_homogeneous/test_2d_acoustics.py:19)
  pressure_classic.txt                         max abs 0.1  (apps/acoustics_2d_variable/test_acoustics_2d_variable.py:21-22)
  fort.q0020 / fort.a0000                      L2 rel 1e-4  (test_acoustics_2d_variable_io.py:51-55)
  verify_shockbubble_classic.txt               L2 abs 1e-12 (apps/euler_2d/test_euler2d_shockbubble.py:17-19)
"""
import os

import numpy as np
import pytest

import apps
import oracle as O
from oracle import driver

GOLD = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'reference_fixtures.npz'))
RPID = {'acoustics': O.RP_ACOUSTICS, 'vc_acoustics': O.RP_VC_ACOUSTICS, 'euler': O.RP_EULER,
        'shallow': O.RP_SHALLOW}


def oracle_run(q0, aux, opts, step_source=None, user_bc_lower=None):
    P = driver.Problem(RPID[opts['rp']], opts['num_waves'], opts['problem_data'], opts['dx'], opts['dy'],
                       opts['bc_lower'], opts['bc_upper'], aux=aux,
                       aux_bc_lower=opts.get('aux_bc_lower', (1, 1)), aux_bc_upper=opts.get('aux_bc_upper', (1, 1)),
                       user_bc_lower=user_bc_lower, kind=opts['kind'],
                       dimensional_split=opts.get('dimensional_split', True),
                       transverse=opts.get('transverse', 1), limiters=opts['limiters'],
                       cfl_max=opts['cfl_max'], cfl_desired=opts['cfl_desired'], dt_initial=opts['dt_initial'],
                       step_source=step_source, source_split=opts.get('source_split', 1),
                       time_integrator=opts.get('time_integrator', 'SSP104'))
    run = driver.Run(P, q0)
    frames = run.run(opts['tfinal'], opts['num_output_times'])
    return frames, run


@pytest.fixture(scope='module')
def vc_classic_run():
    q0, aux, opts = apps.acoustics_variable('classic')
    return oracle_run(q0, aux, opts)


def test_homogeneous_acoustics_classic_fixture():
    q0, aux, opts = apps.acoustics_homogeneous('classic')
    frames, run = oracle_run(q0, aux, opts)
    ok, err = driver.check_diff(GOLD['acoustics_homog_classic_p'], frames[-1][0], reltol=1e-3)
    print("rel L2 err", err, "steps", run.nsteps)
    assert ok


def test_homogeneous_acoustics_sharpclaw_fixture():
    q0, aux, opts = apps.acoustics_homogeneous('sharpclaw')
    frames, run = oracle_run(q0, aux, opts)
    ok, err = driver.check_diff(GOLD['acoustics_homog_sharpclaw_p'], frames[-1][0], reltol=1e-3)
    print("rel L2 err", err, "steps", run.nsteps)
    assert ok


def test_variable_acoustics_classic_pressure_fixture(vc_classic_run):
    frames, run = vc_classic_run
    err = np.abs(GOLD['acoustics_vc_pressure_classic'] - frames[20][0]).max()
    print("max abs err", err, "steps", run.nsteps, "rejected", run.nrejected)
    assert err < 1e-1


def test_variable_acoustics_io_fixture(vc_classic_run):
    frames, run = vc_classic_run
    q20 = GOLD['acoustics_vc_fortq0020']
    ok, err = driver.check_diff(q20, frames[20], reltol=1e-4)
    print("rel L2 err vs fort.q0020", err)
    assert ok
    _, aux, _ = apps.acoustics_variable('classic')
    ok, err = driver.check_diff(GOLD['acoustics_vc_forta0000'], aux, reltol=1e-4)
    assert ok


def test_variable_acoustics_sharpclaw_pressure_fixture():
    q0, aux, opts = apps.acoustics_variable('sharpclaw')
    frames, run = oracle_run(q0, aux, opts)
    err = np.abs(GOLD['acoustics_vc_pressure_classic'] - frames[20][0]).max()
    print("max abs err", err, "steps", run.nsteps)
    assert err < 1e-1


def test_shockbubble_classic_fixture():
    q0, aux, opts = apps.shockbubble()
    rad = aux[0]
    frames, run = oracle_run(q0, aux, opts, step_source=lambda q, dt: apps.step_euler_radial_numpy(q, dt, rad),
                             user_bc_lower=apps.shockbubble_inflow_numpy)
    err = np.linalg.norm(GOLD['shockbubble_density'] - frames[-1][0])
    print("L2 abs err", err, "steps", run.nsteps, "rejected", run.nrejected)
    assert err < 1e-12
