"""Deeper DEM parity against the oracle's step: the C1 scene over 100 steps
with a soft spring (contacts throughout, rotations in expmap's series
branch), the C1 scene with its configured stiff spring until its rotations
leave the series branch, the funnel and the rotating box with more particles
and steps than the engine tests.  Every step's candidate pairs, contacts
and triangle checks equal the oracle's, and the whole state equals it bit
for bit.  The oracle runs with OpenMP over the rows of each particle pair."""
import os

import numpy as np
import pytest

import oracle_lib as ol
from delta_amd import _lib, engine, tuning, workloads
from test_gpu_engine import STATE, compare, sincos_branch_used

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def run(hs, k, damper, friction, dt, steps, cut_at_sincos=False):
    ds = engine.DeviceScene(hs)
    p = engine.step_params(method=_lib.DH_BF, dt=dt, spring=k, damper=damper, friction=friction,
                           cell=hs.cell_size())
    total = done = 0
    for s in range(steps):
        if sincos_branch_used(hs, dt):
            assert cut_at_sincos, f"step {s}: rotation in expmap's sin/cos branch"
            break
        info = engine.step(ds, p)
        nc, checks, npairs = ol.or_step(hs, _lib.DH_BF, k, damper, friction, dt,
                                        (0.0, -9.81, 0.0), omp_threads=THREADS)
        assert (info.contacts, info.candidate_pairs, info.triangle_checks) == (nc, npairs, checks), s
        compare(ds, hs, s)
        total += nc
        done += 1
    return total, done


def test_c1_soft_100_steps():
    """64 icospheres and the floor, k = 10 N/m: the overlapping lattice
    pushes apart slowly, with contacts in every step."""
    total, done = run(engine.c1_scene(), 1e1, 0.1, 0.0, 1e-6, 100)
    assert done == 100 and total > 100 * 1000


def test_c1_configured_until_the_sincos_branch():
    """The configured C1 (k = 1e5 N/m, dt = 1e-6 s): its first step flings the
    particles to rotations of several radians per step, where expmap calls
    sin and cos (and the device's may differ from the host libm's in the last
    bit, on 0.15% of arguments); the steps before are bit for bit."""
    total, done = run(engine.c1_scene(), 1e5, 0.1, 0.0, 1e-6, 30, cut_at_sincos=True)
    assert done >= 1 and total >= 14025


def test_funnel_120_particles():
    tuning.set(detect_queue_min_items=0)
    try:
        total, done = run(workloads.funnel_scene(n=120, pitch=1.3e-3), 1e1, 0.1, 0.5, 5e-7, 4)
    finally:
        tuning.set(**tuning.default())
    assert done == 4 and total > 0


def test_rotating_box_1000_particles():
    hs = workloads.drum_scene(n=1000, pitch=1.2e-3)
    total, done = run(hs, 1e1, 0.1, 0.5, 5e-7, 6)
    assert done == 6 and total > 0
    assert hs.rot[0][1] != 0.0  # the box turned
