"""Pins of the adaptive time-step restatement (Collision.cpp:188-299,
State.cpp:26-35,137-199): head-on cases worked by hand, one per branch of
triggerParticleTooClose and of adaptiveTimeStep."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from delta_amd.engine import DhAdaptiveState


def state(dt, mesh=1.0, min_dt=1e-4, step=0, travel=0.0):
    s = DhAdaptiveState()
    s.dt, s.min_dt, s.increase, s.mesh_width = dt, min_dt, 1.1, mesh
    s.max_dt = float("inf")
    s.step = step
    s.max_vel_travel = travel
    return s


def pairs_step(s, vA, vB, radius=0.3, eps=0.01):
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0]]).ravel()
    vel = np.array([vA, vB], float).ravel()
    rad = np.array([radius, radius])
    ep = np.array([eps, eps])
    pr = np.array([0, 1], np.int32)
    o = ol.oracle()
    o.or_adaptive_pairs(pr.ctypes.data_as(ol.IP), 1, ol.dptr(pos), ol.dptr(vel), ol.dptr(rad),
                        ol.dptr(ep), C.byref(s))
    return vel


def finish(s, vel):
    obst = np.zeros(2, np.int32)
    ol.oracle().or_adaptive_finish(ol.dptr(vel), obst.ctypes.data_as(ol.IP), 2, C.byref(s))


def test_head_on_close_keeps_dt():
    """A at 0 moving +1, B at 1 moving -1, radius 0.3, eps 0.01, dt 0.1:
    pdt = 0.11, pd = 0.78, vBA = -2, pvBA = -1.56; halo radius 0.341, so
    the predicted halo gap per step is 0.098/0.11 < 1.56: close, remaining
    1 - 0.62 = 0.38, proposal 0.3116.  The proposal exceeds 1.1 dt, so dt
    stays 0.1 and the upper bound (0.8 / (2.2 * 2) = 0.1818) is not applied."""
    s = state(0.1)
    vel = pairs_step(s, [1, 0, 0], [-1, 0, 0])
    assert s.close == pytest.approx(0.3116, rel=1e-12)
    assert s.max_vel_approach == 2.0 and s.separate == 0
    finish(s, vel)
    assert s.dt == 0.1 and s.step == 1 and s.time == 0.1
    assert s.max_dt == pytest.approx(0.8 / 4.4, rel=1e-15)
    assert s.max_vel_travel == 1.0


def test_head_on_close_cuts_dt():
    """The same pair at dt 0.5: pdt 0.55, pd = 1 - 1.1 = -0.1 in signed terms,
    |pdAB| = 0.1, predicted halo gap (0.1 - 0.682) / 0.55 < 0.2: close with
    the same proposal 0.3116 <= 1.1 * 0.5, so dt becomes 0.3116."""
    s = state(0.5)
    vel = pairs_step(s, [1, 0, 0], [-1, 0, 0])
    finish(s, vel)
    assert s.dt == pytest.approx(0.3116, rel=1e-12)


def test_separating_pair_grows_dt_after_step_two():
    """Moving apart: vBA = 2 > 0, pvBA > 0, separate; no proposal, so from
    step 3 on dt grows by 1.1, then is capped by mesh / (2.2 vmax) 0.8 with
    vmax the previous travel speed (3 here)."""
    s = state(0.1, step=3, travel=3.0)
    vel = pairs_step(s, [-1, 0, 0], [1, 0, 0])
    assert s.separate == 1 and s.close == 0.0 and s.max_vel_approach == 3.0
    finish(s, vel)
    assert s.dt == pytest.approx(0.11, rel=1e-15)
    s2 = state(0.2, step=3, travel=3.0)
    vel = pairs_step(s2, [-1, 0, 0], [1, 0, 0])
    finish(s2, vel)
    assert s2.dt == pytest.approx(0.8 / 6.6, rel=1e-15)  # 0.22 capped at 0.1212


def test_below_min_dt_jumps_to_min_dt():
    """Speeds +-20 at dt 0.01: pdt 0.011, |pdAB| = 0.56, halo gap -0.122 <
    0, so close; dt < mindt 0.05 gives dt = 0.05, then the upper bound
    mesh / (2.2 * 40) * 0.8 applies: 0.909 with mesh 100, 0.00909 with 1."""
    s = state(0.01, mesh=100.0, min_dt=0.05)
    vel = pairs_step(s, [20, 0, 0], [-20, 0, 0])
    assert s.close == pytest.approx(0.3116, rel=1e-12) and s.max_vel_approach == 40.0
    finish(s, vel)
    assert s.dt == 0.05
    s = state(0.01, mesh=1.0, min_dt=0.05)
    vel = pairs_step(s, [20, 0, 0], [-20, 0, 0])
    finish(s, vel)
    assert s.dt == pytest.approx(0.8 / 88, rel=1e-15)
