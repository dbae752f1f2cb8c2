"""The C++ model driver on the GPU against the oracle driver.

Per-substep tendencies: both drivers start from the same state (the GPU
state copied into the oracle), run one substep with a fixed dt, and their
stored tendencies (cA times the total tendency, or the running sum) agree
to 1e-12 normwise per field.  Short runs: fixed dt, fields after n steps
within the tolerance stated in each test.  Adaptive dt: the dt sequences
are identical.
"""
import numpy as np
import pytest

from microhh_amd import model as mm
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _pair(**kw):
    case = mm.dry_cbl_case(**kw)
    g = mm.Model(case)
    o = po.Model(mm.oracle_case(case), *mm.profiles(case))
    return case, g, o


def _interior(o):
    d = o.d
    return np.s_[d.kstart:d.kend, d.jstart:d.jend, d.istart:d.iend]


def _nw(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / den if den > 0 else np.max(np.abs(a - b))


def test_initial_fields_identical():
    case, g, o = _pair(itot=32, jtot=24, ktot=16, xsize=3200.0, ysize=2400.0, zsize=800.0, adaptive=False, dt=1.0)
    sl = _interior(o)
    for n in ("u", "v", "w", "th"):
        assert np.array_equal(g.field(n)[sl], o.field(n)[sl]), n
    g.close(); o.close()


@pytest.mark.parametrize("advec,diff,fluxlimit", [(0, 0, False), (1, 1, False), (0, 1, False), (1, 0, False),
                                                  (0, 0, True)])
def test_substep_tendencies(advec, diff, fluxlimit):
    case, g, o = _pair(itot=48, jtot=32, ktot=24, xsize=4800.0, ysize=3200.0, zsize=1200.0,
                       adaptive=False, dt=2.0, advec=advec, diff=diff, rndamp_uvw=0.5, rndamp_th=0.5,
                       fluxlimit=fluxlimit)
    assert o.d.kgc == (2 if fluxlimit else 1)
    sl = _interior(o)
    # spin up a few substeps on the GPU so that the state is not the noise field
    for _ in range(4):
        g.substep()
    names = ("u", "v", "w", "th", "ut", "vt", "wt", "tht")
    for n in names:
        o.field(n)[...] = g.field(n)
    # the oracle's time loop must be at the same substep: advance its counter with zero-dt copies
    while o.substep_index != g.time()[2]:
        o.substep()
        for n in names:
            o.field(n)[...] = g.field(n)
    for s in range(3):
        g.substep()
        o.substep()
        for n in ("ut", "vt", "wt", "tht", "u", "v", "w", "th"):
            err = _nw(g.field(n)[sl], o.field(n)[sl])
            tol = 1e-12
            assert err < tol, f"substep {s} {n}: {err:.3e}"
    g.close(); o.close()


def test_fused_matches_unfused():
    kw = dict(itot=64, jtot=48, ktot=40, xsize=6400.0, ysize=4800.0, zsize=2000.0,
              adaptive=False, dt=2.0, rndamp_uvw=0.5, rndamp_th=0.5)
    a = mm.Model(mm.dry_cbl_case(fused=True, **kw))
    b = mm.Model(mm.dry_cbl_case(fused=False, **kw))
    for _ in range(6):
        a.substep(); b.substep()
    for n in ("u", "v", "w", "th", "ut", "vt", "wt", "tht"):
        err = _nw(a.field(n), b.field(n))
        assert err < 1e-13, f"{n}: {err:.3e}"
    a.close(); b.close()


def test_short_run_fixed_dt():
    case, g, o = _pair(itot=64, jtot=64, ktot=32, xsize=6400.0, ysize=6400.0, zsize=1600.0,
                       adaptive=False, dt=2.0, rndamp_uvw=0.5, rndamp_th=0.5)
    sl = _interior(o)
    for step in range(5):
        for _ in range(3):
            g.substep(); o.substep()
    # five RK3 steps of a turbulent state: chaotic growth of rounding stays far below 1e-9
    for n in ("u", "v", "w", "th"):
        err = _nw(g.field(n)[sl], o.field(n)[sl])
        assert err < 1e-9, f"{n}: {err:.3e}"
    assert g.div() < 1e-10
    g.close(); o.close()


def test_adaptive_dt_sequence():
    case, g, o = _pair(itot=32, jtot=32, ktot=16, xsize=3200.0, ysize=3200.0, zsize=800.0,
                       adaptive=True, dt=5.0, dtmax=60.0, rndamp_uvw=1.0, rndamp_th=1.0)
    for step in range(4):
        for _ in range(3):
            g.substep(); o.substep()
        assert g.time()[1] == o.dt, f"step {step}: dt {g.time()[1]} vs {o.dt}"
    g.close(); o.close()
