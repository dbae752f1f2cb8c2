"""HIP graphs of the fused substep (mhh_graph_*): one model, and an ensemble
of models of different sizes advanced by one launch per substep, against the
same substeps issued call by call.  The graphs hold the same kernels with the
same arguments, so the fields agree bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(32, 24, 16, 5), (48, 40, 24, 6), (64, 32, 20, 7)]
# the member grids of the batched dry-CBL ensemble (FP32), one of each
ENSEMBLE = [(32, 32, 32, 11), (64, 64, 48, 12), (96, 96, 64, 13), (128, 128, 96, 14)]
STEPS = [(0, 0.41), (1, 0.41), (2, 0.41), (0, 0.37), (1, 0.37)]   # (substep, dt)
CB = (1. / 3., 15. / 16., 8. / 15.)


def _desc(substep, dt):
    from microhh_amd import abi
    smag = abi.SmagParams(0.23, 1.0 / 3.0, 1.5e-5, (C.c_double * 8)(*([2e-5] * 8)), 1, 0, 1)
    bc = abi.BcDesc(abi.BC["noslip"], abi.BC["freeslip"], (C.c_int * 8)(*([abi.BC["dirichlet"]] * 8)),
                    (C.c_int * 8)(*([abi.BC["neumann"]] * 8)))
    return abi.SubstepDesc(smag, bc, 1, CB[substep] * dt, substep, dt,
                           abi.MHH_RK3_SCALARS | abi.MHH_RK3_DEFER)


def _case(itot, jtot, ktot, seed, stream, fp32=False):
    from microhh_amd import abi
    from tests.golden.make_golden import make_inputs
    from tests._gpu import GpuCase
    p, i = make_inputs(dict(itot=itot, jtot=jtot, ktot=ktot, stretch=ktot <= 24, anelastic=True, dx=50.0,
                            dy=40.0, seed=seed))
    kc = ktot + 2 * p["kgc"]
    thref = np.full(kc, 300.0) + 0.01 * np.arange(kc)
    g = GpuCase(itot, jtot, ktot, p["igc"], p["jgc"], p["kgc"], p["dx"], p["dy"], i, nscalars=1,
                precision=abi.MHH_FP32 if fp32 else abi.MHH_FP64,
                thref=thref, threfh=thref + 0.005, stream=stream)
    for n, src in (("u", "u"), ("v", "v"), ("w", "w"), ("s0", "s")):
        g.put(n, i[src])
    for n in ("ut", "vt", "wt", "st0"):
        g.put(n, np.zeros_like(i["u"]))
    return g


def _direct(g, d):
    g.call("mhh_boundary_set_prognostic_cyclic_bcs")
    g.call("mhh_boundary_set_ghost_cells", C.byref(d.bc))
    g.call("mhh_diff_smag2_exec_viscosity", C.byref(d.smag))
    g.call("mhh_boundary_set_ghost_cells", C.byref(d.bc))
    g.call("mhh_tendency_2i5_smag2", C.byref(d.smag), d.buoyancy)
    g.call("mhh_pres_2_exec_rk3", d.subdt, d.substep, d.rkdt, d.flags)


def _state(g):
    return {n: g.get(n) for n in ("u", "v", "w", "s0", "p")}


@pytest.fixture
def streams():
    from microhh_amd import abi
    L = abi.lib()
    made = []

    def make():
        s = C.c_void_p()
        abi.check(L.mhh_stream_create(C.byref(s)), "mhh_stream_create")
        made.append(s)
        return s
    yield make
    for s in made:
        L.mhh_stream_destroy(s)


def test_graph_substeps_match_direct_calls(streams):
    from microhh_amd import abi
    L = abi.lib()
    it, jt, kt, seed = SIZES[0]
    a = _case(it, jt, kt, seed, streams())
    b = _case(it, jt, kt, seed, streams())
    graph = C.c_void_p()
    try:
        for n, (sub, dt) in enumerate(STEPS):
            d = _desc(sub, dt)
            _direct(a, d)
            if n == 0:
                abi.check(L.mhh_graph_capture_substep(b.ctx, C.byref(d), C.byref(graph)), "capture")
            else:
                abi.check(L.mhh_graph_update(graph, C.byref(d)), "update")
            abi.check(L.mhh_graph_launch(graph), "launch")
            sa, sb = _state(a), _state(b)
            for name in sa:
                assert np.array_equal(sa[name], sb[name]), f"step {n} {name}"
        # settled tendencies agree too
        a.call("mhh_tendency_settle"); b.call("mhh_tendency_settle")
        for name in ("ut", "vt", "wt", "st0"):
            assert np.array_equal(a.get(name), b.get(name)), name
    finally:
        L.mhh_graph_destroy(graph)
        a.close(); b.close()


@pytest.mark.parametrize("members", ["fp64", "fp32-ensemble-grids"])
def test_graph_ensemble_one_launch_per_substep(streams, members):
    """Models of different sizes, each on its own stream, in one graph: every
    member as its call-by-call run."""
    from microhh_amd import abi
    L = abi.lib()
    sizes, fp32 = (SIZES, False) if members == "fp64" else (ENSEMBLE, True)
    direct = [_case(*sz, streams(), fp32) for sz in sizes]
    batch = [_case(*sz, streams(), fp32) for sz in sizes]
    ctxs = (C.c_void_p * len(batch))(*[g.ctx for g in batch])
    graph = C.c_void_p()
    try:
        for n, (sub, dt) in enumerate(STEPS):
            # each member its own time step
            descs = (abi.SubstepDesc * len(batch))(*[_desc(sub, dt * (1. + 0.1 * m)) for m in range(len(batch))])
            for m, g in enumerate(direct):
                _direct(g, descs[m])
            if n == 0:
                abi.check(L.mhh_graph_capture_substeps(ctxs, descs, len(batch), C.byref(graph)), "capture")
            else:
                abi.check(L.mhh_graph_update(graph, descs), "update")
            abi.check(L.mhh_graph_launch(graph), "launch")
            for m in range(len(batch)):
                sa, sb = _state(direct[m]), _state(batch[m])
                for name in sa:
                    assert np.array_equal(sa[name], sb[name]), f"step {n} member {m} {name}"
    finally:
        L.mhh_graph_destroy(graph)
        for g in direct + batch:
            g.close()


def test_graph_launch_checks_deferred_state(streams):
    """A graph recorded with a deferred update pending cannot be launched once
    the update has been completed (another call settled it)."""
    from microhh_amd import abi
    L = abi.lib()
    it, jt, kt, seed = SIZES[0]
    g = _case(it, jt, kt, seed, streams())
    graph = C.c_void_p()
    try:
        d0 = _desc(0, 0.4)
        _direct(g, d0)                      # leaves a deferred update pending
        d1 = _desc(1, 0.4)
        abi.check(L.mhh_graph_capture_substep(g.ctx, C.byref(d1), C.byref(graph)), "capture")
        g.call("mhh_tendency_settle")       # state changes
        assert L.mhh_graph_launch(graph) == abi.MHH_ESTATE
        abi.check(L.mhh_graph_update(graph, C.byref(d1)), "update")
        abi.check(L.mhh_graph_launch(graph), "launch")
    finally:
        L.mhh_graph_destroy(graph)
        g.close()


def test_graph_needs_own_stream():
    from microhh_amd import abi
    L = abi.lib()
    it, jt, kt, seed = SIZES[0]
    g = _case(it, jt, kt, seed, None)
    graph = C.c_void_p()
    try:
        d = _desc(0, 0.4)
        assert L.mhh_graph_capture_substep(g.ctx, C.byref(d), C.byref(graph)) == abi.MHH_EINVAL
    finally:
        g.close()


@pytest.mark.parametrize("mode", ["streams", "graph"])
@pytest.mark.parametrize("adaptive", [False, True])
def test_ensemble_matches_models_run_alone(adaptive, mode):
    """mhh_ensemble_*: four driver models of different sizes (one of them
    in the mixed mode) advanced together against the same models stepped
    one at a time.  With an
    adaptive time step the first substep of every step runs through the
    model's own substep and the other two as graph launches (graph mode) or
    as mhh_substep_exec on the member's stream (streams mode): bitwise the
    same.  With a fixed time step every substep takes the second route; the
    first substep of a step then runs the tendency pass without the
    time-step maxima (a kernel variant compiled apart, whose sums may
    contract differently): within the 1e-12 parity bar in FP64, and within
    a few FP32 roundings (1e-5) for the mixed member."""
    from microhh_amd import model as mm
    # three FP64 members and one in the mixed mode (FP32 fields, FP64 mean pressure)
    sizes = [(32, 24, 16, mm.MHH_FP64), (48, 40, 24, mm.MHH_FP64), (64, 32, 20, mm.MHH_FP64),
             (40, 32, 16, mm.MHH_MIXED)]
    cases = [mm.dry_cbl_case(itot=i, jtot=j, ktot=k, xsize=100.0 * i, ysize=100.0 * j, zsize=50.0 * k,
                             dt=2.0, adaptive=adaptive, rndseed=3 + n, precision=prec)
             for n, (i, j, k, prec) in enumerate(sizes)]
    streams = mm.Ensemble.streams(len(cases))
    alone = [mm.Model(c) for c in cases]
    members = [mm.Model(c, stream=s) for c, s in zip(cases, streams)]
    ens = mm.Ensemble(members, mode=mode)
    try:
        nsteps = 3
        for m in alone:
            m.steps(nsteps)
        ens.steps(nsteps)
        st = ens.stats()
        if mode == "streams":
            assert st == dict(launches=0, recordings=0, direct=3 * nsteps), st
        elif adaptive:
            assert st == dict(launches=2 * nsteps, recordings=st["recordings"], direct=nsteps), st
        else:
            assert st["launches"] == 3 * nsteps and st["direct"] == 0, st
            assert st["recordings"] <= 4, st      # one per substep index, plus substep 0's state after step 1
        for n, (a, b) in enumerate(zip(alone, members)):
            assert a.time() == b.time(), n
            for name in ("u", "v", "w", "th", "p", "ut", "tht"):
                ra, rb = a.field(name), b.field(name)
                if adaptive:
                    assert np.array_equal(ra, rb), f"member {n} {name}"
                else:
                    # the mixed member's variants differ by FP32 roundings
                    tol = 1e-12 if sizes[n][3] == mm.MHH_FP64 else 1e-5
                    err = np.max(np.abs(ra - rb)) / max(np.max(np.abs(ra)), 1e-300)
                    assert err < tol, f"member {n} {name}: {err:.3e}"
    finally:
        ens.close()
        for m in alone + members:
            m.close()
        mm.Ensemble.destroy_streams(streams)
