"""The oracle against the golden vectors written by the reference's own
translation units (tests/golden/make_golden.py, oracle/_ref).

CPU only.  Tolerance 1e-13 normwise: both sides are the same FP64 expression
trees compiled with -ffp-contract=off; differences are rounding only.
"""
import copy

import numpy as np
import pytest

from oracle import pyoracle as po
from tests._golden import GRIDS, load, normwise

TOL = 1e-13


def _dims(p):
    return po.dims(p["itot"], p["jtot"], p["ktot"], p["igc"], p["jgc"], p["kgc"],
                   p["dx"] * p["itot"], p["dy"] * p["jtot"], 1.0)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64).copy()


@pytest.fixture(params=GRIDS)
def case(request):
    p, i, o = load(request.param)
    return _dims(p), p, i, o


def test_dims_match_fixture(case):
    d, p, i, o = case
    assert d.ncells == i["u"].size
    assert abs(d.dx - p["dx"]) < 1e-12 and abs(d.dy - p["dy"]) < 1e-12


@pytest.mark.parametrize("comp", ["u", "v", "w", "s"])
@pytest.mark.parametrize("scheme", ["2i5", "2"])
def test_advec(case, comp, scheme):
    d, p, i, o = case
    L = po.lib()
    tname = {"u": "ut", "v": "vt", "w": "wt", "s": "st"}[comp]
    t = _c(i[tname])
    fn = getattr(L, f"mo_advec_{scheme}_{comp}")
    dz = i["dzhi"] if comp == "w" else i["dzi"]
    if comp == "s":
        fn(d, po.ptr(t), po.ptr(_c(i["s"])), po.ptr(_c(i["u"])), po.ptr(_c(i["v"])), po.ptr(_c(i["w"])),
           po.ptr(_c(dz)), po.ptr(_c(i["rhoref"])), po.ptr(_c(i["rhorefh"])))
    else:
        fn(d, po.ptr(t), po.ptr(_c(i["u"])), po.ptr(_c(i["v"])), po.ptr(_c(i["w"])),
           po.ptr(_c(dz)), po.ptr(_c(i["rhoref"])), po.ptr(_c(i["rhorefh"])))
    ref = o[f"advec_{scheme}_{comp}_{tname}"]
    assert normwise(t, ref) < TOL
    # ghost cells untouched
    assert np.array_equal(t == i[tname], ref == i[tname])


def test_advec_2i5_s_lim(case):
    d, p, i, o = case
    t = _c(i["st"])
    po.lib().mo_advec_2i5_s_lim(d, po.ptr(t), po.ptr(_c(i["s"])), po.ptr(_c(i["u"])), po.ptr(_c(i["v"])),
                                po.ptr(_c(i["w"])), po.ptr(_c(i["dzi"])), po.ptr(_c(i["rhoref"])),
                                po.ptr(_c(i["rhorefh"])))
    ref = o["advec_2i5_s_lim_st"]
    assert normwise(t, ref) < TOL
    assert np.array_equal(t == i["st"], ref == i["st"])
    # the limiter differs from the unlimited scheme on these rough fields
    assert normwise(ref, o["advec_2i5_s_st"]) > 1e-3


@pytest.mark.parametrize("scheme", ["2i5", "2"])
def test_cfl(case, scheme):
    d, p, i, o = case
    fn = getattr(po.lib(), f"mo_advec_{scheme}_cfl")
    cfl = fn(d, po.ptr(_c(i["u"])), po.ptr(_c(i["v"])), po.ptr(_c(i["w"])), po.ptr(_c(i["dzi"])), p["dt"])
    ref = o[f"advec_{scheme}_cfl_cfl"].item()
    assert abs(cfl - ref) <= TOL * abs(ref)


@pytest.mark.parametrize("sfc", [0, 1])
def test_strain2(case, sfc):
    d, p, i, o = case
    s2 = _c(i["strain2"])
    po.lib().mo_calc_strain2(d, sfc, po.ptr(s2), po.ptr(_c(i["u"])), po.ptr(_c(i["v"])), po.ptr(_c(i["w"])),
                             po.ptr(_c(i["ugradbot"])), po.ptr(_c(i["vgradbot"])),
                             po.ptr(_c(i["z"])), po.ptr(_c(i["dzi"])), po.ptr(_c(i["dzhi"])))
    ref = o["strain2_sfc_strain2" if sfc else "strain2_nosfc_strain2"]
    assert normwise(s2, ref) < TOL


@pytest.mark.parametrize("sfc", [0, 1])
@pytest.mark.parametrize("comp", ["u", "v"])
def test_diff_uv(case, comp, sfc):
    d, p, i, o = case
    at = _c(i["at"])
    fn = getattr(po.lib(), f"mo_diff_{comp}")
    fn(d, sfc, po.ptr(at), po.ptr(_c(i["u"])), po.ptr(_c(i["v"])), po.ptr(_c(i["w"])),
       po.ptr(_c(i["dzi"])), po.ptr(_c(i["dzhi"])), po.ptr(_c(i["evisc"])),
       po.ptr(_c(i["fluxbot"])), po.ptr(_c(i["fluxtop"])),
       po.ptr(_c(i["rhoref"])), po.ptr(_c(i["rhorefh"])), p["visc"])
    ref = o[f"diff_{comp}_{'sfc' if sfc else 'nosfc'}_at"]
    assert normwise(at, ref) < TOL


def test_diff_w(case):
    d, p, i, o = case
    at = _c(i["at"])
    po.lib().mo_diff_w(d, po.ptr(at), po.ptr(_c(i["u"])), po.ptr(_c(i["v"])), po.ptr(_c(i["w"])),
                       po.ptr(_c(i["dzi"])), po.ptr(_c(i["dzhi"])), po.ptr(_c(i["evisc"])),
                       po.ptr(_c(i["rhoref"])), po.ptr(_c(i["rhorefh"])), p["visc"])
    assert normwise(at, o["diff_w_at"]) < TOL


@pytest.mark.parametrize("sfc", [0, 1])
def test_diff_c(case, sfc):
    d, p, i, o = case
    at = _c(i["at"])
    po.lib().mo_diff_c(d, sfc, po.ptr(at), po.ptr(_c(i["a"])), po.ptr(_c(i["dzi"])), po.ptr(_c(i["dzhi"])),
                       po.ptr(_c(i["evisc"])), po.ptr(_c(i["fluxbot"])), po.ptr(_c(i["fluxtop"])),
                       po.ptr(_c(i["rhoref"])), po.ptr(_c(i["rhorefh"])), p["tPr"], p["visc"])
    assert normwise(at, o[f"diff_c_{'sfc' if sfc else 'nosfc'}_at"]) < TOL


def test_dnmul(case):
    d, p, i, o = case
    v = po.lib().mo_calc_dnmul(d, po.ptr(_c(i["evisc"])), po.ptr(_c(i["dzi"])), p["tPr"])
    ref = o["dnmul_dnmul"].item()
    assert abs(v - ref) <= TOL * ref


@pytest.mark.parametrize("comp", ["c", "w"])
def test_diff_2(case, comp):
    d, p, i, o = case
    at = _c(i["at"])
    getattr(po.lib(), f"mo_diff_2_{comp}")(d, po.ptr(at), po.ptr(_c(i["a"])), p["visc"],
                                           po.ptr(_c(i["dzi"])), po.ptr(_c(i["dzhi"])))
    assert normwise(at, o[f"diff_2_{comp}_at"]) < TOL


@pytest.mark.parametrize("substep", [0, 1, 2])
def test_rk3(case, substep):
    d, p, i, o = case
    a, at = _c(i["a"]), _c(i["at"])
    po.lib().mo_rk3(d, po.ptr(a), po.ptr(at), substep, p["dt"])
    assert np.array_equal(a, o[f"rk3{substep}_a"])
    assert np.array_equal(at, o[f"rk3{substep}_at"])
