"""GPU test plumbing: a C-ABI context whose fields live in device buffers
allocated through the library (mhh_device_alloc).

Only the GPU tests import this module.  Every computation goes through
libmicrohh_amd.so.
"""
import ctypes as C

import numpy as np

from microhh_amd import abi

FIELDS3 = ("u", "v", "w", "ut", "vt", "wt", "p", "evisc", "tmp1", "tmp2")
PLANES = ("u_bot", "u_top", "u_gradbot", "u_gradtop", "u_fluxbot", "u_fluxtop",
          "v_bot", "v_top", "v_gradbot", "v_gradtop", "v_fluxbot", "v_fluxtop",
          "dudz", "dvdz", "dbdz", "z0m")
SPLANES = ("s_bot", "s_top", "s_gradbot", "s_gradtop", "s_fluxbot", "s_fluxtop")


def dptr(a):
    return a.ctypes.data_as(abi.DP)


class GpuCase:
    """One context with all fields allocated (zero) as float64/float32 tensors."""

    def __init__(self, itot, jtot, ktot, igc, jgc, kgc, dx, dy, grid, nscalars=1,
                 precision=abi.MHH_FP64, thref=None, threfh=None, stream=None):
        self.dtype = "float64" if precision == abi.MHH_FP64 else "float32"
        self.icells, self.jcells, self.kcells = itot + 2 * igc, jtot + 2 * jgc, ktot + 2 * kgc
        self.shape3 = (self.kcells, self.jcells, self.icells)
        self.shape2 = (self.jcells, self.icells)
        L = abi.lib()
        self._grid = {k: np.ascontiguousarray(grid[k], dtype=np.float64)
                      for k in ("z", "zh", "dz", "dzh", "dzi", "dzhi", "rhoref", "rhorefh")}
        gd = abi.GridDesc(itot, jtot, ktot, igc, jgc, kgc, dx, dy, float(grid["zh"][kgc + ktot]),
                          *[dptr(self._grid[k]) for k in ("z", "zh", "dz", "dzh", "dzi", "dzhi", "rhoref", "rhorefh")])
        h = C.c_void_p()
        abi.check(L.mhh_ctx_create(C.byref(gd), precision, stream, C.byref(h)), "mhh_ctx_create")
        self.ctx = h
        self.t = {}
        for n in FIELDS3:
            self.t[n] = abi.DeviceArray(self.shape3, self.dtype)
        for n in PLANES:
            self.t[n] = abi.DeviceArray(self.shape2, self.dtype)
        self.nscalars = nscalars
        for q in range(nscalars):
            self.t[f"s{q}"] = abi.DeviceArray(self.shape3, self.dtype)
            self.t[f"st{q}"] = abi.DeviceArray(self.shape3, self.dtype)
            for n in SPLANES:
                self.t[f"{n}{q}"] = abi.DeviceArray(self.shape2, self.dtype)
        fs = abi.FieldSet()
        for n in FIELDS3 + PLANES:
            setattr(fs, n, self.t[n].ptr)
        fs.nscalars = nscalars
        for q in range(nscalars):
            fs.s[q] = self.t[f"s{q}"].ptr
            fs.st[q] = self.t[f"st{q}"].ptr
            for n in SPLANES:
                getattr(fs, n)[q] = self.t[f"{n}{q}"].ptr
        self.fs = fs
        abi.check(L.mhh_fields_bind(self.ctx, C.byref(fs)), "mhh_fields_bind")
        if thref is not None:
            self._thref = np.ascontiguousarray(thref, dtype=np.float64)
            self._threfh = np.ascontiguousarray(threfh, dtype=np.float64)
            abi.check(L.mhh_set_base_state(self.ctx, dptr(self._thref), dptr(self._threfh)), "mhh_set_base_state")

    def put(self, name, arr):
        abi.check(abi.lib().mhh_sync(self.ctx), "mhh_sync")
        self.t[name].upload(arr)

    def get(self, name):
        abi.check(abi.lib().mhh_sync(self.ctx), "mhh_sync")
        return self.t[name].download().astype(np.float64)

    def ptr(self, name):
        return self.t[name].ptr

    def call(self, fname, *args):
        abi.check(getattr(abi.lib(), fname)(self.ctx, *args), fname)

    def close(self):
        if self.ctx:
            abi.lib().mhh_ctx_destroy(self.ctx)
            self.ctx = None
            for a in self.t.values():
                a.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
