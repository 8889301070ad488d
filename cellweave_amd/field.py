"""Python handle on the GPU field solve (thin layer over the C ABI).

Field mirrors cellweave's ScalarField (crates/engine/src/pde.rs:30-53) with
its state on the device: bulk upload, or the gathers from a Lattice (the
cell-id lattice, crates/engine/src/cpm.rs:198) as crates/engine/src/coupling.rs
does them; relax() and step() run on the GPU, download() brings the values
back.
Errors map onto the reference's two variants: ConfigError for
CellweaveError::Config and SimulationError for CellweaveError::Simulation
(crates/core/src/error.rs:7-14).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import (CRIT_MAX_CHANGE, CRIT_REL_RESIDUAL, PREC_FP32_REFINE, PREC_FP64,
                   RelaxOpts, RelaxStats)


class CellweaveError(RuntimeError):
    pass


class ConfigError(CellweaveError):
    """CellweaveError::Config."""


class SimulationError(CellweaveError):
    """CellweaveError::Simulation: the field did not settle."""


class DeviceError(CellweaveError):
    pass


def _check(rc: int):
    if rc == _abi.CW_OK:
        return
    msg = _abi.last_error()
    if rc == _abi.CW_ERR_CONFIG:
        raise ConfigError(msg)
    if rc == _abi.CW_ERR_NOT_CONVERGED:
        raise SimulationError(msg)
    if rc == _abi.CW_ERR_NONFINITE:
        raise SimulationError(msg)
    if rc == _abi.CW_ERR_ARG:
        raise ValueError(msg)
    raise DeviceError(msg)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_opts(omega=0.0, criterion=CRIT_MAX_CHANGE, tolerance=1e-8, max_sweeps=2_000_000,
              precision=PREC_FP64, sweeps_per_pass=0, check_every=0, inner_tolerance=0.0,
              max_outer=0, tile_shape=-1, use_graph=0) -> RelaxOpts:
    o = RelaxOpts()
    _abi.lib().cw_relax_opts_default(C.byref(o))
    o.omega = omega
    o.criterion = criterion
    o.tolerance = tolerance
    o.max_sweeps = max_sweeps
    o.precision = precision
    o.sweeps_per_pass = sweeps_per_pass
    o.check_every = check_every
    o.inner_tolerance = inner_tolerance
    o.max_outer = max_outer
    o.tile_shape = tile_shape
    o.use_graph = use_graph
    return o


def _u8(a, n):
    a = np.ascontiguousarray(a, dtype=np.uint8).reshape(-1)
    assert a.size == n, (a.size, n)
    return a


class Lattice:
    """Device copy of CpmLattice::grid (cpm.rs:198): u32 cell ids, 0 the
    medium, 0xFFFFFFFF an obstacle (cpm.rs:137)."""

    def __init__(self, width, height, depth=1):
        self._h = C.c_void_p()
        _check(_abi.lib().cw_lattice_create(width, height, depth, C.byref(self._h)))
        self.w, self.h, self.d = width, height, depth
        self.n = width * height * depth

    def upload(self, ids):
        a = np.ascontiguousarray(ids, dtype=np.uint32).reshape(-1)
        assert a.size == self.n
        _check(_abi.lib().cw_lattice_upload(self._h, a.ctypes.data_as(C.POINTER(C.c_uint32))))

    def set(self, voxels, ids):
        """Set the listed voxels (flat indices) to the listed ids, in order:
        a lattice step's changes without re-uploading the grid."""
        v = np.ascontiguousarray(voxels, dtype=np.uint64).reshape(-1)
        i = np.ascontiguousarray(ids, dtype=np.uint32).reshape(-1)
        assert v.size == i.size
        _check(_abi.lib().cw_lattice_set(self._h, v.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         i.ctypes.data_as(C.POINTER(C.c_uint32)), v.size))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _abi.lib().cw_lattice_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field:
    def __init__(self, width, height, depth=1, diff=0.2, decay=0.0, initial=0.0):
        L = _abi.lib()
        self._h = C.c_void_p()
        _check(L.cw_field_create(width, height, depth, diff, decay, initial, C.byref(self._h)))
        self.w, self.h, self.d = width, height, depth
        self.n = width * height * depth

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _abi.lib().cw_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_source(self, source: float):
        _check(_abi.lib().cw_field_set_source(self._h, source))

    def upload(self, values=None, held=None, uptake=None, production=None):
        keep = []

        def dbl(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            assert a.size == self.n, (a.size, self.n)
            keep.append(a)
            return _dptr(a)

        hp = None
        if held is not None:
            hb = np.ascontiguousarray(held, dtype=np.uint8).reshape(-1)
            assert hb.size == self.n
            keep.append(hb)
            hp = hb.ctypes.data_as(C.POINTER(C.c_uint8))
        _check(_abi.lib().cw_field_upload(self._h, dbl(values), hp, dbl(uptake), dbl(production)))

    def fill(self, value: float):
        """ScalarField::fill (pde.rs:212-215), on the device."""
        _check(_abi.lib().cw_field_fill(self._h, value))

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        _check(_abi.lib().cw_field_download(self._h, _dptr(out)))
        return out

    # ---- state on the device: gathers, holds, step (coupling.rs, pde.rs) ----
    def apply_uptake(self, lattice: Lattice, consumes, rate: float):
        """coupling.rs:26-39; consumes[id] != 0 for a living cell kind."""
        c = _u8(consumes, len(consumes))
        _check(_abi.lib().cw_field_apply_uptake(self._h, lattice._h,
                                                c.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                c.size, rate))

    def apply_uptake_per_cell(self, lattice: Lattice, rates, consumes=None):
        r = np.ascontiguousarray(rates, dtype=np.float64).reshape(-1)
        cp = None
        if consumes is not None:
            c = _u8(consumes, r.size)
            cp = c.ctypes.data_as(C.POINTER(C.c_uint8))
        _check(_abi.lib().cw_field_apply_uptake_per_cell(self._h, lattice._h, cp, _dptr(r),
                                                         r.size))

    def apply_background_uptake(self, lattice: Lattice, rate: float):
        """coupling.rs:47-53."""
        _check(_abi.lib().cw_field_apply_background_uptake(self._h, lattice._h, rate))

    def apply_production(self, lattice: Lattice, per_cell):
        """coupling.rs:61-73."""
        r = np.ascontiguousarray(per_cell, dtype=np.float64).reshape(-1)
        _check(_abi.lib().cw_field_apply_production(self._h, lattice._h, _dptr(r), r.size))

    def hold_obstacles_at(self, lattice: Lattice, value: float):
        """coupling.rs:80-82."""
        _check(_abi.lib().cw_field_hold_obstacles_at(self._h, lattice._h, value))

    def hold_where(self, mask, value: float):
        """pde.rs:491-498 with the predicate as a mask."""
        m = _u8(mask, self.n)
        _check(_abi.lib().cw_field_hold_where(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8)),
                                              value))

    def hold_bath_at(self, open_mask, value: float):
        """pde.rs:550-631 with is_open as a mask."""
        m = _u8(open_mask, self.n)
        _check(_abi.lib().cw_field_hold_bath_at(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                value))

    def hold_medium_bath(self, lattice: Lattice, value: float):
        """coupling.rs:156-158."""
        _check(_abi.lib().cw_field_hold_medium_bath(self._h, lattice._h, value))

    def release_all(self):
        """pde.rs:468-470."""
        _check(_abi.lib().cw_field_release_all(self._h))

    def free_count(self) -> int:
        """pde.rs:460-463."""
        v = C.c_uint64()
        _check(_abi.lib().cw_field_free_count(self._h, C.byref(v)))
        return int(v.value)

    def inflow(self) -> float:
        """pde.rs:415-456 (fixed-tree sum)."""
        v = C.c_double()
        _check(_abi.lib().cw_field_inflow(self._h, C.byref(v)))
        return v.value

    def step(self, dt: float):
        """DiffusionField::step (pde.rs:705-728)."""
        _check(_abi.lib().cw_field_step(self._h, dt))

    def mean_per_cell(self, lattice: Lattice, ncells: int) -> np.ndarray:
        """coupling.rs:164-182."""
        out = np.zeros(max(int(ncells), 1), dtype=np.float64)
        _check(_abi.lib().cw_field_mean_per_cell(self._h, lattice._h, int(ncells), _dptr(out)))
        return out[:ncells]

    def state(self):
        """(held, uptake, production) as host arrays."""
        held = np.empty(self.n, dtype=np.uint8)
        up = np.empty(self.n, dtype=np.float64)
        pr = np.empty(self.n, dtype=np.float64)
        _check(_abi.lib().cw_field_download_state(self._h, held.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                  _dptr(up), _dptr(pr)))
        return held, up, pr

    def relax(self, tolerance: float, max_sweeps: int) -> int:
        """ScalarField::relax (pde.rs:245-288): returns the sweep count."""
        n = C.c_uint64()
        _check(_abi.lib().cw_field_relax(self._h, tolerance, max_sweeps, C.byref(n)))
        return int(n.value)

    def relax_ex(self, **kw) -> dict:
        o = make_opts(**kw)
        s = RelaxStats()
        _check(_abi.lib().cw_field_relax_ex(self._h, C.byref(o), C.byref(s)))
        return s.as_dict()

    def adjoint(self, target, grid=None, ncells=0, tolerance=1e-12, omega=0.0,
                max_sweeps=2_000_000) -> dict:
        """Gradients of 0.5*sum((c - target)^2) with respect to D, the held
        value and the per-cell uptake rates (cw_field_adjoint)."""
        t = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        assert t.size == self.n
        gp = None
        if grid is not None:
            g = np.ascontiguousarray(grid, dtype=np.uint32).reshape(-1)
            assert g.size == self.n
            gp = g.ctypes.data_as(C.POINTER(C.c_uint32))
        du = np.zeros(max(int(ncells), 1), dtype=np.float64)
        o = make_opts(omega=omega, criterion=CRIT_REL_RESIDUAL, tolerance=tolerance,
                      max_sweeps=max_sweeps)
        r = _abi.AdjointResult()
        _check(_abi.lib().cw_field_adjoint(self._h, _dptr(t), gp, int(ncells), C.byref(o),
                                           _dptr(du), C.byref(r)))
        out = r.as_dict()
        out["du"] = du[:ncells]
        return out

    def residual(self) -> float:
        r = C.c_double()
        _check(_abi.lib().cw_field_residual(self._h, C.byref(r)))
        return r.value


def batch_relax(fields, check=True, **kw):
    """Solve every field; returns per-field stats. Two to four fields run as
    concurrent solves on their own streams, more in one launch sequence.
    Under the max-change rule each field gets the bits it gets alone; under
    the relative-residual rule and in FP32 refinement, the bits it gets alone
    at the same tile shape (the default shape counts the whole batch's tiles)."""
    o = make_opts(**kw)
    n = len(fields)
    arr = (C.c_void_p * n)(*[f._h.value for f in fields])
    stats = (RelaxStats * n)()
    rc = _abi.lib().cw_batch_relax(arr, n, C.byref(o), stats)
    if check:
        _check(rc)
    return [s.as_dict() for s in stats], rc


def campaign_relax(fields, groups, shapes=None, check=True, **kw):
    """The device-driven batch of planes (cw_campaign_relax): field i belongs
    to group groups[i], each group swept with its own tile shape shapes[g]
    (-1: the library's choice). Returns (per-field stats, status)."""
    o = make_opts(**kw)
    n = len(fields)
    groups = [int(g) for g in groups]
    assert len(groups) == n
    ng = max(groups) + 1
    shapes = [-1] * ng if shapes is None else [int(x) for x in shapes]
    assert len(shapes) == ng
    arr = (C.c_void_p * n)(*[f._h.value for f in fields])
    garr = (C.c_uint32 * n)(*groups)
    sarr = (C.c_int32 * ng)(*shapes)
    stats = (RelaxStats * n)()
    rc = _abi.lib().cw_campaign_relax(arr, n, garr, sarr, ng, C.byref(o), stats)
    if check:
        _check(rc)
    return [s.as_dict() for s in stats], rc


def time_passes(fields, passes, **kw) -> float:
    """Average milliseconds of one sweep-kernel launch as cw_batch_relax
    would run it (HIP events on the launching stream). With two to four
    fields, their launches run concurrently and this is the longest
    per-field average interval."""
    o = make_opts(**kw)
    n = len(fields)
    arr = (C.c_void_p * n)(*[f._h.value for f in fields])
    ms = C.c_double()
    _check(_abi.lib().cw_batch_time_passes(arr, n, C.byref(o), passes, C.byref(ms)))
    return ms.value


def autotune(fields, launches=10, **kw):
    """Rank every sweep-kernel shape for these fields: [(shape, us per sweep), ...]."""
    o = make_opts(**kw)
    n = len(fields)
    arr = (C.c_void_p * n)(*[f._h.value for f in fields])
    cap = 64
    shapes = (C.c_int32 * cap)()
    us = (C.c_double * cap)()
    got = C.c_uint32()
    _check(_abi.lib().cw_autotune(arr, n, C.byref(o), launches, shapes, us, cap, C.byref(got)))
    return [(int(shapes[i]), float(us[i])) for i in range(got.value)]


__all__ = ["autotune", "time_passes", "Field", "batch_relax", "campaign_relax", "make_opts", "CellweaveError", "ConfigError",
           "SimulationError", "DeviceError", "CRIT_MAX_CHANGE", "CRIT_REL_RESIDUAL",
           "PREC_FP64", "PREC_FP32_REFINE"]
