"""ctypes wrapper of the CPU FP64 oracle (oracle/liborf.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, and only there as the checker or the timed CPU
baseline. The product library never imports this module.

The wrapped C functions restate crates/engine/src/pde.rs and
crates/engine/src/coupling.rs; see oracle/pde_oracle.h for the line-by-line
citations.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborf.so")

OK, ERR_CONFIG, ERR_NOT_SETTLED = 0, 1, 2

_lib = None


def build() -> None:
    """Compile liborf.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        d, u32, sz, i32 = C.c_double, C.c_uint32, C.c_size_t, C.c_int
        dp = C.POINTER(C.c_double)
        u8p = C.POINTER(C.c_uint8)
        u32p = C.POINTER(C.c_uint32)
        szp = C.POINTER(C.c_size_t)
        sig = {
            "orf_create": (i32, [u32, u32, u32, d, d, d, C.POINTER(P)]),
            "orf_destroy": (None, [P]),
            "orf_values": (dp, [P]),
            "orf_next": (dp, [P]),
            "orf_uptake": (dp, [P]),
            "orf_production": (dp, [P]),
            "orf_held": (u8p, [P]),
            "orf_len": (sz, [P]),
            "orf_set_source": (None, [P, d]),
            "orf_set_uptake": (None, [P, u32, u32, d]),
            "orf_set_production": (None, [P, u32, u32, d]),
            "orf_set_uptake_at": (None, [P, sz, d]),
            "orf_set_production_at": (None, [P, sz, d]),
            "orf_clear_uptake": (None, [P]),
            "orf_clear_production": (None, [P]),
            "orf_fill": (None, [P, d]),
            "orf_free_span": (i32, [P, u32p]),
            "orf_reference_omega": (d, [P, u32p]),
            "orf_half_sweep": (d, [P, d, u32, u32p, i32]),
            "orf_sor_sweep": (d, [P, d, u32p, i32]),
            "orf_relax": (i32, [P, d, sz, d, i32, szp]),
            "orf_inflow": (d, [P]),
            "orf_free_count": (sz, [P]),
            "orf_release_all": (None, [P]),
            "orf_hold_where": (None, [P, d, u8p]),
            "orf_hold_bath_at": (None, [P, d, u8p]),
            "orf_step": (i32, [P, d]),
            "orf_apply_uptake": (None, [P, u32p, u8p, sz, d]),
            "orf_apply_uptake_per_cell": (None, [P, u32p, u8p, dp, sz]),
            "orf_apply_background_uptake": (None, [P, u32p, d]),
            "orf_apply_production": (None, [P, u32p, dp, sz]),
            "orf_mean_per_cell": (None, [P, u32p, sz, dp]),
            "orf_rel_residual": (d, [P, dp, dp]),
            "orf_residual_vector": (None, [P, dp]),
            "orf_relax_residual": (i32, [P, d, sz, d, i32, szp]),
            "orf_adjoint": (i32, [P, dp, d, d, sz, u32p, sz, dp, dp, dp, dp, szp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class RelaxError(RuntimeError):
    """The oracle's relax ran out of sweeps (CellweaveError::Simulation)."""


class ConfigError(ValueError):
    """The oracle refused a configuration (CellweaveError::Config)."""


class OracleField:
    """One ScalarField held by the C oracle (pde.rs:30-53)."""

    def __init__(self, w, h, d=1, diff=0.2, decay=0.0, initial=0.0):
        L = lib()
        self._h = C.c_void_p()
        rc = L.orf_create(w, h, d, diff, decay, initial, C.byref(self._h))
        if rc != OK:
            raise ConfigError(f"field {w}x{h}x{d} D={diff} decay={decay} refused")
        self.w, self.h, self.d = w, h, d
        self.n = w * h * d
        self.diff, self.decay = diff, decay

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().orf_destroy(h)
            self._h = None

    # Views onto the C arrays, valid until the next sweep swaps values and
    # next. Each view holds a reference to this field, so that an expression
    # such as `make_field().values.copy()` cannot free the memory under it.
    def _view(self, ptr, ctype) -> np.ndarray:
        buf = (ctype * self.n).from_address(C.cast(ptr, C.c_void_p).value)
        buf._owner = self
        return np.ctypeslib.as_array(buf)

    @property
    def values(self) -> np.ndarray:
        return self._view(lib().orf_values(self._h), C.c_double)

    @property
    def held(self) -> np.ndarray:
        return self._view(lib().orf_held(self._h), C.c_uint8)

    @property
    def uptake(self) -> np.ndarray:
        return self._view(lib().orf_uptake(self._h), C.c_double)

    @property
    def production(self) -> np.ndarray:
        return self._view(lib().orf_production(self._h), C.c_double)

    def set_source(self, s):
        lib().orf_set_source(self._h, s)

    def set_uptake(self, x, y, r):
        lib().orf_set_uptake(self._h, x, y, r)

    def set_production(self, x, y, r):
        lib().orf_set_production(self._h, x, y, r)

    def set_uptake_at(self, i, r):
        lib().orf_set_uptake_at(self._h, i, r)

    def set_production_at(self, i, r):
        lib().orf_set_production_at(self._h, i, r)

    def clear_uptake(self):
        lib().orf_clear_uptake(self._h)

    def clear_production(self):
        lib().orf_clear_production(self._h)

    def fill(self, v):
        lib().orf_fill(self._h, v)

    def concentration_at(self, x, y):
        if x >= self.w or y >= self.h:
            return 0.0
        return float(self.values[y * self.w + x])

    def value_at(self, i):
        return float(self.values[i]) if i < self.n else 0.0

    def free_span(self):
        s = (C.c_uint32 * 6)()
        if not lib().orf_free_span(self._h, s):
            return None
        return tuple(int(v) for v in s)

    def reference_omega(self):
        span = self.free_span()
        if span is None:
            return float("nan")
        s = (C.c_uint32 * 6)(*span)
        return lib().orf_reference_omega(self._h, s)

    def relax(self, tolerance, max_sweeps, omega=float("nan"), threads=1):
        done = C.c_size_t()
        rc = lib().orf_relax(self._h, tolerance, max_sweeps, omega, threads, C.byref(done))
        if rc == ERR_NOT_SETTLED:
            raise RelaxError(
                f"diffusion did not settle within {max_sweeps} sweeps at tolerance {tolerance}"
            )
        return int(done.value)

    def relax_residual(self, rel_tol, max_sweeps, omega=float("nan"), threads=1):
        done = C.c_size_t()
        rc = lib().orf_relax_residual(self._h, rel_tol, max_sweeps, omega, threads, C.byref(done))
        if rc == ERR_NOT_SETTLED:
            raise RelaxError(f"residual {rel_tol} not reached in {max_sweeps} sweeps")
        return int(done.value)

    def sor_sweep(self, omega, threads=1, span=None):
        """One red-black sweep over the free span (pass `span`, from
        free_span(), to skip finding it again when the holds do not change)."""
        span = self.free_span() if span is None else span
        s = (C.c_uint32 * 6)(*span)
        return lib().orf_sor_sweep(self._h, omega, s, threads)

    def copy_next_from_values(self):
        nxt = self._view(lib().orf_next(self._h), C.c_double)
        nxt[:] = self.values

    def inflow(self):
        return lib().orf_inflow(self._h)

    def free_count(self):
        return int(lib().orf_free_count(self._h))

    def release_all(self):
        lib().orf_release_all(self._h)

    def hold_where(self, value, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        assert m.size == self.n
        lib().orf_hold_where(self._h, value, _ptr(m, C.c_uint8))

    def hold_bath_at(self, value, is_open):
        m = np.ascontiguousarray(is_open, dtype=np.uint8).reshape(-1)
        assert m.size == self.n
        lib().orf_hold_bath_at(self._h, value, _ptr(m, C.c_uint8))

    def hold_bath(self, value, is_open_xy):
        """pde.rs:535-538: is_open asked by (x, y) on a flat field."""
        m = np.zeros(self.n, dtype=np.uint8)
        for i in range(self.n):
            m[i] = bool(is_open_xy(i % self.w, (i // self.w) % self.h))
        self.hold_bath_at(value, m)

    def hold_pixels(self, value, is_source_xy):
        m = np.zeros(self.n, dtype=np.uint8)
        for y in range(self.h):
            for x in range(self.w):
                m[y * self.w + x] = bool(is_source_xy(x, y))
        self.hold_where(value, m)

    def step(self, dt):
        rc = lib().orf_step(self._h, dt)
        if rc != OK:
            raise ConfigError(f"time step {dt} refused")

    def apply_uptake(self, grid, consumes, rate):
        g = np.ascontiguousarray(grid, dtype=np.uint32).reshape(-1)
        c = np.ascontiguousarray(consumes, dtype=np.uint8)
        lib().orf_apply_uptake(self._h, _ptr(g, C.c_uint32), _ptr(c, C.c_uint8), c.size, rate)

    def apply_uptake_per_cell(self, grid, consumes, per_cell):
        g = np.ascontiguousarray(grid, dtype=np.uint32).reshape(-1)
        c = np.ascontiguousarray(consumes, dtype=np.uint8)
        p = np.ascontiguousarray(per_cell, dtype=np.float64)
        assert c.size == p.size
        lib().orf_apply_uptake_per_cell(
            self._h, _ptr(g, C.c_uint32), _ptr(c, C.c_uint8), _ptr(p, C.c_double), p.size
        )

    def apply_background_uptake(self, grid, rate):
        g = np.ascontiguousarray(grid, dtype=np.uint32).reshape(-1)
        lib().orf_apply_background_uptake(self._h, _ptr(g, C.c_uint32), rate)

    def apply_production(self, grid, per_cell):
        g = np.ascontiguousarray(grid, dtype=np.uint32).reshape(-1)
        p = np.ascontiguousarray(per_cell, dtype=np.float64)
        lib().orf_apply_production(self._h, _ptr(g, C.c_uint32), _ptr(p, C.c_double), p.size)

    def mean_per_cell(self, grid, ncells):
        g = np.ascontiguousarray(grid, dtype=np.uint32).reshape(-1)
        out = np.zeros(ncells, dtype=np.float64)
        lib().orf_mean_per_cell(self._h, _ptr(g, C.c_uint32), ncells, _ptr(out, C.c_double))
        return out

    def rel_residual(self):
        r2, b2 = C.c_double(), C.c_double()
        rel = lib().orf_rel_residual(self._h, C.byref(r2), C.byref(b2))
        return rel, r2.value, b2.value

    def residual_vector(self):
        out = np.zeros(self.n, dtype=np.float64)
        lib().orf_residual_vector(self._h, _ptr(out, C.c_double))
        return out

    def adjoint(self, target, omega=float("nan"), rel_tol=1e-13, max_sweeps=1_000_000,
                grid=None, ncells=0):
        t = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        lam = np.zeros(self.n, dtype=np.float64)
        gd, gg = C.c_double(), C.c_double()
        du = np.zeros(max(ncells, 1), dtype=np.float64)
        sweeps = C.c_size_t()
        gp = None
        if grid is not None:
            g = np.ascontiguousarray(grid, dtype=np.uint32).reshape(-1)
            gp = _ptr(g, C.c_uint32)
        rc = lib().orf_adjoint(
            self._h, _ptr(t, C.c_double), omega, rel_tol, max_sweeps, gp, ncells,
            _ptr(lam, C.c_double), C.byref(gd), C.byref(gg), _ptr(du, C.c_double),
            C.byref(sweeps),
        )
        if rc != OK:
            raise RelaxError("adjoint did not settle")
        return {"lambda": lam, "dD": gd.value, "dg": gg.value, "du": du[:ncells],
                "sweeps": int(sweeps.value)}


def optimal_omega(longest: int) -> float:
    """2 / (1 + sin(pi / L)): the uncapped rule pde.rs:267-273 uses in a volume."""
    return 2.0 / (1.0 + math.sin(math.pi / max(longest, 2)))
