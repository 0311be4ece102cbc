"""Python handle over the C ABI (the HIP engine) — also reused, with another
library and prefix, by the test harness to drive the CPU oracle through the
same calls.

Loading is strict: if the in-tree library is missing, `load_library` raises;
there is no fallback to any other implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import abi
from .case import Spec

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "liblbflow_hip.so")
_lib = None


class LbfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lbflow_hip error {code}: {msg}")
        self.code = code


def library_path() -> str:
    return _LIB_PATH


def load_library(path: Optional[str] = None):
    """Loads liblbflow_hip.so (built by lbflow_amd.build) and declares its ABI."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or _LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP engine library not found at {p}; build it with `python -m lbflow_amd.build`")
    lib = C.CDLL(p)
    abi.declare(lib, abi.ENTRY_POINTS)
    abi.declare(lib, abi.SERVER_ENTRY_POINTS)
    abi.declare(lib, abi.TUNING_ENTRY_POINTS)
    if path is None:
        _lib = lib
    return lib


class Case:
    """One lattice case resident on the device (or in the oracle)."""

    def __init__(self, spec: Spec, lib=None, prefix: str = "lbf_"):
        self.spec = spec
        self.lib = lib if lib is not None else load_library()
        self.prefix = prefix
        self.q, self.d, self.N = spec.q, spec.d, spec.num_nodes
        self._desc, self._keep = spec.to_c()
        h = C.c_void_p()
        self._call("create", C.byref(self._desc), C.byref(h))
        self.h = h

    # -- plumbing
    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _call(self, name, *args):
        rc = self._fn(name)(*args)
        if rc != 0:
            msg = self._fn("last_error")()
            raise LbfError(rc, msg.decode() if msg else "")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference Lattice API
    def initialize_nodes(self):
        self._call("initialize_nodes", self.h)
        return self

    def main_steps(self, nsteps: int = 1, residuals: bool = False, stop: bool = False):
        flags = (abi.STEP_RESIDUALS if residuals else 0) | (abi.STEP_STOP_ON_CONVERGENCE if stop else 0)
        self._call("main_steps", self.h, int(nsteps), flags)
        return self

    def synchronize(self):
        if hasattr(self.lib, self.prefix + "synchronize"):
            self._call("synchronize", self.h)

    def run(self, max_iterations: int) -> int:
        out = C.c_uint64()
        self._call("run", self.h, C.c_uint64(max_iterations), C.byref(out))
        return int(out.value)

    def residuals(self):
        r = C.c_double()
        ru = (C.c_double * 3)()
        rc = (C.c_double * abi.MAX_SCALARS)()
        self._call("residuals", self.h, C.byref(r), ru, rc)
        return r.value, np.array(ru[: self.d]), np.array(rc[: len(self.spec.scalars)])

    def time_step(self) -> int:
        t = C.c_uint64()
        self._call("time_step", self.h, C.byref(t))
        return int(t.value)

    def reduce(self, which: int, scalar: int = 0) -> float:
        out = C.c_double()
        self._call("reduce", self.h, C.c_int32(which | (scalar << 8)), C.byref(out))
        return out.value

    def fields(self):
        rho = np.empty(self.N)
        u = np.empty(self.N * self.d)
        self._call("get_fields", self.h, rho.ctypes.data_as(C.POINTER(C.c_double)),
                   u.ctypes.data_as(C.POINTER(C.c_double)))
        return rho, u.reshape(self.N, self.d)

    def scalar(self, s: int = 0):
        c = np.empty(self.N)
        q = np.empty(self.N)
        self._call("get_scalar", self.h, C.c_uint32(s), c.ctypes.data_as(C.POINTER(C.c_double)),
                   q.ctypes.data_as(C.POINTER(C.c_double)))
        return c, q

    def pops(self, field: int = 0):
        out = np.empty(self.N * self.q)
        self._call("get_pops", self.h, C.c_int32(field), out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(self.N, self.q)

    def pops_noalloc(self, field: int = 0):
        """lbf_get_pops_noalloc: the same result, staged through the library's
        fixed device array instead of an allocation."""
        out = np.empty(self.N * self.q)
        self._call("get_pops_noalloc", self.h, C.c_int32(field),
                   out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(self.N, self.q)

    def set_pops(self, f, field: int = 0):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        assert a.size == self.N * self.q
        self._call("set_pops", self.h, C.c_int32(field), a.ctypes.data_as(C.POINTER(C.c_double)))
        return self

    def link_tables(self):
        t = np.empty(self.N, np.uint8)
        m = np.empty(self.N, np.uint32)
        f = np.empty(self.N, np.uint8)
        self._call("get_link_tables", self.h, t.ctypes.data_as(C.POINTER(C.c_uint8)),
                   m.ctypes.data_as(C.POINTER(C.c_uint32)), f.ctypes.data_as(C.POINTER(C.c_uint8)))
        return t, m, f


def device_bytes(spec: Spec, lib=None) -> int:
    lib = lib if lib is not None else load_library()
    desc, keep = spec.to_c()
    out = C.c_uint64()
    rc = lib.lbf_device_bytes(C.byref(desc), C.byref(out))
    if rc != 0:
        raise LbfError(rc, lib.lbf_last_error().decode())
    return int(out.value)
