"""Python handle over the serving tier (include/lbflow_serve.h).

    srv = Server(hbm_budget_bytes=8 << 30)
    t = srv.submit(spec, steps=2000, keep_pops=True)
    srv.wait(t)
    rho, u, pops = srv.result(t)
    srv.stats()          # p50/p99 time to solution, cases/s, MLUPS, HBM peak
    srv.close()

A request runs `steps` lattice steps exactly as Case.main_steps(steps) would;
the server batches it with every other resident case on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .case import Spec
from .engine import LbfError, load_library

_dp = C.POINTER(C.c_double)


class Server:
    def __init__(self, hbm_budget_bytes: int = 0, max_slots: int = 0, max_steps_per_round: int = 0,
                 autotune: bool = True, grid: int = 0, tile: int = 0, lib=None):
        self.lib = lib if lib is not None else load_library()
        cfg = abi.ServerConfig(abi_version=abi.ABI_VERSION, max_slots=max_slots,
                               hbm_budget_bytes=hbm_budget_bytes,
                               max_steps_per_round=max_steps_per_round,
                               autotune=1 if autotune else 0, grid=grid, tile=tile)
        h = C.c_void_p()
        self._call(self.lib.lbf_server_create, C.byref(cfg), C.byref(h))
        self.h = h
        self._specs = {}

    def _call(self, fn, *args):
        rc = fn(*args)
        if rc != 0:
            msg = self.lib.lbf_last_error()
            raise LbfError(rc, msg.decode() if msg else "")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbf_server_destroy(self.h)
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

    def submit(self, spec: Spec, steps: int, keep_pops: bool = False) -> int:
        desc, keep = spec.to_c()
        t = C.c_uint64()
        flags = abi.SERVE_KEEP_POPS if keep_pops else 0
        self._call(self.lib.lbf_server_submit, self.h, C.byref(desc), C.c_uint64(int(steps)),
                   C.c_uint32(flags), C.byref(t))
        self._specs[t.value] = (spec.num_nodes, spec.d, spec.q, keep_pops)
        return int(t.value)

    def info(self, ticket: int) -> abi.RequestInfo:
        out = abi.RequestInfo()
        self._call(self.lib.lbf_server_info, self.h, C.c_uint64(ticket), C.byref(out))
        return out

    def wait(self, ticket: int, timeout: float = 600.0):
        self._call(self.lib.lbf_server_wait, self.h, C.c_uint64(ticket), C.c_double(timeout))

    def drain(self, timeout: float = 3600.0):
        self._call(self.lib.lbf_server_drain, self.h, C.c_double(timeout))

    def result(self, ticket: int, release: bool = True):
        n, d, q, keep = self._specs[ticket]
        rho = np.empty(n)
        u = np.empty(n * d)
        pops = np.empty(n * q) if keep else None
        self._call(self.lib.lbf_server_result, self.h, C.c_uint64(ticket),
                   rho.ctypes.data_as(_dp), u.ctypes.data_as(_dp),
                   pops.ctypes.data_as(_dp) if keep else None, C.c_int32(1 if release else 0))
        if release:
            del self._specs[ticket]
        return rho, u.reshape(n, d), (pops.reshape(n, q) if keep else None)

    def stats(self) -> dict:
        s = abi.ServerStats()
        self._call(self.lib.lbf_server_get_stats, self.h, C.byref(s))
        return {name: getattr(s, name) for name, _ in s._fields_ if name != "reserved"}

    def now(self) -> float:
        return float(self.lib.lbf_server_now(self.h))
