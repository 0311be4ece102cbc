"""Loader of the CPU oracle (oracle/lbf_oracle.c) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / the CPU baseline. The product
(lbflow_amd) never imports it.

The oracle exports the same calls as the HIP engine with the prefix `lbo_`
instead of `lbf_`, so `lbflow_amd.engine.Case(spec, lib=load(), prefix="lbo_")`
drives it exactly like the device engine.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from lbflow_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liblbf_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

EXTRA = [
    ("lbo_set_threads", None, [C.c_int]),
    ("lbo_max_threads", C.c_int, []),
    ("lbo_sizeof", C.c_size_t, [C.c_int]),
    ("lbo_tables", C.c_int, [C.c_int, _ip, _ip, _ip, _dp, _ip, _ip, _ip, _ip]),
    ("lbo_mrt_matrices", C.c_int, [C.c_int, _dp, _dp]),
    ("lbo_k_equilibrium", C.c_int, [C.c_int, C.c_double, _dp, _dp]),
    ("lbo_k_bgk", C.c_int, [C.c_int, _dp, _dp, C.c_double, _dp]),
    ("lbo_k_trt", C.c_int, [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp]),
    ("lbo_k_mrt", C.c_int, [C.c_int, C.c_double, _dp, _dp, _dp, _dp]),
    ("lbo_k_meq", C.c_int, [C.c_int, C.c_double, _dp, _dp]),
    ("lbo_k_momentum_source", C.c_int, [C.c_int, _dp, _dp, C.c_double, _dp]),
    ("lbo_k_scalar_source", C.c_int, [C.c_int, C.c_double, C.c_double, _dp]),
    ("lbo_k_zou_he", C.c_int, [C.c_int, _dp, C.c_uint, C.c_double, _dp]),
    ("lbo_k_zou_he_general", C.c_int, [C.c_int, C.c_int, _dp, C.c_uint, C.c_double, _dp]),
    ("lbo_k_velocity_explicit", C.c_int, [C.c_int, C.c_double, _dp, _dp]),
    ("lbo_neighbor", C.c_long, [C.c_void_p, C.c_long, C.c_int]),
    ("lbo_node_set", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_long, _dp]),
    ("lbo_node_get", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_long, _dp]),
    ("lbo_node_op", C.c_int, [C.c_void_p, C.c_int, C.c_long, C.c_int]),
    ("lbo_node_moments", C.c_int, [C.c_void_p, C.c_long, _dp, _dp]),
    ("lbo_node_set_moments", C.c_int, [C.c_void_p, C.c_long, C.c_double, _dp]),
]


def build() -> str:
    """Compiles the oracle (gcc, contraction off) into oracle/build/."""
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stderr[-4000:])
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError(f"oracle library missing at {LIB}; run `make -C oracle`")
    lib = C.CDLL(LIB)
    eps = [(n.replace("lbf_", "lbo_", 1), r, a) for n, r, a in abi.ENTRY_POINTS
           if n not in ("lbf_synchronize", "lbf_device_bytes", "lbf_get_pops_noalloc")]
    for name, res, args in eps:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args  # uint64_t and unsigned long long share the ABI
    for name, res, args in EXTRA:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def arr(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def tables(vs):
    lib = load()
    d, q, nqf = C.c_int(), C.c_int(), C.c_int()
    c = (C.c_int * 81)()
    w = (C.c_double * 27)()
    qbar = (C.c_int * 27)()
    qf = (C.c_int * 54)()
    nrm = (C.c_int * 6)()
    assert lib.lbo_tables(vs, C.byref(d), C.byref(q), c, w, qbar, qf, C.byref(nqf), nrm) == 0
    Q = q.value
    return dict(d=d.value, q=Q, c=np.array(c[: 3 * Q]).reshape(Q, 3), w=np.array(w[:Q]),
                qbar=np.array(qbar[:Q]), qfaces=np.array(qf[:]).reshape(6, 9)[:, : nqf.value],
                normal=np.array(nrm[:]))


def mrt_matrices(vs):
    Q = abi.Q_OF[vs]
    M = np.zeros(Q * Q)
    Mi = np.zeros(Q * Q)
    assert load().lbo_mrt_matrices(vs, M.ctypes.data_as(_dp), Mi.ctypes.data_as(_dp)) == 0
    return M.reshape(Q, Q), Mi.reshape(Q, Q)


def k_equilibrium(vs, value, u):
    out = np.zeros(abi.Q_OF[vs])
    ua, up = arr(u)
    load().lbo_k_equilibrium(vs, value, up, out.ctypes.data_as(_dp))
    return out


def k_bgk(vs, f, feq, tau):
    out = np.zeros(abi.Q_OF[vs])
    fa, fp = arr(f)
    ea, ep = arr(feq)
    load().lbo_k_bgk(vs, fp, ep, tau, out.ctypes.data_as(_dp))
    return out


def k_trt(vs, f, feq, wp, wm):
    out = np.zeros(abi.Q_OF[vs])
    fa, fp = arr(f)
    ea, ep = arr(feq)
    load().lbo_k_trt(vs, fp, ep, wp, wm, out.ctypes.data_as(_dp))
    return out


def k_mrt(vs, value, u, f, s):
    out = np.zeros(abi.Q_OF[vs])
    ua, up = arr(u)
    fa, fp = arr(f)
    sa, sp = arr(s)
    assert load().lbo_k_mrt(vs, value, up, fp, sp, out.ctypes.data_as(_dp)) == 0
    return out


def k_meq(vs, rho, u):
    out = np.zeros(abi.Q_OF[vs])
    ua, up = arr(u)
    load().lbo_k_meq(vs, rho, up, out.ctypes.data_as(_dp))
    return out


def k_momentum_source(vs, u, F, tau):
    out = np.zeros(abi.Q_OF[vs])
    ua, up = arr(u)
    Fa, Fp = arr(F)
    load().lbo_k_momentum_source(vs, up, Fp, tau, out.ctypes.data_as(_dp))
    return out


def k_scalar_source(vs, S, tau_g):
    out = np.zeros(abi.Q_OF[vs])
    load().lbo_k_scalar_source(vs, S, tau_g, out.ctypes.data_as(_dp))
    return out


def k_zou_he(face, f, known, density=0.0, vel=(0.0, 0.0, 0.0)):
    fa = np.array(f, dtype=np.float64).copy()
    va, vp = arr(list(vel) + [0.0] * (3 - len(vel)))
    rc = load().lbo_k_zou_he(face, fa.ctypes.data_as(_dp), known, density, vp)
    return rc, fa


def k_zou_he_general(vs, face, f, known, density=0.0, vel=(0.0, 0.0, 0.0)):
    fa = np.array(f, dtype=np.float64).copy()
    va, vp = arr(list(vel) + [0.0] * (3 - len(vel)))
    rc = load().lbo_k_zou_he_general(vs, face, fa.ctypes.data_as(_dp), known, density, vp)
    return rc, fa


def k_velocity_explicit(vs, rho, f):
    out = np.zeros(3)
    fa, fp = arr(f)
    load().lbo_k_velocity_explicit(vs, rho, fp, out.ctypes.data_as(_dp))
    return out[: abi.DIM_OF[vs]]


def case(spec):
    from lbflow_amd.engine import Case
    return OracleCase(spec)


class OracleCase:
    """Oracle handle with the interface of lbflow_amd.engine.Case plus node hooks."""

    def __new__(cls, spec):
        from lbflow_amd.engine import Case

        class _O(Case):
            def synchronize(self):
                pass

            def neighbor(self, node, i):
                return int(self.lib.lbo_neighbor(self.h, node, i))

            def node_set(self, node, values, field=0, which=0):
                a, p = arr(values)
                assert self.lib.lbo_node_set(self.h, field, which, node, p) == 0

            def node_get(self, node, field=0, which=0):
                out = np.zeros(self.q)
                assert self.lib.lbo_node_get(self.h, field, which, node, out.ctypes.data_as(_dp)) == 0
                return out

            def node_op(self, op, node, arg=0):
                assert self.lib.lbo_node_op(self.h, op, node, arg) == 0

            def node_moments(self, node):
                r = C.c_double()
                u = np.zeros(3)
                self.lib.lbo_node_moments(self.h, node, C.byref(r), u.ctypes.data_as(_dp))
                return r.value, u[: self.d]

            def node_set_moments(self, node, rho, u):
                a, p = arr(list(u) + [0.0] * (3 - len(u)))
                self.lib.lbo_node_set_moments(self.h, node, rho, p)

        return _O(spec, lib=load(), prefix="lbo_")
