"""Python builder for the case descriptor of the C ABI.

`Spec` is a plain-Python description of one lattice case (the fields of the
reference's `momentum::Parameters`, src/momentum.rs:22-37, and
`passive_scalar::Parameters`, src/passive_scalar.rs:16-25, minus the closures,
which become uniform values or static fields). `Spec.to_c()` produces the
`lbf_case_desc` struct plus the numpy buffers it points into; the same
descriptor drives the HIP library and the CPU oracle.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from . import abi


@dataclasses.dataclass
class FaceBC:
    """Momentum face condition (src/momentum/bc.rs:11-26)."""
    kind: int = abi.BC_PERIODIC
    density: Optional[float] = None
    velocity: Sequence[Optional[float]] = (None, None, None)

    @staticmethod
    def no_slip():
        return FaceBC(abi.BC_NO_SLIP)

    @staticmethod
    def periodic():
        return FaceBC(abi.BC_PERIODIC)

    @staticmethod
    def bounce_back(density, velocity):
        return FaceBC(abi.BC_BOUNCE_BACK, density, tuple(velocity))

    @staticmethod
    def anti_bounce_back(density):
        return FaceBC(abi.BC_ANTI_BOUNCE_BACK, density)

    @staticmethod
    def zou_he(density=None, velocity=(None, None, None)):
        return FaceBC(abi.BC_ZOU_HE, density, tuple(velocity))

    def to_c(self) -> abi.FaceBC:
        c = abi.FaceBC()
        c.kind = self.kind
        known = 0
        if self.density is not None:
            known |= abi.KNOWN_DENSITY
            c.density = float(self.density)
        for x, v in enumerate(list(self.velocity)[:3]):
            if v is not None:
                known |= abi.KNOWN_UX << x
                c.velocity[x] = float(v)
        c.known = known
        return c


@dataclasses.dataclass
class ScalarFaceBC:
    """Scalar face condition (src/passive_scalar/bc.rs:9-30)."""
    kind: int = abi.SBC_PERIODIC
    scalar_value: float = 0.0
    unknown_populations_only: bool = False

    def to_c(self) -> abi.ScalarFaceBC:
        c = abi.ScalarFaceBC()
        c.kind = self.kind
        c.scalar_value = float(self.scalar_value)
        c.unknown_populations_only = int(bool(self.unknown_populations_only))
        return c


@dataclasses.dataclass
class Collision:
    """`CollisionOperator` (src/lib.rs:37-42)."""
    op: int = abi.BGK
    tau: float = 0.5
    omega_plus: float = 0.0
    omega_minus: float = 0.0
    relaxation: Sequence[float] = ()

    @staticmethod
    def bgk(tau):
        return Collision(abi.BGK, tau=tau)

    @staticmethod
    def trt(omega_plus, omega_minus):
        return Collision(abi.TRT, omega_plus=omega_plus, omega_minus=omega_minus)

    @staticmethod
    def trt_magic(tau, magic):
        """TRT from tau and the magic parameter Lambda = (1/w+ - 1/2)(1/w- - 1/2)."""
        wp = 1.0 / tau
        wm = 1.0 / (magic / (tau - 0.5) + 0.5)
        return Collision(abi.TRT, omega_plus=wp, omega_minus=wm)

    @staticmethod
    def mrt(relaxation):
        return Collision(abi.MRT, relaxation=tuple(relaxation))

    def to_c(self) -> abi.Collision:
        c = abi.Collision()
        c.op = self.op
        c.tau = float(self.tau)
        c.omega_plus = float(self.omega_plus)
        c.omega_minus = float(self.omega_minus)
        for k, s in enumerate(self.relaxation):
            c.relaxation[k] = float(s)
        return c


@dataclasses.dataclass
class ScalarSpec:
    """One passive scalar (src/passive_scalar.rs:16-25)."""
    collision: Collision = dataclasses.field(default_factory=lambda: Collision.bgk(0.6))
    faces: List[ScalarFaceBC] = dataclasses.field(default_factory=list)
    inner_bc: int = abi.INNER_ANTI_BOUNCE_BACK
    initial_value: object = 0.0  # float or array[N]
    source: object = None        # None, float or array[N]
    adsorption: Optional[dict] = None  # {adsorption_rate, desorption_rate, q_max, initial_q_ads}


@dataclasses.dataclass
class Spec:
    velocity_set: int
    n: Sequence[int]
    collision: Collision
    faces: List[FaceBC]
    node_type: Optional[np.ndarray] = None  # uint8[N] (0 fluid, 1 solid)
    force: object = None                     # None, (Fx, Fy[, Fz]) or array[N, d]
    initial_density: object = 1.0
    initial_velocity: object = None          # None -> zeros, tuple or array[N, d]
    scalars: List[ScalarSpec] = dataclasses.field(default_factory=list)
    precision: int = abi.F64
    streaming: int = abi.STREAM_AUTO
    shift_scale: float = 0.0
    freeze_momentum: bool = False
    zou_he_3d: bool = False                  # own Zou-He closure on the 3-D lattices
    multiphase: object = None                # None or (G, wall_density): Shan-Chen force

    @property
    def d(self) -> int:
        return abi.DIM_OF[self.velocity_set]

    @property
    def q(self) -> int:
        return abi.Q_OF[self.velocity_set]

    @property
    def shape(self):
        return tuple(int(v) for v in self.n[: self.d])

    @property
    def num_nodes(self) -> int:
        return int(np.prod(self.shape))

    def replace(self, **kw) -> "Spec":
        return dataclasses.replace(self, **kw)

    def to_c(self):
        """Returns (desc, keepalive): the descriptor and the objects its pointers reference."""
        keep = []
        d = abi.CaseDesc()
        d.abi_version = abi.ABI_VERSION
        d.velocity_set = self.velocity_set
        dims = list(self.n) + [1] * (3 - len(self.n))
        for i in range(3):
            d.n[i] = int(dims[i]) if i < self.d else 1
        d.flags = (abi.CASE_FREEZE_MOMENTUM if self.freeze_momentum else 0) | \
            (abi.CASE_ZOU_HE_3D if self.zou_he_3d else 0)
        d.collision = self.collision.to_c()
        N = self.num_nodes
        if self.node_type is not None:
            nt = np.ascontiguousarray(self.node_type, dtype=np.uint8).reshape(-1)
            if nt.size != N:
                raise ValueError(f"node_type has {nt.size} entries, lattice has {N}")
            keep.append(nt)
            d.node_type = nt.ctypes.data_as(C.POINTER(C.c_uint8))
        nf = 4 if self.d == 2 else 6
        if len(self.faces) != nf:
            raise ValueError(f"need {nf} face conditions")
        for f in range(nf):
            d.faces[f] = self.faces[f].to_c()
        if self.force is None:
            d.force_kind = abi.SOURCE_NONE
        elif np.ndim(self.force) == 1 and len(self.force) == self.d:
            d.force_kind = abi.SOURCE_UNIFORM
            for x in range(self.d):
                d.force[x] = float(self.force[x])
        else:
            ff = np.ascontiguousarray(self.force, dtype=np.float64).reshape(-1)
            if ff.size != N * self.d:
                raise ValueError("force field must hold N*d values")
            keep.append(ff)
            d.force_kind = abi.SOURCE_FIELD
            d.force_field = ff.ctypes.data_as(C.POINTER(C.c_double))
        if np.ndim(self.initial_density) == 0:
            d.initial_density_uniform = float(self.initial_density)
        else:
            a = np.ascontiguousarray(self.initial_density, dtype=np.float64).reshape(-1)
            keep.append(a)
            d.initial_density = a.ctypes.data_as(C.POINTER(C.c_double))
        if self.initial_velocity is None:
            pass
        elif np.ndim(self.initial_velocity) == 1 and len(self.initial_velocity) == self.d:
            for x in range(self.d):
                d.initial_velocity_uniform[x] = float(self.initial_velocity[x])
        else:
            a = np.ascontiguousarray(self.initial_velocity, dtype=np.float64).reshape(-1)
            if a.size != N * self.d:
                raise ValueError("initial velocity field must hold N*d values")
            keep.append(a)
            d.initial_velocity = a.ctypes.data_as(C.POINTER(C.c_double))
        d.precision = self.precision
        d.streaming = self.streaming
        d.shift_scale = float(self.shift_scale)
        if self.multiphase is not None:
            d.multiphase = 1
            d.g_function = float(self.multiphase[0])
            d.wall_density = float(self.multiphase[1])
        d.n_scalars = len(self.scalars)
        if self.scalars:
            arr = (abi.ScalarDesc * len(self.scalars))()
            for s, sp in enumerate(self.scalars):
                sd = arr[s]
                sd.collision = sp.collision.to_c()
                sd.velocity_set = self.velocity_set
                sd.inner_bc = sp.inner_bc
                if len(sp.faces) != nf:
                    raise ValueError(f"scalar {s}: need {nf} face conditions")
                for f in range(nf):
                    sd.faces[f] = sp.faces[f].to_c()
                if np.ndim(sp.initial_value) == 0:
                    sd.initial_value_uniform = float(sp.initial_value)
                else:
                    a = np.ascontiguousarray(sp.initial_value, dtype=np.float64).reshape(-1)
                    keep.append(a)
                    sd.initial_value = a.ctypes.data_as(C.POINTER(C.c_double))
                if sp.source is None:
                    sd.source_kind = abi.SOURCE_NONE
                elif np.ndim(sp.source) == 0:
                    sd.source_kind = abi.SOURCE_UNIFORM
                    sd.source_uniform = float(sp.source)
                else:
                    a = np.ascontiguousarray(sp.source, dtype=np.float64).reshape(-1)
                    keep.append(a)
                    sd.source_kind = abi.SOURCE_FIELD
                    sd.source_field = a.ctypes.data_as(C.POINTER(C.c_double))
                if sp.adsorption:
                    sd.has_adsorption = 1
                    sd.adsorption_rate = float(sp.adsorption.get("adsorption_rate", 0.0))
                    sd.desorption_rate = float(sp.adsorption.get("desorption_rate", 0.0))
                    sd.q_max = float(sp.adsorption.get("q_max", 1.0))
                    sd.initial_q_ads = float(sp.adsorption.get("initial_q_ads", 0.0))
            keep.append(arr)
            d.scalars = C.cast(arr, C.POINTER(abi.ScalarDesc))
        keep.append(d)
        return d, keep
