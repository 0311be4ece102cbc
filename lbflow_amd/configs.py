"""Case builders for the benchmark configurations (sizes are parameters, so
the same builders give the small parity cases and the full-size runs)."""
from __future__ import annotations

import numpy as np

from . import abi
from .case import Collision, FaceBC, ScalarFaceBC, ScalarSpec, Spec
from .geometry import bernoulli_mask, discs_mask, sphere_mask


def cavity(n=128, tau=0.6, lid=0.1, precision=abi.F64, op=None) -> Spec:
    """D2Q9 lid-driven cavity: the reference's default parameters
    (src/momentum.rs:39-69) with n and tau changed — W/E/S no-slip, N a
    moving bounce-back wall (rho 1, u = (lid, 0)), rest-state start."""
    return Spec(
        velocity_set=abi.D2Q9, n=(n, n), collision=op or Collision.bgk(tau),
        faces=[FaceBC.no_slip(), FaceBC.no_slip(), FaceBC.no_slip(),
               FaceBC.bounce_back(1.0, (lid, 0.0))],
        initial_density=1.0, initial_velocity=(0.0, 0.0), precision=precision)


def porous_channel(n=256, precision=abi.F64, p_solid=0.3, seed=2, tau=0.8, magic=3.0 / 16.0,
                   force=1e-6, op=None, velocity_set=abi.D3Q19, streaming=abi.STREAM_AUTO) -> Spec:
    """D3Q19 TRT body-force channel through a Bernoulli(p_solid) porous solid:
    periodic in x, no-slip walls on y/z, force along x."""
    shape = (n, n, n) if np.ndim(n) == 0 else tuple(n)
    mask = bernoulli_mask(shape, p_solid, seed) if p_solid > 0 else None
    return Spec(
        velocity_set=velocity_set, n=shape, collision=op or Collision.trt_magic(tau, magic),
        faces=[FaceBC.periodic(), FaceBC.periodic(), FaceBC.no_slip(), FaceBC.no_slip(),
               FaceBC.no_slip(), FaceBC.no_slip()],
        node_type=mask, force=(force, 0.0, 0.0), initial_density=1.0,
        initial_velocity=(0.0, 0.0, 0.0), precision=precision, streaming=streaming)


def periodic_box(n=256, precision=abi.F64, tau=0.8, magic=3.0 / 16.0, velocity_set=abi.D3Q19,
                 streaming=abi.STREAM_AUTO, op=None) -> Spec:
    """All-fluid fully periodic box (the clean roofline point), TRT."""
    shape = (n, n, n) if np.ndim(n) == 0 else tuple(n)
    return Spec(
        velocity_set=velocity_set, n=shape, collision=op or Collision.trt_magic(tau, magic),
        faces=[FaceBC.periodic()] * 6, initial_density=1.0, initial_velocity=(0.0, 0.0, 0.0),
        precision=precision, streaming=streaming)


def coupled_discs(nx=1024, ny=512, count=200, rmin=4.0, rmax=12.0, seed=4, tau=0.6, tau_g=0.6,
                  u_in=0.05, k_a=1e-3, q_max=1.0, precision=abi.F32,
                  zeroth_unknown_only=False) -> Spec:
    """D2Q9 BGK flow + coupled D2Q9 passive scalar with Langmuir adsorption on
    the walls of random disc obstacles: Zou-He velocity inlet (W) and
    pressure outlet (E), no-slip S/N; scalar c = 1 at the inlet
    (anti-bounce-back), zeroth-order outflow, no-flux S/N, inner bounce-back."""
    mask = discs_mask(nx, ny, count, rmin, rmax, seed)
    scalar = ScalarSpec(
        collision=Collision.bgk(tau_g),
        faces=[ScalarFaceBC(abi.SBC_ANTI_BOUNCE_BACK, 1.0),
               ScalarFaceBC(abi.SBC_ZEROTH_ORDER_NO_FLUX, unknown_populations_only=zeroth_unknown_only),
               ScalarFaceBC(abi.SBC_BB_NO_FLUX), ScalarFaceBC(abi.SBC_BB_NO_FLUX)],
        inner_bc=abi.INNER_BOUNCE_BACK, initial_value=0.0,
        adsorption=dict(adsorption_rate=k_a, desorption_rate=0.0, q_max=q_max, initial_q_ads=0.0))
    return Spec(
        velocity_set=abi.D2Q9, n=(nx, ny), collision=Collision.bgk(tau),
        faces=[FaceBC.zou_he(None, (u_in, 0.0)), FaceBC.zou_he(1.0, (None, 0.0)),
               FaceBC.no_slip(), FaceBC.no_slip()],
        node_type=mask, initial_density=1.0, initial_velocity=(0.0, 0.0), scalars=[scalar],
        precision=precision)


def sphere_channel(nx=384, ny=128, nz=128, radius=16, u_in=0.05, tau=0.6, precision=abi.F64,
                   velocity_set=abi.D3Q27, op=None, inlet_outlet=True) -> Spec:
    """Flow past a bounce-back sphere of the given radius centred at
    (nx/4, ny/2, nz/2). inlet_outlet: Zou-He velocity inlet u = (u_in, 0, 0)
    on the west face and pressure outlet rho = 1 (zero tangential velocity) on
    the east face, periodic lateral faces; this uses the engine's own 3-D
    Zou-He closure, which the reference does not have. Otherwise periodic
    along x and no-slip walls on y/z, starting from a uniform u_in."""
    mask = sphere_mask(nx, ny, nz, (nx // 4, ny // 2, nz // 2), radius)
    if inlet_outlet:
        faces = [FaceBC.zou_he(None, (u_in, 0.0, 0.0)), FaceBC.zou_he(1.0, (None, 0.0, 0.0)),
                 FaceBC.periodic(), FaceBC.periodic(), FaceBC.periodic(), FaceBC.periodic()]
    else:
        faces = [FaceBC.periodic()] * 2 + [FaceBC.no_slip()] * 4
    return Spec(
        velocity_set=velocity_set, n=(nx, ny, nz), collision=op or Collision.bgk(tau),
        faces=faces, node_type=mask, initial_density=1.0, initial_velocity=(u_in, 0.0, 0.0),
        precision=precision, zou_he_3d=inlet_outlet)


def serving_case(velocity_set=abi.D2Q9, side=64, tau=0.8, porosity=0.8, seed=0,
                 precision=abi.F32, force=1e-5) -> Spec:
    """One case of the serving mix: BGK(tau), all faces periodic, a
    Bernoulli(1 - porosity) solid mask, a small body force along x so the
    flow is not at rest (force is this build's choice; the mix leaves it open)."""
    d = abi.DIM_OF[velocity_set]
    shape = (side,) * d
    mask = bernoulli_mask(shape, 1.0 - porosity, seed)
    return Spec(
        velocity_set=velocity_set, n=shape, collision=Collision.bgk(tau),
        faces=[FaceBC.periodic()] * (2 * d), node_type=mask,
        force=(force,) + (0.0,) * (d - 1), initial_density=1.0,
        initial_velocity=(0.0,) * d, precision=precision)


def serving_mix(count=2000, seed=5, side_2d=(32, 256), side_3d=(32, 256),
                steps=(1000, 20000)):
    """The serving mix: lattice D2Q9 / D3Q19 50/50, sides log-uniform in the
    given ranges, tau ~ U[0.55, 1.2], steps ~ U{steps}, porosity ~ U[0.6, 0.9],
    storage FP32 or FP16-shifted 50/50. Returns [(spec, steps)] lazily built
    from per-case parameters (cases are generated when iterated)."""
    rng = np.random.default_rng(seed)
    params = []
    for i in range(count):
        vs = abi.D2Q9 if rng.random() < 0.5 else abi.D3Q19
        lo, hi = side_2d if vs == abi.D2Q9 else side_3d
        side = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        tau = float(rng.uniform(0.55, 1.2))
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        porosity = float(rng.uniform(0.6, 0.9))
        prec = abi.F32 if rng.random() < 0.5 else abi.F16_SHIFTED
        params.append(dict(velocity_set=vs, side=side, tau=tau, porosity=porosity,
                           seed=seed * 100003 + i, precision=prec, steps=nsteps))
    return params
