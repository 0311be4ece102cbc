"""Captured step batches (lbf_main_steps / lbf_run replay HIP graphs of 16
steps on cases of at most 2^21 nodes): each replay gives the bits of the same
steps launched one by one. Every call below replays at least one batch, and
the twin case (LBF_STEP_GRAPHS=0 at creation: no capture) runs the same calls;
reads and reductions run between the replays. FP64 cases are also checked
against the oracle."""
import os

import numpy as np
import pytest

from lbflow_amd import abi, configs
from lbflow_amd.case import Collision
from lbflow_amd.engine import Case
from oracle import oracle

pytestmark = pytest.mark.gpu


def uncaptured(spec):
    os.environ["LBF_STEP_GRAPHS"] = "0"
    try:
        return Case(spec)
    finally:
        del os.environ["LBF_STEP_GRAPHS"]


CASES = {
    "cavity_f64_two_array": configs.cavity(n=32, tau=0.6).replace(streaming=abi.STREAM_TWO_ARRAY),
    "cavity_f64_in_place": configs.cavity(n=32, tau=0.6).replace(streaming=abi.STREAM_AA),
    "porous_f32_in_place": configs.porous_channel(n=(24, 20, 16), precision=abi.F32, force=1e-4),
    "porous_f16s_in_place": configs.porous_channel(n=(24, 20, 16), precision=abi.F16_SHIFTED,
                                                   force=1e-4),
    "discs_f32_scalar": configs.coupled_discs(nx=48, ny=24, count=6, rmin=1.5, rmax=3, seed=1,
                                              k_a=0.05),
    "sphere_d3q27_mrt_f64": configs.sphere_channel(nx=24, ny=12, nz=12, radius=3,
                                                   op=Collision.mrt([1 / 0.6] * 10 + [1.2] * 17)),
}


def state(c, spec):
    out = [c.pops(), *c.fields()]
    for s in range(len(spec.scalars)):
        out += [c.pops(1 + s), *c.scalar(s)]
    return out


def same(a, b):
    return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))


@pytest.mark.parametrize("name", list(CASES))
def test_replayed_batches_equal_single_launches(name):
    spec = CASES[name]
    print(f"[graphs] {name}", flush=True)
    g, e = Case(spec), uncaptured(spec)
    o = oracle.OracleCase(spec) if spec.precision == abi.F64 else None
    for c in (g, e) + ((o,) if o else ()):
        c.initialize_nodes()
        c.main_steps(1)
    for k, n in enumerate((17, 17, 40, 33)):  # 1, 1, 2 and 2 replays
        for c in (g, e) + ((o,) if o else ()):
            c.main_steps(n)
        sg, se = state(g, spec), state(e, spec)
        assert same(sg, se), f"{name}: call {k}"
        if o:
            assert same(sg, state(o, spec)), f"{name}: call {k} vs the oracle"
        # eager work between replays
        assert g.reduce(abi.REDUCE_NONFINITE) == 0
        assert g.reduce(abi.REDUCE_TOTAL_MASS) == e.reduce(abi.REDUCE_TOTAL_MASS)


@pytest.mark.parametrize("name", ["cavity_f64_two_array", "porous_f32_in_place"])
def test_replayed_run_loop_equals_single_launches(name):
    # lbf_run replays batches with residuals and the stop rule every step
    spec = CASES[name]
    g, e = Case(spec), uncaptured(spec)
    for c in (g, e):
        c.initialize_nodes()
    assert g.run(120) == e.run(120) == 122
    assert same(state(g, spec), state(e, spec))
    assert g.residuals()[0] == e.residuals()[0]
    g.main_steps(5)
    e.main_steps(5)
    assert g.run(200) == e.run(200)
    assert same(state(g, spec), state(e, spec))
