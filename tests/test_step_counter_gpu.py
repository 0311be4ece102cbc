"""Steps without residuals or the stop rule advance the device step counter
once per batch (each launch reads the counter plus its offset in the batch;
a captured 16-step batch carries its own advance). These runs mix such
batches of ragged lengths with residual steps, the stop-rule loop and reads
in between, on in-place and two-array cases (with a scalar, and from a fresh
array), and check after every call that the step counter equals the steps
run and that the populations are the oracle's bits (FP64) or the per-step
twin's (FP32 / shifted: the same case with step graphs off)."""
import os

import numpy as np
import pytest

from lbflow_amd import abi, configs
from lbflow_amd.engine import Case
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = {
    "cavity_f64_in_place": configs.cavity(n=24, tau=0.6).replace(streaming=abi.STREAM_AA),
    "cavity_f64_two_array": configs.cavity(n=24, tau=0.6).replace(streaming=abi.STREAM_TWO_ARRAY),
    "porous_f64_in_place": configs.porous_channel(n=(12, 10, 8), force=1e-4).replace(
        streaming=abi.STREAM_AA),
    "discs_f64_scalar": configs.coupled_discs(nx=40, ny=20, count=5, rmin=1.5, rmax=3, seed=2,
                                              precision=abi.F64, k_a=0.05),
    "porous_f16s_in_place": configs.porous_channel(n=(24, 20, 16), precision=abi.F16_SHIFTED,
                                                   force=1e-4),
}

# (kind, steps): plain batches of ragged length (some cross 16-step graph
# batches), residual steps, and the stop-rule loop
PLAN = [("steps", 1), ("steps", 3), ("resid", 2), ("steps", 16), ("steps", 17), ("read", 0),
        ("steps", 33), ("resid", 1), ("steps", 5), ("run", 0), ("steps", 2)]


def twin(spec):
    os.environ["LBF_STEP_GRAPHS"] = "0"
    try:
        return Case(spec)
    finally:
        del os.environ["LBF_STEP_GRAPHS"]


@pytest.mark.parametrize("name", list(CASES))
def test_counter_and_bits_over_mixed_batches(name):
    sp = CASES[name]
    print(f"[step-counter] {name}", flush=True)
    g = Case(sp)
    ref = oracle.OracleCase(sp) if sp.precision == abi.F64 else twin(sp)
    for c in (g, ref):
        c.initialize_nodes()
    done = 0
    for kind, n in PLAN:
        if kind == "steps":
            for c in (g, ref):
                c.main_steps(n)
            done += n
        elif kind == "resid":
            for c in (g, ref):
                c.main_steps(n, residuals=True)
            done += n
        elif kind == "run":
            # the stop-rule loop up to a step index past the current one
            target = done + 6
            got = [c.run(target) for c in (g, ref)]
            assert got[0] == got[1]
            done = got[0]
        else:
            g.fields()
            g.reduce(abi.REDUCE_TOTAL_MASS)
        assert g.time_step() == done and ref.time_step() == done, (kind, n)
        for field in range(1 + len(sp.scalars)):
            pg, pr = g.pops(field), ref.pops(field)
            assert np.array_equal(pg, pr), (kind, n, field, np.abs(pg - pr).max())
    rg, ug = g.fields()
    rr, ur = ref.fields()
    assert np.array_equal(rg, rr) and np.array_equal(ug, ur)
    if sp.scalars:
        assert np.array_equal(g.scalar(0), ref.scalar(0))
