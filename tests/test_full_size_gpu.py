"""GPU FP64 paths (two-array and in-place schemes) against the CPU oracle at
the workloads' real sizes.

- D2Q9 BGK lid-driven cavity, 128x128, tau 0.6, lid 0.1: the whole run of
  20,001 steps (the reference's `--max-iterations 20000` loop runs steps
  0..20000).
- D3Q19 TRT porous channel, 256^3, Bernoulli(0.3) mask seed 2, force 1e-6:
  the node classes and bounce-back link masks bit for bit, then 12 steps.
- D2Q9 BGK flow + passive scalar with Langmuir adsorption, 1024x512 with 200
  discs: 300 steps (the full 50,000 would take the oracle about 20 minutes).
- D3Q27 MRT flow past a sphere, 384x128x128 (no reference oracle: the two FP64
  schemes against each other, bit for bit, 20 steps).

Tolerance as in the small parity cases: 1e-12 relative on populations,
moments, scalar values and adsorbed amounts.
"""
import numpy as np
import pytest

from lbflow_amd import abi, configs
from lbflow_amd.engine import Case
from oracle import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def rel_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


SCHEMES = [abi.STREAM_TWO_ARRAY, abi.STREAM_AA]


def run_both(sp, steps):
    g = Case(sp)
    o = oracle.OracleCase(sp)
    g.initialize_nodes()
    o.initialize_nodes()
    g.main_steps(steps)
    o.main_steps(steps)
    g.synchronize()
    return g, o


def check_fields(g, o, sp, what):
    assert g.time_step() == o.time_step()
    assert rel_err(g.pops(), o.pops()) <= RTOL, f"{what}: populations"
    rg, ug = g.fields()
    ro, uo = o.fields()
    assert rel_err(rg, ro) <= RTOL and rel_err(ug, uo) <= RTOL, f"{what}: moments"
    for s in range(len(sp.scalars)):
        cg, qg = g.scalar(s)
        co, qo = o.scalar(s)
        assert rel_err(cg, co) <= RTOL, f"{what}: scalar {s}"
        if sp.scalars[s].adsorption:
            assert rel_err(qg, qo) <= RTOL, f"{what}: adsorbed amount {s}"
    assert g.reduce(abi.REDUCE_NONFINITE) == 0


@pytest.mark.parametrize("streaming", SCHEMES)
def test_cavity_128_whole_run(streaming):
    sp = configs.cavity(n=128, tau=0.6, lid=0.1).replace(streaming=streaming)
    print(f"[full] cavity 128^2, 20001 steps, streaming={streaming}", flush=True)
    g, o = run_both(sp, 20001)
    check_fields(g, o, sp, "cavity")
    _, u = o.fields()
    assert np.abs(u).max() > 1e-3  # the lid has driven a flow
    g.close()
    o.close()


@pytest.mark.parametrize("streaming", SCHEMES)
def test_porous_256_masks_and_steps(streaming):
    sp = configs.porous_channel(n=256, precision=abi.F64, streaming=streaming)
    print(f"[full] porous 256^3: link tables, 12 steps, streaming={streaming}", flush=True)
    g = Case(sp)
    o = oracle.OracleCase(sp)
    t1, m1, f1 = g.link_tables()
    t2, m2, f2 = o.link_tables()
    assert np.array_equal(t1, t2) and np.array_equal(m1, m2) and np.array_equal(f1, f2)
    del t1, m1, f1, t2, m2, f2
    g.initialize_nodes()
    o.initialize_nodes()
    g.main_steps(12)
    o.main_steps(12)
    check_fields(g, o, sp, "porous 256^3")
    g.close()
    o.close()


def test_coupled_discs_1024x512():
    sp = configs.coupled_discs(precision=abi.F64)
    print("[full] coupled discs 1024x512, 300 steps", flush=True)
    g, o = run_both(sp, 300)
    check_fields(g, o, sp, "coupled discs")
    assert g.reduce(abi.REDUCE_ADSORBED, 0) > 0.0  # adsorption has started
    g.close()
    o.close()


def test_sphere_d3q27_mrt_both_schemes_agree():
    # the reference has no D3Q27 MRT, so the oracle cannot run this case; its
    # two FP64 schemes (dense M / M^-1 products in both) must give the same bits
    sp = configs.sphere_channel(precision=abi.F64)
    print("[full] D3Q27 MRT sphere 384x128x128, 20 steps, both schemes", flush=True)
    out = []
    for streaming in SCHEMES:
        g = Case(sp.replace(streaming=streaming))
        g.initialize_nodes()
        g.main_steps(20)
        out.append((g.pops(), *g.fields()))
        g.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert np.isfinite(out[0][2]).all() and np.abs(out[0][2]).max() > 1e-3
