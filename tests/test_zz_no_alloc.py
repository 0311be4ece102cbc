"""lbf_get_pops_noalloc (read-back through the library's fixed device array,
no allocation) against lbf_get_pops, through the C ABI: bit for bit, on two
calls in a row of which the second has the smaller input, over every
read-back kernel (in-place momentum, two-array momentum, two-array scalar)
and a lattice larger than one chunk of the array."""
import numpy as np
import pytest

from lbflow_amd import abi, configs
from lbflow_amd.case import Collision
from lbflow_amd.engine import Case

pytestmark = pytest.mark.gpu

CHUNK_BYTES = 16 << 20  # size of the library's scratch array


def stepped(sp, steps):
    g = Case(sp)
    g.initialize_nodes()
    g.main_steps(steps)
    return g


def check(g, field=0):
    a = g.pops(field)
    b = g.pops_noalloc(field)
    assert np.array_equal(a, b)
    return a


def test_noalloc_matches_get_pops_large_then_small():
    # 80^3 D3Q19 FP32 in place: 9.7 million f64 values, five chunks of the array
    big = configs.porous_channel(n=80, precision=abi.F32, force=1e-5)
    assert big.num_nodes * big.q * 8 > 4 * CHUNK_BYTES
    small = configs.porous_channel(n=(12, 10, 8), precision=abi.F16_SHIFTED, force=1e-4)
    gb = stepped(big, 7)
    gs = stepped(small, 5)
    print("[noalloc] 80^3 D3Q19 FP32, then 12x10x8 FP16-shifted", flush=True)
    fb = check(gb)
    fs = check(gs)
    assert np.abs(fb).max() > 0 and np.abs(fs).max() > 0
    assert np.all(fb[big.node_type.astype(bool)] == 0)  # solid nodes read as zeros
    gb.close()
    gs.close()


def test_noalloc_two_array_momentum_and_scalar():
    # two-array kernels: coupled momentum + scalar with adsorption (FP64), and
    # D3Q27 MRT on two arrays, read back after an odd and an even step count
    sp = configs.coupled_discs(nx=64, ny=40, count=8, rmin=1.5, rmax=3, seed=3,
                               precision=abi.F64, k_a=0.05)
    mrt = configs.porous_channel(n=(12, 10, 9), velocity_set=abi.D3Q27, force=1e-4,
                                 op=Collision.mrt([1.0] * 4 + [1.3] * 6 + [1.1] * 17),
                                 streaming=abi.STREAM_TWO_ARRAY)
    for steps in (5, 6):
        print(f"[noalloc] two-array cases after {steps} steps", flush=True)
        g = stepped(sp, steps)
        check(g, 0)
        check(g, 1)
        g.close()
        h = stepped(mrt, steps)
        check(h, 0)
        h.close()
