"""The reference's time loop with I/O driven by the GPU engine writes the same
case directory, byte for byte, as the oracle (FP64 parity path)."""
import os

import pytest

from lbflow_amd import configs
from lbflow_amd import io as lio
from lbflow_amd.engine import Case
from oracle import oracle

pytestmark = pytest.mark.gpu


def tree(root):
    out = {}
    for d, _, files in os.walk(root):
        for f in files:
            p = os.path.join(d, f)
            out[os.path.relpath(p, root)] = open(p).read()
    return out


@pytest.mark.parametrize("name,spec,steps,freq", [
    ("cavity", configs.cavity(n=16, tau=0.6), 30, 10),
    ("porous", configs.porous_channel(n=(12, 10, 8), force=1e-4), 12, 5),
    ("discs_scalar", configs.coupled_discs(nx=40, ny=20, count=5, rmin=1.5, rmax=3, seed=2,
                                           precision=0, k_a=0.05), 15, 7),
])
def test_gpu_case_directory_equals_oracle(tmp_path, name, spec, steps, freq):
    print(f"[io-gpu] {name}", flush=True)
    names = ["c"] if spec.scalars else []
    post = [("mean_density.csv", 3, lio.mean_density)]
    g_root, o_root = str(tmp_path / "gpu"), str(tmp_path / "oracle")
    n_g = lio.run_case(g_root, Case(spec), steps, freq, scalar_names=names, post_functions=post)
    n_o = lio.run_case(o_root, oracle.OracleCase(spec), steps, freq, scalar_names=names,
                       post_functions=post)
    assert n_g == n_o
    tg, to = tree(g_root), tree(o_root)
    assert sorted(tg) == sorted(to)
    diff = [k for k in tg if tg[k] != to[k]]
    # residual sums are not pinned to an order (rayon): those two files may
    # differ in the last digits only; every field file must match byte for byte
    assert all(k.endswith("residuals.csv") or k.endswith("mean_density.csv") for k in diff), diff
    for k in tg:
        if k.endswith("residuals.csv") or k.endswith("mean_density.csv"):
            assert_csv_close(tg[k], to[k], 1e-12)


def assert_csv_close(a, b, rtol):
    """Same header, same time-step column, values equal to rtol relative. The
    files print 9 significant digits ({:.8e}), so two values within rtol may
    still round to neighbouring last digits: one unit there is allowed too."""
    la, lb = a.strip().splitlines(), b.strip().splitlines()
    assert len(la) == len(lb) and la[0] == lb[0]
    for ra, rb in zip(la[1:], lb[1:]):
        va, vb = ra.split(","), rb.split(",")
        assert len(va) == len(vb) and va[0] == vb[0]  # time step
        for x, y in zip(va[1:], vb[1:]):
            last_digit = 10.0 ** (int(y.lower().split("e")[1]) - 8)
            x, y = float(x), float(y)
            assert abs(x - y) <= max(rtol * abs(y), 1.0001 * last_digit), (ra, rb)
