"""Two time steps per pass (LBF_STREAM_TWO_STEP, tb.hpp) against the two-array
scheme it accelerates and against the FP64 oracle.

A pass runs the two-array kernel's node arithmetic twice, keeping the first
step's post-collision values in LDS, so the bar is bit for bit:
  - populations, densities and velocities equal to a LBF_STREAM_TWO_ARRAY twin
    run with the same step schedule, in FP64 and in FP32;
  - in FP64 also equal to the oracle (the two-array FP64 path's own bar);
  - in FP32 also equal to the in-place (AA) kernel: all three FP32 schemes
    collide through the same pair-form code (pair_collide.inc);
  - in shifted storage (FP32 / FP16 / BF16) equal to the in-place kernel;
over even and odd step counts (an odd count ends with a single step), split
batches, batches with residual steps in between (single steps), lbf_run, a
restart from lbf_set_pops (the pass then starts from post-stream values),
tiles cut at the domain's edge and z segments of 1 to nz planes.
"""
import os

import numpy as np
import pytest

from lbflow_amd import abi, configs, geometry
from lbflow_amd.case import Collision, FaceBC, Spec
from lbflow_amd.engine import Case, LbfError
from oracle import oracle

pytestmark = pytest.mark.gpu

MRT19 = [1.0, 1.19, 1.4, 1.0, 1.2, 1.0, 1.2, 1.0, 1.2, 1.25, 1.4, 1.25, 1.4, 1.25, 1.25, 1.25,
         1.98, 1.98, 1.98]


def cases():
    frng = np.random.default_rng(31)
    moving_lid = Spec(velocity_set=abi.D3Q19, n=(14, 11, 9), collision=Collision.bgk(0.7),
                      faces=[FaceBC.periodic(), FaceBC.periodic(), FaceBC.no_slip(),
                             FaceBC.no_slip(), FaceBC.no_slip(),
                             FaceBC.bounce_back(1.0, (0.05, 0.01, 0.0))],
                      node_type=geometry.bernoulli_mask((14, 11, 9), 0.1, 4),
                      initial_density=1.0, initial_velocity=(0.0, 0.0, 0.0))
    walls_x = Spec(velocity_set=abi.D3Q15, n=(9, 12, 10), collision=Collision.trt(1.1, 0.85),
                   faces=[FaceBC.bounce_back(1.0, (0.0, 0.02, 0.01)), FaceBC.no_slip(),
                          FaceBC.periodic(), FaceBC.periodic(), FaceBC.no_slip(),
                          FaceBC.bounce_back(1.0, (0.03, 0.0, 0.0))],
                   node_type=geometry.bernoulli_mask((9, 12, 10), 0.15, 3),
                   initial_density=1.0, initial_velocity=(0.0, 0.0, 0.0))
    lid_d3q27 = Spec(velocity_set=abi.D3Q27, n=(10, 9, 8), collision=Collision.trt(1.1, 0.85),
                     faces=[FaceBC.periodic(), FaceBC.periodic(), FaceBC.no_slip(),
                            FaceBC.no_slip(), FaceBC.no_slip(),
                            FaceBC.bounce_back(1.0, (0.04, 0.02, 0.0))],
                     node_type=geometry.bernoulli_mask((10, 9, 8), 0.1, 6),
                     initial_density=1.0, initial_velocity=(0.0, 0.0, 0.0))
    return {
        "porous_trt_force": configs.porous_channel(n=(24, 20, 16), force=1e-4),
        "porous_bgk_guo": configs.porous_channel(n=(16, 12, 20), op=Collision.bgk(0.7),
                                                 force=1e-4, seed=5),
        # several tiles in x and y, cut at the domain's edge
        "porous_ragged_tiles": configs.porous_channel(n=(70, 37, 13), force=1e-4, seed=9),
        "periodic_box_velocity": configs.periodic_box(n=(12, 10, 8)).replace(
            initial_velocity=frng.uniform(-0.03, 0.03, (960, 3))),
        "porous_mrt_d3q19": configs.porous_channel(n=(12, 10, 9), op=Collision.mrt(MRT19),
                                                   force=1e-4),
        "force_field_d3q19": configs.porous_channel(n=(10, 9, 8), p_solid=0.2, seed=13).replace(
            force=1e-4 * frng.standard_normal((720, 3))),
        "moving_lid_d3q19_bgk": moving_lid,
        "moving_walls_d3q15_trt": walls_x,
        "porous_mrt_d3q15": configs.porous_channel(
            n=(11, 9, 10), velocity_set=abi.D3Q15,
            op=Collision.mrt([1.0, 1.2, 1.3, 1.0, 1.1, 1.0, 1.1, 1.0, 1.1, 1.25, 1.25, 1.25,
                              1.25, 1.25, 1.1]), force=1e-4, seed=17),
        "moving_lid_d3q27_trt": lid_d3q27,
        "d3q27_bgk": configs.porous_channel(n=(10, 8, 9), velocity_set=abi.D3Q27,
                                            op=Collision.bgk(0.8), force=1e-4),
        "d3q27_mrt_sphere": configs.sphere_channel(
            nx=20, ny=10, nz=12, radius=3, op=Collision.mrt([1 / 0.6] * 10 + [1.2] * 17),
            inlet_outlet=False),
    }


CASES = cases()


def pair(sp, precision):
    a = Case(sp.replace(precision=precision, streaming=abi.STREAM_TWO_STEP))
    b = Case(sp.replace(precision=precision, streaming=abi.STREAM_TWO_ARRAY))
    a.initialize_nodes()
    b.initialize_nodes()
    return a, b


def assert_same(a, b):
    assert np.array_equal(a.pops(), b.pops())
    ra, ua = a.fields()
    rb, ub = b.fields()
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub)
    assert a.time_step() == b.time_step()


@pytest.mark.parametrize("precision", [abi.F64, abi.F32])
@pytest.mark.parametrize("name", list(CASES))
def test_passes_equal_two_array_steps(name, precision):
    print(f"[two-step] {name} precision={precision}", flush=True)
    a, b = pair(CASES[name], precision)
    # even and odd batches; the odd ones end with a single step
    for n in (8, 3, 5, 2):
        a.main_steps(n)
        b.main_steps(n)
        assert_same(a, b)


@pytest.mark.parametrize("name", list(CASES))
def test_f32_schemes_agree(name):
    print(f"[two-step-f32-aa] {name}", flush=True)
    sp = CASES[name].replace(precision=abi.F32)
    cs = [Case(sp.replace(streaming=s)) for s in
          (abi.STREAM_TWO_STEP, abi.STREAM_AA, abi.STREAM_TWO_ARRAY)]
    for c in cs:
        c.initialize_nodes()
        c.main_steps(6)
        c.main_steps(3, residuals=True)
        c.main_steps(5)
    for c in cs[1:]:
        assert_same(cs[0], c)


@pytest.mark.parametrize("name", ["porous_trt_force", "moving_lid_d3q19_bgk", "d3q27_bgk",
                                  "moving_walls_d3q15_trt", "porous_mrt_d3q15",
                                  "moving_lid_d3q27_trt"])
def test_passes_f64_equal_oracle(name):
    sp = CASES[name].replace(precision=abi.F64)
    print(f"[two-step-oracle] {name}", flush=True)
    g = Case(sp.replace(streaming=abi.STREAM_TWO_STEP))
    o = oracle.OracleCase(sp)
    g.initialize_nodes()
    o.initialize_nodes()
    g.main_steps(9)
    o.main_steps(9)
    assert np.array_equal(g.pops(), o.pops())
    rg, ug = g.fields()
    ro, uo = o.fields()
    assert np.array_equal(rg, ro) and np.array_equal(ug, uo)


@pytest.mark.parametrize("precision", [abi.F64, abi.F32])
def test_residual_steps_and_run_between_passes(precision):
    print(f"[two-step-mixed] precision={precision}", flush=True)
    a, b = pair(CASES["porous_ragged_tiles"], precision)
    for c in (a, b):
        c.main_steps(4)
        c.main_steps(3, residuals=True)
        c.main_steps(6)
    assert_same(a, b)
    for x, y in zip(a.residuals(), b.residuals()):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    # the drop-in loop runs single steps with residuals and the stop rule
    assert a.run(30) == b.run(30)
    assert_same(a, b)
    a.main_steps(6)
    b.main_steps(6)
    assert_same(a, b)


@pytest.mark.parametrize("precision", [abi.F64, abi.F32])
def test_restart_from_set_pops(precision):
    print(f"[two-step-restart] precision={precision}", flush=True)
    a, b = pair(CASES["porous_bgk_guo"], precision)
    a.main_steps(5)
    b.main_steps(5)
    f = a.pops()
    # post-stream populations written back: the next pass starts from them
    a.set_pops(f)
    b.set_pops(f)
    a.main_steps(6)
    b.main_steps(6)
    assert_same(a, b)


@pytest.mark.parametrize("zs", ["1", "4", "13"])
def test_z_segments(zs, monkeypatch):
    print(f"[two-step-segments] zs={zs}", flush=True)
    monkeypatch.setenv("LBF_TB_ZS", zs)
    a, b = pair(CASES["porous_ragged_tiles"], abi.F32)
    a.main_steps(10)
    b.main_steps(10)
    assert_same(a, b)


def tiny_cases():
    rng = np.random.default_rng(17)
    m3 = np.zeros((3, 2, 2), np.uint8)
    m3[1, 0, 1] = 1
    box = Spec(velocity_set=abi.D3Q19, n=(5, 4, 6), collision=Collision.trt(1.1, 0.9),
               faces=[FaceBC.no_slip()] * 5 + [FaceBC.bounce_back(1.0, (0.04, 0.0, 0.0))],
               initial_density=1.0, initial_velocity=(0.0, 0.0, 0.0))
    return {
        # one-node-wide and two-node axes wrap onto themselves
        "d3q19_2x2x3_solid": configs.periodic_box(n=(2, 2, 3)).replace(
            node_type=m3.reshape(-1), initial_velocity=rng.uniform(-0.02, 0.02, (12, 3))),
        "d3q27_1x1x4": configs.periodic_box(n=(1, 1, 4), velocity_set=abi.D3Q27).replace(
            initial_velocity=rng.uniform(-0.02, 0.02, (4, 3))),
        "d3q19_nz1": configs.periodic_box(n=(9, 7, 1)).replace(
            initial_velocity=rng.uniform(-0.02, 0.02, (63, 3))),
        "d3q15_nz2_walls": configs.porous_channel(n=(11, 6, 2), velocity_set=abi.D3Q15,
                                                  force=1e-4, seed=3),
        # every face non-periodic, a moving lid on top
        "closed_box_lid": box,
    }


@pytest.mark.parametrize("precision", [abi.F64, abi.F32])
@pytest.mark.parametrize("name", list(tiny_cases()))
def test_tiny_lattices(name, precision):
    print(f"[two-step-tiny] {name} precision={precision}", flush=True)
    a, b = pair(tiny_cases()[name], precision)
    for n in (6, 3, 4):
        a.main_steps(n)
        b.main_steps(n)
        assert_same(a, b)


@pytest.mark.parametrize("precision", [abi.F32_SHIFTED, abi.F16_SHIFTED, abi.BF16_SHIFTED])
@pytest.mark.parametrize("name", list(CASES))
def test_shifted_passes_equal_in_place(name, precision):
    # shifted storage has no plain two-array scheme to compare with: the
    # in-place kernel runs the same pair-form arithmetic on the same stored values
    print(f"[two-step-shifted] {name} precision={precision}", flush=True)
    sp = CASES[name].replace(precision=precision)
    a = Case(sp.replace(streaming=abi.STREAM_TWO_STEP))
    b = Case(sp.replace(streaming=abi.STREAM_AA))
    for c in (a, b):
        c.initialize_nodes()
    for n in (8, 3, 5, 2):
        a.main_steps(n)
        b.main_steps(n)
        assert_same(a, b)
    a.main_steps(3, residuals=True)
    b.main_steps(3, residuals=True)
    assert_same(a, b)
    f = a.pops()
    a.set_pops(f)
    b.set_pops(f)
    a.main_steps(4)
    b.main_steps(4)
    assert_same(a, b)


@pytest.mark.parametrize("precision", [abi.F32, abi.F16_SHIFTED])
def test_in_place_restart_at_odd_step(precision):
    # a restart after an odd step count: the in-place kernel starts its
    # pattern with the even (own-slot) form and gives the two-step kernel's
    # bits, on every repetition
    print(f"[two-step-restart-odd] precision={precision}", flush=True)
    sp = CASES["porous_trt_force"].replace(precision=precision)
    t = Case(sp.replace(streaming=abi.STREAM_TWO_STEP))
    t.initialize_nodes()
    t.main_steps(7)
    f = t.pops()
    t.set_pops(f)
    t.main_steps(6)
    for _ in range(3):
        a = Case(sp.replace(streaming=abi.STREAM_AA))
        a.initialize_nodes()
        a.main_steps(7)
        a.set_pops(f)
        a.main_steps(6)
        assert_same(a, t)
        a.close()


def test_unsupported_cases_are_refused():
    anti = configs.porous_channel(n=(12, 10, 8))
    anti.faces = anti.faces[:4] + [FaceBC.anti_bounce_back(1.01), FaceBC.anti_bounce_back(1.0)]
    for sp in (configs.cavity(n=16),                                     # 2-D
               anti,                                                     # anti-bounce-back
               configs.sphere_channel(nx=20, ny=10, nz=12, radius=3)):  # Zou-He faces
        with pytest.raises(LbfError) as e:
            Case(sp.replace(streaming=abi.STREAM_TWO_STEP))
        assert e.value.code == abi.EUNSUPPORTED


@pytest.mark.parametrize("name,sp", [
    ("porous_256", configs.porous_channel(n=256, precision=abi.F32)),
    ("sphere_d3q27_mrt", configs.sphere_channel(
        precision=abi.F32, op=Collision.mrt([1 / 0.6] * 10 + [1.2] * 17), inlet_outlet=False)),
])
def test_full_size_passes_equal_two_array(name, sp):
    # the benchmark geometries: 256^3 D3Q19 TRT porous channel and the
    # 384x128x128 D3Q27 MRT sphere (periodic x, no-slip y/z)
    print(f"[two-step-full] {name}", flush=True)
    a, b = pair(sp, abi.F32)
    a.main_steps(6)
    b.main_steps(6)
    assert_same(a, b)
