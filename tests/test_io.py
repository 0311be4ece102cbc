"""Case-directory I/O in the reference's formats (CPU): Rust `{:.8e}`
formatting, file layouts and headers, restart read-back, and the time loop
with I/O driven through the oracle (the GPU run is compared with it in
test_io_gpu.py)."""
import os

import numpy as np
import pytest

from lbflow_amd import configs, geometry
from lbflow_amd import io as lio
from oracle import oracle


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()
    oracle.load()


def test_rust_lower_exp_format():
    # Rust `format!("{:.8e}", x)`: no '+', no exponent padding (LowerExp)
    cases = [(1.0, "1.00000000e0"), (0.0, "0.00000000e0"), (-0.0, "-0.00000000e0"),
             (-2.5e-7, "-2.50000000e-7"), (123456.789, "1.23456789e5"),
             (1e-300, "1.00000000e-300"), (1.5e10, "1.50000000e10"), (0.1, "1.00000000e-1"),
             (float("inf"), "inf"), (-float("inf"), "-inf"), (float("nan"), "NaN"),
             (9.999999999e-1, "1.00000000e0")]
    got = lio.fmt_e(np.array([x for x, _ in cases]))
    assert got.tolist() == [s for _, s in cases]
    assert lio.fmt_e(np.array([0.5, 12.0]), 4).tolist() == ["5.0000e-1", "1.2000e1"]


def test_coordinates_file(tmp_path):
    lio.case_setup(str(tmp_path))
    nt = np.array([0, 1, 0, 0, 0, 1], np.uint8)
    lio.write_coordinates(str(tmp_path), (3, 2), nt, delta_x=0.5)
    text = open(tmp_path / "data" / "coordinates.csv").read().splitlines()
    assert text[0] == "i,j,x,y,node_type"
    assert text[1] == "0,0,2.50000000e-1,2.50000000e-1,fluid"
    assert text[2] == "1,0,7.50000000e-1,2.50000000e-1,solid"
    assert text[6] == "2,1,1.25000000e0,7.50000000e-1,solid"
    lio.write_coordinates(str(tmp_path), (2, 1, 2), None)
    t3 = open(tmp_path / "data" / "coordinates.csv").read().splitlines()
    assert t3[0] == "i,j,k,x,y,z,node_type" and t3[4] == "1,0,1,1.50000000e0,5.00000000e-1,1.50000000e0,fluid"


def test_field_files_round_trip(tmp_path):
    root = str(tmp_path)
    lio.case_setup(root)
    rng = np.random.default_rng(0)
    rho = 1 + 1e-3 * rng.standard_normal(50)
    u = 1e-2 * rng.standard_normal((50, 3))
    lio.write_density(root, 7, rho)
    lio.write_velocity(root, 7, u)
    lio.write_scalar(root, 7, "c", rho - 1)
    assert open(tmp_path / "data" / "7" / "density.csv").readline().strip() == "density"
    assert open(tmp_path / "data" / "7" / "velocity.csv").readline().strip() == \
        "velocity_x,velocity_y,velocity_z"
    assert np.allclose(lio.density_from_time_step(root, 7), rho, rtol=6e-9, atol=0)
    assert np.allclose(lio.velocity_from_time_step(root, 7), u, rtol=6e-9, atol=1e-300)
    assert np.allclose(lio.scalar_from_time_step(root, 7, "c"), rho - 1, rtol=6e-9, atol=0)
    assert lio.collect_time_steps(root) == [7]


def test_residuals_and_post_files(tmp_path):
    root = str(tmp_path)
    lio.case_setup(root)
    lio.append_residuals(root, 0, 1e-3, [2e-4, 3e-5], [4e-6], ["c"])
    lio.append_residuals(root, 1, 5e-4, [1e-4, 1e-5], [2e-6], ["c"])
    lines = open(tmp_path / "data" / "residuals.csv").read().splitlines()
    assert lines == ["time_step,density,velocity_x,velocity_y,c",
                     "0,1.00000000e-3,2.00000000e-4,3.00000000e-5,4.00000000e-6",
                     "1,5.00000000e-4,1.00000000e-4,1.00000000e-5,2.00000000e-6"]
    lio.append_post(root, "mean.csv", 0, [("mean_density", 1.0)])
    lio.append_post(root, "mean.csv", 10, [("mean_density", 0.5)])
    assert open(tmp_path / "post_processing" / "mean.csv").read().splitlines() == \
        ["time_step,mean_density", "0,1.00000000e0", "10,5.00000000e-1"]


def test_vtk_files(tmp_path):
    p = str(tmp_path / "f.vtk")
    lio.write_vtk_momentum(p, (2, 2), np.array([1.0, 1.0, 0.5, 2.0]), np.zeros((4, 2)),
                           physical=dict(reference_pressure=101325.0, pressure_factor=1.0,
                                         velocity_factor=2.0))
    lines = open(p).read().splitlines()
    assert lines[:6] == ["# vtk DataFile Version 3.0", "lbflow simulation data", "ASCII",
                         "DATASET STRUCTURED_GRID", "DIMENSIONS 2 2 1", "POINTS 4 float"]
    assert lines[6] == "5.0000e-1 5.0000e-1 0.0000e0"
    assert lines[10] == "POINT_DATA 4" and lines[11] == "SCALARS density float 1"
    assert lines[13] == "1.00000000e0" and lines[17] == "VECTORS velocity float"
    assert "SCALARS physical_pressure float 1" in lines
    lio.write_vtk_node_type(str(tmp_path / "n.vtk"), (2, 2), np.array([0, 1, 0, 0]))
    assert open(tmp_path / "n.vtk").read().splitlines()[-4:] == ["0", "1", "0", "0"]


def test_map_xyz_round_trip(tmp_path):
    m = geometry.bernoulli_mask((5, 4, 3), 0.3, 1)
    p = str(tmp_path / "map.xyz")
    geometry.write_map_xyz(p, m, (5, 4, 3))
    assert np.array_equal(geometry.read_map_xyz(p), m.reshape(-1))


def test_time_loop_with_io_on_the_oracle(tmp_path):
    root = str(tmp_path)
    sp = configs.cavity(n=8, tau=0.6)
    o = oracle.OracleCase(sp)
    steps = lio.run_case(root, o, max_iterations=12, write_data_frequency=5,
                         post_functions=[("mean_density.csv", 4, lio.mean_density)])
    assert steps == 14  # t > max_iterations stops after step index 13
    assert lio.collect_time_steps(root) == [0, 5, 10]
    res = open(tmp_path / "data" / "residuals.csv").read().splitlines()
    assert res[0] == "time_step,density,velocity_x,velocity_y" and len(res) == 15
    post = open(tmp_path / "post_processing" / "mean_density.csv").read().splitlines()
    assert post[0] == "time_step,mean_density" and [l.split(",")[0] for l in post[1:]] == \
        ["0", "4", "8", "12"]
    # the step-10 files restart a case (InitialDensity/Velocity::FromTimeStep)
    rho10 = lio.density_from_time_step(root, 10)
    u10 = lio.velocity_from_time_step(root, 10)
    r = oracle.OracleCase(sp)
    r.initialize_nodes()
    r.main_steps(11)
    rho, u = r.fields()
    assert np.allclose(rho10, rho, rtol=6e-9) and np.allclose(u10, u, rtol=6e-9, atol=1e-12)
    restarted = oracle.OracleCase(sp.replace(initial_density=rho10, initial_velocity=u10))
    restarted.initialize_nodes()
    assert np.allclose(restarted.fields()[0], rho10)
