"""Case-directory I/O in the reference's formats, fed from device downloads.

Same directories, file names, headers and number formatting as lbflow, so a
run on the GPU writes a case directory the reference's own post-processing
(and its `FromTimeStep` / `FromFile` restarts) reads unchanged:

  data/coordinates.csv           i,j[,k],x,y[,z],node_type   (src/momentum/io.rs:191-224)
  data/<t>/density.csv           density                     (src/momentum/io.rs:113-143)
  data/<t>/velocity.csv          velocity_x,velocity_y[,velocity_z] (io.rs:145-189)
  data/<t>/<scalar>.csv          <scalar>                    (src/passive_scalar/io.rs:76-108)
  data/residuals.csv             time_step,density,velocity_x,...[,<scalar>...]
                                                             (src/momentum/io.rs:53-72,
                                                              src/passive_scalar/io.rs:112-152)
  post_processing/<file>         time_step,<name>...         (src/momentum/io.rs:226-248)
  post_processing/vtk_files/*.vtk  ASCII STRUCTURED_GRID     (src/momentum/post/vtk.rs:89-190)
  pre_processing/map.xyz         solid mask                  (src/io.rs:121-143, geometry.py)

Numbers use Rust's `{:.8e}` (LowerExp): mantissa with 8 decimals and an
exponent without sign padding or leading zeros ("1.00000000e0",
"-2.50000000e-7"); NaN prints as "NaN". Reference files are written per
thread chunk and unified afterwards (src/io.rs:232-248); the unified file is
what is written here directly.

`run_case` mirrors the reference's time loop with its I/O (src/momentum.rs:
262-292): coordinates, then per step the data files every
`write_data_frequency` steps (and at step 0), post-processing, the residuals
line, and the stop rule.
"""
from __future__ import annotations

import os
import re
from typing import Callable, Iterable, Optional, Sequence

import numpy as np

DATA_PATH = "data"
PRE_PROCESSING_PATH = "pre_processing"
POST_PROCESSING_PATH = "post_processing"
VTK_PATH = os.path.join("post_processing", "vtk_files")
COORDINATES_FILE = "coordinates.csv"
DENSITY_FILE = "density.csv"
VELOCITY_FILE = "velocity.csv"
RESIDUALS_FILE = "residuals.csv"
BOUNCE_BACK_MAP_FILE = "map.xyz"

MIN_ITER = 10          # src/constants.rs:15
TOLERANCE = 1e-7       # src/constants.rs:17-21

_EXP = re.compile(r"e\+?(-?)0*(\d+)")
_AXES = ("x", "y", "z")


def fmt_e(values, prec: int = 8) -> np.ndarray:
    """Rust `{:.<prec>e}` of every value (array of str)."""
    a = np.asarray(values, dtype=np.float64)
    s = np.char.mod(f"%.{prec}e", a.reshape(-1))
    out = np.array([_EXP.sub(r"e\1\2", v) for v in s.tolist()], dtype=object)
    out[np.isnan(a.reshape(-1))] = "NaN"
    return out.reshape(a.shape)


def _join_rows(cols: Sequence[np.ndarray]) -> str:
    if len(cols) == 1:
        return "\n".join(cols[0].tolist())
    return "\n".join(",".join(row) for row in zip(*[c.tolist() for c in cols]))


def case_setup(root: str) -> None:
    """Creates data/, pre_processing/, post_processing/vtk_files/ (src/io.rs:37-58)."""
    for p in (DATA_PATH, PRE_PROCESSING_PATH, POST_PROCESSING_PATH, VTK_PATH):
        os.makedirs(os.path.join(root, p), exist_ok=True)


def node_indices(shape) -> np.ndarray:
    """[N, d] integer indices in reference node order (src/momentum/lattice.rs:70-85)."""
    d = len(shape)
    N = int(np.prod(shape))
    i = np.arange(N)
    idx = [i % shape[0]]
    if d == 2:
        idx.append(i // shape[0])
    else:
        idx.append((i // shape[0]) % shape[1])
        idx.append(i // (shape[0] * shape[1]))
    return np.stack(idx, axis=1)


def write_coordinates(root: str, shape, node_type: Optional[np.ndarray], delta_x: float = 1.0):
    """data/coordinates.csv: indices, cell-centre coordinates (x + 0.5) * delta_x
    (src/momentum/lattice.rs:86-89) and node type."""
    d = len(shape)
    idx = node_indices(shape)
    N = idx.shape[0]
    types = np.zeros(N, np.uint8) if node_type is None else np.asarray(node_type).reshape(-1)
    coords = (idx.astype(np.float64) + 0.5) * delta_x
    header = "i,j,x,y,node_type" if d == 2 else "i,j,k,x,y,z,node_type"
    cols = [idx[:, a].astype(str).astype(object) for a in range(d)]
    cols += [fmt_e(coords[:, a]) for a in range(d)]
    cols.append(np.where(types.astype(bool), "solid", "fluid").astype(object))
    path = os.path.join(root, DATA_PATH, COORDINATES_FILE)
    with open(path, "w") as fh:
        fh.write(header + "\n" + _join_rows(cols) + "\n")
    return path


def step_dir(root: str, time_step: int) -> str:
    p = os.path.join(root, DATA_PATH, str(int(time_step)))
    os.makedirs(p, exist_ok=True)
    return p


def write_density(root: str, time_step: int, rho: np.ndarray) -> str:
    path = os.path.join(step_dir(root, time_step), DENSITY_FILE)
    with open(path, "w") as fh:
        fh.write("density\n" + _join_rows([fmt_e(np.asarray(rho).reshape(-1))]) + "\n")
    return path


def write_velocity(root: str, time_step: int, u: np.ndarray) -> str:
    u = np.asarray(u)
    d = u.shape[1]
    header = ",".join(f"velocity_{_AXES[a]}" for a in range(d))
    path = os.path.join(step_dir(root, time_step), VELOCITY_FILE)
    with open(path, "w") as fh:
        fh.write(header + "\n" + _join_rows([fmt_e(u[:, a]) for a in range(d)]) + "\n")
    return path


def write_scalar(root: str, time_step: int, name: str, values: np.ndarray) -> str:
    path = os.path.join(step_dir(root, time_step), f"{name}.csv")
    with open(path, "w") as fh:
        fh.write(name + "\n" + _join_rows([fmt_e(np.asarray(values).reshape(-1))]) + "\n")
    return path


def append_residuals(root: str, time_step: int, r_rho: float, r_u: Sequence[float],
                     r_c: Sequence[float] = (), scalar_names: Sequence[str] = ()) -> None:
    """One line of data/residuals.csv, the header at step 0 (src/io.rs:262-271)."""
    path = os.path.join(root, DATA_PATH, RESIDUALS_FILE)
    with open(path, "a") as fh:
        if time_step == 0:
            head = "time_step,density," + ",".join(f"velocity_{_AXES[a]}" for a in range(len(r_u)))
            head += "".join(f",{n}" for n in scalar_names)
            fh.write(head + "\n")
        vals = fmt_e(np.array([r_rho, *r_u, *r_c], dtype=np.float64)).tolist()
        fh.write(f"{int(time_step)}," + ",".join(vals) + "\n")


def append_post(root: str, file_name: str, time_step: int, results: Sequence[tuple]) -> None:
    """post_processing/<file_name>: `time_step,<name>...` at step 0, then one
    line per call (src/momentum/io.rs:226-248). results: [(name, value)]."""
    path = os.path.join(root, POST_PROCESSING_PATH, file_name)
    with open(path, "a") as fh:
        if time_step == 0:
            fh.write("time_step" + "".join(f",{n}" for n, _ in results) + "\n")
        vals = fmt_e(np.array([v for _, v in results], dtype=np.float64)).tolist()
        fh.write(f"{int(time_step)}" + "".join(f",{v}" for v in vals) + "\n")


# ---------------------------------------------------------------- reading back
def read_csv_values(path: str) -> list:
    """Data lines of a CSV file without its header (src/io.rs:150-160)."""
    with open(path) as fh:
        lines = fh.read().strip().split("\n")
    return lines[1:]


def read_scalars(path: str) -> np.ndarray:
    return np.array([float(s) for s in read_csv_values(path)], dtype=np.float64)


def read_vectors(path: str) -> np.ndarray:
    return np.array([[float(v) for v in s.split(",")] for s in read_csv_values(path)],
                    dtype=np.float64)


def density_from_time_step(root: str, time_step: int) -> np.ndarray:
    """`InitialDensity::FromTimeStep` (src/momentum/io.rs:266-281)."""
    return read_scalars(os.path.join(root, DATA_PATH, str(int(time_step)), DENSITY_FILE))


def velocity_from_time_step(root: str, time_step: int) -> np.ndarray:
    """`InitialVelocity::FromTimeStep` (src/momentum/io.rs:283-300)."""
    return read_vectors(os.path.join(root, DATA_PATH, str(int(time_step)), VELOCITY_FILE))


def scalar_from_time_step(root: str, time_step: int, name: str) -> np.ndarray:
    """`InitialScalarValue::FromTimeStep` (src/passive_scalar/io.rs:154-171)."""
    return read_scalars(os.path.join(root, DATA_PATH, str(int(time_step)), f"{name}.csv"))


def collect_time_steps(root: str) -> list:
    """Numeric step directories under data/, sorted (src/io.rs:162-178)."""
    out = []
    for e in os.listdir(os.path.join(root, DATA_PATH)):
        if e.isdigit() and os.path.isdir(os.path.join(root, DATA_PATH, e)):
            out.append(int(e))
    return sorted(out)


# ------------------------------------------------------------------------- VTK
def _vtk_head(fh, shape, coords):
    n = list(shape) + [1] * (3 - len(shape))
    N = int(np.prod(shape))
    fh.write("# vtk DataFile Version 3.0\nlbflow simulation data\nASCII\n"
             "DATASET STRUCTURED_GRID\n")
    fh.write(f"DIMENSIONS {n[0]} {n[1]} {n[2]}\nPOINTS {N} float\n")
    c = np.zeros((N, 3))
    c[:, : coords.shape[1]] = coords
    fh.write(_join_rows([fmt_e(c[:, 0], 4), fmt_e(c[:, 1], 4), fmt_e(c[:, 2], 4)])
             .replace(",", " ") + "\n")
    fh.write(f"POINT_DATA {N}\n")


def write_vtk_momentum(path: str, shape, rho: np.ndarray, u: np.ndarray, delta_x: float = 1.0,
                       physical: Optional[dict] = None) -> None:
    """ASCII structured-grid VTK with density and velocity (src/momentum/post/vtk.rs:131-190);
    physical = {"reference_pressure", "pressure_factor", "velocity_factor"} adds the
    physical pressure p_ref + c_s^2 (rho - 1) * factor and the scaled velocity."""
    coords = (node_indices(shape).astype(np.float64) + 0.5) * delta_x
    u = np.asarray(u)
    uu = np.zeros((u.shape[0], 3))
    uu[:, : u.shape[1]] = u
    with open(path, "w") as fh:
        _vtk_head(fh, shape, coords)
        fh.write("SCALARS density float 1\nLOOKUP_TABLE default\n")
        fh.write(_join_rows([fmt_e(np.asarray(rho).reshape(-1))]) + "\n")
        fh.write("VECTORS velocity float\n")
        fh.write(_join_rows([fmt_e(uu[:, a]) for a in range(3)]).replace(",", " ") + "\n")
        if physical:
            p = physical["reference_pressure"] + (np.asarray(rho) - 1.0) / 3.0 * physical["pressure_factor"]
            fh.write("SCALARS physical_pressure float 1\nLOOKUP_TABLE default\n")
            fh.write(_join_rows([fmt_e(p.reshape(-1))]) + "\n")
            fh.write("VECTORS physical_velocity float\n")
            pv = uu * physical["velocity_factor"]
            fh.write(_join_rows([fmt_e(pv[:, a]) for a in range(3)]).replace(",", " ") + "\n")


def write_vtk_node_type(path: str, shape, node_type: Optional[np.ndarray],
                        delta_x: float = 1.0) -> None:
    """Node-type VTK (src/momentum/post/vtk.rs:89-129)."""
    coords = (node_indices(shape).astype(np.float64) + 0.5) * delta_x
    N = int(np.prod(shape))
    t = np.zeros(N, np.uint8) if node_type is None else np.asarray(node_type).reshape(-1)
    with open(path, "w") as fh:
        _vtk_head(fh, shape, coords)
        fh.write("SCALARS node_type float 1\nLOOKUP_TABLE default\n")
        fh.write("\n".join("1" if v else "0" for v in t.tolist()) + "\n")


# -------------------------------------------------------------------- time loop
def run_case(root: str, case, max_iterations: int, write_data_frequency: int,
             scalar_names: Sequence[str] = (),
             post_functions: Iterable = (), delta_x: float = 1.0,
             log: Optional[Callable[[str], None]] = None) -> int:
    """The reference's `run` loop with its I/O (src/momentum.rs:262-292,
    src/passive_scalar.rs:136-177) over a device case (engine.Case):
    per step t: main_steps (with residuals), data files when t % n == 0 or
    t == 0, post functions, the residuals line, then the stop rule
    (t > MIN_ITER and all residuals <= 1e-7) or t > max_iterations.
    post_functions: [(file_name, interval, fn(case) -> [(name, value)])].
    Returns the number of steps run."""
    spec = case.spec
    case_setup(root)
    write_coordinates(root, spec.shape, spec.node_type, delta_x)
    case.initialize_nodes()
    t = 0
    while True:
        case.main_steps(1, residuals=True)
        if t % write_data_frequency == 0 or t == 0:
            rho, u = case.fields()
            write_density(root, t, rho)
            write_velocity(root, t, u)
            for s, name in enumerate(scalar_names):
                c, _ = case.scalar(s)
                write_scalar(root, t, name, c)
        for file_name, interval, fn in post_functions:
            if t % interval == 0:
                append_post(root, file_name, t, fn(case))
        r_rho, r_u, r_c = case.residuals()
        append_residuals(root, t, r_rho, r_u, r_c[: len(scalar_names)], scalar_names)
        if log:
            log(f"{t:>10d} {r_rho:16.8e}" + "".join(f" {x:16.8e}" for x in r_u))
        converged = r_rho <= TOLERANCE and all(x <= TOLERANCE for x in r_u) and \
            all(x <= TOLERANCE for x in r_c[: len(scalar_names)])
        if (t > MIN_ITER and converged) or t > max_iterations:
            return t + 1
        t += 1


def mean_density(case):
    """`compute_mean_density` as a post function (src/momentum/post.rs:43-58),
    reduced in FP64 on the device."""
    from . import abi
    return [("mean_density", case.reduce(abi.REDUCE_MEAN_DENSITY))]
