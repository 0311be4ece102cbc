"""Node-type masks and the reference's bounce-back map format.

Masks are generated on the host and shipped as node_type[N] (0 fluid,
1 solid) in the reference node order i = x + nx*(y + ny*z), so every run of
a configuration sees bit-identical geometry.
"""
from __future__ import annotations

import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser of (x + golden gamma), elementwise on uint64."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + _G)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int) -> np.ndarray:
    """Counter-based uniforms u_i = splitmix64(seed ^ i) / 2^64, i = 0..n-1."""
    idx = np.arange(n, dtype=np.uint64) ^ np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
    return (splitmix64(idx) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def bernoulli_mask(shape, p_solid: float, seed: int) -> np.ndarray:
    """i.i.d. Bernoulli(p_solid) solid nodes; shape is (nx, ny[, nz])."""
    n = int(np.prod(shape))
    return (uniform01(seed, n) < p_solid).astype(np.uint8)


def discs_mask(nx: int, ny: int, count: int, rmin: float, rmax: float, seed: int) -> np.ndarray:
    """Random disc obstacles: centres uniform in the box, radii uniform in [rmin, rmax];
    a node is solid if (x - cx)^2 + (y - cy)^2 <= r^2 (node coordinates x, y)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cx = rng.uniform(0, nx, count)
    cy = rng.uniform(0, ny, count)
    r = rng.uniform(rmin, rmax, count)
    x, y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    solid = np.zeros((ny, nx), dtype=bool)
    for i in range(count):
        solid |= (x - cx[i]) ** 2 + (y - cy[i]) ** 2 <= r[i] ** 2
    return solid.astype(np.uint8).reshape(-1)


def sphere_mask(nx: int, ny: int, nz: int, center, radius: float) -> np.ndarray:
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    return (d2 <= radius * radius).astype(np.uint8).reshape(-1)


def write_map_xyz(path: str, node_type: np.ndarray, shape) -> None:
    """Writes the reference's pre_processing/map.xyz layout (src/io.rs:121-143):
    z-slices separated by a blank line; in a slice, lines top (y = ny-1) to
    bottom, values x ascending, "0" fluid / "1" solid."""
    nx, ny = shape[0], shape[1]
    nz = shape[2] if len(shape) > 2 else 1
    a = np.asarray(node_type).reshape(nz, ny, nx)
    slices = []
    for z in range(nz):
        lines = [" ".join(str(int(v)) for v in a[z, y]) for y in range(ny - 1, -1, -1)]
        slices.append("\n".join(lines))
    with open(path, "w") as fh:
        fh.write("\n\n".join(slices) + "\n")


def read_map_xyz(path: str) -> np.ndarray:
    """Reads map.xyz into node_type[N] in reference node order (src/io.rs:121-143)."""
    text = open(path).read().strip()
    out = []
    for sl in text.split("\n\n"):
        rows = [line.split() for line in sl.strip().splitlines()]
        for row in reversed(rows):
            for v in row:
                if v not in ("0", "1"):
                    raise ValueError(f"Invalid node type: {v}")
                out.append(int(v))
    return np.array(out, dtype=np.uint8)
