"""Scene generators and settings for the MPS step.

The 2D lattice generator restates the reference's pre-processing programs
(inputs/dambreak/main.cpp:26-113, same file in inputs/dambreak_fine and
inputs/hydrostatic with other constants): particles sit at ((i+0.5) l0,
(j+0.5) l0), classified dummy wall / wall / empty / fluid by overlapping
boxes with a 0.01 l0 tolerance, ids assigned in generation order (x outer,
y inner), open top.  The 3D generator extends the same rules with a z axis
(walls on five sides).  Everything is FP64 IEEE arithmetic in the same
operation order as the C++ generator, so coordinates are bit-identical.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Tuple

import numpy as np

GHOST, FLUID, WALL, DUMMY_WALL = -1, 0, 1, 2
BC_IGNORED, BC_SURFACE, BC_INNER = 0, 1, 2


class MpsSettings(ctypes.Structure):
    """ctypes mirror of ``mps_settings`` in include/mps_amd.h."""

    _fields_ = [
        ("dim", ctypes.c_int32),
        ("qc_on", ctypes.c_int32),
        ("hos_on", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("particle_distance", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("finish_time", ctypes.c_double),
        ("output_interval", ctypes.c_double),
        ("cfl_condition", ctypes.c_double),
        ("domain_min", ctypes.c_double * 3),
        ("domain_max", ctypes.c_double * 3),
        ("density", ctypes.c_double),
        ("kinematic_viscosity", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("re_pressure", ctypes.c_double),
        ("re_viscosity", ctypes.c_double),
        ("re_surface", ctypes.c_double),
        ("re_max", ctypes.c_double),
        ("surface_threshold", ctypes.c_double),
        ("compressibility", ctypes.c_double),
        ("hos_gamma", ctypes.c_double),
        ("relaxation", ctypes.c_double),
        ("collision_distance", ctypes.c_double),
        ("restitution", ctypes.c_double),
    ]


@dataclasses.dataclass
class Settings:
    """The reference's Settings after Settings::load (src/settings.cpp:21-84):
    radius ratios are kept as ratios here and multiplied by l0 in
    :meth:`to_c`, exactly as settings.cpp:43-56 does."""

    dim: int = 2
    particle_distance: float = 0.01
    dt: float = 0.001
    finish_time: float = 2.0
    output_interval: float = 0.05
    cfl_condition: float = 0.3
    domain_x: Tuple[float, float] = (-0.1, 0.4)
    domain_y: Tuple[float, float] = (-0.1, 0.35)
    domain_z: Tuple[float, float] = (0.0, 0.0)
    density: float = 1000.0
    kinematic_viscosity: float = 1.0e-6
    gravity: Tuple[float, float, float] = (0.0, -9.8, 0.0)
    ratio_pressure: float = 3.1
    ratio_viscosity: float = 3.1
    ratio_surface: float = 3.1
    surface_threshold: float = 0.97
    qc_on: bool = True
    compressibility: float = 0.45e-9
    hos_on: bool = False
    hos_gamma: float = 1.0e-3
    relaxation: float = 0.2
    collision_ratio: float = 0.5
    restitution: float = 0.2
    input_csv: str = ""

    def to_c(self) -> MpsSettings:
        s = MpsSettings()
        l0 = float(self.particle_distance)
        s.dim = int(self.dim)
        s.qc_on = int(bool(self.qc_on))
        s.hos_on = int(bool(self.hos_on))
        s.particle_distance = l0
        s.dt = self.dt
        s.finish_time = self.finish_time
        s.output_interval = self.output_interval
        s.cfl_condition = self.cfl_condition
        for k, rng in enumerate((self.domain_x, self.domain_y, self.domain_z)):
            s.domain_min[k] = float(rng[0])
            s.domain_max[k] = float(rng[1])
        s.density = self.density
        s.kinematic_viscosity = self.kinematic_viscosity
        for k in range(3):
            s.gravity[k] = float(self.gravity[k])
        s.re_pressure = float(self.ratio_pressure) * l0
        s.re_viscosity = float(self.ratio_viscosity) * l0
        s.re_surface = float(self.ratio_surface) * l0
        s.re_max = max(s.re_pressure, s.re_viscosity, s.re_surface)
        s.surface_threshold = self.surface_threshold
        s.compressibility = self.compressibility if self.qc_on else 0.0
        s.hos_gamma = self.hos_gamma if self.hos_on else 0.0
        s.relaxation = self.relaxation
        s.collision_distance = float(self.collision_ratio) * l0
        s.restitution = self.restitution
        return s


def _c_round(x: float) -> int:
    """C round(): half away from zero (Python's round() is half-even)."""
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def _inside(v, lo, hi, eps):
    # isInside (inputs/dambreak/main.cpp:103-113): lo - eps < v < hi + eps
    return (lo - eps < v) & (v < hi + eps)


@dataclasses.dataclass
class Scene:
    type: np.ndarray  # int8 (N,)
    pos: np.ndarray   # float64 (N, 3)
    vel: np.ndarray   # float64 (N, 3)

    @property
    def n(self) -> int:
        return int(self.type.shape[0])

    def counts(self):
        return {t: int(np.count_nonzero(self.type == t)) for t in (FLUID, WALL, DUMMY_WALL)}


def lattice_2d(l0: float, tank: Tuple[float, float], fluid_x: Tuple[float, float],
               fluid_y: Tuple[float, float], wall_layers: int = 1,
               dummy_layers: int = 2) -> Scene:
    """2D tank (inner box [0,Lx]x[0,Ly]) with a fluid block, per
    inputs/dambreak/main.cpp:26-80."""
    eps = 0.01 * l0
    all_layers = dummy_layers + wall_layers
    x_max, y_max = float(tank[0]), float(tank[1])
    ix0 = _c_round(0.0 / l0) - all_layers
    ix1 = _c_round(x_max / l0) + all_layers
    iy0 = _c_round(0.0 / l0) - all_layers
    iy1 = _c_round(y_max / l0) + all_layers
    ix = np.arange(ix0, ix1 + 1, dtype=np.int64)
    iy = np.arange(iy0, iy1 + 1, dtype=np.int64)
    gx, gy = np.meshgrid(ix, iy, indexing="ij")  # x outer, y inner
    px = (gx.astype(np.float64) + 0.5) * l0
    py = (gy.astype(np.float64) + 0.5) * l0
    t = np.full(px.shape, GHOST, dtype=np.int8)
    fx0, fx1 = sorted(fluid_x)
    fy0, fy1 = sorted(fluid_y)
    # dummy wall, wall, empty, fluid regions in the generator's order
    m = _inside(px, -all_layers * l0, x_max + all_layers * l0, eps) & _inside(py, -all_layers * l0, y_max, eps)
    t[m] = DUMMY_WALL
    m = _inside(px, -wall_layers * l0, x_max + wall_layers * l0, eps) & _inside(py, -wall_layers * l0, y_max, eps)
    t[m] = WALL
    m = _inside(px, 0.0, x_max, eps) & _inside(py, 0.0, y_max, eps)
    t[m] = GHOST
    m = _inside(px, fx0, fx1, eps) & _inside(py, fy0, fy1, eps)
    t[m] = FLUID
    keep = (t != GHOST).reshape(-1)
    typ = t.reshape(-1)[keep]
    pos = np.zeros((typ.shape[0], 3), dtype=np.float64)
    pos[:, 0] = px.reshape(-1)[keep]
    pos[:, 1] = py.reshape(-1)[keep]
    return Scene(typ.astype(np.int8), pos, np.zeros_like(pos))


def lattice_3d(l0: float, tank: Tuple[float, float, float], fluid_x, fluid_y, fluid_z,
               wall_layers: int = 2, dummy_layers: int = 2) -> Scene:
    """3D tank (inner box [0,Lx]x[0,Ly]x[0,Lz], y vertical, open top) with the
    2D generator's rules extended to z (walls on five sides)."""
    eps = 0.01 * l0
    all_layers = dummy_layers + wall_layers
    x_max, y_max, z_max = (float(v) for v in tank)
    ix = np.arange(_c_round(0.0) - all_layers, _c_round(x_max / l0) + all_layers + 1, dtype=np.int64)
    iy = np.arange(_c_round(0.0) - all_layers, _c_round(y_max / l0) + all_layers + 1, dtype=np.int64)
    iz = np.arange(_c_round(0.0) - all_layers, _c_round(z_max / l0) + all_layers + 1, dtype=np.int64)
    fx0, fx1 = sorted(fluid_x)
    fy0, fy1 = sorted(fluid_y)
    fz0, fz1 = sorted(fluid_z)
    types, poss = [], []
    # x outer, then y, then z innermost; chunk over x to bound memory
    pyz_y, pyz_z = np.meshgrid((iy.astype(np.float64) + 0.5) * l0, (iz.astype(np.float64) + 0.5) * l0,
                               indexing="ij")
    pyz_y = pyz_y.reshape(-1)
    pyz_z = pyz_z.reshape(-1)
    for gx in ix:
        px = np.full(pyz_y.shape, (float(gx) + 0.5) * l0)
        t = np.full(pyz_y.shape, GHOST, dtype=np.int8)
        m = (_inside(px, -all_layers * l0, x_max + all_layers * l0, eps)
             & _inside(pyz_y, -all_layers * l0, y_max, eps)
             & _inside(pyz_z, -all_layers * l0, z_max + all_layers * l0, eps))
        t[m] = DUMMY_WALL
        m = (_inside(px, -wall_layers * l0, x_max + wall_layers * l0, eps)
             & _inside(pyz_y, -wall_layers * l0, y_max, eps)
             & _inside(pyz_z, -wall_layers * l0, z_max + wall_layers * l0, eps))
        t[m] = WALL
        m = _inside(px, 0.0, x_max, eps) & _inside(pyz_y, 0.0, y_max, eps) & _inside(pyz_z, 0.0, z_max, eps)
        t[m] = GHOST
        m = _inside(px, fx0, fx1, eps) & _inside(pyz_y, fy0, fy1, eps) & _inside(pyz_z, fz0, fz1, eps)
        t[m] = FLUID
        keep = t != GHOST
        types.append(t[keep])
        poss.append(np.stack([px[keep], pyz_y[keep], pyz_z[keep]], axis=1))
    typ = np.concatenate(types).astype(np.int8)
    pos = np.concatenate(poss).astype(np.float64)
    return Scene(typ, pos, np.zeros_like(pos))


# ---- the reference's tutorial inputs (inputs/*/main.cpp constants) ----

def tutorial_dambreak() -> Scene:
    return lattice_2d(0.01, (0.3, 0.25), (0.0, 0.1), (0.0, 0.2), 1, 2)


def tutorial_dambreak_fine() -> Scene:
    return lattice_2d(0.004, (1.0, 0.25), (0.0, 0.4), (0.0, 0.2), 1, 2)


def tutorial_hydrostatic() -> Scene:
    return lattice_2d(0.005, (0.2, 0.35), (0.0, 0.2), (0.0, 0.3), 1, 2)


# ---- benchmark / parity workloads ----

def dambreak_settings(l0: float, dt: float, finish: float, nu: float = 1.0e-6, dim: int = 2,
                      restitution: float = 0.2, relaxation: float = 0.2) -> Settings:
    """Dam-break physics shared by the workloads: rho 1000, g=(0,-9.8,0),
    radii 2.1 l0 (pressure: density, Laplacian, gradient), 3.1 l0
    (viscosity), 2.1 l0 (surface detection), beta 0.97, quasi
    compressibility 0.45e-9, collision 0.5 l0.  The reference has a single
    pressure radius (src/settings.cpp:43-46), so 2.1 l0 also serves the PPE
    Laplacian; re_max = 3.1 l0 sizes the bucket grid."""
    s = Settings(dim=dim, particle_distance=l0, dt=dt, finish_time=finish, output_interval=1.0,
                 cfl_condition=0.3, domain_x=(-0.1, 1.1), domain_y=(-0.1, 0.7),
                 domain_z=(0.0, 0.0) if dim == 2 else (-0.1, 0.4), kinematic_viscosity=nu,
                 ratio_pressure=2.1, ratio_viscosity=3.1, ratio_surface=2.1, restitution=restitution,
                 relaxation=relaxation)
    return s


def dambreak_2d(l0: float, wall_layers: int = 2, dummy_layers: int = 2) -> Scene:
    """Tank 1.0 x 0.6 m, water column 0.25 x 0.5 m against the left wall."""
    return lattice_2d(l0, (1.0, 0.6), (0.0, 0.25), (0.0, 0.5), wall_layers, dummy_layers)


def workload(name: str):
    """(settings, scene, steps) for a named workload.

    dambreak2d_584:    l0 25 mm, 584 particles, dt 1e-3, 1000 steps.
    dambreak2d_2M:     l0 0.25 mm, 2,035,232 particles, dt 1e-5, 500 steps.
    dambreak3d_3.6M:   l0 2.5 mm, 3,622,656 particles, dt 1e-5, 300 steps.
    dambreak2d_35k:    l0 2 mm, 35,682 particles, nu 1e-3, dt 1e-4, 2000 steps.
    """
    if name == "dambreak2d_584":
        return dambreak_settings(0.025, 1.0e-3, 1.0005), dambreak_2d(0.025), 1000
    if name == "dambreak2d_2M":
        return dambreak_settings(0.00025, 1.0e-5, 0.005), dambreak_2d(0.00025), 500
    if name == "dambreak3d_3.6M":
        sc = lattice_3d(0.0025, (1.0, 0.6, 0.3), (0.0, 0.25), (0.0, 0.5), (0.0, 0.3), 2, 2)
        return dambreak_settings(0.0025, 1.0e-5, 0.003005, dim=3), sc, 300
    if name == "dambreak2d_35k":
        return dambreak_settings(0.002, 1.0e-4, 0.2, nu=1.0e-3), dambreak_2d(0.002), 2000
    raise KeyError(name)


def steps_between(t0: float, dt: float, finish: float) -> int:
    """Number of steps the driver loop takes (src/simulation.cpp:39-43):
    while (time + dt <= finish) { time += dt; ... }."""
    t, k = float(t0), 0
    while t + dt <= finish:
        t += dt
        k += 1
    return k


# ---- the reference's CSV format (src/exporter.cpp:10-55, src/mps.cpp:38-94) ----

def read_csv(path: str):
    """Reads an input/output CSV: (time, Scene, pressure, ndr).  Row 1 col 2
    is the time, row 3 the header, then one row per particle."""
    with open(path, "r") as f:
        first = f.readline().strip().split(",")
        t0 = float(first[1])
        f.readline()
        header = f.readline().strip().split(",")
        data = np.loadtxt(f, delimiter=",", ndmin=2)
    col = {h: k for k, h in enumerate(header)}
    ids = data[:, col["ID"]].astype(np.int64)
    if not np.array_equal(ids, np.arange(ids.shape[0])):
        raise ValueError("particle ids must be 0..N-1 in order")
    typ = data[:, col["Type"]].astype(np.int8)
    pos = np.stack([data[:, col["Position.x (m)"]], data[:, col["Position.y (m)"]],
                    data[:, col["Position.z (m)"]]], axis=1)
    vel = np.stack([data[:, col["Velocity.x (m/s)"]], data[:, col["Velocity.y (m/s)"]],
                    data[:, col["Velocity.z (m/s)"]]], axis=1)
    p = data[:, col["Pressure (Pa)"]] if "Pressure (Pa)" in col else np.zeros(ids.shape[0])
    ndr = data[:, col["Number Density Ratio"]] if "Number Density Ratio" in col else np.zeros(ids.shape[0])
    return t0, Scene(typ, pos.astype(np.float64), vel.astype(np.float64)), p, ndr


# ---- ensembles of independent scenes ----

class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): a documented 64-bit generator, so
    an ensemble drawn from a seed is reproducible from any language."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))


def ensemble_members(seed: int = 2024, count: int = 64, l0_choices=(0.004, 0.002, 0.001)):
    """Independent 2D scenes: half dam breaks (tank 1.0 x 0.6, column against
    the left wall), half hydrostatic tanks (the column fills the tank's width,
    tank 0.1 m taller than the column).  Column width U[0.15, 0.35] m, height
    U[0.3, 0.6] m (rounded to whole particle rows), l0 drawn from l0_choices,
    nu log-uniform in [1e-6, 1e-3], dt = 0.2 l0 / 3.5 m/s.  Returns a list of
    (settings, scene, kind)."""
    rng = SplitMix64(seed)
    out = []
    for k in range(count):
        kind = "dambreak" if k % 2 == 0 else "hydrostatic"
        w = 0.15 + 0.2 * rng.uniform()
        hgt = 0.3 + 0.3 * rng.uniform()
        l0 = l0_choices[min(int(rng.uniform() * len(l0_choices)), len(l0_choices) - 1)]
        nu = 10.0 ** (-6.0 + 3.0 * rng.uniform())
        nx = max(2, int(round(w / l0)))
        ny = max(2, int(round(hgt / l0)))
        w, hgt = nx * l0, ny * l0
        dt = 0.2 * l0 / 3.5
        if kind == "dambreak":
            tank = (1.0, 0.6)
        else:
            tank = (w, hgt + 0.1)
        sc = lattice_2d(l0, tank, (0.0, w), (0.0, hgt), 2, 2)
        s = Settings(dim=2, particle_distance=l0, dt=dt, finish_time=400 * dt, output_interval=1.0,
                     cfl_condition=0.3, domain_x=(-0.1, tank[0] + 0.1), domain_y=(-0.1, tank[1] + 0.1),
                     kinematic_viscosity=nu, ratio_pressure=2.1, ratio_viscosity=3.1, ratio_surface=2.1)
        out.append((s, sc, kind))
    return out
