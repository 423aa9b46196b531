"""Scan workloads and analytic phantoms (host side).

A workload is a reference ``.prm`` (Src/FileParams.cpp:10 keys) plus the
synthetic scan that feeds it: per-view tube angles and couch positions in the
reference's file formats (float32, Src/DataStructs.cpp:138-160), an analytic
ellipsoid phantom and a Poisson noise level.  The reference ships no data, so
every input of the tests and the benchmark is made from these definitions.

Geometry conventions (shared with the GPU projector ``lct_project_ellipsoids``):
  * tube angle of view g:  b_g = (b0 + g * 2 pi / NProjTurn) mod 2 pi  (float32)
  * couch position:        z_g = z_start + g * PitchValue / NProjTurn  (float32)
  * source of view g at (-dso sin b_g, dso cos b_g, z_g); channel `is` looks
    along in-plane angle b_g - (is - ws) ds / dso; row `it` climbs
    (it - wt) dt / dso per unit in-plane length (curved, source-centred
    detector; ds, dt are pitches at the isocentre: detector pitch * dso / SDD).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _lib

SDD = 949.0          # source-to-detector distance of the listed workloads (mm)
DSO = 541.0          # source-to-isocentre distance (mm)
PITCH_ISO = 1.0 * DSO / SDD   # 1.0 mm detector pitch seen at the isocentre = 0.570074 mm
MU_WATER = 0.019     # 1/mm
RECON_SCALE = 2.0    # reference BP scale: 2 pi / npt over a 2 pi window of blended
                     # conjugate pairs = 2 x textbook parallel-beam FBP (see DESIGN)
ELL_STRIDE = 16


def hu(v: np.ndarray, scale: float = RECON_SCALE) -> np.ndarray:
    """Reference-scale reconstruction value -> Hounsfield units."""
    return 1000.0 * (np.asarray(v, np.float64) / scale - MU_WATER) / MU_WATER


def hu_per_unit(scale: float = RECON_SCALE) -> float:
    """HU per unit of reconstruction value (for error conversion)."""
    return 1000.0 / (scale * MU_WATER)


# ------------------------------------------------------------------ workload

@dataclasses.dataclass
class Workload:
    name: str
    ns: int
    nt: int
    n_proj_turn: int
    pitch: float          # dimensionless helical pitch (table feed / collimation)
    nx: int
    dx: float
    dz: float
    readings: int
    phantom: str          # "shepp_logan" | "random" | "water" | "random_spheres"
    seed: int
    n_ellipsoids: int = 10
    cylinder_radius: float = 0.0
    n_spheres: int = 0
    i0: float = 1e5
    noise_seed: int = 0
    window: int = _lib.WINDOW_NONE
    ds: float = PITCH_ISO
    dt: float = PITCH_ISO
    dso: float = DSO
    angle0: float = 0.25
    z_start: float = 0.0
    z_dir: int = 1
    view_offset: int = 0
    data_offset: int = 0
    offset_x: float = 0.0
    offset_y: float = 0.0
    phantom_scale: float = 95.0

    @property
    def central_channel(self) -> float:
        return (self.ns - 1) / 2.0

    @property
    def pitch_value(self) -> float:
        """Reference PitchValue: table feed per rotation in mm."""
        return float(np.float32(self.pitch * self.nt * self.dt))

    def prm(self) -> _lib.Prm:
        p = _lib.Prm()
        p.ns, p.nt = self.ns, self.nt
        p.ds, p.dt, p.pitch_value = self.ds, self.dt, self.pitch_value
        p.nx = p.ny = self.nx
        p.dx, p.dz = self.dx, self.dz
        p.offset_x, p.offset_y = self.offset_x, self.offset_y
        p.readings, p.view_offset, p.data_offset = self.readings, self.view_offset, self.data_offset
        p.r_src_to_iso, p.central_channel = self.dso, self.central_channel
        p.n_proj_turn = self.n_proj_turn
        return p

    def prm_text(self, raw_dir: str, out_dir: str, data_file="proj.raw", angle_file="angles.raw",
                 couch_file="couch.raw", out_name="image.raw") -> str:
        """A .prm file for the reference parser (all 23 keys)."""
        kv = [
            ("RawDataDir", raw_dir), ("DataInterpFile", data_file), ("ViewAngleFile", angle_file),
            ("CouchPosFile", couch_file), ("OutputDir", out_dir), ("OutputName", out_name),
            ("Ns", self.ns), ("Nt", self.nt), ("Ds", repr(float(np.float32(self.ds)))),
            ("Dt", repr(float(np.float32(self.dt)))), ("PitchValue", repr(self.pitch_value)),
            ("Nx", self.nx), ("Ny", self.nx), ("Dx", self.dx), ("Dz", self.dz),
            ("OffsetX", self.offset_x), ("OffsetY", self.offset_y), ("Readings", self.readings),
            ("ViewOffset", self.view_offset), ("DataOffset", self.data_offset),
            ("RSrcToIso", self.dso), ("CentralChannel", self.central_channel),
            ("NProjTurn", self.n_proj_turn),
        ]
        return "".join(f"{k}: {v}\n" for k, v in kv)

    def total_views(self) -> int:
        """Views in the files (angles/couch past ViewOffset, data past DataOffset)."""
        return self.readings

    def angles_couch(self):
        """(angles, couch) float32 arrays of `readings` views (already past ViewOffset)."""
        g = np.arange(self.readings, dtype=np.float64)
        ang = np.mod(self.angle0 + g * (2.0 * math.pi / self.n_proj_turn), 2.0 * math.pi)
        z = self.z_start + self.z_dir * g * (self.pitch_value / self.n_proj_turn)
        return ang.astype(np.float32), z.astype(np.float32)

    def z_center(self) -> float:
        _, z = self.angles_couch()
        return 0.5 * (float(z[0]) + float(z[-1]))

    def ellipsoids(self) -> np.ndarray:
        zc = self.z_center()
        if self.phantom == "shepp_logan":
            return shepp_logan_3d(self.seed, self.phantom_scale, zc)
        if self.phantom == "water":
            return water_cylinder(self.cylinder_radius or 80.0, MU_WATER)
        _, z = self.angles_couch()
        zr = (float(min(z[0], z[-1])), float(max(z[0], z[-1])))
        ell = random_ellipsoids(self.seed, self.n_ellipsoids, zr,
                                self.cylinder_radius or 200.0)
        parts = [ell]
        if self.n_spheres:
            parts.append(random_spheres(self.seed + 7, self.n_spheres, zr,
                                        0.8 * (self.cylinder_radius or 200.0)))
        if self.cylinder_radius:
            parts.append(water_cylinder(self.cylinder_radius, MU_WATER))
        return np.concatenate(parts, axis=0)


def write_scan_files(w: Workload, directory: str, fan: np.ndarray | None = None,
                     recon_type: str | None = None) -> str:
    """Writes the reference's input files for workload `w` into `directory`
    (trailing '/' added: the reference joins paths by concatenation):
      angles.raw / couch.raw   float32, ViewOffset leading values then Readings
      proj.raw                 float32 [DataOffset + Readings][nt][ns] (zeros if fan is None)
      scan.prm                 all 23 keys (plus ReconType when given)
    Returns the .prm path."""
    import os
    d = directory.rstrip("/") + "/"
    os.makedirs(d, exist_ok=True)
    a, c = w.angles_couch()
    pad = np.zeros(w.view_offset, np.float32)
    np.concatenate([pad + 9.0, a]).astype(np.float32).tofile(d + "angles.raw")
    np.concatenate([pad - 9.0, c]).astype(np.float32).tofile(d + "couch.raw")
    if fan is None:
        fan = np.zeros((w.data_offset + w.readings, w.nt, w.ns), np.float32)
    np.ascontiguousarray(fan, np.float32).tofile(d + "proj.raw")
    text = w.prm_text(d, d)
    if recon_type:
        text += f"ReconType: {recon_type}\n"
    with open(d + "scan.prm", "w") as f:
        f.write(text)
    return d + "scan.prm"


def _rot(phi: float, theta: float, psi: float) -> np.ndarray:
    """Body->world rotation Rz(phi) Rx(theta) Rz(psi) (radians)."""
    def rz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])

    def rx(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])

    return rz(phi) @ rx(theta) @ rz(psi)


def make_ellipsoid(c, axes, R, mu) -> np.ndarray:
    e = np.zeros(ELL_STRIDE)
    e[0:3] = c
    e[3:6] = axes
    e[6:15] = np.asarray(R, np.float64).reshape(9)
    e[15] = mu
    return e


# 3D Shepp-Logan head (Kak & Slaney ellipses extended to 3D, with Toft's
# higher-contrast densities): density, semi-axes (a, b, c), centre
# (x0, y0, z0) in units of the head radius, Euler angles (phi, theta, psi) deg.
_SL3D = [
    (1.0, .6900, .920, .810, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
    (-.8, .6624, .874, .780, 0.0, -.0184, 0.0, 0.0, 0.0, 0.0),
    (-.2, .1100, .310, .220, .22, 0.0, 0.0, -18.0, 0.0, 10.0),
    (-.2, .1600, .410, .280, -.22, 0.0, 0.0, 18.0, 0.0, 10.0),
    (.1, .2100, .250, .410, 0.0, .35, -.15, 0.0, 0.0, 0.0),
    (.1, .0460, .046, .050, 0.0, .1, .25, 0.0, 0.0, 0.0),
    (.1, .0460, .046, .050, 0.0, -.1, .25, 0.0, 0.0, 0.0),
    (.1, .0460, .023, .050, -.08, -.605, 0.0, 0.0, 0.0, 0.0),
    (.1, .0230, .023, .020, 0.0, -.606, 0.0, 0.0, 0.0, 0.0),
    (.1, .0230, .046, .020, .06, -.605, 0.0, 0.0, 0.0, 0.0),
]


def shepp_logan_3d(seed: int, scale_mm: float, z_centre: float, mu_max: float = 0.02,
                   perturb: float = 0.05) -> np.ndarray:
    """10-ellipsoid 3D Shepp-Logan; centres shifted by up to +-perturb of the
    semi-axes, semi-axes scaled by 1 +- perturb and Euler angles shifted by up
    to +-perturb * 180 deg, all uniform from `seed`; densities x mu_max so the
    attenuation spans 0 .. mu_max per mm."""
    rng = np.random.default_rng(seed)
    out = []
    for (A, a, b, c, x0, y0, z0, phi, th, psi) in _SL3D:
        axes = np.array([a, b, c]) * (1.0 + perturb * rng.uniform(-1, 1, 3))
        ctr = np.array([x0, y0, z0]) + perturb * rng.uniform(-1, 1, 3) * np.array([a, b, c])
        ang = np.radians([phi, th, psi]) + perturb * math.pi * rng.uniform(-1, 1, 3)
        R = _rot(*ang)
        out.append(make_ellipsoid(ctr * scale_mm + np.array([0.0, 0.0, z_centre]),
                                  axes * scale_mm, R, A * mu_max))
    return np.stack(out)


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def random_ellipsoids(seed: int, n: int, z_range, radius: float, semi=(5.0, 120.0),
                      mu=(0.0, 0.025)) -> np.ndarray:
    """n ellipsoids: semi-axes U(semi), uniform random orientation, additive
    attenuation U(mu), centres uniform in the disk of radius 0.8 * radius and
    in z_range."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        r = 0.8 * radius * math.sqrt(rng.uniform())
        a = rng.uniform(0, 2 * math.pi)
        zc = rng.uniform(*z_range)
        axes = rng.uniform(semi[0], semi[1], 3)
        out.append(make_ellipsoid([r * math.cos(a), r * math.sin(a), zc], axes,
                                  _random_rotation(rng), rng.uniform(*mu)))
    return np.stack(out) if out else np.zeros((0, ELL_STRIDE))


def random_spheres(seed: int, n: int, z_range, radius: float, r=(0.5, 2.0),
                   mu: float = 0.05) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        rr = radius * math.sqrt(rng.uniform())
        a = rng.uniform(0, 2 * math.pi)
        s = rng.uniform(*r)
        out.append(make_ellipsoid([rr * math.cos(a), rr * math.sin(a), rng.uniform(*z_range)],
                                  [s, s, s], np.eye(3), mu))
    return np.stack(out) if out else np.zeros((0, ELL_STRIDE))


def water_cylinder(radius: float, mu: float, half_length: float = 1.0e5) -> np.ndarray:
    """A cylinder along z as an ellipsoid with a very long z semi-axis."""
    return make_ellipsoid([0.0, 0.0, 0.0], [radius, radius, half_length], np.eye(3), mu)[None, :]


# ------------------------------------------------------------- the workloads

def _w(**kw) -> Workload:
    return Workload(**kw)


WORKLOADS = {
    # Plain FDK: 16 x 368 curved detector, 576 views/turn, pitch 1.0, 256^2 x 32
    # at 0.8 mm, Shepp-Logan, I0 = 1e5, Ram-Lak x Hann.  Six turns (3456
    # views): the planner needs more than the two turns of the prose
    # (Src/DataStructs.cpp:181-218).
    "helical_16x368_256x32": _w(name="helical_16x368_256x32", ns=368, nt=16, n_proj_turn=576,
                                pitch=1.0, nx=256, dx=0.8, dz=0.8, readings=3456,
                                phantom="shepp_logan", seed=1, i0=1e5, noise_seed=101,
                                window=_lib.WINDOW_HANN),
    # Clinical: 64 x 736, 1152 views/turn, pitch 0.8, 512^2 x 256 at 0.7 / 0.6 mm,
    # 40 random ellipsoids in a 250 mm water cylinder, I0 = 2e5, Ram-Lak x
    # Shepp-Logan.  8192 views give exactly 256 slices (8 blocks).
    "helical_64x736_512x256": _w(name="helical_64x736_512x256", ns=736, nt=64,
                                 n_proj_turn=1152, pitch=0.8, nx=512, dx=0.7, dz=0.6,
                                 readings=8192, phantom="random", seed=2, n_ellipsoids=40,
                                 cylinder_radius=250.0, i0=2e5, noise_seed=102,
                                 window=_lib.WINDOW_SHEPP_LOGAN),
    # Large detector: 96 x 896, 2304 views/turn, pitch 1.0, 1024^2 x 512 at 0.4 mm.
    "helical_96x896_1024x512": _w(name="helical_96x896_1024x512", ns=896, nt=96,
                                  n_proj_turn=2304, pitch=1.0, nx=1024, dx=0.4, dz=0.4,
                                  readings=12288, phantom="random_spheres", seed=4,
                                  n_ellipsoids=80, n_spheres=200, cylinder_radius=200.0,
                                  i0=1e5, noise_seed=104, window=_lib.WINDOW_SHEPP_LOGAN),
    # Learned-correction geometry: plain-FDK geometry at pitch 1.5, 128 slices.
    "helical_16x368_p15_256x128": _w(name="helical_16x368_p15_256x128", ns=368, nt=16,
                                     n_proj_turn=576, pitch=1.5, nx=256, dx=0.8, dz=0.8,
                                     readings=6000, phantom="random", seed=1000,
                                     n_ellipsoids=40, cylinder_radius=0.0, i0=2e5,
                                     noise_seed=1100, window=_lib.WINDOW_HANN),
    # Small geometry for oracle-speed tests: 8 x 64 detector at 2 mm, 96 views/turn.
    "tiny": _w(name="tiny", ns=64, nt=8, n_proj_turn=96, pitch=1.0, nx=32, dx=3.0, dz=2.0,
               readings=1100, phantom="shepp_logan", seed=5, i0=1e5, noise_seed=105,
               window=_lib.WINDOW_NONE, ds=2.0, dt=2.0, phantom_scale=30.0),
}


def workload(name: str, **overrides) -> Workload:
    return dataclasses.replace(WORKLOADS[name], **overrides)
