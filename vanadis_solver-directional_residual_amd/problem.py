"""Host-side problem setup for the Vanadis DR-HEX8 solve path.

Structured tensor-grid meshes exactly as the reference builds them, the
reference's own default case, and the seeded synthetic advection-diffusion
problems used for parity tests and benchmarks.

Reference routines restated (paths relative to the reference tree):
  siatka            src_v2026_2_1/VANADIS.FOR:226-244  (x-fastest node grid)
  w_el              VANADIS.FOR:246-276                 (ISS(E,8), 1-based)
  kod_el_dr         VANADIS.FOR:278-320                 (face codes)
  build_bc_nodes /
  mark_face_nodes   VANADIS.FOR:2026-2117 (V1 faces);
                    src_updates/vanadis_full_DR_OpenMP_timing.for:1719-1811 (V3)
  wsp_na_o          VANADIS.FOR:104-166                 (default 7x20x20 mesh)
  read_cfg defaults src_v2026_2_1/vanadis.cfg:6
  atmosfera         VANADIS.FOR:188-224                 (source in element 1083)
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# --------------------------------------------------------------------------
# deterministic seeded generator (SplitMix64, counter based, vectorised)
# --------------------------------------------------------------------------
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """`count` 64-bit outputs of SplitMix64 for (seed, stream), element i being
    mix(state0 + (i+1) * golden).  Identical to the sequential generator."""
    base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)
    i = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = base + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, count, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# --------------------------------------------------------------------------
# mesh (reference conventions)
# --------------------------------------------------------------------------
def siatka(dx: np.ndarray, dy: np.ndarray, dz: np.ndarray):
    """Node coordinates xx, yy, zz [nax*nay*naz], x fastest (VANADIS.FOR:226-244)."""
    nax, nay, naz = len(dx), len(dy), len(dz)
    xx = np.tile(np.asarray(dx, np.float64), nay * naz)
    yy = np.tile(np.repeat(np.asarray(dy, np.float64), nax), naz)
    zz = np.repeat(np.asarray(dz, np.float64), nax * nay)
    return xx, yy, zz


def w_el(nax: int, nay: int, naz: int) -> np.ndarray:
    """ISS as the Fortran column-major ISS(E,8) flattened: iss[j*E + e] is the
    1-based node of local node j+1 of element e+1 (VANADIS.FOR:246-276)."""
    ex, ey, ez = nax - 1, nay - 1, naz - 1
    i, j, k = np.meshgrid(np.arange(ex), np.arange(ey), np.arange(ez), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))  # x fastest
    n0 = 1 + i + nax * j + nax * nay * k
    off = np.array([0, 1, 1 + nax, nax, nax * nay, 1 + nax * nay, 1 + nax * nay + nax,
                    nax * nay + nax], dtype=np.int64)
    iss = (n0[None, :] + off[:, None]).astype(np.int32)
    return np.ascontiguousarray(iss.reshape(-1))


def kod_el_dr(nax: int, nay: int, naz: int) -> np.ndarray:
    """Face codes kod[f*E + e] (VANADIS.FOR:278-320): 1 z-min, 2 z-max, 3 x-min,
    4 x-max, 5 y-max, 6 y-min."""
    ex, ey, ez = nax - 1, nay - 1, naz - 1
    E = ex * ey * ez
    e = np.arange(E)
    i, j, k = e % ex, (e // ex) % ey, e // (ex * ey)
    kod = np.zeros((6, E), dtype=np.int32)
    kod[0, k == 0] = 1
    kod[1, k == ez - 1] = 2
    kod[2, i == 0] = 3
    kod[3, i == ex - 1] = 4
    kod[4, j == ey - 1] = 5
    kod[5, j == 0] = 6
    return np.ascontiguousarray(kod.reshape(-1))


_FACE_NODES_V1 = {1: (1, 2, 3, 4), 2: (5, 6, 7, 8), 3: (1, 4, 8, 5), 4: (2, 3, 7, 6),
                  5: (4, 3, 7, 8), 6: (1, 2, 6, 5)}


def build_bc_nodes(iss: np.ndarray, kod: np.ndarray, nax: int, nay: int, naz: int,
                   facemap: int = 1) -> np.ndarray:
    """Dirichlet mask on faces 1,2,4,5,6 (VANADIS.FOR:2026-2117).  facemap=3
    reproduces the V3 table where faces 5 and 6 carry each other's node sets
    (omp_timing.for:1793-1807)."""
    E = (nax - 1) * (nay - 1) * (naz - 1)
    iss2 = iss.reshape(8, E)
    kod2 = kod.reshape(6, E)
    mask = np.zeros(nax * nay * naz, dtype=np.int32)
    for f in (1, 2, 4, 5, 6):
        nodes = _FACE_NODES_V1[f]
        if facemap == 3 and f in (5, 6):
            nodes = _FACE_NODES_V1[11 - f]
        sel = kod2[f - 1] == f
        for ln in nodes:
            mask[iss2[ln - 1, sel] - 1] = 1
    return mask


# --------------------------------------------------------------------------
# the reference's default case (VANADIS.FOR:30-32, 104-166, vanadis.cfg)
# --------------------------------------------------------------------------
REF_DX = [0.0, 20.0, 35.0, 85.0, 135.0, 230.0, 420.0]
REF_DY = [-4000.0, -1500.0, -900.0, -700.0, -500.0, -400.0, -270.0, -180.0, -90.0, -25.0,
          25.0, 90.0, 180.0, 270.0, 400.0, 500.0, 700.0, 900.0, 1500.0, 4000.0]
REF_DZ = list(REF_DY)


@dataclasses.dataclass
class Problem:
    """One discrete problem: mesh, per-element fields, mask, sources."""
    name: str
    nax: int
    nay: int
    naz: int
    xx: np.ndarray
    yy: np.ndarray
    zz: np.ndarray
    iss: np.ndarray
    kod: np.ndarray
    mask: np.ndarray          # int32 [N] in {0,1}
    lam: np.ndarray           # [E]
    vel: np.ndarray           # [3E] SoA
    q_in: np.ndarray          # [E]
    P: float = 0.0
    alpha: float = 0.0
    qs: float = 0.0
    variant: int = 1
    scheme: int = 1           # 0 theta, 1 steady, 2 implicit Euler
    dt: float = 1.0

    @property
    def N(self) -> int:
        return self.nax * self.nay * self.naz

    @property
    def E(self) -> int:
        return (self.nax - 1) * (self.nay - 1) * (self.naz - 1)

    def axes(self):
        return (self.xx[: self.nax], self.yy[:: self.nax][: self.nay],
                self.zz[:: self.nax * self.nay][: self.naz])


def reference_default_case(Q=1.0, K=40.0, P=7e-4, V=(0.0, -5.0, 0.0), alpha=0.01,
                           variant=1, facemap=1, dt=100.0) -> Problem:
    """The reference's shipped case: 7x20x20 nodes, cfg values of
    vanadis.cfg:6, source Q in element 1083, Robin alpha on the ground face 3,
    theta scheme with dt = 100 s."""
    nax, nay, naz = len(REF_DX), len(REF_DY), len(REF_DZ)
    xx, yy, zz = siatka(REF_DX, REF_DY, REF_DZ)
    iss = w_el(nax, nay, naz)
    kod = kod_el_dr(nax, nay, naz)
    mask = build_bc_nodes(iss, kod, nax, nay, naz, facemap)
    E = (nax - 1) * (nay - 1) * (naz - 1)
    lam = np.full(E, K)
    vel = np.concatenate([np.full(E, V[0]), np.full(E, V[1]), np.full(E, V[2])])
    q_in = np.zeros(E)
    q_in[1083 - 1] = Q
    return Problem("reference_default", nax, nay, naz, xx, yy, zz, iss, kod, mask, lam, vel,
                   q_in, P=P, alpha=alpha, variant=variant, scheme=0, dt=dt)


# --------------------------------------------------------------------------
# seeded unit-box advection-diffusion problems
# --------------------------------------------------------------------------
def velocity_field(cx, cy, cz, seed: int, amp: float = 0.2, base=(1.0, 0.5, 0.1)):
    """v = base + amp * (1/3) sum_m a_m sin(2 pi k_m . x + phi_m), three modes
    with integer k_m in {1,2,3}^3, unit a_m and phase phi_m, all seed-drawn;
    amp bounds the total perturbation."""
    u = uniform01(seed, 3 * 7, stream=1)
    vx = np.full_like(cx, base[0])
    vy = np.full_like(cx, base[1])
    vz = np.full_like(cx, base[2])
    for m in range(3):
        r = u[7 * m: 7 * m + 7]
        kv = 1 + np.floor(3 * r[0:3]).astype(np.int64)
        th = math.acos(1 - 2 * r[3])
        ph = 2 * math.pi * r[4]
        a = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        phase = 2 * math.pi * r[5]
        s = np.sin(2 * math.pi * (kv[0] * cx + kv[1] * cy + kv[2] * cz) + phase) * (amp / 3.0)
        vx = vx + a[0] * s
        vy = vy + a[1] * s
        vz = vz + a[2] * s
    return vx, vy, vz


def unit_box(ex: int, ey: int, ez: int, *, vel_seed: int = 0, lam=1e-3,
             lam_range=None, lam_seed: int = 1, src_center=(0.3, 0.5, 0.5),
             src_sigma=0.1, scheme=1, dt=1.0, variant=1, P=0.0, name="unit_box",
             extent=(1.0, 1.0, 1.0)) -> Problem:
    """Unit box [0,1]^3 with ex x ey x ez elements (uniform tensor grid),
    per-element velocity and diffusivity at centroids, Gaussian source
    q_in,e = exp(-|c_e - c|^2 / (2 sigma^2)) vol_e, Dirichlet c = 0 on the inflow
    faces x=0, y=0, z=0.  lam_range=(lo, hi) draws log-uniform per element."""
    nax, nay, naz = ex + 1, ey + 1, ez + 1
    ax = np.arange(nax) * (extent[0] / ex)
    ay = np.arange(nay) * (extent[1] / ey)
    az = np.arange(naz) * (extent[2] / ez)
    xx, yy, zz = siatka(ax, ay, az)
    iss = w_el(nax, nay, naz)
    kod = kod_el_dr(nax, nay, naz)
    E = ex * ey * ez
    e = np.arange(E)
    i, j, k = e % ex, (e // ex) % ey, e // (ex * ey)
    cx = 0.5 * (ax[i] + ax[i + 1])
    cy = 0.5 * (ay[j] + ay[j + 1])
    cz = 0.5 * (az[k] + az[k + 1])
    vol = (ax[i + 1] - ax[i]) * (ay[j + 1] - ay[j]) * (az[k + 1] - az[k])
    vx, vy, vz = velocity_field(cx, cy, cz, vel_seed)
    if lam_range is None:
        lamv = np.full(E, float(lam))
    else:
        lo, hi = math.log10(lam_range[0]), math.log10(lam_range[1])
        lamv = 10.0 ** (lo + (hi - lo) * uniform01(lam_seed, E, stream=2))
    d2 = (cx - src_center[0]) ** 2 + (cy - src_center[1]) ** 2 + (cz - src_center[2]) ** 2
    q_in = np.exp(-d2 / (2.0 * src_sigma ** 2)) * vol
    n = np.arange(nax * nay * naz)
    gi, gj, gk = n % nax, (n // nax) % nay, n // (nax * nay)
    mask = ((gi == 0) | (gj == 0) | (gk == 0)).astype(np.int32)
    return Problem(name, nax, nay, naz, xx, yy, zz, iss, kod, mask, lamv,
                   np.concatenate([vx, vy, vz]), q_in, P=P, alpha=0.0,
                   variant=variant, scheme=scheme, dt=dt)


# Workloads (elements per axis).  The 32x32x16 case has 33*33*17 = 18,513
# nodes; its "35,937 nodes" description matches 32^3 elements instead.
CONFIG_SHAPES = {
    "c1": (32, 32, 16),
    "c2": (128, 128, 64),
    "c3": (256, 256, 64),
    "c4a": (96, 96, 48),
    "c5": (128, 128, 64),
}


def config_problem(name: str, **kw) -> Problem:
    if name in ("c1", "c2"):
        ex, ey, ez = CONFIG_SHAPES[name]
        return unit_box(ex, ey, ez, vel_seed=0, lam=1e-3, name=name, **kw)
    if name == "c3":
        ex, ey, ez = CONFIG_SHAPES[name]
        return unit_box(ex, ey, ez, vel_seed=0, lam_range=(1e-4, 1e-2), lam_seed=1,
                        name=name, **kw)
    raise KeyError(name)


def _batch_fields(seed: int):
    """Scalar diffusivity log-uniform in [1e-4, 1e-2], source centre in
    [0.2, 0.8]^3 and sigma in [0.05, 0.2], drawn from `seed`."""
    u = uniform01(seed, 5, stream=3)
    lam = 10.0 ** (-4.0 + 2.0 * u[0])
    centre = tuple(0.2 + 0.6 * u[1:4])
    sigma = 0.05 + 0.15 * u[4]
    return lam, centre, sigma


def batch_rhs_problems(shape=(96, 96, 48), count=32, seed0=100, **kw):
    """`count` systems sharing the mesh and the velocity field (seed 0); system
    b has its own scalar diffusivity and Gaussian source from seed seed0 + b."""
    out = []
    for b in range(count):
        lam, c, sig = _batch_fields(seed0 + b)
        out.append(unit_box(*shape, vel_seed=0, lam=lam, src_center=c, src_sigma=sig,
                            name=f"rhs{b}", **kw))
    return out


def batch_subdomain_problems(count=16, seed0=200, lo=24, hi=96, **kw):
    """`count` independent boxes of n^3 elements, n uniform in {lo..hi}, each
    with the field recipe of batch_rhs_problems, all drawn from seed0 + b."""
    out = []
    for b in range(count):
        n = lo + int(uniform01(seed0 + b, 1, stream=4)[0] * (hi - lo + 1))
        lam, c, sig = _batch_fields(seed0 + b)
        out.append(unit_box(n, n, n, vel_seed=0, lam=lam, src_center=c, src_sigma=sig,
                            name=f"sub{b}-{n}", **kw))
    return out

