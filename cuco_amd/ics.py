"""Seeded synthetic initial conditions for the gravity-path workloads.

Units are the reference's (begrun.c:33-35,65-70): kpc/h, 1e10 Msun/h, km/s,
so G = 43007.1 and H0 = 0.1 internal.  Internal velocities are a^2 dx/dt,
the quantity the reference holds after init.c:30 (Vel *= a^(3/2)).

Workloads (lengths quoted in Mpc/h, stored in kpc/h):

* ``lattice32``  -- 32^3 particles, 64^3 mesh, L=100.  Zel'dovich displacement
  of a lattice by a Gaussian field with P(k) ~ k^-2 (white noise filtered in
  k-space, so Hermitian by construction), rms |displacement| 0.2 mean spacing,
  a0 = 0.02, 10 KDK steps to a = 0.05.
* ``halo128``    -- 128^3, 256^3 mesh, L=250: 50% uniform, 50% in 300 NFW
  halos (dN/dM ~ M^-1.9 over two decades, c = 8, truncated at r200 defined
  against 200x the mean density), isotropic Gaussian velocities of
  dispersion V200/sqrt(2) per component.
* ``halo256``    -- 256^3, 512^3 mesh, L=500, same generator with 2000 halos.
* ``ensemble64`` -- 64 boxes of N_b in {32,48,64,80,96}^3, mesh 2 N_b^(1/3),
  L_b = N_b^(1/3) * 3.125 Mpc/h, Zel'dovich ICs, seed 4004*1000 + b.
* ``halo512``    -- 512^3, 1024^3 mesh, L=1000, 20000 halos.

Every box uses equal masses m = Omega_m rho_crit L^3 / N, split scale
asmth = 1.25 mesh cells, rcut = 4.5 asmth, Plummer softening 1/30 of the
mean spacing (spline h = 2.8 eps, gravlist.cpp:36).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Physical constants and units, allvars.h:15-18 and begrun.c:33-35.
GRAVITY = 6.673e-11
HUBBLE = 3.2407789e-18
UNIT_LENGTH_M = 3.085678e19
UNIT_MASS_KG = 1.98892e40
UNIT_VELOCITY = 1e3
UNIT_TIME_S = UNIT_LENGTH_M / UNIT_VELOCITY
G_INTERNAL = GRAVITY / UNIT_LENGTH_M ** 3 * UNIT_MASS_KG * UNIT_TIME_S ** 2
HUBBLE_INTERNAL = HUBBLE * UNIT_TIME_S
MPC = 1000.0  # kpc/h per Mpc/h

OMEGA_M = 0.3
OMEGA_L = 0.7
H_PARAM = 0.7


@dataclass
class Box:
    name: str
    n_side: int
    M: int
    L: float                 # kpc/h
    seed: int
    kind: str                # "zeldovich" or "halos"
    n_halos: int = 0
    a_start: float = 0.02
    a_end: float = 0.05
    nsteps: int = 10
    extra: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return self.n_side ** 3

    @property
    def mass(self) -> float:
        rho_crit = 3 * HUBBLE_INTERNAL ** 2 / (8 * math.pi * G_INTERNAL)
        # (L / n_side)^3 rather than L^3 / N: boxes of equal mean spacing get
        # bit-identical masses.
        return OMEGA_M * rho_crit * (self.L / self.n_side) ** 3

    @property
    def asmth(self) -> float:
        return 1.25 * self.L / self.M      # pm_periodic.c:27

    @property
    def rcut(self) -> float:
        return 4.5 * self.asmth            # pm_periodic.c:28

    @property
    def softening(self) -> float:
        return self.L / self.n_side / 30.0

    @property
    def soft_h(self) -> float:
        return 2.8 * self.softening        # gravlist.cpp:36

    def engine_kwargs(self) -> dict:
        return dict(pm_grid=self.M, box_size=self.L, G=G_INTERNAL,
                    asmth=self.asmth, rcut=self.rcut, soft_h=self.soft_h)


WORKLOADS = {
    "lattice32": Box("lattice32", 32, 64, 100 * MPC, 1001, "zeldovich"),
    "halo128": Box("halo128", 128, 256, 250 * MPC, 2002, "halos", n_halos=300,
                   a_start=0.5, a_end=0.55, nsteps=1),
    "halo256": Box("halo256", 256, 512, 500 * MPC, 3003, "halos", n_halos=2000,
                   a_start=0.5, a_end=0.55, nsteps=1),
    "halo512": Box("halo512", 512, 1024, 1000 * MPC, 5005, "halos",
                   n_halos=20000, a_start=0.5, a_end=0.55, nsteps=5),
}


def small_box(name: str, n_side: int, M: int, L_mpc: float, seed: int,
              kind: str = "zeldovich", n_halos: int = 0) -> Box:
    """A reduced box with the same generators (for fast parity tests)."""
    return Box(name, n_side, M, L_mpc * MPC, seed, kind, n_halos=n_halos)


def ensemble_boxes(nboxes: int = 64, seed: int = 4004) -> list:
    rng = np.random.default_rng(seed)
    sides = rng.choice([32, 48, 64, 80, 96], size=nboxes)
    boxes = []
    for b, s in enumerate(sides):
        s = int(s)
        boxes.append(Box(f"ensemble64_{b}", s, 2 * s, s * 3.125 * MPC,
                         seed * 1000 + b, "zeldovich", a_start=0.02,
                         a_end=0.05, nsteps=20))
    return boxes


def hubble_a(a: float) -> float:
    return HUBBLE_INTERNAL * math.sqrt(OMEGA_M / a ** 3 +
                                       (1 - OMEGA_M - OMEGA_L) / a ** 2 + OMEGA_L)


def growth_rate(a: float) -> float:
    om_a = OMEGA_M / a ** 3 / (hubble_a(a) / HUBBLE_INTERNAL) ** 2
    return om_a ** 0.55


def zeldovich(box: Box):
    """Lattice displaced by a k^-2 Gaussian field; returns pos, vel, mass."""
    n = box.n_side
    rng = np.random.default_rng(box.seed)
    white = rng.standard_normal((n, n, n))
    wk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(n) * n
    kz = np.fft.rfftfreq(n) * n
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[0, 0, 0] = 1.0
    delta = wk / np.sqrt(k2)          # amplitude ~ k^-1 => P(k) ~ k^-2
    delta[0, 0, 0] = 0.0
    psi = []
    for K in (KX, KY, KZ):
        psi.append(np.fft.irfftn(1j * K / k2 * delta, s=(n, n, n), axes=(0, 1, 2)))
    psi = np.stack([p.reshape(-1) for p in psi], axis=1)
    spacing = box.L / n
    rms = math.sqrt(np.mean(np.sum(psi ** 2, axis=1)))
    psi *= 0.2 * spacing / rms
    q = (np.indices((n, n, n)).reshape(3, -1).T + 0.5) * spacing
    pos = np.mod(q + psi, box.L)
    a = box.a_start
    vel = a * a * hubble_a(a) * growth_rate(a) * psi
    mass = np.full(box.N, box.mass)
    return pos, vel, mass


def _nfw_mass(x):
    return np.log1p(x) - x / (1.0 + x)


def halos(box: Box):
    """50% uniform, 50% in NFW halos; returns pos, vel, mass."""
    rng = np.random.default_rng(box.seed)
    N = box.N
    n_halo_parts = N // 2
    n_uniform = N - n_halo_parts
    nh = box.n_halos
    # Masses from dN/dM ~ M^-1.9 on [1, 100] (inverse CDF), particles ~ mass.
    alpha = 1.9
    u = rng.random(nh)
    lo, hi = 1.0, 100.0
    e = 1.0 - alpha
    masses = (lo ** e + u * (hi ** e - lo ** e)) ** (1.0 / e)
    share = masses / masses.sum() * n_halo_parts
    counts = np.floor(share).astype(np.int64)
    rem = n_halo_parts - counts.sum()
    order = np.argsort(-(share - counts))
    counts[order[:rem]] += 1
    centers = rng.random((nh, 3)) * box.L
    # r200 against 200x the mean density: n = 200 (N/L^3) 4/3 pi r^3.
    r200 = np.cbrt(3.0 * counts * box.L ** 3 / (800.0 * math.pi * N))
    conc = 8.0
    xs = np.concatenate([[0.0], np.geomspace(1e-6, conc, 20000)])
    ms = _nfw_mass(xs)
    pos = np.empty((N, 3))
    vel = np.empty((N, 3))
    a = box.a_start
    start = 0
    for h in range(nh):
        c = int(counts[h])
        if c == 0:
            continue
        ur = rng.random(c) * _nfw_mass(conc)
        x = np.interp(ur, ms, xs)
        r = x * r200[h] / conc
        d = rng.standard_normal((c, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pos[start:start + c] = centers[h] + r[:, None] * d
        v200 = math.sqrt(G_INTERNAL * c * box.mass / r200[h])
        vel[start:start + c] = rng.standard_normal((c, 3)) * (v200 / math.sqrt(2.0)) * a
        start += c
    pos[start:] = rng.random((n_uniform, 3)) * box.L
    vel[start:] = 0.0
    pos = np.mod(pos, box.L)
    # mod can return L for tiny negatives; fold those back.
    pos[pos >= box.L] -= box.L
    mass = np.full(N, box.mass)
    return pos, vel, mass


def generate(box: Box):
    if box.kind == "zeldovich":
        return zeldovich(box)
    if box.kind == "halos":
        return halos(box)
    raise ValueError(box.kind)
