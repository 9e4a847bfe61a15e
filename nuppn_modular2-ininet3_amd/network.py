"""Synthetic reaction networks and zone states for the network-integration path.

The reference builds its network from REACLIB files (physics/source/ppn_physics.F90:624-1021)
whose rate blobs are not distributed, so every workload here is a seeded synthetic network
that follows the reference's topology conventions exactly:

* reaction ``l`` reads ``k2(l) x species k1(l) + k4(l) x k3(l) -> k6(l) x k5(l) + k8(l) x k7(l)``
  (physics/source/rates.F90:13-16);
* a capture is ``(k1=target, k3=projectile, k5=gamma, k7=product)``, a photodisintegration is
  ``(k1=target, k3=gamma, k5=ejectile, k7=residual)``, a beta decay is
  ``(k1=parent, k3=gamma, k5=gamma, k7=daughter)`` (physics/source/ppn_physics.F90:753-795);
* identical reactants / products are folded: ``k1==k3 -> k2+=k4, k4=0, k3=gamma`` and
  ``k5==k7 -> k8+=k6, k6=0, k5=gamma`` (physics/source/ppn_physics.F90:805-815);
* the gamma pseudo-species ``'OOOOO'`` has A=0 and molar abundance 1 inside dX/dt and J
  (solver/source/jac_rhs.F90:52-57,239-240).

Species layout (0-based): index 0 is gamma, index 1 the neutron, then the (Z, N) grid in
Z-major, N-ascending order, trimmed to the requested species count.  Rate coefficients are
REACLIB 7-parameter sets ``a0..a6`` stored as structure-of-arrays.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

# Reaction classes: they decide the density scaling of reaclib_calculate_rates
# (physics/source/reaclib.F90:562-576): two-body captures are multiplied by rho,
# photodisintegrations and decays are not.
RTYPE_CAPTURE = 0
RTYPE_PHOTO = 1
RTYPE_DECAY = 2

GAMMA = 0
NEUTRON = 1

SEED_BASE = 20261203


@dataclasses.dataclass
class Network:
    """Topology and coefficients of one reaction network (0-based species indices)."""

    name: str
    nsp: int
    nre: int
    aw: np.ndarray        # float64 [nsp] mass numbers (anetw)
    zw: np.ndarray        # float64 [nsp] charges (znetw)
    k: np.ndarray         # int32 [8, nre]: k1..k8, species entries 0-based, counts as is
    rtype: np.ndarray     # int32 [nre]
    coef: np.ndarray      # float64 [7, nre]: a0..a6
    zn: np.ndarray        # int32 [nsp, 2] (Z, N) of each species (gamma: (0,0), n: (0,1))

    @property
    def k1(self):
        return self.k[0]

    @property
    def k3(self):
        return self.k[2]

    @property
    def k5(self):
        return self.k[4]

    @property
    def k7(self):
        return self.k[6]

    def lightest(self, count: int) -> np.ndarray:
        """Indices of the ``count`` lightest species by mass number (ties by index), gamma excluded."""
        idx = np.arange(1, self.nsp)
        order = np.lexsort((idx, self.aw[1:]))
        return idx[order[:count]]


def _grid(z_max: int, n_lo: int, n_hi: int, target: int):
    """(Z, N) grid points Z=1..z_max, N in [Z+n_lo, Z+n_hi], N >= 0, Z-major order.

    When the band holds fewer than ``target`` points the upper edge of the N band is widened
    one step at a time until it does; the list is then trimmed to ``target`` points.
    """
    hi = n_hi
    while True:
        pts = [(z, n) for z in range(1, z_max + 1) for n in range(max(0, z + n_lo), z + hi + 1)]
        if len(pts) >= target or target <= 0:
            break
        hi += 1
    if target > 0:
        pts = pts[:target]
    return pts, hi


def build_network(name: str, z_max: int, n_lo: int, n_hi: int, nsp_total: Optional[int],
                  seed: int) -> Network:
    """Build the synthetic network.

    ``nsp_total`` counts gamma and the neutron; ``None`` keeps the full grid.
    Links per grid species (Z, N): captures of n, p, alpha to (Z, N+1), (Z+1, N), (Z+2, N+2),
    each with its reverse photodisintegration, and a beta decay to (Z+1, N-1); links whose
    product is not in the network are dropped, so every reaction conserves mass number.
    """
    target = -1 if nsp_total is None else nsp_total - 2
    pts, _ = _grid(z_max, n_lo, n_hi, target)
    zn = [(0, 0), (0, 1)] + pts
    nsp = len(zn)
    index = {p: i for i, p in enumerate(zn)}
    aw = np.array([0.0] + [float(z + n) for (z, n) in zn[1:]], dtype=np.float64)
    zw = np.array([float(z) for (z, n) in zn], dtype=np.float64)
    iprot = index.get((1, 0))
    ihe4 = index.get((2, 2))

    rows = []  # (k1..k8, rtype)
    for s in range(2, nsp):
        z, n = zn[s]
        for proj, dz, dn in ((NEUTRON, 0, 1), (iprot, 1, 0), (ihe4, 2, 2)):
            if proj is None:
                continue
            prod = index.get((z + dz, n + dn))
            if prod is None:
                continue
            rows.append([s, 1, proj, 1, GAMMA, 0, prod, 1, RTYPE_CAPTURE])
            rows.append([prod, 1, GAMMA, 0, proj, 1, s, 1, RTYPE_PHOTO])
        dau = index.get((z + 1, n - 1))
        if dau is not None:
            rows.append([s, 1, GAMMA, 0, GAMMA, 0, dau, 1, RTYPE_DECAY])
    arr = np.array(rows, dtype=np.int64)
    k = arr[:, :8].T.copy()
    rtype = arr[:, 8].astype(np.int32)
    # identical-particle folding (physics/source/ppn_physics.F90:805-815)
    for l in range(k.shape[1]):
        if k[0, l] == k[2, l]:
            k[1, l] += k[3, l]
            k[3, l] = 0
            k[2, l] = GAMMA
        if k[4, l] == k[6, l]:
            k[7, l] += k[5, l]
            k[5, l] = 0
            k[4, l] = GAMMA
    nre = k.shape[1]

    rng = np.random.default_rng([seed, 1])
    coef = np.zeros((7, nre), dtype=np.float64)
    coef[0] = rng.uniform(-10.0, 25.0, nre)
    coef[1] = rng.uniform(-40.0, 0.0, nre)
    coef[2] = rng.uniform(-30.0, 0.0, nre)
    coef[3] = rng.normal(0.0, 0.5, nre)
    coef[4] = rng.normal(0.0, 0.5, nre)
    coef[5] = rng.normal(0.0, 0.5, nre)
    coef[6] = rng.uniform(-1.5, 0.0, nre)
    lam = np.exp(rng.uniform(math.log(1e-12), math.log(1e2), nre))
    dec = rtype == RTYPE_DECAY
    coef[:, dec] = 0.0
    coef[0, dec] = np.log(lam[dec])

    return Network(name=name, nsp=nsp, nre=nre, aw=aw, zw=zw, k=k.astype(np.int32),
                   rtype=rtype, coef=coef, zn=np.array(zn, dtype=np.int32))


def _loguniform(rng, lo, hi, size):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def initial_abundances(net: Network, nz: int, n_light: int, alpha: float,
                       seed_level: float, rng) -> np.ndarray:
    """Mass fractions [nz, nsp]: Dirichlet(alpha) over the n_light lightest species plus
    ``seed_level``-sized seeds on the other species, normalised to sum 1; gamma stays 0."""
    x = np.zeros((nz, net.nsp), dtype=np.float64)
    light = net.lightest(n_light)
    x[:, light] = rng.dirichlet(np.full(n_light, alpha), size=nz)
    if seed_level > 0.0:
        rest = np.setdiff1d(np.arange(1, net.nsp), light)
        x[:, rest] = seed_level * rng.uniform(0.5, 1.5, size=(nz, rest.size))
    x /= x.sum(axis=1, keepdims=True)
    x[:, GAMMA] = 0.0
    return x


@dataclasses.dataclass
class Workload:
    """One configuration: the network plus per-zone thermodynamic state and step sizes.

    ``t9``/``rho`` are [nz, 2] (start and end of the cycle, interpolated linearly inside
    integrate_network); ``dt`` is [nz]; ``x0`` is [nz, nsp].  ``cycles`` is how many times the
    same cycle is repeated (abundances carry over).  ``method`` is 0 (backward Euler) or 1
    (Bader-Deuflhard).
    """

    name: str
    net: Network
    t9: np.ndarray
    rho: np.ndarray
    dt: np.ndarray
    x0: np.ndarray
    cycles: int = 1
    method: int = 0

    @property
    def nz(self):
        return self.x0.shape[0]


def config(cfg: int, nz: Optional[int] = None, seed: Optional[int] = None) -> Workload:
    """Synthetic workloads 0..4.

    0: one ppn zone, ~270-species network (Z=1..30, N in [Z-2, Z+6]); a 1,000-step trajectory
       of log-uniform T9 in 0.05..0.5 and rho in 1e1..1e4 (consecutive trajectory points are the
       cycle end points), dt = 1e4 s per step; Dirichlet(0.5) over the 10 lightest species.
       ``nz`` here is the number of trajectory steps kept.
    1: 2,048 zones, 1,100-species network (Z=1..83, N in [Z-3, Z+10]); T9 0.05..3, rho 1e0..1e8,
       dt = 1e4 s, one cycle; Dirichlet(0.3) over the 20 lightest plus 1e-10 seeds.
    2: 8,192 zones, same network; T9 0.02..4, dt 1e0..1e8 s, rho 1e0..1e8, 50 cycles.
    3: 256 zones, 5,000-species network (Z=1..100, N band from [Z-5, Z+25] widened until it
       holds 5,000 species); T9 0.1..2, rho 1e2..1e6, dt = 1e4 s.
    4: 4,096 zones, 1,100-species network, Bader-Deuflhard; T9 1..4, rho 1e5..1e9,
       dt 1e-3..1e2 s.
    """
    s = SEED_BASE + cfg if seed is None else seed
    rng = np.random.default_rng([s, 2])
    if cfg == 0:
        net = build_network("c0_270", 30, -2, 6, None, s)
        steps = 1000 if nz is None else nz
        t9p = _loguniform(rng, 0.05, 0.5, steps + 1)
        rhop = _loguniform(rng, 1e1, 1e4, steps + 1)
        x0 = initial_abundances(net, 1, 10, 0.5, 0.0, rng)
        t9 = np.stack([t9p[:-1], t9p[1:]], axis=1)
        rho = np.stack([rhop[:-1], rhop[1:]], axis=1)
        return Workload("c0", net, t9, rho, np.full(steps, 1e4), x0)
    if cfg in (1, 2, 4):
        net = build_network("net_1100", 83, -3, 10, 1100, SEED_BASE + 1)
        if cfg == 1:
            n = 2048 if nz is None else nz
            t9 = _loguniform(rng, 0.05, 3.0, n)
            rho = _loguniform(rng, 1e0, 1e8, n)
            dt = np.full(n, 1e4)
            cycles, method = 1, 0
        elif cfg == 2:
            n = 8192 if nz is None else nz
            t9 = _loguniform(rng, 0.02, 4.0, n)
            rho = _loguniform(rng, 1e0, 1e8, n)
            dt = _loguniform(rng, 1e0, 1e8, n)
            cycles, method = 50, 0
        else:
            n = 4096 if nz is None else nz
            t9 = _loguniform(rng, 1.0, 4.0, n)
            rho = _loguniform(rng, 1e5, 1e9, n)
            dt = _loguniform(rng, 1e-3, 1e2, n)
            cycles, method = 1, 1
        x0 = initial_abundances(net, n, 20, 0.3, 1e-10, rng)
        return Workload(f"c{cfg}", net, np.stack([t9, t9], 1), np.stack([rho, rho], 1), dt, x0,
                        cycles, method)
    if cfg == 3:
        net = build_network("net_5000", 100, -5, 25, 5000, SEED_BASE + 3)
        n = 256 if nz is None else nz
        t9 = _loguniform(rng, 0.1, 2.0, n)
        rho = _loguniform(rng, 1e2, 1e6, n)
        x0 = initial_abundances(net, n, 20, 0.3, 1e-10, rng)
        return Workload("c3", net, np.stack([t9, t9], 1), np.stack([rho, rho], 1),
                        np.full(n, 1e4), x0)
    raise ValueError(f"unknown workload {cfg}")


@dataclasses.dataclass
class StellarModel:
    """The per-zone structure an mppnp cycle reads besides the abundances (frames/mppnp/source/
    mppnp.F90:600-640, from the stellar-evolution input): mass coordinate ``xm`` (solar masses,
    increasing outward), shell masses ``dq`` (MESA convention, dq_reverse), radius ``rad`` (solar
    radii), density ``rho`` (g/cc), temperature ``t9`` and the Eulerian diffusion coefficient
    ``d_eul`` (cm^2/s: large in the convective zones, 0 in the radiative ones)."""

    xm: np.ndarray
    dq: np.ndarray
    rad: np.ndarray
    rho: np.ndarray
    t9: np.ndarray
    d_eul: np.ndarray

    @property
    def m(self):
        return self.xm.size


def stellar_model(m: int, seed: int = SEED_BASE + 7, t9_center: float = 2.0,
                  t9_surface: float = 0.01) -> StellarModel:
    """A synthetic star of ``m`` zones: a hot, dense, convective core, a radiative envelope with
    one convective shell, and a cool convective surface layer.  T9 falls from ``t9_center`` to
    ``t9_surface`` so that zoneselect keeps the inner part of the star; shell masses are uneven
    (log-normal scatter around an even split of one solar mass)."""
    rng = np.random.default_rng([seed, m])
    dq = np.exp(rng.normal(0.0, 0.3, m))
    dq /= dq.sum()
    xm = np.cumsum(dq)
    q = xm / xm[-1]
    rad = 0.02 + 0.98 * q ** 1.5
    rho = 10.0 ** (2.5 - 6.0 * q)
    t9 = t9_center * (t9_surface / t9_center) ** q
    d_eul = np.zeros(m)
    d_eul[q < 0.15] = 10.0 ** rng.uniform(14.0, 15.0, int((q < 0.15).sum()))
    shell = (q > 0.45) & (q < 0.55)
    d_eul[shell] = 10.0 ** rng.uniform(12.0, 13.0, int(shell.sum()))
    d_eul[q > 0.9] = 10.0 ** rng.uniform(15.0, 16.0, int((q > 0.9).sum()))
    return StellarModel(xm, dq, rad, rho, t9, d_eul)


# REACLIB's partition-function temperature grid (the 24 points of its winvn table)
PF_T9 = np.array([0.1, 0.15, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.5, 2.0, 2.5, 3.0,
                  3.5, 4.0, 4.5, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0])


@dataclasses.dataclass
class RatePhysics:
    """Nuclear data for the rate physics beyond the bare REACLIB sum (physics/source/
    evaluate_rates.F90:14-172): partition functions on REACLIB's T9 grid, the REACLIB reverse
    ('v') rates with their chapter and participants (reaclib.F90:512-557), spin factors, nuclear
    masses and binding energies for detailed balance (reverse.F90), and the switches (Chugunov
    screening and detailed balance are the reference's defaults, physics_knobs.F90:110-111)."""

    pf_t9: np.ndarray        # [24]
    pf_tab: np.ndarray       # [nsp, 24]
    rev_ch: np.ndarray       # int32 [nre]: chapter of a REACLIB reverse rate, 0 otherwise
    rev_part: np.ndarray     # int32 [6, nre]: participants e1..e6 (species, -1 none)
    g_spin: np.ndarray       # [nsp] 2J + 1
    mass_g: np.ndarray       # [nsp] nuclear masses (g)
    q_mev: np.ndarray        # [nsp] binding energies (MeV)
    screen: int = 1
    detailed_balance: int = 1


def rate_physics(net: Network, seed: int = SEED_BASE + 11, screen: int = 1,
                 detailed_balance: int = 1) -> RatePhysics:
    """Synthetic nuclear data for ``net``: partition functions 1 + c (T9 / 10)^q rising with
    temperature (c log-uniform up to A / 5, flat 1 for gamma, n, p, alpha); ground-state spins
    by the parity of Z and N; binding energies 8 MeV per nucleon +- 1 MeV (alpha's measured
    28.296 MeV, 0 for n and p); masses Z m_p + N m_n - B / c^2.  The photodisintegrations
    are REACLIB reverse rates of chapter 2 (target -> ejectile + residual)."""
    rng = np.random.default_rng([seed, net.nsp])
    nsp, nre = net.nsp, net.nre
    z = net.zn[:, 0].astype(np.float64)
    n = net.zn[:, 1].astype(np.float64)
    a = z + n
    pf = np.ones((nsp, PF_T9.size))
    heavy = a > 4
    cfac = np.exp(rng.uniform(math.log(1e-3), 0.0, nsp)) * np.maximum(a, 1.0) / 5.0
    qexp = rng.uniform(1.5, 4.0, nsp)
    pf[heavy] = 1.0 + cfac[heavy, None] * (PF_T9[None, :] / 10.0) ** qexp[heavy, None]
    g = np.ones(nsp)
    zi, ni = net.zn[:, 0], net.zn[:, 1]
    odd_a = ((zi + ni) % 2) == 1
    odd_odd = ((zi % 2) == 1) & ((ni % 2) == 1)
    g[odd_a] = 2.0 * rng.integers(1, 6, nsp)[odd_a]
    g[odd_odd] = 2.0 * rng.integers(0, 5, nsp)[odd_odd] + 1.0
    g[0] = 1.0
    g[(a == 1)] = 2.0       # n and p: J = 1/2
    # binding energies of about 8 MeV per nucleon with a +-1 MeV scatter: n, p and alpha
    # captures release a few MeV (separation energies 6..10 MeV, alpha Q ~ 4 MeV)
    b = 8.0 * a + rng.uniform(-1.0, 1.0, nsp)
    b[a <= 1] = 0.0
    b[(zi == 2) & (ni == 2)] = 28.2957
    b[0] = 0.0
    mev2erg = 1.602177e-12 * 1.0e6
    c2 = 2.99792458e10 ** 2
    mass = (z * 938.27231 + n * 939.56563 - b) * mev2erg / c2
    mass[0] = 0.0
    rev_ch = np.zeros(nre, np.int32)
    rev_part = np.full((6, nre), -1, np.int32)
    photo = net.rtype == RTYPE_PHOTO
    rev_ch[photo] = 2
    k = net.k
    rev_part[0, photo] = k[0, photo]
    # the ejectile and residual (a folded identical pair keeps both on k7)
    rev_part[1, photo] = np.where(k[4, photo] == GAMMA, k[6, photo], k[4, photo])
    rev_part[2, photo] = k[6, photo]
    return RatePhysics(PF_T9.copy(), pf, rev_ch, rev_part, g, mass, b, screen, detailed_balance)
