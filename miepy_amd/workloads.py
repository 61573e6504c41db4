"""Synthetic cluster workloads (seeded, numpy PCG64).

- hexagonal_lattice_layers: the L-layer hexagonal lattice of the reference's
  benchmarks/bench_hexagonal.py:12-27 (L = 3 gives 37 sites).
- random_cluster: centers uniform in a cube, rejection-sampled so that every
  center is at least its radius from the faces and every pair satisfies
  |r_i - r_j| >= 1.1 (a_i + a_j) (2.2 a for equal radii).

Named workloads (sizes of the cluster-solve benchmark configurations):
  hex37   37 spheres, eps 2.56, r 75 nm, hex lattice 600 nm, water,
          x-polarized plane wave along +z, 800 nm, lmax 2 (n = 592)
  rand1000  1000 spheres in a 10 um cube, r 100 nm, eps -20+1.2j, vacuum,
          600 nm, lmax 3 (n = 30000)
  rand2500  2500 spheres in a 16 um cube, r ~ U(80,120) nm,
          eps ~ U(2,6) + i U(0,0.1), vacuum, 600 nm, lmax 3 (n = 75000),
          16 plane waves at seeded incidence angles
  batch256  256 clusters: N log-uniform in [10, 400], lmax in {2,3,4},
          wavelength U(400,1000) nm, radius U(50,150) nm, dielectric or
          metal eps, 8 plane waves each
  serving   Poisson arrivals; N log-uniform in [5, 3000], lmax in {2..5},
          1-32 wavelengths in 400-1000 nm, 1-4 sources
"""

import math

import numpy as np

from .materials import constant_material
from .sources import plane_wave

nm = 1e-9
um = 1e-6


def hexagonal_lattice_layers(L):
    k1 = np.array([1.0, 0.0, 0.0])
    k2 = np.array([math.cos(math.pi / 3), math.sin(math.pi / 3), 0.0])
    lattice = [np.zeros(3)]
    for step in range(1, L + 1):
        lattice.append(step * k1)
        for direc in [k2 - k1, -k1, -k2, -k2 + k1, k1, k2]:
            for _ in range(step):
                lattice.append(lattice[-1] + direc)
        lattice.pop()
    return np.asarray(lattice, dtype=float)


def random_cluster(N, box, radius, rng, gap=1.1, max_tries=10_000_000):
    """N centers in [0, box]^3 with per-particle radius (scalar or [N])."""
    radius = np.broadcast_to(np.asarray(radius, dtype=float), (N,))
    pos = np.empty((N, 3))
    cell = 2 * gap * radius.max()
    grid = {}
    count = 0
    tries = 0
    while count < N:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could only place {count} of {N} spheres")
        a = radius[count]
        c = a + rng.random(3) * (box - 2 * a)
        key = tuple((c // cell).astype(int))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        if np.sum((pos[j] - c) ** 2) < (gap * (a + radius[j])) ** 2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pos[count] = c
            grid.setdefault(key, []).append(count)
            count += 1
    return pos


def random_plane_waves(S, rng, random_pol=False):
    out = []
    for _ in range(S):
        theta = math.acos(rng.uniform(-1, 1))
        phi = rng.uniform(0, 2 * math.pi)
        if random_pol:
            chi = rng.uniform(0, math.pi)
            delta = rng.uniform(0, 2 * math.pi)
            pol = [math.cos(chi), complex(math.cos(delta), math.sin(delta)) * math.sin(chi)]
        else:
            pol = [1, 0]
        out.append(plane_wave(pol, theta=theta, phi=phi))
    return out


class Workload:
    """One cluster system: geometry, materials, sources, solve settings."""

    def __init__(self, name, position, radius, eps, wavelength, lmax, sources, medium_eps=1.0, tol=1e-10,
                 conducting=None):
        self.name = name
        self.position = np.asarray(position, dtype=float)
        self.N = self.position.shape[0]
        self.radius = np.broadcast_to(np.asarray(radius, dtype=float), (self.N,)).copy()
        self.eps = np.broadcast_to(np.asarray(eps, dtype=complex), (self.N,)).copy()
        self.wavelength = float(wavelength)
        self.lmax = int(lmax)
        self.sources = list(sources)
        self.medium_eps = medium_eps
        self.tol = tol
        self.conducting = conducting

    @property
    def rmax(self):
        return self.lmax * (self.lmax + 2)

    @property
    def n(self):
        return 2 * self.rmax * self.N

    @property
    def k(self):
        return 2 * math.pi * math.sqrt(np.real(self.medium_eps)) / self.wavelength

    def materials(self):
        return [constant_material(eps=e) for e in self.eps]

    def medium(self):
        return constant_material(eps=self.medium_eps)

    def __repr__(self):
        return f"Workload({self.name}: N={self.N}, lmax={self.lmax}, n={self.n}, S={len(self.sources)})"


def hex37():
    pos = hexagonal_lattice_layers(3) * 600 * nm
    return Workload("hex37", pos, 75 * nm, 2.56 + 0j, 800 * nm, 2, [plane_wave.from_string("x")],
                    medium_eps=1.33 ** 2, tol=1e-12)


def rand1000(seed=2):
    rng = np.random.default_rng(seed)
    pos = random_cluster(1000, 10 * um, 100 * nm, rng)
    return Workload("rand1000", pos, 100 * nm, -20 + 1.2j, 600 * nm, 3, [plane_wave.from_string("x")], tol=1e-10)


def rand2500(seed=3, nsrc=16):
    rng = np.random.default_rng(seed)
    radius = rng.uniform(80 * nm, 120 * nm, 2500)
    pos = random_cluster(2500, 16 * um, radius, rng)
    eps = rng.uniform(2, 6, 2500) + 1j * rng.uniform(0, 0.1, 2500)
    return Workload("rand2500", pos, radius, eps, 600 * nm, 3, random_plane_waves(nsrc, rng), tol=1e-10)


def scaled_random(N, lmax, seed=0, density_box=10 * um / 1000 ** (1 / 3), radius=100 * nm, eps=-20 + 1.2j,
                  wavelength=600 * nm, nsrc=1, tol=1e-10):
    """A rand1000-like cluster with N spheres at the same number density."""
    rng = np.random.default_rng(seed)
    box = density_box * N ** (1 / 3)
    pos = random_cluster(N, box, radius, rng)
    srcs = [plane_wave.from_string("x")] if nsrc == 1 else random_plane_waves(nsrc, rng)
    return Workload(f"random{N}", pos, radius, eps, wavelength, lmax, srcs, tol=tol)


def _batch_cluster(rng, Nmin, Nmax, lmax_choices, nsrc, tol, name):
    N = int(math.floor(math.exp(rng.uniform(math.log(Nmin), math.log(Nmax)))))
    lmax = int(rng.choice(lmax_choices))
    wl = rng.uniform(400, 1000) * nm
    a = rng.uniform(50, 150) * nm
    if rng.random() < 0.5:
        eps = complex(rng.uniform(2, 6), rng.uniform(0, 0.1))
    else:
        eps = complex(-20, 1.2)
    # box sized for a volume fraction of ~1 % so the gap rule is easy to meet
    box = (N * (4 / 3) * math.pi * a ** 3 / 0.01) ** (1 / 3)
    box = max(box, 4 * a)
    pos = random_cluster(N, box, a, rng)
    return Workload(name, pos, a, eps, wl, lmax, random_plane_waves(nsrc, rng, random_pol=True), tol=tol)


def batch256(seed=4, count=256):
    rng = np.random.default_rng(seed)
    return [_batch_cluster(rng, 10, 400, (2, 3, 4), 8, 1e-8, f"batch{i}") for i in range(count)]


def serving_requests(seed=5, rate=200.0, duration=60.0, Nmax=3000):
    """Yield (arrival_time, Workload list over the request's wavelengths)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    idx = 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return
        nwl = int(rng.integers(1, 33))
        nsrc = int(rng.integers(1, 5))
        base = _batch_cluster(rng, 5, Nmax, (2, 3, 4, 5), nsrc, 1e-8, f"req{idx}")
        wls = np.sort(rng.uniform(400, 1000, nwl)) * nm
        reqs = [Workload(f"req{idx}_wl{i}", base.position, base.radius, base.eps, wl, base.lmax, base.sources,
                         tol=base.tol) for i, wl in enumerate(wls)]
        yield t, reqs
        idx += 1
