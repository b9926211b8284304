"""Seeded synthetic inputs for the Green-function action and solve.

Everything is derived from one SplitMix64 stream (Steele, Lea & Flood 2014), with uniform
deviates from the top 53 bits and normal deviates by the Box-Muller transform, so the same
seed gives the same system on every machine without depending on numpy's generators.

Geometry follows the reference's frame: the origin is the cell centre and grid point i of
block ib along one axis sits at ((ib - nb/2)*4 + i + 0.5)*h
(src/green_function.cxx:813,863 for the block coordinates, include/green_dyadic.hxx:71 for
the grid point inside a block).  Atoms are given as xyzZinso rows
(x, y, z, Z, id, numax, sigma, 0) as in src/green_input.cxx:89-122.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 stream: state += golden gamma; output = mix(state)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next_u64() >> 11) * (1.0 / (1 << 53))
        return lo + (hi - lo) * u

    def normal(self, sigma: float = 1.0) -> float:
        # Box-Muller, first of the pair; 1 - u keeps the log argument in (0, 1]
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        return sigma * math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def hash_uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Counter-based deviates in [-1, 1): the SplitMix64 output function applied to seed + (i+1)*gamma,
    top 53 bits.  Vectorised; used for seeded test vectors (identical on every platform)."""
    i = np.arange(offset, offset + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (i + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def nsho(numax: int) -> int:
    """Number of 3D SHO functions up to numax (include/sho_tools.hxx:21-27)."""
    return (numax + 1) * (numax + 2) * (numax + 3) // 6


@dataclass
class System:
    """One synthetic Hamiltonian: grid, potential, atoms and their matrices."""
    nb: tuple                   # blocks per direction
    h: float                    # grid spacing (Bohr)
    bc: tuple                   # boundary conditions per direction (0 iso, 1 periodic, 2 vacuum, 3 repeat)
    V: np.ndarray               # potential on the full grid, shape (4nb_z, 4nb_y, 4nb_x), Hartree
    xyzZinso: np.ndarray        # (natoms, 8)
    hs: list = field(default_factory=list)  # per atom: (2, nc, nc) h then s, L2-normalised SHO basis

    @property
    def ng(self):
        return tuple(4 * n for n in self.nb)

    @property
    def natoms(self):
        return self.xyzZinso.shape[0]

    @property
    def dvol(self):
        return self.h ** 3


def make_system(seed: int, nb, h: float = 0.25, ngauss: int = 8, natoms: int = 4,
                numax: int = 3, sigma: float = 0.5, bc=(1, 1, 1), depth=(-2.0, 0.0),
                width: float = 1.0, h_sigma: float = 0.1, s_sigma: float = 0.01) -> System:
    """Build a seeded system.

    Draw order (fixed): Gaussian centres (x, y, z) and depth per Gaussian; then atom positions
    (x, y, z) per atom; then per atom the upper triangle (row-major, diagonal included) of h,
    then of s - 1.
    """
    nb = tuple(int(n) for n in nb)
    rng = SplitMix64(seed)
    cell = np.array([4 * n * h for n in nb])
    centres = np.zeros((ngauss, 3))
    depths = np.zeros(ngauss)
    for g in range(ngauss):
        for d in range(3):
            centres[g, d] = rng.uniform(-0.5 * cell[d], 0.5 * cell[d])
        depths[g] = rng.uniform(depth[0], depth[1])

    # grid point coordinates relative to the cell centre
    axes = [((np.arange(4 * nb[d]) - 2 * nb[d]) + 0.5) * h for d in range(3)]
    V = np.zeros((4 * nb[2], 4 * nb[1], 4 * nb[0]))
    zz, yy, xx = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    for g in range(ngauss):
        diff = [xx - centres[g, 0], yy - centres[g, 1], zz - centres[g, 2]]
        r2 = np.zeros_like(V)
        for d in range(3):
            dd = diff[d]
            if bc[d] in (1, 3):  # minimum-image distance in periodic directions
                dd = dd - cell[d] * np.round(dd / cell[d])
            r2 += dd * dd
        V += depths[g] * np.exp(-r2 / (2.0 * width * width))

    xyz = np.zeros((natoms, 8))
    for ia in range(natoms):
        for d in range(3):
            xyz[ia, d] = rng.uniform(-0.5 * cell[d], 0.5 * cell[d])
        xyz[ia, 3] = 1.0          # Z (not used by the action)
        xyz[ia, 4] = ia           # global atom id
        xyz[ia, 5] = numax
        xyz[ia, 6] = sigma
    hs = []
    nc = nsho(numax)
    for ia in range(natoms):
        m = np.zeros((2, nc, nc))
        for k, (sd, diag) in enumerate(((h_sigma, 0.0), (s_sigma, 1.0))):
            for i in range(nc):
                for j in range(i, nc):
                    v = rng.normal(sd)
                    m[k, i, j] = v
                    m[k, j, i] = v
            m[k] += diag * np.eye(nc)
        hs.append(m)
    return System(nb=nb, h=h, bc=tuple(bc), V=V, xyzZinso=xyz, hs=hs)


# The workloads, with the seed 0xA43C0000 + i of workload i.
def small_periodic(seed=0xA43C0000):
    """8x8x8 blocks (32^3 grid), periodic, 8 Gaussians, 4 atoms numax 3, r_trunc 6 (untruncated)."""
    return make_system(seed, (8, 8, 8), ngauss=8, natoms=4, numax=3)


def wrap_32(seed=0xA43C0001, natoms=64, ngauss=64):
    """32x32x32 blocks (128^3 grid), periodic (wrapped by a 10 Bohr truncation), 64 atoms numax 5."""
    return make_system(seed, (32, 32, 32), ngauss=ngauss, natoms=natoms, numax=5)


def contour_32(seed=0xA43C0003, natoms=128, ngauss=64):
    """32x32x32 blocks (128^3 grid), periodic (wrapped by truncation), 128 atoms numax 5: the system of the
    batched energy-contour workload."""
    return make_system(seed, (32, 32, 32), ngauss=ngauss, natoms=natoms, numax=5)


def contour_16(seed=0xA43C0004, natoms=16, ngauss=16):
    """16x16x16 blocks (64^3 grid), periodic (wrapped by a 6 Bohr truncation), 16 atoms numax 3: the system of the
    adjoint-gradient checks."""
    return make_system(seed, (16, 16, 16), ngauss=ngauss, natoms=natoms, numax=3)


def contour_instances(n_energies=32, n_kpoints=4, radii=(6.0, 8.0, 10.0, 12.0), seed=0xA43C0003):
    """(E, k, radius) instances of a batched contour integration.

    Energies: Gauss-Legendre nodes x_i of [-1, 1] mapped to angles phi_i = pi (x_i + 1) / 2 on a half ellipse
    from E = -1 Ha to E_F = 0 Ha, E_i = -1/2 - cos(phi_i)/2 + i (0.005 + 0.295 sin(phi_i)), so Im E runs from
    0.005 Ha (slow solves) to 0.3 Ha (fast ones); the weights are the Gauss-Legendre weights times dE/dphi.
    k-points: n_kpoints seeded points in the first Brillouin zone (units of 2 pi / cell).  Radius: drawn per
    instance uniformly from `radii`.  Returns a list of dicts with keys E, weight, k, radius."""
    x, w = np.polynomial.legendre.leggauss(n_energies)
    phi = 0.5 * np.pi * (x + 1.0)
    E = -0.5 - 0.5 * np.cos(phi) + 1j * (0.005 + 0.295 * np.sin(phi))
    dE = (0.5 * np.sin(phi) + 1j * 0.295 * np.cos(phi)) * 0.5 * np.pi
    rng = SplitMix64(seed ^ 0x6B70)
    kpts = [tuple(rng.uniform(-0.5, 0.5) for _ in range(3)) for _ in range(n_kpoints)]
    out = []
    for k in kpts:
        for i in range(n_energies):
            r = radii[min(len(radii) - 1, int(rng.uniform() * len(radii)))]
            out.append(dict(E=complex(E[i]), weight=complex(w[i] * dE[i]), k=k, radius=float(r)))
    return out
