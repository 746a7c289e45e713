"""Synthetic waveguide workloads for tests and the benchmark.

Every workload is an air-filled PEC rectangular waveguide cross-section meshed
with axis-aligned quads, refined with seeded randomness. The random draws come
from SplitMix64 (one stream per workload): u = (next >> 11) * 2^-53, taken per
h-refineable Elem in ascending id order, one level at a time.

* ``waveguide_small`` ([0,1]x[0,0.5], 4x2 quads, orders (4,4), one global T
  refinement, then U on the leaves touching the centre node (0.5, 0.25);
  8x8 Gauss points): 48 Elems, 1096 DoFs. Lowest non-gradient eigenvalue near
  the TE10 cutoff pi^2.
* ``waveguide_64x32`` ([0,1]x[0,0.5], 64x32 quads, orders (6,6), two levels
  of T refinement on a random 25 % of the leaves per level (seed 0); 8x8 Gauss).
* ``hp_irregular`` ([0,2]x[0,1], 128x64 quads, per base Elem orders (i,j)
  drawn iid from {1..8}^2 (seed 1, inherited by children), three levels of
  T / U / V refinement with probabilities 0.2 / 0.1 / 0.1 per leaf and level;
  Legendre basis, 10x10 Gauss).
* ``precision_study``: the ``waveguide_small`` mesh at orders (8,8), 12x12 Gauss.
* ``many_waveguides``: 256 independent waveguides (width a ~ U[1,2], height 0.5,
  nx ~ U{4..16}, ny = nx // 2, uniform order p ~ U{2..6}, one T refinement on a
  random 30 % of Elems), placed side by side in one Mesh so that one assembly
  covers all of them (block-diagonal GEP).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from .domain import ContinuityCondition, Domain
from .galerkin import HierPoly, KOLShapeFn
from .mesh import HRef, Mesh, Orders

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + min(hi - lo, int(self.uniform() * (hi - lo + 1)))


@dataclass
class Workload:
    name: str
    domain: Domain
    glq: Tuple[int, int]
    basis: type = HierPoly
    meta: dict = field(default_factory=dict)


def _leaf_random_refine(mesh: Mesh, rng: SplitMix64, pick) -> None:
    """One level: draw per h-refineable Elem (ascending id) and apply pick(u) -> HRef or None."""
    refs = []
    for e in range(mesh.num_elems):
        if mesh.elem_is_h_refineable(e):
            r = pick(rng.uniform())
            if r is not None:
                refs.append((e, r))
    mesh.execute_h_refinements(refs)


def waveguide_small_mesh(orders: int = 4) -> Mesh:
    mesh = Mesh.rect_grid(4, 2, 0.0, 1.0, 0.0, 0.5)
    mesh.set_global_expansion_orders(Orders(orders, orders))
    mesh.global_h_refinement(HRef.T)
    centre = 7  # node (0.5, 0.25) of the 5x3 node grid
    mesh.h_refine_with_filter(lambda e: HRef.u() if centre in e.nodes else None)
    return mesh


def waveguide_small() -> Workload:
    return Workload("waveguide_small", Domain.from_mesh(waveguide_small_mesh(4), ContinuityCondition.HCurl), (8, 8),
                    meta=dict(a=1.0, b=0.5))


def precision_study() -> Workload:
    return Workload("precision_study", Domain.from_mesh(waveguide_small_mesh(8), ContinuityCondition.HCurl), (12, 12),
                    meta=dict(a=1.0, b=0.5))


def waveguide_64x32_mesh(nx: int = 64, ny: int = 32, order: int = 6, levels: int = 2, prob: float = 0.25,
                         seed: int = 0) -> Mesh:
    mesh = Mesh.rect_grid(nx, ny, 0.0, 1.0, 0.0, 0.5)
    rng = SplitMix64(seed)
    for _ in range(levels):
        _leaf_random_refine(mesh, rng, lambda u: HRef.T if u < prob else None)
    mesh.set_global_expansion_orders(Orders(order, order))
    return mesh


def waveguide_64x32(nx: int = 64, ny: int = 32) -> Workload:
    mesh = waveguide_64x32_mesh(nx, ny)
    return Workload(f"waveguide_{nx}x{ny}", Domain.from_mesh(mesh, ContinuityCondition.HCurl), (8, 8),
                    meta=dict(a=1.0, b=0.5, nx=nx, ny=ny, order=6, levels=2, prob=0.25, seed=0))


def hp_irregular_mesh(nx: int = 128, ny: int = 64, levels: int = 3, seed: int = 1) -> Mesh:
    mesh = Mesh.rect_grid(nx, ny, 0.0, 2.0, 0.0, 1.0)
    rng = SplitMix64(seed)
    items = [(e, Orders(rng.randint(1, 8), rng.randint(1, 8))) for e in range(nx * ny)]
    mesh.set_expansion_orders(items)

    def pick(u):
        if u < 0.2:
            return HRef.T
        if u < 0.3:
            return HRef.u()
        if u < 0.4:
            return HRef.v()
        return None

    for _ in range(levels):
        _leaf_random_refine(mesh, rng, pick)
    return mesh


def hp_irregular(nx: int = 128, ny: int = 64) -> Workload:
    mesh = hp_irregular_mesh(nx, ny)
    return Workload(f"hp_irregular_{nx}x{ny}", Domain.from_mesh(mesh, ContinuityCondition.HCurl), (10, 10),
                    basis=KOLShapeFn, meta=dict(a=2.0, b=1.0, nx=nx, ny=ny, levels=3, seed=1))


def many_waveguides_mesh(count: int = 256, seed: int = 2, gap: float = 0.5):
    """Returns (mesh, instances) with per-instance (x offset, width, nx, ny, order, elem range)."""
    rng = SplitMix64(seed)
    pts: List[List[float]] = []
    ids: List[List[int]] = []
    inst = []
    x0 = 0.0
    for k in range(count):
        a = 1.0 + rng.uniform()
        nx = rng.randint(4, 16)
        ny = nx // 2
        p = rng.randint(2, 6)
        base_node = len(pts)
        base_elem = len(ids)
        xs = np.linspace(x0, x0 + a, nx + 1)
        ys = np.linspace(0.0, 0.5, ny + 1)
        for j in range(ny + 1):
            for i in range(nx + 1):
                pts.append([xs[i], ys[j]])
        for j in range(ny):
            for i in range(nx):
                n0 = base_node + j * (nx + 1) + i
                ids.append([n0, n0 + 1, n0 + nx + 1, n0 + nx + 2])
        inst.append(dict(x0=x0, a=a, nx=nx, ny=ny, order=p, elems=(base_elem, len(ids))))
        x0 += a + gap
    mats = np.tile(np.array([1.0, 0.0, 1.0, 0.0]), (len(ids), 1))
    mesh = Mesh.from_parts(np.array(pts), np.array(ids, dtype=np.int32), mats)
    items = []
    for d in inst:
        o = Orders(d["order"], d["order"])
        items.extend((e, o) for e in range(*d["elems"]))
    mesh.set_expansion_orders(items)
    refs = []
    for e in range(mesh.num_elems):
        if rng.uniform() < 0.3:
            refs.append((e, HRef.T))
    mesh.execute_h_refinements(refs)
    return mesh, inst


def many_waveguides(count: int = 256) -> Workload:
    mesh, inst = many_waveguides_mesh(count)
    return Workload(f"many_waveguides_{count}", Domain.from_mesh(mesh, ContinuityCondition.HCurl), (8, 8),
                    meta=dict(instances=inst))


def te10_cutoff(a: float) -> float:
    """Lowest TE cutoff k^2 = (pi / a)^2 of an a x b guide with a > b."""
    return (math.pi / a) ** 2
