"""Synthetic atomistic systems matching the benchmark workloads (seeded).

All generators place atoms uniformly with a minimum-distance rejection
(minimum image for periodic cells) and draw atomic types uniformly or with the
given probabilities.
"""

import math

import numpy as np


def _reject_place(rng, n, sample, min_dist, periodic_cell=None, max_tries=200):
    """Place n points drawn by `sample()` with pairwise distance >= min_dist."""
    pts = np.zeros((n, 3))
    inv = np.linalg.inv(periodic_cell) if periodic_cell is not None else None
    count = 0
    md2 = min_dist * min_dist
    tries = 0
    while count < n:
        p = sample()
        tries += 1
        if count:
            d = pts[:count] - p
            if inv is not None:
                f = d @ inv
                f -= np.round(f)
                d = f @ periodic_cell
            if np.min(np.einsum("ij,ij->i", d, d)) < md2:
                if tries > max_tries * n:
                    raise RuntimeError("could not place atoms with the requested minimum distance")
                continue
        pts[count] = p
        count += 1
    return pts


def triclinic_cell(a, b, c, alpha, beta, gamma):
    """Cell rows from lengths and angles in degrees (featomic/src/systems/cell.rs:93-115)."""
    ca, cb = math.cos(math.radians(alpha)), math.cos(math.radians(beta))
    sg, cg = math.sin(math.radians(gamma)), math.cos(math.radians(gamma))
    bx, by = b * cg, b * sg
    cx = c * cb
    cy = c * (ca - cb * cg) / sg
    cz2 = c * c - cy * cy - cx * cx
    if cz2 <= 0:
        return None
    return np.array([[a, 0, 0], [bx, by, 0], [cx, cy, math.sqrt(cz2)]])


def batched_uneven(seed=0, n_structures=4096, n_min=8, n_max=256, types=(1, 6, 7, 8), min_dist=0.8):
    """Half periodic (cubic / triclinic, density U[0.08, 0.12] atoms/A^3), half open
    clusters in 15 A boxes; atom counts U{n_min..n_max}; types uniform."""
    rng = np.random.default_rng(seed)
    systems = []
    for s in range(n_structures):
        n = int(rng.integers(n_min, n_max + 1))
        t = rng.choice(np.array(types, dtype=np.int32), size=n)
        if s % 2 == 0:
            rho = rng.uniform(0.08, 0.12)
            volume = n / rho
            if rng.random() < 0.5:
                L = volume ** (1.0 / 3.0)
                cell = np.eye(3) * L
            else:
                while True:
                    angles = rng.uniform(60.0, 120.0, 3)
                    unit = triclinic_cell(1.0, 1.0, 1.0, *angles)
                    if unit is None:
                        continue
                    v1 = abs(np.linalg.det(unit))
                    if v1 < 0.3:
                        continue
                    cell = unit * (volume / v1) ** (1.0 / 3.0)
                    break
            pos = _reject_place(rng, n, lambda: rng.uniform(0, 1, 3) @ cell, min_dist, cell)
        else:
            cell = np.zeros((3, 3))
            pos = _reject_place(rng, n, lambda: rng.uniform(0, 15.0, 3), min_dist)
        systems.append((pos, t, cell))
    return systems


def periodic_cubic_cells(seed=0, n_cells=32, L=10.0, n_min=80, n_max=120, types=(1, 6, 8), probabilities=(0.5, 0.3, 0.2),
                         min_dist=0.8):
    """Periodic cubic cells of side L with U{n_min..n_max} atoms."""
    rng = np.random.default_rng(seed)
    systems = []
    cell = np.eye(3) * L
    for _ in range(n_cells):
        n = int(rng.integers(n_min, n_max + 1))
        t = rng.choice(np.array(types, dtype=np.int32), size=n, p=np.array(probabilities))
        pos = _reject_place(rng, n, lambda: rng.uniform(0, L, 3), min_dist, cell)
        systems.append((pos, t, cell))
    return systems


def bulk_cell(seed=0, n_atoms=20000, density=0.1, types=(1, 6, 7, 8, 16), min_dist=0.8):
    """One periodic cubic cell at the given density (grid-accelerated rejection)."""
    rng = np.random.default_rng(seed)
    L = (n_atoms / density) ** (1.0 / 3.0)
    cell = np.eye(3) * L
    nb = max(1, int(L / min_dist))
    h = L / nb
    grid = {}
    pos = np.zeros((n_atoms, 3))
    count = 0
    md2 = min_dist * min_dist
    while count < n_atoms:
        p = rng.uniform(0, L, 3)
        b = tuple((p // h).astype(int) % nb)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = ((b[0] + dx) % nb, (b[1] + dy) % nb, (b[2] + dz) % nb)
                    for q in grid.get(key, ()):
                        d = pos[q] - p
                        d -= L * np.round(d / L)
                        if d @ d < md2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pos[count] = p
            grid.setdefault(b, []).append(count)
            count += 1
    t = rng.choice(np.array(types, dtype=np.int32), size=n_atoms)
    return [(pos, t, cell)]


def lode_cells(seed=0, n_structures=256, n_min=500, n_max=1500, density=0.1, types=(1, 6, 8), min_dist=0.8):
    """Periodic triclinic cells (angles U[60, 120] degrees, volume N / density) with
    U{n_min..n_max} atoms and types uniform: the LODE workload."""
    rng = np.random.default_rng(seed)
    systems = []
    for _ in range(n_structures):
        n = int(rng.integers(n_min, n_max + 1))
        t = rng.choice(np.array(types, dtype=np.int32), size=n)
        volume = n / density
        while True:
            angles = rng.uniform(60.0, 120.0, 3)
            unit = triclinic_cell(1.0, 1.0, 1.0, *angles)
            if unit is None or abs(np.linalg.det(unit)) < 0.3:
                continue
            cell = unit * (volume / abs(np.linalg.det(unit))) ** (1.0 / 3.0)
            break
        pos = _reject_place(rng, n, lambda: rng.uniform(0, 1, 3) @ cell, min_dist, cell)
        systems.append((pos, t, cell))
    return systems
