"""Seeded synthetic colony inputs.

All randomness comes from a counter-based splitmix64 stream (Steele, Lea &
Flood 2014), so every input is a pure function of (seed, stream, index) and is
identical on any machine; std::*_distribution is avoided on purpose because its
output is implementation-defined.

Bodies are returned sorted by x with gid = index, as the reference numbers them
(Domain::assignGlobalIDs, code/cpp/src/spatial/Domain.cpp:492-508).
"""
import math

import numpy as np

from .abi_types import Bodies

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from stream `stream` of seed `seed`."""
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64((seed * 0x100000001B3 + stream * 0x9E37 + 1) & 0xFFFFFFFFFFFFFFFF))
        ctr = np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN + key
        z = _mix64(ctr)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller on two uniform streams."""
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 7919, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def _planar_quat(theta):
    q = np.zeros((len(theta), 4))
    q[:, 0] = np.cos(theta / 2)
    q[:, 3] = np.sin(theta / 2)
    return q


def _sorted_bodies(pos, theta, length, diameter=0.5, l0=1.0):
    order = np.argsort(pos[:, 0], kind="stable")
    n = len(length)
    return Bodies(pos[order], _planar_quat(theta[order]), length[order], np.full(n, l0),
                  np.full(n, diameter), np.zeros(n))


def random_disc(n: int, seed: int = 0, area_fraction: float = 0.9, diameter: float = 0.5,
                lmin: float = 1.0, lmax: float = 2.0) -> Bodies:
    """Single-step hard-model workload: n spherocylinders with lengths U[lmin,lmax],
    orientations U[0, pi), centres uniform in a disc sized so the summed
    spherocylinder area is `area_fraction` of the disc (z = 0, in-plane)."""
    length = lmin + (lmax - lmin) * uniform(seed, 1, n)
    theta = math.pi * uniform(seed, 2, n)
    area = (length - diameter) * diameter + math.pi * diameter * diameter / 4.0
    R = math.sqrt(area.sum() / (area_fraction * math.pi))
    r = R * np.sqrt(uniform(seed, 3, n))
    phi = 2.0 * math.pi * uniform(seed, 4, n)
    pos = np.zeros((n, 3))
    pos[:, 0] = r * np.cos(phi)
    pos[:, 1] = r * np.sin(phi)
    return _sorted_bodies(pos, theta, length, diameter)


def disc_radius(bodies: Bodies) -> float:
    return float(np.sqrt((bodies.pos[:, 0] ** 2 + bodies.pos[:, 1] ** 2).max())) if bodies.n else 0.0


def packed_colony(n: int, seed: int = 0, diameter: float = 0.5, gap: float = 0.02,
                  angle_noise: float = 0.03, patch: float = 8.0) -> Bodies:
    """Colony-like packing of n cells built directly instead of by growth: the
    plane is tiled by disc-shaped domains of diameter `patch` on a triangular
    lattice, each with its own nematic direction; inside a domain cells of
    length U[1,2] sit end to end in rows spaced diameter+gap apart, and only
    cells lying wholly inside their domain are kept, so domains never cross.
    Each cell is then tilted by N(0, angle_noise), which produces the small
    overlaps of a grown colony.  The n cells nearest the origin are returned."""
    R = math.sqrt(n * 0.696 / (0.55 * math.pi)) + patch
    rad = patch / 2
    out_pos, out_theta, out_len = [], [], []
    pid = 0
    rowh = patch * math.sqrt(3) / 2
    nrow = int(math.ceil(2 * R / rowh)) + 1
    ncol = int(math.ceil(2 * R / patch)) + 1
    for py in range(nrow):
        for px in range(ncol):
            pid += 1
            cx = -R + px * patch + (0.5 * patch if py % 2 else 0.0)
            cy = -R + py * rowh
            if math.hypot(cx, cy) > R + patch:
                continue
            th = math.pi * uniform(seed, 10_000 + pid, 1)[0]
            ct, st = math.cos(th), math.sin(th)
            rows = int(patch / (diameter + gap)) + 2
            per_row = int(patch / 1.0) + 3
            lens = 1.0 + uniform(seed, 20_000 + pid, rows * per_row).reshape(rows, per_row)
            jit = uniform(seed, 30_000 + pid, rows)
            for j in range(rows):
                v = -rad + (j + 0.5) * (diameter + gap)
                u = -rad - jit[j] * 1.5
                for k in range(per_row):
                    L = lens[j, k]
                    uc = u + L / 2
                    u += L + gap
                    # both cap tips inside the domain disc
                    reach = math.hypot(abs(uc) + L / 2, abs(v) + diameter / 2)
                    if reach > rad:
                        continue
                    x = cx + uc * ct - v * st
                    y = cy + uc * st + v * ct
                    out_pos.append((x, y))
                    out_theta.append(th)
                    out_len.append(L)
    pos2 = np.array(out_pos)
    theta = np.array(out_theta)
    length = np.array(out_len)
    # keep the n cells nearest the origin
    rr = np.hypot(pos2[:, 0], pos2[:, 1])
    keep = np.argsort(rr, kind="stable")[:n]
    pos = np.zeros((len(keep), 3))
    pos[:, :2] = pos2[keep]
    noise = angle_noise * normal(seed, 5, len(keep))
    return _sorted_bodies(pos, np.mod(theta[keep] + noise, math.pi), length[keep], diameter)


def dense_colony(n: int, seed: int = 0, diameter: float = 0.5, domain: float = 30.0,
                 row_gap: float = 0.002, end_gap: float = 0.002, angle_noise: float = 0.003,
                 lmin: float = 1.0, lmax: float = 2.0) -> Bodies:
    """Dense colony-like packing of exactly n cells (area fraction ~0.85 at
    300k cells, overlaps up to ~1e-2 from the tilt noise, no crossing rods).

    Nematic domains: sites on a jittered triangular lattice of spacing
    `domain`, each with a direction U[0, pi) (so orientations are uniform over
    the colony).  In its Voronoi region a domain is filled with rows of cells
    of length U[lmin, lmax] laid end to end, rows spaced diameter+row_gap apart
    (negative gaps = slight compression, as in a growing colony); a cell is
    kept only if its whole spherocylinder lies inside its region, so cells of
    different domains never cross.  Each kept cell is tilted by
    N(0, angle_noise).  The n cells nearest the origin are returned, sorted by
    x.  Requires scipy (k-d tree for the Voronoi test)."""
    from scipy.spatial import cKDTree

    area_per_cell = (0.5 * (lmin + lmax) - diameter) * diameter + math.pi * diameter ** 2 / 4
    R = math.sqrt(n * area_per_cell / (0.80 * math.pi)) + domain
    rowh = domain * math.sqrt(3) / 2
    ys = np.arange(-R - domain, R + domain + rowh, rowh)
    xs = np.arange(-R - domain, R + domain + domain, domain)
    sx, sy = np.meshgrid(xs, ys)
    sx = sx + (np.arange(len(ys))[:, None] % 2) * domain / 2
    sites = np.stack([sx.ravel(), sy.ravel()], 1)
    ns = len(sites)
    sites = sites + (uniform(seed, 40, 2 * ns).reshape(ns, 2) - 0.5) * 0.3 * domain
    keep_site = np.hypot(sites[:, 0], sites[:, 1]) < R + domain
    sites = sites[keep_site]
    ns = len(sites)
    theta_s = math.pi * uniform(seed, 41, ns)
    tree = cKDTree(sites)
    half = 0.9 * domain
    pitch_v = diameter + row_gap
    nrows = int(2 * half / pitch_v) + 2
    per_row = int(2 * half / lmin) + 3
    pos_l, th_l, len_l = [], [], []
    for s in range(ns):
        L = lmin + (lmax - lmin) * uniform(seed, 100_000 + s, nrows * per_row).reshape(nrows, per_row)
        u0 = -half - 1.5 * uniform(seed, 200_000 + s, nrows)[:, None]
        ends = u0 + np.cumsum(L + end_gap, axis=1)
        uc = ends - (L + end_gap) / 2 - end_gap / 2
        v = (-half + (np.arange(nrows) + 0.5) * pitch_v)[:, None] * np.ones((1, per_row))
        ok = np.abs(uc) + L / 2 <= half
        uc, v, L = uc[ok], v[ok], L[ok]
        ct, st = math.cos(theta_s[s]), math.sin(theta_s[s])
        x = sites[s, 0] + uc * ct - v * st
        y = sites[s, 1] + uc * st + v * ct
        hc = L / 2 - diameter / 2
        # both core end points inside the own Voronoi region with margin d/2
        good = np.ones(len(x), bool)
        for sgn in (-1.0, 1.0):
            P = np.stack([x + sgn * hc * ct, y + sgn * hc * st], 1)
            d, idx = tree.query(P, k=4)
            good &= idx[:, 0] == s
            own = sites[s]
            for j in range(1, 4):
                other = sites[idx[:, j]]
                sep = np.linalg.norm(other - own, axis=1)
                margin = ((P - other) ** 2).sum(1) - ((P - own) ** 2).sum(1)
                good &= margin >= diameter * sep
        pos_l.append(np.stack([x[good], y[good]], 1))
        th_l.append(np.full(good.sum(), theta_s[s]))
        len_l.append(L[good])
    pos2 = np.concatenate(pos_l)
    theta = np.concatenate(th_l)
    length = np.concatenate(len_l)
    rr = np.hypot(pos2[:, 0], pos2[:, 1])
    if len(rr) < n:
        raise ValueError("dense_colony: generated too few cells")
    keep = np.argsort(rr, kind="stable")[:n]
    pos = np.zeros((n, 3))
    pos[:, :2] = pos2[keep]
    noise = angle_noise * normal(seed, 42, n)
    return _sorted_bodies(pos, np.mod(theta[keep] + noise, math.pi), length[keep], diameter)


def ensemble(n_colonies: int, seed: int = 0, lo: int = 500, hi: int = 50_000):
    """Colony packings (packed_colony) with sizes log-uniform in [lo, hi],
    colony i from seed + 1000 + i: the ensemble of independent colonies."""
    u = np.array(uniform(seed + 977, 0, n_colonies))
    sizes = np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo))).astype(int)
    return [packed_colony(int(n), seed=seed + 1000 + i) for i, n in enumerate(sizes)]


def area_fraction(bodies: Bodies) -> float:
    """Summed spherocylinder area over the area of the enclosing disc."""
    d = bodies.diameter
    a = ((bodies.length - d) * d + math.pi * d * d / 4).sum()
    R = disc_radius(bodies) + 0.5 * bodies.length.max()
    return float(a / (math.pi * R * R))


def single_cell(length: float = 1.0, l0: float = 1.0) -> Bodies:
    """The reference's initial colony: one cell at the origin, q = (1,0,0,0),
    length l0, diameter l0/2 (code/cpp/src/main.cpp:27-32)."""
    return Bodies(np.zeros((1, 3)), np.array([[1.0, 0.0, 0.0, 0.0]]), np.array([length]),
                  np.array([l0]), np.array([l0 / 2]), np.zeros(1))


def divide(bodies: Bodies, seed: int, step: int, noise: float = 0.01) -> Bodies:
    """Seeded division outside the step (replaces Particle::divide,
    code/cpp/src/simulation/Particle.cpp:122-150, whose split uses an unseeded
    std::random_device): every cell with length >= 2*l0 splits into two halves
    of equal length along its axis, each rotated in-plane by N(0, noise) rad.
    Children of cell i are placed at i (left) and appended (right); the result
    is re-sorted by x with gid = index (Domain.cpp:492-508)."""
    L = bodies.length
    div = np.nonzero(L >= 2 * bodies.l0)[0]
    if len(div) == 0:
        return bodies
    pos = bodies.pos.copy()
    quat = bodies.quat.copy()
    length = L.copy()
    ldot = bodies.ldot.copy()
    theta = 2.0 * np.arctan2(quat[:, 3], quat[:, 0])
    t = np.stack([np.cos(theta), np.sin(theta)], axis=1)
    half = L[div] / 2.0
    left = pos[div, :2] - (half / 2)[:, None] * t[div]
    right = pos[div, :2] + (half / 2)[:, None] * t[div]
    nz = noise * normal(seed, 1000 + step, 2 * len(div))
    new_pos = np.zeros((len(div), 3))
    new_pos[:, :2] = right
    new_pos[:, 2] = pos[div, 2]
    pos[div, :2] = left
    th_left = theta[div] + nz[: len(div)]
    th_right = theta[div] + nz[len(div):]
    quat[div] = _planar_quat(th_left)
    length[div] = half
    allpos = np.concatenate([pos, new_pos])
    allquat = np.concatenate([quat, _planar_quat(th_right)])
    alllen = np.concatenate([length, half])
    allldot = np.concatenate([ldot, ldot[div]])
    n = len(alllen)
    order = np.argsort(allpos[:, 0], kind="stable")
    l0 = np.concatenate([bodies.l0, bodies.l0[div]])
    d = np.concatenate([bodies.diameter, bodies.diameter[div]])
    return Bodies(allpos[order], allquat[order], alllen[order], l0[order], d[order],
                  allldot[order])


def two_rods(sep: float, angle: float = 0.0, la: float = 1.0, lb: float = 1.0,
             offset_y: float = 0.0, d: float = 0.5) -> Bodies:
    """Two in-plane rods: A at the origin along x, B at (sep, offset_y) at
    `angle` radians (gid order = x order)."""
    pos = np.array([[0.0, 0.0, 0.0], [sep, offset_y, 0.0]])
    th = np.array([0.0, angle])
    return Bodies(pos, _planar_quat(th), np.array([la, lb]), np.ones(2), np.full(2, d),
                  np.zeros(2))
