"""Seeded synthetic tube volumes for the five measured workloads.

Every generator is deterministic for a given seed (numpy PCG64).  Tubes are
line segments rendered binary by "voxel-centre distance to segment <= r";
trees grow each child from the end of an existing segment at 25-60 degrees to
its parent, random azimuth, length U(8, 0.4*N).  Volumes are float32 arrays of
shape (Z, Y, X).
"""
import numpy as np


def _unit(v):
    return v / np.linalg.norm(v)


def _perp(v, rng):
    a = np.array([1.0, 0.0, 0.0]) if abs(v[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    p = _unit(np.cross(v, a))
    q = np.cross(v, p)
    phi = rng.uniform(0, 2 * np.pi)
    return np.cos(phi) * p + np.sin(phi) * q


def tree_segments(shape, n_tubes, radius_fn, rng, margin=2.0):
    """Grow a branching tree of `n_tubes` segments inside `shape` (Z, Y, X).

    radius_fn(level, rng) gives the radius of a segment at tree depth `level`.
    Returns a list of (A, B, r, level) with A, B in (x, y, z) voxel coordinates.
    """
    Z, Y, X = shape
    hi = np.array([X, Y, Z], float) - 1.0
    N = min(X, Y, Z)
    segs = []
    r0 = min(radius_fn(0, rng), 0.2 * N)
    lo_b = min(r0 + margin, 0.5 * N - 1.0)
    A = np.array([rng.uniform(lo_b, hi[0] - lo_b), rng.uniform(lo_b, hi[1] - lo_b), lo_b + 1.0])
    d = _unit(np.array([rng.normal(0, 0.3), rng.normal(0, 0.3), 1.0]))
    Lmin = min(8.0, 0.2 * N)
    L = rng.uniform(Lmin, max(0.4 * N, Lmin + 1))
    B = np.clip(A + L * d, lo_b, hi - lo_b)
    segs.append((A, B, r0, 0))
    tries = 0
    while len(segs) < n_tubes and tries < 50 * n_tubes:
        tries += 1
        pa, pb, pr, lvl = segs[rng.integers(len(segs))]
        axis = _unit(pb - pa) if np.linalg.norm(pb - pa) > 1e-6 else np.array([0, 0, 1.0])
        ang = np.deg2rad(rng.uniform(25, 60))
        dirn = _unit(np.cos(ang) * axis + np.sin(ang) * _perp(axis, rng))
        r = min(radius_fn(lvl + 1, rng), 0.2 * N)
        L = rng.uniform(Lmin, max(0.4 * N, Lmin + 1))
        lo_b = min(r + margin, 0.5 * N - 1.0)
        A = pb.copy()
        B = np.clip(A + L * dirn, lo_b, hi - lo_b)
        if np.linalg.norm(B - A) < min(4.0, 0.2 * N):
            continue
        segs.append((A, B, r, lvl + 1))
    return segs


def _render(vol, A, B, r, value):
    """Set vol to `value` where the distance to segment AB is <= r (in place)."""
    Z, Y, X = vol.shape
    lo = np.maximum(np.floor(np.minimum(A, B) - r - 1), 0).astype(int)
    hi = np.minimum(np.ceil(np.maximum(A, B) + r + 1), np.array([X, Y, Z]) - 1).astype(int)
    AB = B - A
    L2 = float(AB @ AB)
    xs = np.arange(lo[0], hi[0] + 1, dtype=np.float64)
    ys = np.arange(lo[1], hi[1] + 1, dtype=np.float64)
    for z in range(lo[2], hi[2] + 1):
        # restrict to the part of the segment whose z is within r of this slice
        if abs(AB[2]) > 1e-9:
            t0, t1 = sorted(((z - r - A[2]) / AB[2], (z + r - A[2]) / AB[2]))
            t0, t1 = max(t0, 0.0), min(t1, 1.0)
            if t0 > t1:
                continue
            P0, P1 = A + t0 * AB, A + t1 * AB
            xl, xh = min(P0[0], P1[0]) - r - 1, max(P0[0], P1[0]) + r + 1
            yl, yh = min(P0[1], P1[1]) - r - 1, max(P0[1], P1[1]) + r + 1
            sx = xs[(xs >= xl) & (xs <= xh)]
            sy = ys[(ys >= yl) & (ys <= yh)]
        else:
            if abs(z - A[2]) > r:
                continue
            sx, sy = xs, ys
        if sx.size == 0 or sy.size == 0:
            continue
        px = sx[None, :] - A[0]
        py = sy[:, None] - A[1]
        pz = z - A[2]
        if L2 > 0:
            t = np.clip((px * AB[0] + py * AB[1] + pz * AB[2]) / L2, 0.0, 1.0)
        else:
            t = np.zeros(np.broadcast(px, py).shape)
        dx = px - t * AB[0]
        dy = py - t * AB[1]
        dz = pz - t * AB[2]
        inside = dx * dx + dy * dy + dz * dz <= r * r
        if inside.any():
            yi = (sy - 0).astype(int)
            xi = (sx - 0).astype(int)
            sub = vol[z, yi[0]:yi[-1] + 1, xi[0]:xi[-1] + 1]
            sub[inside] = value


def render_tubes(shape, segs, value=1.0, dtype=np.float32, radius_add=0.0):
    vol = np.zeros(shape, dtype)
    for A, B, r, _ in segs:
        _render(vol, A, B, r + radius_add, value)
    return vol


def tube_volume(shape, n_tubes, rmin, rmax, seed, noise=0.0):
    """Binary tree of tubes (1.0 on 0.0) plus N(0, noise^2)."""
    rng = np.random.default_rng(seed)
    segs = tree_segments(shape, n_tubes, lambda lvl, g: g.uniform(rmin, rmax), rng)
    vol = render_tubes(shape, segs)
    if noise > 0:
        vol += rng.standard_normal(shape, dtype=np.float32) * np.float32(noise)
    return vol, segs


def config1(seed=101):
    """64^3, 12 tubes r ~ U(1,4), 1.0 on 0.0, + N(0, 0.1^2)."""
    return tube_volume((64, 64, 64), 12, 1.0, 4.0, seed, noise=0.1)


def airway(shape=(256, 256, 256), levels=5, seed=202, noise=0.05):
    """Airway-like tree: radius 12 halving per level (>= 1), lumen 0.0, wall
    band U(1,2) at 1.0, background 0.15, + N(0, noise^2)."""
    rng = np.random.default_rng(seed)
    rad = lambda lvl, g: max(12.0 / (2 ** min(lvl, levels - 1)), 1.0)
    n = 2 ** levels - 1
    segs = tree_segments(shape, n, rad, rng, margin=4.0)
    vol = np.full(shape, 0.15, np.float32)
    walls = [rng.uniform(1, 2) for _ in segs]
    for (A, B, r, _), w in zip(segs, walls):
        _render(vol, A, B, r + w, 1.0)
    for A, B, r, _ in segs:
        _render(vol, A, B, r, 0.0)
    vol += rng.standard_normal(shape, dtype=np.float32) * np.float32(noise)
    return vol, segs


def config2(seed=202):
    return airway((256, 256, 256), 5, seed)


def ct_vessels(shape=(400, 512, 512), n_tubes=40, seed=303):
    """Tubes 1.0 r ~ U(0.8, 8) over tissue N(0.3, 0.05^2) * (1 + 0.2 B),
    B = sum of 3 random low-frequency cosines normalised to [-1, 1]."""
    rng = np.random.default_rng(seed)
    segs = tree_segments(shape, n_tubes, lambda lvl, g: g.uniform(0.8, 8.0), rng)
    Z, Y, X = shape
    zz, yy, xx = np.meshgrid(np.arange(Z) / Z, np.arange(Y) / Y, np.arange(X) / X, indexing="ij",
                             sparse=True)
    B = np.zeros(shape, np.float32)
    for _ in range(3):
        k = rng.uniform(0.5, 2.0, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        B += (np.cos(2 * np.pi * k[0] * xx + ph[0]) * np.cos(2 * np.pi * k[1] * yy + ph[1]) *
              np.cos(2 * np.pi * k[2] * zz + ph[2])).astype(np.float32)
    B /= max(float(np.abs(B).max()), 1e-6)
    vol = (0.3 + 0.05 * rng.standard_normal(shape, dtype=np.float32)) * (1 + 0.2 * B)
    tubes = render_tubes(shape, segs)
    vol[tubes > 0] = 1.0
    return vol.astype(np.float32), segs


def config3(seed=303):
    return ct_vessels((400, 512, 512), 40, seed)


def config4(seeds=(401, 402, 403, 404), size=128):
    return [tube_volume((size, size, size), 12, 1.0, 5.0, s, noise=0.1) for s in seeds]


def mra_thin(shape=(512, 512, 512), n_tubes=120, seed=505, scale=0.08):
    """Thin vessels r ~ U(0.5, 3) at 1.0 with Rician noise: sqrt((s+n1)^2+n2^2)."""
    rng = np.random.default_rng(seed)
    segs = tree_segments(shape, n_tubes, lambda lvl, g: g.uniform(0.5, 3.0), rng)
    s = render_tubes(shape, segs)
    n1 = rng.standard_normal(shape, dtype=np.float32) * np.float32(scale)
    s += n1
    del n1
    n2 = rng.standard_normal(shape, dtype=np.float32) * np.float32(scale)
    np.square(s, out=s)
    np.square(n2, out=n2)
    s += n2
    del n2
    np.sqrt(s, out=s)
    return s, segs


def config5(seed=505, shape=(512, 512, 512)):
    return mra_thin(shape, 120, seed)
