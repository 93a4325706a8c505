"""Seeded synthetic inputs for the five workloads (numpy PCG64; deterministic per seed).

sphere   : N points on the unit sphere + isotropic Gaussian noise (100,000 pts, sigma 0.005, seed 1)
torus    : torus R=1, r=0.3, density varying 1-10x along the major angle, noise sigma 0.003 (4M pts, seed 2)
icosphere: icosahedron subdivided L times, vertices displaced radially by U(-0.01, 0.01) (L=8, seed 3)
mixture  : 12 anisotropic Gaussian clusters + 4 noisy planar patches, local density spanning ~300x (seed 4)
superquadric: random superquadric surface, exponents in [0.3, 3], noise 1e-3 of its extent (seeds 5000+i)
"""
import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _split(p):
    p = np.ascontiguousarray(p, dtype=np.float32)
    return np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1]), np.ascontiguousarray(p[:, 2])


def sphere(n=100_000, sigma=0.005, seed=1):
    r = _rng(seed)
    g = r.standard_normal((n, 3))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    return _split(u + sigma * r.standard_normal((n, 3)))


def torus(n=4_000_000, R=1.0, rr=0.3, sigma=0.003, seed=2):
    r = _rng(seed)
    out_t = np.empty(0)
    while out_t.size < n:  # major angle: density 1 + 9 (1 - cos t) / 2 in [1, 10]
        t = r.uniform(0, 2 * np.pi, 2 * n)
        keep = r.uniform(0, 10, 2 * n) < 1 + 4.5 * (1 - np.cos(t))
        out_t = np.concatenate([out_t, t[keep]])
    t = out_t[:n]
    out_p = np.empty(0)
    while out_p.size < n:  # minor angle: area element (R + r cos p)
        p = r.uniform(0, 2 * np.pi, 2 * n)
        keep = r.uniform(0, R + rr, 2 * n) < R + rr * np.cos(p)
        out_p = np.concatenate([out_p, p[keep]])
    p = out_p[:n]
    pts = np.stack([(R + rr * np.cos(p)) * np.cos(t), (R + rr * np.cos(p)) * np.sin(t), rr * np.sin(p)], axis=1)
    return _split(pts + sigma * r.standard_normal((n, 3)))


def icosphere(levels=8, noise=0.01, seed=3):
    """Returns (x, y, z, tri) with tri an (F, 3) uint32 array; 10*4^L+2 vertices, 20*4^L triangles."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(levels):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        es = np.sort(e, axis=1)
        uniq, inv = np.unique(es[:, 0] * (1 << 32) + es[:, 1], return_inverse=True)
        a, b = uniq >> 32, uniq & 0xFFFFFFFF
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        m = f.shape[0]
        m01, m12, m20 = base + inv[:m], base + inv[m:2 * m], base + inv[2 * m:]
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    r = _rng(seed)
    v = v * (1.0 + r.uniform(-noise, noise, (v.shape[0], 1)))
    x, y, z = _split(v)
    return x, y, z, np.ascontiguousarray(f, dtype=np.uint32)


def _rotation(r):
    q, rr = np.linalg.qr(r.standard_normal((3, 3)))
    return q * np.sign(np.diag(rr))


def mixture(n=16_000_000, seed=4):
    """12 anisotropic Gaussians (axis ratios up to 20:1) + 4 noisy planar patches; per-component counts and
    scales are drawn so local densities span roughly 300x."""
    r = _rng(seed)
    comps = 16
    w = np.exp(r.uniform(np.log(1.0), np.log(20.0), comps))
    counts = np.floor(w / w.sum() * n).astype(np.int64)
    counts[0] += n - counts.sum()
    parts = []
    for c in range(comps):
        m = int(counts[c])
        centre = r.uniform(-1, 1, 3)
        rot = _rotation(r)
        if c < 12:
            scale = np.exp(r.uniform(np.log(0.02), np.log(0.3)))
            axes = scale * np.array([1.0, np.exp(r.uniform(np.log(1 / 20), 0)), np.exp(r.uniform(np.log(1 / 20), 0))])
            p = r.standard_normal((m, 3)) * axes
        else:
            half = r.uniform(0.2, 0.8, 2)
            p = np.stack([r.uniform(-half[0], half[0], m), r.uniform(-half[1], half[1], m),
                          0.002 * r.standard_normal(m)], 1)
        parts.append(p @ rot.T + centre)
    pts = np.concatenate(parts)
    pts = pts[r.permutation(pts.shape[0])]
    return _split(pts)


def superquadric(n, seed):
    r = _rng(seed)
    e1, e2 = r.uniform(0.3, 3.0, 2)
    a = r.uniform(0.5, 2.0, 3)
    eta = r.uniform(-np.pi / 2, np.pi / 2, n)
    om = r.uniform(-np.pi, np.pi, n)

    def sp(v, e):
        return np.sign(v) * np.abs(v) ** e

    ce, se = np.cos(eta), np.sin(eta)
    pts = np.stack([a[0] * sp(ce, e1) * sp(np.cos(om), e2), a[1] * sp(ce, e1) * sp(np.sin(om), e2),
                    a[2] * sp(se, e1)], 1)
    pts = pts @ _rotation(r).T
    ext = float(np.max(pts.max(0) - pts.min(0)))
    pts += 1e-3 * ext * r.standard_normal((n, 3))
    return _split(pts)


def batched_clouds(count=512, lo=5_000, hi=500_000, seed=5000):
    """Uneven batch: count clouds with sizes log-uniform in [lo, hi], k drawn from {16, 40, 64}."""
    r = _rng(seed)
    sizes = np.floor(np.exp(r.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    ks = r.choice(np.array([16, 40, 64]), count)
    clouds = [superquadric(int(sizes[i]), seed + i) for i in range(count)]
    return clouds, ks


def tiny_lattice(m=6, spacing=1.0):
    """Exactly equidistant cubic lattice (exercises the distance tie rule)."""
    g = np.arange(m, dtype=np.float32) * np.float32(spacing)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return _split(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1))
