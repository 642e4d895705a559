"""Seeded synthetic GWR workloads.

Every random number comes from SplitMix64 -> 53-bit uniform -> Box-Muller, in
numpy uint64 arithmetic, so any host produces the same inputs for the same
seed.  Each generator takes n so that small parity cases keep the shape of the
full workload.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class Stream:
    """Counter-based SplitMix64 stream."""

    def __init__(self, seed):
        self.seed = np.uint64(seed)
        self.counter = 0

    def u64(self, count):
        with np.errstate(over="ignore"):
            idx = np.arange(self.counter + 1, self.counter + 1 + count, dtype=np.uint64)
            z = self.seed + idx * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.counter += count
        return z

    def uniform(self, count, lo=0.0, hi=1.0):
        u = (self.u64(count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, count, mu=0.0, sigma=1.0):
        m = (count + 1) // 2
        u1 = 1.0 - self.uniform(m)          # (0, 1]
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:count]
        return mu + sigma * z


def config1(n=2000, seed=1001):
    """n points on U[0,100]^2, k = 4: beta0 = 1+(u+v)/50, beta1 = sin(u/20),
    beta2 = cos(v/25), beta3 = 0.5, noise N(0, 0.5^2)."""
    s = Stream(seed)
    u = s.uniform(n, 0, 100)
    v = s.uniform(n, 0, 100)
    X = np.column_stack([np.ones(n), s.normal(n), s.normal(n), s.normal(n)])
    B = np.column_stack([1 + (u + v) / 50, np.sin(u / 20), np.cos(v / 25), np.full(n, 0.5)])
    y = np.sum(X * B, axis=1) + s.normal(n, 0, 0.5)
    return dict(coords=np.column_stack([u, v]), X=X, y=y, beta=B)


def config2(n=100_000, seed=1002, k=8):
    """lon ~ U[-10,30], lat ~ U[35,60]; X = [1, 7 covariates N(0,1) with
    equicorrelation 0.3]; each beta is a sum of 3 Gaussian bumps (centre in the
    box, width U[2,8] degrees, amplitude N(0,1)); noise N(0, 0.3^2)."""
    s = Stream(seed)
    lon = s.uniform(n, -10, 30)
    lat = s.uniform(n, 35, 60)
    p = k - 1
    C = np.full((p, p), 0.3) + 0.7 * np.eye(p)
    L = np.linalg.cholesky(C)
    Zc = s.normal(n * p).reshape(n, p) @ L.T
    X = np.column_stack([np.ones(n), Zc])
    B = np.zeros((n, k))
    for c in range(k):
        for _ in range(3):
            cx, cy = s.uniform(1, -10, 30)[0], s.uniform(1, 35, 60)[0]
            w = s.uniform(1, 2, 8)[0]
            a = s.normal(1)[0]
            B[:, c] += a * np.exp(-((lon - cx) ** 2 + (lat - cy) ** 2) / (2 * w * w))
    y = np.sum(X * B, axis=1) + s.normal(n, 0, 0.3)
    return dict(coords=np.column_stack([lon, lat]), X=X, y=y, beta=B)


# Bandwidth interval of config 2: [0.5, 20] degrees of arc in the reference's
# great-circle unit (km), 1 degree = 111.3195 km.
CONFIG2_BW_KM = (0.5 * 111.3195, 20 * 111.3195)


def config3(n=50_000, seed=1003, extent=1000.0, scales=(20, 50, 100, 200, 400, 1000)):
    """MGWR workload: U[0,extent]^2, k = 6; beta_j = sum of 16 random cosines
    with length scale scales[j]; noise N(0, 0.5^2)."""
    s = Stream(seed)
    k = len(scales)
    u = s.uniform(n, 0, extent)
    v = s.uniform(n, 0, extent)
    X = np.column_stack([np.ones(n)] + [s.normal(n) for _ in range(k - 1)])
    B = np.zeros((n, k))
    for c, ell in enumerate(scales):
        for _ in range(16):
            a = s.normal(1)[0] / 4.0
            om = s.normal(2)
            ph = s.uniform(1, 0, 2 * np.pi)[0]
            B[:, c] += a * np.cos((om[0] * u + om[1] * v) / ell + ph)
    y = np.sum(X * B, axis=1) + s.normal(n, 0, 0.5)
    return dict(coords=np.column_stack([u, v]), X=X, y=y, beta=B)
