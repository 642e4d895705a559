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


def uniform_square(n=2000, seed=1001):
    """n points on U[0,100]^2, k = 4: beta0 = 1+(u+v)/50, beta1 = sin(u/20),
    beta2 = cos(v/25), beta3 = 0.5, noise N(0, 0.5^2)."""
    s = Stream(seed)
    u = s.uniform(n, 0, 100)
    v = s.uniform(n, 0, 100)
    X = np.column_stack([np.ones(n), s.normal(n), s.normal(n), s.normal(n)])
    B = np.column_stack([1 + (u + v) / 50, np.sin(u / 20), np.cos(v / 25), np.full(n, 0.5)])
    y = np.sum(X * B, axis=1) + s.normal(n, 0, 0.5)
    return dict(coords=np.column_stack([u, v]), X=X, y=y, beta=B)


def sphere_box(n=100_000, seed=1002, k=8):
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


# Bandwidth interval of the sphere-box workload: [0.5, 20] degrees of arc in the reference's
# great-circle unit (km), 1 degree = 111.3195 km.
SPHERE_BOX_BW_KM = (0.5 * 111.3195, 20 * 111.3195)


def multiscale_cosines(n=50_000, seed=1003, extent=1000.0, scales=(20, 50, 100, 200, 400, 1000)):
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


UNEVEN_BATCH_KERNELS = (0, 2, 3)   # Gaussian, bisquare, tricube


def uneven_batch(num=512, seed=1004):
    """Uneven batch: num problems, n ~ logU[200, 20000], k ~ U{3..12},
    coordinates U[0,100]^2, intercept + N(0,1) covariates, each beta linear
    plus one sinusoid (frequency U[0.01, 0.1]), noise N(0, 0.2^2); kernel from
    {Gaussian, bisquare, tricube}, adaptive or fixed with equal probability."""
    s = Stream(seed)
    probs = []
    for p in range(num):
        n = int(np.exp(s.uniform(1, np.log(200), np.log(20000))[0]))
        k = 3 + int(s.uniform(1, 0, 10)[0])
        u = s.uniform(n, 0, 100)
        v = s.uniform(n, 0, 100)
        X = np.column_stack([np.ones(n)] + [s.normal(n) for _ in range(k - 1)])
        B = np.zeros((n, k))
        for c in range(k):
            a0, a1, a2 = s.normal(3)
            fr = s.uniform(1, 0.01, 0.1)[0]
            th = s.uniform(1, 0, 2 * np.pi)[0]
            B[:, c] = a0 + 0.01 * (a1 * u + a2 * v) + np.sin(2 * np.pi * fr * (u * np.cos(th) + v * np.sin(th)))
        y = np.sum(X * B, axis=1) + s.normal(n, 0, 0.2)
        kernel = UNEVEN_BATCH_KERNELS[int(s.uniform(1, 0, 3)[0])]
        adaptive = bool(s.uniform(1)[0] < 0.5)
        probs.append(dict(coords=np.column_stack([u, v]), X=X, y=y, kernel=kernel, adaptive=adaptive))
    return probs


def serving_stream(num=64, rate=2.0, seed=1005, n_range=(500, 200_000)):
    """Serving stream: Poisson arrivals at `rate` requests/s; 70% fits at a
    given bandwidth, 30% CV bandwidth searches; n ~ logU[n_range], k ~ U{2..16};
    coordinates from 8 Gaussian clusters (sigma 5) on [0,100]^2; N(0,1)
    covariates, beta as in multiscale_cosines (random cosines), noise N(0, 0.5^2)."""
    s = Stream(seed)
    centres = s.uniform(16, 10, 90).reshape(8, 2)
    t = 0.0
    reqs = []
    for r in range(num):
        t += -np.log(1.0 - s.uniform(1)[0]) / rate
        n = int(np.exp(s.uniform(1, np.log(n_range[0]), np.log(n_range[1]))[0]))
        k = 2 + int(s.uniform(1, 0, 15)[0])
        cl = (s.uniform(n, 0, 8)).astype(int)
        coords = centres[cl] + s.normal(2 * n, 0, 5).reshape(n, 2)
        X = np.column_stack([np.ones(n)] + [s.normal(n) for _ in range(k - 1)])
        B = np.zeros((n, k))
        for c in range(k):
            ell = 10.0 * (1 + c)
            for _ in range(4):
                a = s.normal(1)[0] / 2.0
                om = s.normal(2)
                ph = s.uniform(1, 0, 2 * np.pi)[0]
                B[:, c] += a * np.cos((om[0] * coords[:, 0] + om[1] * coords[:, 1]) / ell + ph)
        y = np.sum(X * B, axis=1) + s.normal(n, 0, 0.5)
        search = bool(s.uniform(1)[0] < 0.3)
        reqs.append(dict(t=t, coords=coords, X=X, y=y, search=search))
    return reqs


def reference_benchmark(n=50_000, seed=1006):
    """The reference's own GWRBasic benchmark shape (test/testGWRBasic.cpp:380-419):
    u, v ~ U[0,1], X = three N(0,1) columns (no intercept), beta0 = u + v,
    beta1 = u^2 + v^2, beta2 = sin(u) + cos(v), noise N(0,1); fitted with a
    Euclidean Gaussian kernel at the fixed bandwidth 0.2."""
    s = Stream(seed)
    u = s.uniform(n)
    v = s.uniform(n)
    X = np.column_stack([s.normal(n), s.normal(n), s.normal(n)])
    B = np.column_stack([u + v, u * u + v * v, np.sin(u) + np.cos(v)])
    y = np.sum(X * B, axis=1) + s.normal(n)
    return dict(coords=np.column_stack([u, v]), X=X, y=y, beta=B)
