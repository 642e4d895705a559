"""Spatial pruning of the sweeps is exact.

The device keeps the points in a spatial order in 64-point chunks with
bounding boxes; a sweep skips the chunks whose points are all farther from a
workgroup's focus points than the distance at which the kernel weight is
exactly zero in the sweep's arithmetic.  Skipped chunks contribute exactly
zero, so criteria, betas, standard errors, qDiag and the hat traces are the
same bits with pruning on and off, for every kernel, fixed and adaptive
bandwidths, both distances and both precisions.
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def clustered(seed, n, k, geographic):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(10, 90, (6, 2))
    c = centres[rng.integers(0, 6, n)] + rng.normal(0, 4, (n, 2))
    if geographic:
        c = np.column_stack([c[:, 0] * 0.4 - 10, c[:, 1] * 0.3 + 30])   # lon, lat degrees
    X = np.column_stack([np.ones(n)] + [rng.normal(size=n) for _ in range(k - 1)])
    y = X @ rng.normal(size=k) + np.sin(c[:, 0] / 7) + rng.normal(0, 0.3, n)
    return c, X, y


KERNELS = [capi.GAUSSIAN, capi.EXPONENTIAL, capi.BISQUARE, capi.TRICUBE, capi.BOXCAR]


@pytest.mark.parametrize("geographic", [False, True])
@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_pruning_is_exact(geographic, precision):
    n, k = 3000, 4
    c, X, y = clustered(7, n, k, geographic)
    dist = capi.GREAT_CIRCLE if geographic else capi.EUCLIDEAN
    fixed = {capi.GAUSSIAN: 2.0, capi.EXPONENTIAL: 0.5, capi.BISQUARE: 6.0, capi.TRICUBE: 6.0, capi.BOXCAR: 6.0}
    scale = 111.0 * 0.35 if geographic else 1.0   # km per unit of the planar cluster layout
    with capi.Context(0) as ctx:
        ctx.set_data(c, X, y, dist)
        for kern in KERNELS:
            for adaptive, bw in ((False, fixed[kern] * scale), (True, 60.0)):
                say("case: geographic=%s precision=%d kernel=%d adaptive=%s bw=%g" % (geographic, precision, kern,
                                                                                    adaptive, bw))
                out = []
                for on in (True, False):
                    ctx.set_pruning(on)
                    cv = ctx.criterion(kern, adaptive, bw, capi.CRIT_CV, precision)
                    aic = ctx.criterion(kern, adaptive, bw, capi.CRIT_AIC, precision)
                    try:
                        fit = ctx.gwr_run(kern, adaptive, bw, precision=precision)
                    except capi.GwmError as e:
                        fit = {"error": e.status}
                    out.append((cv, aic, fit))
                (cv1, aic1, f1), (cv0, aic0, f0) = out
                assert cv1 == cv0 and aic1 == aic0
                assert ("error" in f1) == ("error" in f0)
                if "error" not in f1:
                    for key in ("betas", "betasSE", "qDiag", "shat"):
                        assert np.array_equal(f1[key], f0[key], equal_nan=True), key


@pytest.mark.parametrize("geographic", [False, True])
def test_pruned_adaptive_select_is_exact(geographic, oracle):
    """Adaptive bandwidths (rank floor(bw) / floor(bw)+1 order statistics)
    from the pruned select equal the unpruned select bit for bit and the
    oracle's sort (Euclidean: bit for bit), with duplicated points (ties) and
    ranks up to n - 1."""
    n = 5000
    c, X, y = clustered(9, n, 2, geographic)
    c[::7] = c[3]            # a block of identical points: ties at every rank they cover
    c[1::11] = np.round(c[1::11], 1)
    dist = capi.GREAT_CIRCLE if geographic else capi.EUCLIDEAN
    with capi.Context(0) as ctx:
        ctx.set_data(c, X, y, dist)
        for bw in (20.0, 20.5, 64.0, 100.25, 700.0, 1000.0, 4999.5, 4998.0):
            say("case: geographic=%s bw=%g" % (geographic, bw))
            ctx.set_pruning(True)
            a = ctx.adaptive_bandwidths(bw)
            ctx.set_pruning(False)
            b = ctx.adaptive_bandwidths(bw)
            assert np.array_equal(a, b)
            ref = np.array([oracle.adaptive_threshold(c, bw, f, geographic=geographic) for f in range(0, n, 97)])
            if geographic:
                # the device's great-circle formula agrees with the reference's to rounding, not bitwise
                assert np.allclose(a[::97], ref, rtol=1e-12, atol=0)
            else:
                assert np.array_equal(a[::97], ref)
