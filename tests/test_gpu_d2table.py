"""The great-circle distance table (gwm_hip_set_distance_table) on the GPU
through the C ABI.

Inside one search / fit / backfit call, mixed-precision sweeps of
great-circle data read every pair's FP32 squared distance from a table built
by the call's first such sweep, instead of forming it in every sweep
(GWRBasic.cpp:440-535 forms each distance again per criterion evaluation).
The table holds the values the sweeps form themselves, with the same
roundings, so every result -- golden-section iterates, criteria, betas, SE,
qDiag, diagnostics -- must be bit-identical with the table on and off.
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    yield c
    c.close()


def run_both(ctx, fn):
    ctx.set_distance_table(0)
    off = fn()
    ctx.set_distance_table(96 * 2 ** 30)
    ctx.reset_timers()
    ctx.set_timing(True)
    on = fn()
    ctx.set_timing(False)
    builds, _ = ctx.kernel_time("d2_table")
    return off, on, builds


def same_fit(a, b):
    for key in ("betas", "betasSE", "shat", "qDiag"):
        assert np.array_equal(a[key], b[key]), key
    assert a["bandwidth"] == b["bandwidth"]
    assert a["trace"] == b["trace"]
    for key in a["diagnostic"]:
        assert a["diagnostic"][key] == b["diagnostic"][key], key


@pytest.mark.parametrize("n", [3000, 20011])
def test_fixed_gaussian_aicc_search_and_fit(ctx, n):
    say("case: table on/off, sphere box n=%d, Gaussian AICc search + fit" % n)
    d = synth.sphere_box(n=n)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    lo, hi = synth.SPHERE_BOX_BW_KM
    off, on, builds = run_both(ctx, lambda: ctx.gwr_run(capi.GAUSSIAN, False, 0.0, select_bw=True,
                                                        crit=capi.CRIT_AIC, precision=capi.MIXED, lower=lo,
                                                        upper=hi))
    assert builds == 1
    same_fit(off, on)


@pytest.mark.parametrize("kernel,adaptive,crit", [(capi.BISQUARE, True, capi.CRIT_CV),
                                                  (capi.EXPONENTIAL, False, capi.CRIT_CV),
                                                  (capi.TRICUBE, True, capi.CRIT_AIC)])
def test_other_kernels_search_and_fit(ctx, kernel, adaptive, crit):
    say("case: table on/off kernel=%d adaptive=%d crit=%d" % (kernel, adaptive, crit))
    d = synth.sphere_box(n=6000, seed=77)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    lo, hi = (20, 3000) if adaptive else synth.SPHERE_BOX_BW_KM
    off, on, builds = run_both(ctx, lambda: ctx.gwr_run(kernel, adaptive, 0.0, select_bw=True, crit=crit,
                                                        precision=capi.MIXED, lower=lo, upper=hi))
    assert builds == 1
    same_fit(off, on)


def test_world_scale_pairs_beyond_sixty_degrees(ctx):
    """Points over the whole sphere: many pairs take the general far form."""
    say("case: table on/off, world-scale points")
    rng = np.random.default_rng(3)
    n = 4000
    lon = rng.uniform(-180, 180, n)
    lat = np.degrees(np.arcsin(rng.uniform(-1, 1, n)))
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    y = X @ np.array([1.0, 0.5, -0.3]) + np.sin(np.radians(lat)) + rng.normal(0, 0.3, n)
    ctx.set_data(np.column_stack([lon, lat]), X, y, capi.GREAT_CIRCLE)
    off, on, builds = run_both(ctx, lambda: ctx.gwr_run(capi.GAUSSIAN, False, 0.0, select_bw=True,
                                                        crit=capi.CRIT_CV, precision=capi.MIXED, lower=300.0,
                                                        upper=12000.0))
    assert builds == 1
    same_fit(off, on)


def test_multiscale_backfit(ctx):
    say("case: table on/off, multiscale backfit on great-circle data")
    d = synth.sphere_box(n=2500, seed=11, k=3)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    k = 3
    args = ([capi.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k, [capi.CRIT_CV] * k, [0, 1, 1], [1e-5] * k)

    def run():
        return ctx.mgwr_run(*args, crit_type=0, crit_threshold=1e-5, precision=capi.MIXED)

    off, on, builds = run_both(ctx, run)
    assert builds == 1
    assert np.array_equal(off["bandwidths"], on["bandwidths"])
    assert np.array_equal(off["betas"], on["betas"])
    assert off["iterations"] == on["iterations"]


def test_table_off_by_cap_and_outside_calls(ctx):
    """A cap below the table's size, FP64 runs and single criterion calls
    never build it; a zero cap and a too-small cap are both accepted."""
    say("case: table cap and scope")
    d = synth.sphere_box(n=3000)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    ctx.reset_timers()
    ctx.set_timing(True)
    ctx.set_distance_table(1e6)   # 4 n^2 bytes > 1 MB
    ctx.gwr_run(capi.GAUSSIAN, False, 300.0, precision=capi.MIXED)
    ctx.set_distance_table(96 * 2 ** 30)
    ctx.gwr_run(capi.GAUSSIAN, False, 0.0, select_bw=True, crit=capi.CRIT_AIC, precision=capi.FP64,
                lower=100.0, upper=2000.0)
    ctx.criterion(capi.GAUSSIAN, False, 300.0, capi.CRIT_AIC, capi.MIXED)
    ctx.set_timing(False)
    builds, _ = ctx.kernel_time("d2_table")
    assert builds == 0
    with pytest.raises(capi.GwmError):
        ctx.set_distance_table(-1.0)
