"""The autotuner of sweep launch shapes changes no result bit: every
candidate shape (one or two focus tiles per wave, tiles in order or by a
stride) sums each focus point's data in the same order.  At n = 3000 the
default grid is small, so the first sweep of each class times every shape."""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED])
def test_shapes_are_bit_identical(prec):
    d = synth.uniform_square(n=3000)
    with capi.Context(0) as ctx:
        ctx.set_data(d["coords"], d["X"], d["y"])
        ctx.set_autotune(False)
        say("case: autotune off, precision %d" % prec)
        ref = [ctx.criterion(capi.BISQUARE, True, bw, capi.CRIT_CV, prec) for bw in (60, 90)]
        fit_ref = ctx.fit(capi.GAUSSIAN, False, 8.0, prec)
        ctx.set_autotune(True)
        say("case: autotune on, precision %d" % prec)
        for rep in range(3):
            got = [ctx.criterion(capi.BISQUARE, True, bw, capi.CRIT_CV, prec) for bw in (60, 90)]
            assert got == ref
        for rep in range(3):
            fit = ctx.fit(capi.GAUSSIAN, False, 8.0, prec)
            for key in ("betas", "betasSE_raw", "shat", "qDiag"):
                assert np.array_equal(fit[key], fit_ref[key]), key
    lines = capi.autotune_report()
    say("\n".join(lines))
    assert any("euclidean cv" in l and "n<2^12" in l for l in lines)
