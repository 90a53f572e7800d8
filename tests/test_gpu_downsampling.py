"""Downsampled observations (getDownsampling() > 1) against the oracle.

The single step computes S / downsampling output samples per DM row with the
shift table scaled by S / downsampling (include/Dedispersion.hpp:76,
src/Shifts.cpp:26), reading the input rows sized from S (the test driver's
S_disp, src/DedispersionTest.cpp:129).  The two-step path is refused for
downsampling > 1: the reference's step one pads its output rows to S1
(include/Dedispersion.hpp:148) while its step two reads them at a pitch of
S1 / downsampling (:169), so the two layouts only agree at downsampling 1."""
import pytest

from dedispersion_amd import DD_SINGLE_STEP, Conf, DDError, Plan, workloads
from tests.gpu_helpers import compare_rows, single_step

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("conf", [None, Conf(64, 4, 4, 2, 4, 1), Conf(64, 2, 8, 4, 4, 0)],
                         ids=["default", "staged", "local0"])
@pytest.mark.parametrize("ds", [2, 3, 8])
def test_single_step_downsampled(torch_cuda, ds, conf):
    w = workloads.tiny(channels=80, samples=1200, dms=21, step=0.7, beams=2, sbeams=2, min_freq=1300.0,
                       bandwidth=1.0, zapped=(1, 40), seed=61)
    w.obs.downsampling = ds
    plan, inp, out, orc, flags = single_step(torch_cuda, w, padding=128, conf=conf, seed=61 + ds)
    z = plan.sizes
    assert z.nr_samples == 1200 // ds
    assert z.nr_dispersed_samples == orc.o.nr_samples_per_dispersed_batch
    want = orc.dedispersion(inp, flags)
    compare_rows(out.reshape(want.shape), want, z.nr_samples, f"downsampling {ds}")
    plan.close()


def test_two_step_downsampled_is_refused(torch_cuda):
    w = workloads.tiny(channels=64, samples=512, dms=4, step=0.5, subbands=4, coarse=3, coarse_step=2.0)
    w.obs.downsampling = 2
    with pytest.raises(DDError, match="downsampling"):
        Plan(w.obs, 128)
