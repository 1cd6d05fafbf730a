"""Recorded batches (vrc_batch_plan_*): the batched stack solve recorded once as a HIP graph
and replayed. Every replay, with the program's own eager work and synchronizations in
between, must give the eager call's results; per-instance failure codes come back from
the replay as from the eager call."""
import numpy as np
import pytest

from tests.test_gpu_failures import rect_instance
from tests.test_gpu_parity_full import metric_instances
from varrcwa_amd import api

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("nh,count,bases", [(5, 6, True), (11, 3, False)])
def test_replays_equal_eager(ctx, nh, count, bases):
    import torch
    insts = metric_instances(40, count)
    eager = api.batch_stack_rcwa(ctx, nh, nh, insts, want_bases=bases)
    plan = api.BatchPlan(ctx, nh, nh, insts, want_bases=bases)
    try:
        for rep in range(3):
            out = plan.launch()
            torch.cuda.synchronize()
            for k in eager:
                assert rel(out[k], eager[k]) <= 1e-13, (rep, k, rel(out[k], eager[k]))
            # eager work between replays (it reuses the context's pool, not the plan's blocks)
            other = api.batch_stack_rcwa(ctx, 3, 3, metric_instances(0, 2))
            torch.cuda.synchronize()
            assert bool(other["S"].isfinite().all())
            out["S"].zero_()  # the next replay rewrites every output
    finally:
        plan.close()


def test_replay_reports_failures(ctx):
    plan = api.BatchPlan(ctx, 3, 3, [rect_instance(0.9), rect_instance(0.6), rect_instance(1.3)])
    try:
        for _ in range(2):
            out = plan.launch(return_info=True)
            assert list(out["info"]) == [0, 4, 0]
        with pytest.raises(api.VrcError, match="instance 1"):
            plan.launch()
    finally:
        plan.close()
