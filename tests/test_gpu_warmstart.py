"""What the trained warm-start network gains over the plain imager, on
held-out seeds of its training distribution (seeds from 1,000,000 up; the
network was trained on seeds below 62,500), with the committed weights
(smili_amd/warmstart/weights/unet128_l1tsv.npz).

For every case: the network's proposal (flux scale fitted along the
objective) has a lower objective than the flat prior the plain imager
starts from; the warm-started L-BFGS-B reaches the plain solve's final
objective within 1e-6 relative; its image then matches the plain image to
NRMSE <= 5e-3; and over the cases it needs fewer iterations in total, with
a median iteration ratio >= 1 and a median time-to-image ratio >= 0.85.

Why 5e-3 and not the 1e-3 image target: a 1e-6 match of an objective that
is quadratic near its minimum fixes the image only to about sqrt(1e-6) =
1e-3 along its flattest directions, so two solves that both stop at that
objective differ by a few 1e-3 (measured up to 5.5e-3 over 1,000 held-out
seeds).  Why 0.85: over 1,000 held-out seeds the median time ratio is
1.05 and over 20 it is 1.1; six cases on a shared box scatter by ~10 %."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS = range(1_000_000, 1_000_006)


@pytest.fixture(scope="module")
def net(gpu_lib):
    from smili_amd.warmstart.unet import UNet
    n, meta = UNet.from_file(batch=1)
    yield n
    n.close()


def test_proposal_beats_prior(net):
    from smili_amd.cost import CostProblem
    from smili_amd.warmstart import data as D
    from smili_amd.warmstart.solve import propose
    for s in SEEDS:
        case = D.make_case(s)
        prep = D.prepare_case(case)
        with CostProblem(**prep.kwargs) as prob:
            x0, _ = propose(net, case, prep, prob)
            f_net, f_prior = prob.cost_grad(x0)[0], prob.cost_grad(prep.x0)[0]
        print(s, case.kind, "objective: prior %.4g proposal %.4g" % (f_prior, f_net))
        assert f_net < f_prior


def test_warm_start_reaches_plain_objective_in_fewer_iterations(net):
    from smili_amd.warmstart.solve import summary, time_to_image
    rows = [time_to_image(net, s) for s in SEEDS]
    for r in rows:
        print(r["seed"], r["kind"], "plain", r["plain_iters"], "it %.3f s" % r["plain_s"],
              "warm", r["warm_iters"], "it", None if r["warm_s"] is None else "%.3f s" % r["warm_s"],
              "nrmse %.2e" % r["nrmse"])
    print(summary(rows))
    assert all(r["reached"] for r in rows)
    assert all(r["warm_start"] == "proposal" for r in rows)
    assert max(r["nrmse"] for r in rows) <= 5e-3
    assert sum(r["warm_iters"] for r in rows) < sum(r["plain_iters"] for r in rows)
    sm = summary(rows)
    assert sm["median_iter_ratio"] >= 1.0
    assert sm["median_time_ratio"] >= 0.85


def test_gap_curves(net):
    """Iterations to each objective gap: both solves reach every gap, in
    order, and the warm start reaches the loose gaps no later."""
    from smili_amd.warmstart.solve import GAPS, gap_curve
    c = gap_curve(net, SEEDS[0])
    print(c)
    for k in ("plain", "warm"):
        assert all(i is not None for i in c[k]) and c[k] == sorted(c[k])
    assert c["warm"][GAPS.index(1.0)] <= c["plain"][GAPS.index(1.0)]
