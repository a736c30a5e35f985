"""The split-precision form against FP64 over 200 steps of the torus at its
full size (192 x 128 x 128, FMKS, a = 0.9375).

FP32 reconstruction / fluxes / EMFs over an FP64 conserved state put a
relative error of about 6e-8 into every face flux, twice a step.  Over 200
steps the state deviates from the FP64 run by the per-variable relative L1
held below, measured on this workload (largest of several runs) with a margin
of 3.  The deviation is of the size the problem itself gives to a one-time
relative perturbation of 1e-6 of the initial state in FP64 for rho, u, B and
the torus body's u^2, which the second test checks on the same run; the
velocities u^1 and u^3 of the floored atmosphere (two thirds of the zones)
carry most of the rest, where the net force on a nearly empty zone is the
small difference of FP32 pressure fluxes that the FP64 geometric source
balances.
"""
import numpy as np
import pytest

from kharma_amd import problems as pr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NSTEPS = 200
NAMES = ["rho", "u", "u1", "u2", "u3", "B1", "B2", "B3"]
# relative L1 after 200 steps, whole interior and torus body (rho > 1e-2, off
# the four polar rows); measured: whole 5.7e-7 1.19e-6 1.29e-5 9.6e-5 1.8e-5
# 5.4e-7 2.2e-7 1.66e-6, body 5.2e-7 1.18e-6 1.7e-7 5.9e-5 9.0e-8 5.4e-7 2.2e-7 1.66e-6.
# The atmosphere's velocity figures (whole u1-u3) are the size of a random walk
# driven by FP32 flux rounding through floored zones: changing only the
# summation order of the FP64 geometric source (in both forms) moved them from
# 7.1e-6 4.8e-5 1.1e-5 to the values above, while the body figures moved < 10 %.
BOUND_WHOLE = np.array([2e-6, 4e-6, 4e-5, 3e-4, 6e-5, 2e-6, 1e-6, 5e-6])
BOUND_BODY = np.array([2e-6, 4e-6, 5e-7, 2e-4, 3e-7, 2e-6, 1e-6, 5e-6])


@pytest.fixture(scope="module")
def runs():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import kharma_amd as kh
    cfg = pr.torus_config(192, 128, 128)
    P0 = pr.torus_prims(cfg)
    sl = (slice(None),) + pr.interior(cfg)

    def run(prec, P):
        print(f"[split drift] {prec}, {NSTEPS} steps", flush=True)
        d = kh.Driver(dict(cfg, precision=prec))
        d.init_prims(torch.from_numpy(P).cuda())
        d.step(NSTEPS)
        Pn, _ = d.state()
        torch.cuda.synchronize()
        d.check()
        return Pn.cpu().numpy()[sl]

    ref = run("fp64", P0)
    split = run("split", P0)
    rng = np.random.default_rng(7)
    pert = run("fp64", P0 * (1 + 1e-6 * rng.standard_normal(P0.shape)))
    j = np.arange(cfg["nx"][1])[None, :, None] * np.ones(ref.shape[1:])
    body = (ref[0] > 1e-2) & ~((j < 4) | (j >= cfg["nx"][1] - 4))
    return ref, split, pert, body


def _rel(P, ref, mask=None):
    d = np.abs(P - ref)
    if mask is None:
        return d.sum(axis=(1, 2, 3)) / np.abs(ref).sum(axis=(1, 2, 3))
    return np.array([d[v][mask].sum() / np.abs(ref[v][mask]).sum() for v in range(8)])


def test_split_drift_within_bound(runs):
    ref, split, _, body = runs
    whole, b = _rel(split, ref), _rel(split, ref, body)
    print("[split drift] whole " + " ".join(f"{n} {x:.2e}" for n, x in zip(NAMES, whole)), flush=True)
    print("[split drift] body  " + " ".join(f"{n} {x:.2e}" for n, x in zip(NAMES, b)), flush=True)
    assert np.all(whole < BOUND_WHOLE), (whole, BOUND_WHOLE)
    assert np.all(b < BOUND_BODY), (b, BOUND_BODY)


def test_split_drift_as_small_as_a_1e6_perturbation(runs):
    """rho, u, B (whole field) and the body's u^2: the split form's deviation is
    within 3x of what a relative 1e-6 perturbation of the initial state gives
    the FP64 run (measured ratios 1.0, 2.2, 0.33, 0.22, 0.09 and 2.0)."""
    ref, split, pert, body = runs
    s, p = _rel(split, ref), _rel(pert, ref)
    sb, pb = _rel(split, ref, body), _rel(pert, ref, body)
    print("[split drift] 1e-6 perturbation whole " + " ".join(f"{n} {x:.2e}" for n, x in zip(NAMES, p)), flush=True)
    idx = [0, 1, 5, 6, 7]
    assert np.all(s[idx] <= 3 * p[idx]), (s[idx], p[idx])
    assert sb[3] <= 3 * pb[3], (sb[3], pb[3])
