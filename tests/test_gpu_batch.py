"""Several simulations in one context: every instance must evolve exactly as it
would alone (bitwise: the kernels see the same data in the same order; only
blockIdx.y selects the instance), whatever the other instances hold, which of
them are active and how many steps each has taken.

The pass variants are pinned to the batch's choice in the one-instance runs:
fields do not depend on the variant, but the FE/RMSE reductions sum
per-workgroup partials, whose grouping follows the tile width.
"""
import numpy as np
import pytest

from xpfc_cuda_amd import BatchContext, Context, XPFC_FP32_MIXED, XPFC_FP64

from cases import growth_case, make

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _mixed_params(shape, prec):
    """Four instances of one grid with different physics and sampling."""
    ps = []
    _, p = growth_case(shape, prec=prec, t_sample=7)
    ps.append(p)
    _, p = make(shape, prec=prec, sigma=0.1, psi0=0.02, t_sample=5, dt=0.3)
    ps.append(p)
    _, p = make(shape, prec=prec, sigma=0.2, psi0=0.05, eps=0.01, mode="sum", t_sample=0)
    ps.append(p)
    _, p = make(shape, prec=prec, sigma=0.08, psi0=0.04, Mo=1.5, int_Mo=0, t_sample=10)
    ps.append(p)
    return ps


def _fields(shape, k, seed=11):
    rng = np.random.default_rng(seed)
    return [0.05 + 0.02 * rng.standard_normal(shape) for _ in range(k)]


def _alone(p, psi, nsteps, variants):
    with Context(p) as c:
        c.set_variants(*variants)
        c.set_field(psi)
        c.step(nsteps)
        return c.get_field(), c.samples()


def _same_samples(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert (x.step, x.has_fe, x.has_sse) == (y.step, y.has_fe, y.has_sse)
        for f in ("F_ref", "f_mean", "mean", "rmse"):
            u, v = getattr(x, f), getattr(y, f)
            assert (np.isnan(u) and np.isnan(v)) or u == v, (f, x.step, u, v)


@pytest.mark.parametrize("shape", [(64, 32), (16, 16, 32)])
@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_batched_equals_alone(shape, prec):
    _flush("batched vs alone %s prec=%d" % (shape, prec))
    ps = _mixed_params(shape, prec)
    psis = _fields(shape, len(ps))
    nsteps = 23
    with BatchContext(ps) as bc:
        assert bc.ninst == len(ps)
        v = bc.autotune()
        for s, x in enumerate(psis):
            bc.set_field(s, x)
        bc.step(nsteps)
        got = [(bc.get_field(s), bc.samples(s)) for s in range(len(ps))]
        for s in range(len(ps)):
            assert bc.current_step(s) == nsteps
    for s, p in enumerate(ps):
        f, smp = _alone(p, psis[s], nsteps, v)
        assert np.all(np.isfinite(f))
        assert np.array_equal(got[s][0], f), "instance %d field differs" % s
        _same_samples(got[s][1], smp)
    # the instances really differ (each reads its own physics)
    assert not np.array_equal(got[0][0], got[2][0])


@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_active_subset_and_slot_reuse(prec):
    shape = (32, 64)
    _flush("active subset prec=%d" % prec)
    ps = _mixed_params(shape, prec)
    psis = _fields(shape, len(ps), seed=5)
    with BatchContext(ps) as bc:
        v = bc.autotune()
        for s, x in enumerate(psis):
            bc.set_field(s, x)
        bc.set_active([3, 1])
        bc.step(12)
        idle0 = bc.get_field(0)
        assert bc.current_step(0) == 0 and bc.current_step(1) == 12
        assert np.array_equal(idle0, bc.get_field(0))
        bc.set_active(None)
        bc.step(8)  # instances now at different step counts: 8 and 20
        f1 = bc.get_field(1)
        f0 = bc.get_field(0)
        # slot 2 handed to a new job with new physics and field
        _, pnew = make(shape, prec=prec, sigma=0.09, psi0=0.03, t_sample=4)
        xnew = _fields(shape, 1, seed=99)[0]
        bc.set_params(2, pnew)
        bc.reset(2)
        bc.set_field(2, xnew)
        bc.step(9)
        f2 = bc.get_field(2)
        s2 = bc.samples(2)
        assert bc.current_step(2) == 9 and bc.current_step(1) == 29
        f1b = bc.get_field(1)
        s1 = bc.samples(1)
    a0, _ = _alone(ps[0], psis[0], 8, v)
    assert np.array_equal(f0, a0)
    a1, _ = _alone(ps[1], psis[1], 20, v)
    assert np.array_equal(f1, a1)
    a1b, smp1 = _alone(ps[1], psis[1], 29, v)
    assert np.array_equal(f1b, a1b)
    _same_samples(s1, smp1)
    a2, smp2 = _alone(pnew, xnew, 9, v)
    assert np.array_equal(f2, a2)
    _same_samples(s2, smp2)



@pytest.mark.parametrize("yv", [16, 17])
def test_large_batch_streaming_passes_equal_alone(yv):
    """A batch whose spectra exceed the 256 MiB last-level cache runs its
    axis-1 passes with streaming cache hints; 32 instances of 128^3 FP32 are
    that large, a single one is not.  Hints change no arithmetic: every
    instance must match its run alone bitwise (yv 17: the axis-1 workgroups
    dealt to the XCDs in pairs)."""
    from xpfc_cuda_amd import instance_bytes
    shape, K = (128, 128, 128), 32
    _flush("large batch %s x %d yv=%d" % (shape, K, yv))
    _, p = growth_case(shape, prec=XPFC_FP32_MIXED, t_sample=0)
    assert K * instance_bytes(p) // 3 >= 256 << 20 > instance_bytes(p) // 3
    psis = _fields(shape, 3, seed=21)
    v = (8, yv, 256)
    with BatchContext([p] * K) as bc:
        bc.set_variants(*v)
        for s in range(K):
            bc.set_field(s, psis[s % 3])
        bc.step(6)
        got = [bc.get_field(s) for s in (0, 1, 2, K - 1)]
    for s, f in zip((0, 1, 2, K - 1), got):
        a, _ = _alone(p, psis[s % 3], 6, v)
        assert np.array_equal(f, a), "instance %d differs" % s


def test_batch_larger_than_the_device_fails_cleanly():
    """A batch whose buffers cannot be allocated (512^3 FP64 x 1000 instances,
    about 1 TB per buffer) fails with XPFC_ERR_NOMEM, and the device stays
    usable for the next context."""
    from xpfc_cuda_amd import XPFCError, api
    _flush("oversized batch")
    _, p = make((512, 512, 512), prec=XPFC_FP64)
    with pytest.raises(XPFCError) as e:
        BatchContext([p] * 1000)
    assert e.value.status == api.XPFC_ERR_NOMEM
    _, q = growth_case((32, 32), t_sample=0)
    with Context(q) as c:
        c.initialize_noise(0.01, 1)
        c.step(5)
        assert np.all(np.isfinite(c.get_field()))


@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_divergence_of_one_instance_does_not_stop_the_others(prec):
    """Instance 0 (64^2, sigma 0.062, n0 0.0564, seed 4: the explicit nonlinear
    term blows up within ~30 steps) beside a healthy liquid, both sampled every
    step for 1100 steps, so the held sums are read back twice in the middle of
    the run.  The healthy instance must take all 1100 steps and end exactly
    as alone; the divergence is reported once by sync, and its records keep
    the non-finite values."""
    from xpfc_cuda_amd import XPFCError
    from xpfc_cuda_amd.api import XPFC_ERR_NONFINITE
    shape, nsteps = (64, 64), 1100
    _flush("divergence in a batch prec=%d" % prec)
    _, bad = make(shape, prec=prec, sigma=0.062, psi0=0.0564, t_sample=1)
    _, good = make(shape, prec=prec, sigma=0.1, psi0=0.05, t_sample=1)
    with BatchContext([bad, good]) as bc:
        v = bc.autotune()
        bc.initialize_noise(0, 0.01, 4)
        bc.initialize_noise(1, 0.01, 5)
        bc.step(nsteps)
        with pytest.raises(XPFCError) as ei:
            bc.sync()
        assert ei.value.status == XPFC_ERR_NONFINITE
        bc.sync()  # reported once
        assert bc.current_step(0) == nsteps and bc.current_step(1) == nsteps
        f_good = bc.get_field(1)
        s_bad, s_good = bc.samples(0), bc.samples(1)
    assert not all(np.isfinite(r.F_ref) for r in s_bad if r.has_fe)
    assert all(np.isfinite(r.F_ref) for r in s_good if r.has_fe)
    assert len(s_good) == len(s_bad)
    with Context(good) as c:
        c.set_variants(*v)
        c.initialize_noise(0.01, 5)
        c.step(nsteps)
        c.sync()
        alone, s_alone = c.get_field(), c.samples()
    assert np.array_equal(f_good, alone)
    _same_samples(s_good, s_alone)
