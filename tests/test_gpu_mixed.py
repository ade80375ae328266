"""Contexts of several grids (one launch per pass over every grid): each
instance evolves exactly as it would alone, bit for bit, whatever the other
grids, physics, sampling periods, active lists and step counts are.

Sampled FE / RMSE records sum per-workgroup partials whose grouping follows
the launch shape, so they are compared to rounding; fe_calc on demand uses a
fixed shape and is compared bitwise.
"""
import numpy as np
import pytest

from xpfc_cuda_amd import Context, MixedContext, XPFC_FP32_MIXED, XPFC_FP64

from cases import growth_case, make

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _mix(prec):
    """Seven instances on four grids (2-D and 3-D, one non-square box), with
    different physics and sampling periods; the 64^2 grid holds three."""
    ps = []
    _, p = growth_case((64, 64), prec=prec, t_sample=7)
    ps.append(p)
    _, p = make((16, 16, 32), prec=prec, sigma=0.1, psi0=0.02, t_sample=5, dt=0.3)
    ps.append(p)
    _, p = make((64, 64), prec=prec, sigma=0.2, psi0=0.05, eps=0.01, mode="sum", t_sample=0)
    ps.append(p)
    _, p = make((32, 64), dx=0.1, prec=prec, sigma=0.08, psi0=0.04, Mo=1.5, int_Mo=0, t_sample=10)
    ps.append(p)
    _, p = make((32, 32, 32), prec=prec, sigma=0.09, psi0=0.03, t_sample=4, powf_fp32=1)
    ps.append(p)
    _, p = make((64, 64), prec=prec, sigma=0.07, psi0=0.05, t_sample=9)
    ps.append(p)
    _, p = make((16, 16, 32), prec=prec, sigma=0.12, psi0=0.06, t_sample=0)
    ps.append(p)
    return ps


def _field(shape, seed):
    return 0.05 + 0.02 * np.random.default_rng(seed).standard_normal(shape)


def _alone(p, psi, nsteps):
    with Context(p) as c:
        c.set_field(psi)
        c.step(nsteps)
        return c.get_field(), c.samples(), c.fe_calc()


def _close_samples(a, b, rel):
    assert [(x.step, x.has_fe, x.has_sse) for x in a] == [(y.step, y.has_fe, y.has_sse) for y in b]
    for x, y in zip(a, b):
        for f in ("F_ref", "f_mean", "mean", "rmse"):
            u, v = getattr(x, f), getattr(y, f)
            assert (np.isnan(u) and np.isnan(v)) or abs(u - v) <= rel * abs(v) + 1e-300, (f, x.step, u, v)


@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
@pytest.mark.parametrize("graphs", [True, False])
def test_mixed_equals_alone(prec, graphs):
    _flush("mixed vs alone prec=%d graphs=%d" % (prec, graphs))
    ps = _mix(prec)
    psis = [_field(tuple(p.n), 30 + s) for s, p in enumerate(ps)]
    nsteps = 61
    with MixedContext(ps) as mc:
        mc.set_graphs(graphs)
        for s, x in enumerate(psis):
            mc.set_field(s, x)
        mc.step(nsteps)
        mc.sync()
        got = [(mc.get_field(s), mc.samples(s), mc.fe_calc(s), mc.current_step(s)) for s in range(len(ps))]
    for s, p in enumerate(ps):
        f, smp, fe = _alone(p, psis[s], nsteps)
        assert np.all(np.isfinite(f))
        assert got[s][3] == nsteps
        assert np.array_equal(got[s][0], f), "instance %d (%s) differs" % (s, p.n)
        assert got[s][2] == fe
        _close_samples(got[s][1], smp, 1e-12 if prec == XPFC_FP64 else 1e-9)
    assert not np.array_equal(got[0][0], got[2][0])  # same grid, own physics


def _families(prec):
    """Grids whose passes fall in every line family (2^k, 3*2^k, 5*2^k
    points), at each family's shortest and longest ragged lines, with square /
    cubic boxes and others, mixed with power-of-two grids in one context."""
    ps = []
    for shape, kw in (((96, 96), dict(sigma=0.08, t_sample=7)),            # 96-point lines, 48-point half lines
                      ((80, 160), dict(sigma=0.1, t_sample=5, dt=0.3)),      # 5*2^k, non-square box
                      ((48, 40, 24), dict(sigma=0.09, t_sample=4)),         # axis 0 3*2^k, axis 1 5*2^k, H = 12
                      ((20, 24, 40), dict(sigma=0.12, t_sample=0)),         # axis 0 5*2^k, axis 1 3*2^k, H = 20
                      ((768, 12), dict(sigma=0.07, t_sample=10)),           # longest 3*2^k line, H = 6
                      ((640, 20), dict(sigma=0.11, t_sample=6)),            # longest 5*2^k line, H = 10
                      ((12, 768), dict(sigma=0.08, t_sample=9)),            # shortest 3*2^k line, H = 384
                      ((20, 640), dict(sigma=0.1, t_sample=3)),             # shortest 5*2^k line, H = 320
                      ((96, 96, 96), dict(sigma=0.09, t_sample=8)),         # cubic, every axis 3*2^k
                      ((64, 64), dict(sigma=0.07, t_sample=7)),
                      ((16, 16, 32), dict(sigma=0.1, t_sample=5))):
        _, p = make(shape, prec=prec, psi0=0.05, **kw)
        ps.append(p)
    return ps


@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
@pytest.mark.parametrize("graphs", [True, False])
def test_mixed_line_families_equal_alone(prec, graphs):
    _flush("mixed line families vs alone prec=%d graphs=%d" % (prec, graphs))
    ps = _families(prec)
    psis = [_field(tuple(p.n), 70 + s) for s, p in enumerate(ps)]
    nsteps = 23
    with MixedContext(ps) as mc:
        mc.set_graphs(graphs)
        for s, x in enumerate(psis):
            mc.set_field(s, x)
        mc.step(nsteps)
        mc.set_active([2, 5, 8])          # three families, some grids idle
        mc.step(4)
        mc.set_active(None)
        mc.step(3)
        mc.sync()
        got = [(mc.get_field(s), mc.current_step(s)) for s in range(len(ps))]
    for s, p in enumerate(ps):
        n = nsteps + 3 + (4 if s in (2, 5, 8) else 0)
        f, _, _ = _alone(p, psis[s], n)
        assert np.all(np.isfinite(f))
        assert got[s][1] == n
        assert np.array_equal(got[s][0], f), "instance %d (%s) differs" % (s, p.n)


@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_mixed_active_subsets_and_slot_reuse(prec):
    _flush("mixed active subsets prec=%d" % prec)
    ps = _mix(prec)
    psis = [_field(tuple(p.n), 50 + s) for s, p in enumerate(ps)]
    with MixedContext(ps) as mc:
        for s, x in enumerate(psis):
            mc.set_field(s, x)
        mc.set_active([4, 0, 3])          # three grids, instances listed out of order
        mc.step(12)
        mc.set_active([1])                # one 3-D grid alone
        mc.step(5)
        mc.set_active(None)
        mc.step(8)
        steps = [mc.current_step(s) for s in range(len(ps))]
        f = [mc.get_field(s) for s in range(len(ps))]
        # slot 5 handed to a new job: new physics and field, step counter reset
        _, pnew = make((64, 64), prec=prec, sigma=0.09, psi0=0.03, t_sample=3)
        xnew = _field((64, 64), 99)
        mc.set_params(5, pnew)
        mc.reset(5)
        mc.set_field(5, xnew)
        mc.step(6)
        f5 = mc.get_field(5)
        f0b = mc.get_field(0)
    assert steps == [20, 13, 8, 20, 20, 8, 8]
    for s, p in enumerate(ps):
        a, _, _ = _alone(p, psis[s], steps[s])
        assert np.array_equal(f[s], a), "instance %d differs" % s
    a5, _, _ = _alone(pnew, xnew, 6)
    assert np.array_equal(f5, a5)
    a0, _, _ = _alone(ps[0], psis[0], 26)
    assert np.array_equal(f0b, a0)


def test_mixed_rejects_operator_calls_and_mixed_precision():
    from xpfc_cuda_amd import XPFCError
    _, a = make((32, 32), prec=XPFC_FP64)
    _, b = make((16, 16, 16), prec=XPFC_FP32_MIXED)
    with pytest.raises(XPFCError):
        MixedContext([a, b])
    _, b = make((16, 16, 16), prec=XPFC_FP64)
    with MixedContext([a, b]) as mc:
        with pytest.raises(XPFCError):
            mc.set_variants(8, 8, 256)
        assert mc.autotune() == (0, 0, 0)
        mc.initialize_noise(0, 0.01, 1)
        with pytest.raises(XPFCError):
            mc.step(1)                    # instance 1 has no field yet
    # the ragged kernels cover extents of 2^k, 3*2^k and 5*2^k points up to
    # 1024 (half lines up to 512): a longer line is refused when the context
    # is made, not at its first step
    for other in ((2048, 64), (1536, 48), (1280, 20), (64, 2048), (16, 8, 1280)):
        _, c = make(other, prec=XPFC_FP64)
        with pytest.raises(XPFCError) as e:
            MixedContext([a, c])
        assert "one-launch form" in str(e.value)


@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_budgeted_blocks_equal_alone(prec):
    """xpfc_step_budget: one block in which each instance takes its own number
    of steps (0 .. the block length), twice in a row with other budgets; every
    field equals its run alone over the same total, and sampled steps inside a
    block are refused."""
    import ctypes as ct
    from xpfc_cuda_amd import XPFCError, lib
    _flush("budgeted blocks prec=%d" % prec)
    ps = []
    for p in _mix(prec):
        p.t_sample = 0
        ps.append(p)
    psis = [_field(tuple(p.n), 70 + s) for s, p in enumerate(ps)]
    budgets = [[50, 0, 17, 32, 1, 50, 9], [20, 32, 32, 3, 32, 0, 31]]
    lens = [50, 32]
    with MixedContext(ps) as mc:
        for s, x in enumerate(psis):
            mc.set_field(s, x)
        for b, n in zip(budgets, lens):
            arr = (ct.c_int * len(ps))(*b)
            assert lib().xpfc_step_budget(mc.handle, n, arr) == 0
        mc.sync()
        totals = [a + b for a, b in zip(*budgets)]
        assert [mc.current_step(s) for s in range(len(ps))] == totals
        got = [mc.get_field(s) for s in range(len(ps))]
        _, q = make((64, 64), prec=prec, t_sample=10)
        mc.set_params(0, q)
        arr = (ct.c_int * len(ps))(*([20] + [0] * (len(ps) - 1)))
        assert lib().xpfc_step_budget(mc.handle, 20, arr) != 0   # steps 70 and 79 sample
    for s, p in enumerate(ps):
        a, _, _ = _alone(p, psis[s], totals[s])
        assert np.array_equal(got[s], a), "instance %d differs" % s


@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
@pytest.mark.parametrize("graphs", [True, False])
def test_budgeted_blocks_one_grid(prec, graphs):
    """xpfc_step_budget on contexts of one grid (a batch, 3-D and 2-D): each
    active instance takes its own number of the block's steps; two blocks
    with other budgets and another active list; every field equals its run
    alone."""
    import ctypes as ct
    from xpfc_cuda_amd import BatchContext, lib
    _flush("budgeted one-grid blocks prec=%d graphs=%d" % (prec, graphs))
    for shape in [(16, 16, 32), (64, 64)]:
        ps = []
        for s in range(5):
            _, p = make(shape, prec=prec, sigma=0.06 + 0.02 * s, psi0=0.03 + 0.01 * s, t_sample=0)
            ps.append(p)
        psis = [_field(shape, 90 + s) for s in range(5)]
        budgets = [[16, 0, 9, 16, 1], [5, 32, 0, 31, 32]]
        lens = [16, 32]
        with BatchContext(ps) as bc:
            bc.set_graphs(graphs)
            for s, x in enumerate(psis):
                bc.set_field(s, x)
            arr = (ct.c_int * 5)(*budgets[0])
            assert lib().xpfc_step_budget(bc.handle, lens[0], arr) == 0
            bc.set_active([4, 1, 3, 0])   # instance 2 sits out the second block
            arr = (ct.c_int * 5)(*budgets[1])
            assert lib().xpfc_step_budget(bc.handle, lens[1], arr) == 0
            bc.sync()
            totals = [budgets[0][s] + (budgets[1][s] if s != 2 else 0) for s in range(5)]
            assert [bc.current_step(s) for s in range(5)] == totals
            got = [bc.get_field(s) for s in range(5)]
        for s, p in enumerate(ps):
            a, _, _ = _alone(p, psis[s], totals[s])
            assert np.array_equal(got[s], a), "%s instance %d differs" % (shape, s)



@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
@pytest.mark.parametrize("graphs", [True, False])
def test_one_and_two_stream_groups_agree(prec, graphs, monkeypatch):
    """A several-grid context steps its instances in two stream groups (each
    grid's instances split between the context stream and a second stream,
    forked and joined every block); XPFC_MIXED_STREAMS=1 keeps one.  Both give
    every instance the same field bit for bit and the same samples, through
    sampled steps and plain blocks, and through budgeted blocks."""
    import ctypes as ct
    from xpfc_cuda_amd import lib
    _flush("one vs two stream groups prec=%d graphs=%d" % (prec, graphs))
    budgets = [[30, 4, 0, 30, 17, 9, 30], [12, 12, 5, 0, 12, 1, 7]]
    out = []
    for groups in ("1", "2"):
        monkeypatch.setenv("XPFC_MIXED_STREAMS", groups)
        ps = _mix(prec)
        psis = [_field(tuple(p.n), 110 + s) for s, p in enumerate(ps)]
        with MixedContext(ps) as mc:
            mc.set_graphs(graphs)
            for s, x in enumerate(psis):
                mc.set_field(s, x)
            mc.step(61)
            mc.sync()
            sampled = ([mc.get_field(s) for s in range(len(ps))], [mc.samples(s) for s in range(len(ps))])
        qs = []
        for p in _mix(prec):
            p.t_sample = 0
            qs.append(p)
        with MixedContext(qs) as mc:
            mc.set_graphs(graphs)
            for s, x in enumerate(psis):
                mc.set_field(s, x)
            for b, n in zip(budgets, (30, 12)):
                assert lib().xpfc_step_budget(mc.handle, n, (ct.c_int * len(qs))(*b)) == 0
            mc.step(9)
            mc.sync()
            steps = [mc.current_step(s) for s in range(len(qs))]
            budgeted = [mc.get_field(s) for s in range(len(qs))]
        out.append((sampled, budgeted, steps))
    assert out[0][2] == out[1][2] == [a + b + 9 for a, b in zip(*budgets)]
    for s in range(len(budgets[0])):
        for k in range(2):
            a = out[0][0][0][s] if k == 0 else out[0][1][s]
            b = out[1][0][0][s] if k == 0 else out[1][1][s]
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, b), "instance %d differs (%s)" % (s, "sampled" if k == 0 else "budgeted")
        _close_samples(out[0][0][1][s], out[1][0][1][s], 1e-12)
