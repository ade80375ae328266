"""The two forms of the state between steps (psi_k, or psi_k transformed back
along axis 0, which saves the axis-0 pass one spectrum of writes): the library
picks the second for 3-D spectra of 16 MiB and more; XPFC_STATE_FORM forces
either.  Both forms against the FP64 oracle (fields, FE / RMSE samples) on
small grids of every line family, against each other, and through the paths
that convert at the boundary (field I/O, FE on demand, checkpoints, batches,
several-grid contexts, graph replays).

Tolerances as test_gpu_parity.py: max|dn| <= 1e-10 (FP64) / 1e-5
(FP32-mixed) after the steps, FE samples 1e-11 / 1e-7 relative.
"""
import numpy as np
import pytest

from oracle import xpfc_oracle as O
from xpfc_cuda_amd import BatchContext, Context, MixedContext, XPFC_FP32_MIXED, XPFC_FP64

from cases import growth_case, make

pytestmark = pytest.mark.gpu

PRECS = [XPFC_FP64, XPFC_FP32_MIXED]


@pytest.fixture
def form(request, monkeypatch):
    monkeypatch.setenv("XPFC_STATE_FORM", str(request.param))
    return request.param


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("form", [0, 1], indirect=True)
@pytest.mark.parametrize("shape,nsteps", [((64, 64), 40), ((96, 80), 30), ((32, 32, 32), 30), ((24, 40, 48), 25),
                                          ((20, 12, 64), 25)])
@pytest.mark.parametrize("prec", PRECS)
def test_steps_vs_oracle(form, shape, nsteps, prec):
    _flush("state form %d %s prec=%d" % (form, shape, prec))
    op, lp = growth_case(shape, prec=prec, t_sample=5)
    psi0 = O.noise_ic(op, 0.01, 3)
    ref, rs = O.run(psi0, op, nsteps, sample=True)
    with Context(lp) as ctx:
        ctx.set_field(psi0)
        ctx.step(nsteps)
        got = ctx.get_field()
        gs = ctx.samples()
        F_now, _ = ctx.fe_calc()
    assert np.max(np.abs(ref - psi0)) > 1e-4  # the field evolved
    err = float(np.max(np.abs(got - ref)))
    assert err <= (1e-10 if prec == XPFC_FP64 else 1e-5), err
    ftol = 1e-11 if prec == XPFC_FP64 else 1e-7
    rfe = {s.step: s.value for s in rs if s.kind == "FE"}
    gfe = {s.step: s.F_ref for s in gs if s.has_fe}
    assert set(rfe) == set(gfe) and rfe
    for st, r in rfe.items():
        assert abs(gfe[st] - r) <= ftol * abs(r), (st, gfe[st], r)
    rsse = {s.step: s.value for s in rs if s.kind == "SSE"}
    gsse = {s.step: s.rmse for s in gs if s.has_sse}
    assert sorted(rsse) == sorted(gsse)
    for st, r in rsse.items():
        assert abs(gsse[st] - r) <= (1e-9 if prec == XPFC_FP64 else 1e-4) * abs(r) + 1e-300, (st, gsse[st], r)
    # FE on demand (psi_k formed at the boundary) = the oracle's at the last step
    assert abs(F_now - rfe[nsteps]) <= ftol * abs(rfe[nsteps])


@pytest.mark.parametrize("shape", [(64, 64), (32, 48, 40)])
@pytest.mark.parametrize("prec", PRECS)
def test_forms_agree(shape, prec, monkeypatch):
    """The two forms differ only by the rounding of one more transform pair of
    the state per step."""
    _flush("forms agree %s prec=%d" % (shape, prec))
    op, lp = growth_case(shape, prec=prec, t_sample=10)
    psi0 = O.noise_ic(op, 0.01, 9)
    out = []
    for f in (0, 1):
        monkeypatch.setenv("XPFC_STATE_FORM", str(f))
        with Context(lp) as ctx:
            ctx.set_field(psi0)
            ctx.step(50)
            out.append((ctx.get_field(), [s.F_ref for s in ctx.samples() if s.has_fe]))
    amp = float(np.max(np.abs(out[0][0] - 0.05)))
    d = float(np.max(np.abs(out[0][0] - out[1][0])))
    _flush("forms agree %s prec=%d amp=%.3g max|dn|=%.3e" % (shape, prec, amp, d))
    assert d <= (1e-12 if prec == XPFC_FP64 else 1e-5) * amp
    for a, b in zip(out[0][1], out[1][1]):
        assert abs(a - b) <= (1e-12 if prec == XPFC_FP64 else 1e-7) * abs(a)


@pytest.mark.parametrize("form", [1], indirect=True)
@pytest.mark.parametrize("prec", PRECS)
def test_axis0_form_batch_mixed_graphs_equal_alone(form, prec):
    """Batches, several-grid contexts and graph replays in the axis-0 form give
    every instance its one-instance field bit for bit."""
    _flush("axis-0 form batch / mixed prec=%d" % prec)
    shapes = [(32, 32, 32), (24, 40, 48), (64, 64), (80, 96)]
    ps = []
    for i, sh in enumerate(shapes):
        _, p = make(sh, prec=prec, sigma=0.07 + 0.01 * i, psi0=0.05, t_sample=6)
        ps.append(p)
    fields = [0.05 + 0.02 * np.random.default_rng(40 + i).standard_normal(p.n) for i, p in enumerate(ps)]
    alone = []
    for p, x in zip(ps, fields):
        with Context(p) as c:
            c.set_graphs(False)
            c.set_field(x)
            c.step(23)
            alone.append(c.get_field())
    with MixedContext(ps) as mc:
        for s, x in enumerate(fields):
            mc.set_field(s, x)
        mc.step(23)
        for s in range(len(ps)):
            assert np.array_equal(mc.get_field(s), alone[s]), s
    _, q = make(shapes[0], prec=prec, sigma=0.09, psi0=0.04, t_sample=6)
    with BatchContext([ps[0], q, ps[0]]) as b:
        for s, x in enumerate([fields[0], 0.04 + 0.01 * np.random.default_rng(7).standard_normal(q.n), fields[0]]):
            b.set_field(s, x)
        b.step(23)
        assert np.array_equal(b.get_field(0), alone[0])
        assert np.array_equal(b.get_field(2), alone[0])


@pytest.mark.parametrize("form", [1], indirect=True)
def test_axis0_form_checkpoints(form, tmp_path, monkeypatch):
    """A checkpoint reloads bit for bit into a context of its own form, and
    into a context of the other form through one axis-0 transform of the
    state (the two runs then differ by that rounding, as the forms do)."""
    _flush("axis-0 form checkpoints")
    _, p = growth_case((32, 32, 32), prec=XPFC_FP32_MIXED, t_sample=10)
    ck = {}
    straight = {}
    for f in (1, 0):
        monkeypatch.setenv("XPFC_STATE_FORM", str(f))
        ck[f] = str(tmp_path / ("s%d.ckpt" % f))
        with Context(p) as c:
            c.initialize_noise(0.01, 5)
            c.step(13)
            c.save_checkpoint(ck[f])
            c.step(20)
            straight[f] = c.get_field()
        with Context(p) as c:
            c.load_checkpoint(ck[f])
            c.step(20)
            assert np.array_equal(c.get_field(), straight[f]), f
    amp = float(np.max(np.abs(straight[1] - 0.05)))
    for f in (1, 0):   # load the other form's checkpoint
        monkeypatch.setenv("XPFC_STATE_FORM", str(f))
        with Context(p) as c:
            c.load_checkpoint(ck[1 - f])
            assert c.current_step == 13
            c.step(20)
            d = float(np.max(np.abs(c.get_field() - straight[1 - f])))
        _flush("checkpoint of form %d loaded in form %d: max|dn| %.3e (amplitude %.3g)" % (1 - f, f, d, amp))
        assert d <= 1e-5 * amp, (f, d, amp)


def test_default_form_choice(monkeypatch):
    """Without the override, at sizes on both sides of the choice (128^3 FP64:
    a 17 MB spectrum, the axis-0 form; 128^3 FP32 and 512^2: psi_k): a field
    round trip and two steps against the oracle."""
    monkeypatch.delenv("XPFC_STATE_FORM", raising=False)
    for shape, prec in (((128, 128, 128), XPFC_FP64), ((128, 128, 128), XPFC_FP32_MIXED), ((512, 512), XPFC_FP64)):
        _flush("default form %s prec=%d" % (shape, prec))
        op, lp = growth_case(shape, prec=prec, t_sample=0)
        psi0 = O.noise_ic(op, 0.01, 1)
        with Context(lp) as c:
            c.set_field(psi0)
            back = c.get_field()
            c.step(2)
            f = c.get_field()
        assert np.max(np.abs(back - psi0)) <= (1e-15 if prec == XPFC_FP64 else 1e-7)
        ref, _ = O.run(psi0, op, 2, sample=False)
        assert np.max(np.abs(f - ref)) <= (1e-12 if prec == XPFC_FP64 else 1e-6)


@pytest.mark.parametrize("form", [1], indirect=True)
@pytest.mark.parametrize("prec", PRECS)
def test_axis0_form_budgeted_blocks_equal_alone(form, prec):
    """The service's budgeted blocks (each instance its own number of steps in
    one captured block) in the axis-0 form: every field equals its run alone."""
    import ctypes as ct
    from xpfc_cuda_amd import lib
    _flush("axis-0 form budgeted blocks prec=%d" % prec)
    ps = []
    for i, sh in enumerate([(32, 32, 32), (24, 40, 48), (32, 32, 32)]):
        _, p = make(sh, prec=prec, sigma=0.08 + 0.01 * i, psi0=0.05, t_sample=0)
        ps.append(p)
    fields = [0.05 + 0.02 * np.random.default_rng(60 + i).standard_normal(p.n) for i, p in enumerate(ps)]
    budgets, lens = [[50, 7, 0], [13, 32, 32]], [50, 32]
    with MixedContext(ps) as mc:
        for s, x in enumerate(fields):
            mc.set_field(s, x)
        for b, n in zip(budgets, lens):
            assert lib().xpfc_step_budget(mc.handle, n, (ct.c_int * len(ps))(*b)) == 0
        mc.sync()
        got = [mc.get_field(s) for s in range(len(ps))]
    for s, p in enumerate(ps):
        with Context(p) as c:
            c.set_field(fields[s])
            c.step(budgets[0][s] + budgets[1][s])
            assert np.array_equal(got[s], c.get_field()), s
