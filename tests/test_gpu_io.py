"""Field I/O on the device path: VTK output byte-identical to the reference
writer's format, VTK restart in both orders, binary checkpoints that resume
bit for bit."""
import numpy as np
import pytest

from oracle import xpfc_oracle as O
from xpfc_cuda_amd import BatchContext, Context, XPFCError, XPFC_FP32_MIXED, XPFC_FP64
from xpfc_cuda_amd.api import XPFC_ERR_ARG, XPFC_ERR_NONFINITE

from cases import growth_case, make

pytestmark = pytest.mark.gpu


def test_write_vtk_matches_reference_format(tmp_path):
    _, p = make((32, 16), dx=0.5)  # Lx = 16, Ly = 8 (integers, as the reference's)
    with Context(p) as c:
        c.initialize_noise(0.01, 3)
        c.step(5)
        psi = c.get_field()
        path = tmp_path / "density_data5.vtk"
        c.write_vtk(str(path))
    assert path.read_text() == O.vtk_text(psi, 16, 8)


@pytest.mark.parametrize("order", [0, 1])
def test_vtk_restart_orders(tmp_path, order):
    _, p = make((32, 32), dx=0.5)
    with Context(p) as c:
        c.initialize_noise(0.02, 8)
        c.step(3)
        psi = c.get_field()
        c.write_vtk(str(tmp_path / "f.vtk"))
    p.restart_order = order
    with Context(p) as c:
        c.initialize_vtk(str(tmp_path / "f.vtk"))
        got = c.get_field()
    six = np.vectorize(lambda x: float("%g" % x))(psi)
    expect = six if order == 0 else six.T  # the reference's fill transposes x and y
    expect = expect + (p.psi0 - expect.mean())  # average-density fix
    assert np.max(np.abs(got - expect)) < 1e-15


def test_vtk_restart_3d(tmp_path):
    _, p = make((16, 8, 32))
    with Context(p) as c:
        c.initialize_noise(0.02, 9)
        psi = c.get_field()
        c.write_vtk(str(tmp_path / "g.vtk"))
        c.initialize_vtk(str(tmp_path / "g.vtk"))
        got = c.get_field()
    six = np.vectorize(lambda x: float("%g" % x))(psi)
    assert np.max(np.abs(got - (six + p.psi0 - six.mean()))) < 1e-15


@pytest.mark.parametrize("shape", [(64, 32), (16, 16, 32)])
@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_checkpoint_resumes_bit_for_bit(tmp_path, shape, prec):
    print("checkpoint %s prec=%d" % (shape, prec), flush=True)
    _, p = growth_case(shape, prec=prec, t_sample=10)
    ck = str(tmp_path / "state.ckpt")
    with Context(p) as c:
        c.initialize_noise(0.01, 5)
        c.step(17)
        c.save_checkpoint(ck)
        c.step(26)
        straight = c.get_field()
        tail = [s for s in c.samples() if s.step > 17]
    with Context(p) as c:
        c.load_checkpoint(ck)
        assert c.current_step == 17
        c.step(26)
        resumed = c.get_field()
        rs = c.samples()
    assert np.array_equal(straight, resumed)
    assert [(s.step, s.has_fe, s.has_sse, s.F_ref, s.rmse) for s in rs] == \
           [(s.step, s.has_fe, s.has_sse, s.F_ref, s.rmse) for s in tail]


def test_checkpoint_into_a_batch_slot(tmp_path):
    _, p = growth_case((32, 64), t_sample=0)
    _, q = make((32, 64), sigma=0.15, psi0=0.03, t_sample=0)
    ck = str(tmp_path / "one.ckpt")
    with Context(p) as c:
        c.initialize_noise(0.01, 1)
        c.step(9)
        c.save_checkpoint(ck)
        c.step(11)
        alone = c.get_field()
    with BatchContext([q, q, q]) as b:
        for s in range(3):
            b.initialize_noise(s, 0.01, 20 + s)
        b.load_checkpoint(1, ck)          # slot 1 takes over the saved job, physics included
        b.step(11)
        assert np.array_equal(b.get_field(1), alone)
        assert b.current_step(1) == 20 and b.current_step(0) == 11


def test_checkpoint_rejects_other_grids(tmp_path):
    _, p = make((32, 32))
    _, q = make((64, 32))
    ck = str(tmp_path / "x.ckpt")
    with Context(p) as c:
        c.initialize_noise(0.01, 1)
        c.save_checkpoint(ck)
    with Context(q) as c:
        with pytest.raises(Exception):
            c.load_checkpoint(ck)
    # same extents, other box length (dx = 0.1 instead of 0.125): the
    # wavenumbers and the C2 table would disagree with the checkpoint's physics
    _, r = make((32, 32), dx=0.1)
    with Context(r) as c:
        with pytest.raises(XPFCError) as ei:
            c.load_checkpoint(ck)
        assert ei.value.status == XPFC_ERR_ARG


def test_checkpoint_rejects_other_precision_and_damaged_files(tmp_path):
    _, p32 = make((32, 32), prec=XPFC_FP32_MIXED)
    _, p64 = make((32, 32), prec=XPFC_FP64)
    ck = tmp_path / "s.ckpt"
    with Context(p32) as c:
        c.initialize_noise(0.01, 1)
        c.save_checkpoint(str(ck))
    with Context(p64) as c:
        with pytest.raises(Exception):
            c.load_checkpoint(str(ck))
    data = ck.read_bytes()
    (tmp_path / "short.ckpt").write_bytes(data[: len(data) // 2])
    (tmp_path / "magic.ckpt").write_bytes(b"NOTACKPT" + data[8:])
    with Context(p32) as c:
        for bad in ("short.ckpt", "magic.ckpt", "missing.ckpt"):
            with pytest.raises(Exception):
                c.load_checkpoint(str(tmp_path / bad))
        c.load_checkpoint(str(ck))  # the intact file still loads
        assert c.current_step == 0
