"""The allocation-free noise initialisation (xpfc_inst_initialize_noise_noalloc)
against the allocating entry point it mirrors (xpfc_inst_initialize_noise):
identical fields, identical after steps, on two calls in a row whose second
needs fewer partial sums than the first (the fixed device array keeps the
first call's values); and its refusal while records of the slot are pending."""
import ctypes

import numpy as np
import pytest

from xpfc_cuda_amd import BatchContext, XPFCError, XPFC_FP32_MIXED, XPFC_FP64, lib
from xpfc_cuda_amd.api import XPFC_ERR_STATE

from cases import growth_case

pytestmark = pytest.mark.gpu


def _noalloc(bc, s, amp, seed):
    return lib().xpfc_inst_initialize_noise_noalloc(bc.handle, s, ctypes.c_double(amp),
                                                    ctypes.c_ulonglong(seed))


@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_noalloc_noise_equals_allocating_entry_point(prec):
    # 128^3: 4096 partial sums (the whole array); then 32^2: 4 of them
    for shape, seed in (((128, 128, 128), 3), ((32, 32), 8)):
        print("noalloc noise %s prec=%d" % (shape, prec), flush=True)
        _, p = growth_case(shape, prec=prec, t_sample=0)
        with BatchContext([p, p]) as bc:
            assert _noalloc(bc, 0, 0.01, seed) == 0
            bc.initialize_noise(1, 0.01, seed)
            a, b = bc.get_field(0), bc.get_field(1)
            assert bc.avg_calc(0) == bc.avg_calc(1)
            bc.step(7)
            a7, b7 = bc.get_field(0), bc.get_field(1)
        assert np.array_equal(a, b), shape
        assert np.array_equal(a7, b7), shape
        assert not np.array_equal(a, a7)


def test_noalloc_noise_refuses_while_records_are_pending():
    print("noalloc pending", flush=True)
    _, p = growth_case((32, 32), prec=XPFC_FP64, t_sample=1)
    with BatchContext([p, p]) as bc:
        bc.initialize_noise(0, 0.01, 1)
        bc.initialize_noise(1, 0.01, 2)
        bc.step(3)                       # sampled steps queued, not read back
        assert _noalloc(bc, 0, 0.01, 5) == XPFC_ERR_STATE
        bc.sync()
        assert _noalloc(bc, 0, 0.01, 5) == 0
        bc.reset(1)
        bc.initialize_noise(1, 0.01, 5)
        assert np.array_equal(bc.get_field(0), bc.get_field(1))
