"""The no-alloc map projection (ds_problem_smp_grid_noalloc, through the C
ABI) against ds_problem_smp_grid: bit-identical maps on two calls in a
row, the second with a smaller problem and grid; a grid that does not fit
the static scratch is rejected with DS_E_ARG and leaves the output
untouched, and the next call still works."""
import ctypes as C

import numpy as np
import pytest

from dachshund_amd._lib import DS_E_ARG, DS_OK

pytestmark = pytest.mark.gpu


def _problem(D, w):
    from dachshund_amd import synth
    pix = synth.make_pixels(w)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
    wf.set_x(np.random.default_rng(len(pix)).standard_normal(len(pix)))
    return wf


def _noalloc(D, wf, box, shape):
    from dachshund_amd._lib import lib
    m = np.full(int(np.prod(shape)), -7.0)
    rc = lib().ds_problem_smp_grid_noalloc(
        wf.h, *map(float, box), *map(int, shape),
        m.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, m


def test_noalloc_map_equals_allocating_map():
    import dachshund_amd as D
    from dachshund_amd import synth
    big = synth.workload("c2")                 # 76,800 px, 60 x 48 x 880
    small = synth.workload("small")            # 2,000 px, 8 x 8 x 32
    wf1 = _problem(D, big)
    print(f"[noalloc] call 1: n={wf1.n}, grid {big.map_shape}", flush=True)
    ref1 = wf1.smp_grid(*big.box, *big.map_shape)
    rc, m1 = _noalloc(D, wf1, big.box, big.map_shape)
    assert rc == DS_OK
    assert np.array_equal(m1, ref1)
    wf2 = _problem(D, small)
    print(f"[noalloc] call 2: n={wf2.n}, grid {small.map_shape}", flush=True)
    rc, m2 = _noalloc(D, wf2, small.box, small.map_shape)
    assert rc == DS_OK
    assert np.array_equal(m2, wf2.smp_grid(*small.box, *small.map_shape))
    # a grid whose map alone exceeds the scratch: rejected, m untouched
    rc, m3 = _noalloc(D, wf2, small.box, (200, 200, 400))
    assert rc == DS_E_ARG
    assert np.all(m3 == -7.0)
    # and the array still serves the next call
    rc, m4 = _noalloc(D, wf1, big.box, (30, 24, 440))
    assert rc == DS_OK
    assert np.array_equal(m4, wf1.smp_grid(*big.box, 30, 24, 440))
    wf1.close()
    wf2.close()
