"""tpf_dccd_mass_correct_noalloc (fixed device scratch, factor formed on the
device) against tpf_dccd_mass_correct (per-call allocation, factor formed on
the host), through the C-ABI: identical outputs and mass_before on two calls
in a row, the second on a smaller grid, then a third with mass_before NULL
(asynchronous).  Both sides run the same reduction kernels, so the sums are
bit-identical; the factor is one IEEE division on either side; the bound
written here is 1e-15 absolute on values in [0, 1] and 1e-14 relative on
the mass."""

import ctypes as C
import sys

import numpy as np
import pytest
import torch

from twophaseflow_amd import _native as nat
from twophaseflow_amd.device import ptr, stream_handle

pytestmark = pytest.mark.gpu

CASES = [
    # (shape, L, periodic): the first has 97^3 nodes, which uses all 2048
    # reduction workgroups; the second is smaller
    ((97, 97, 97), (1.0, 1.2, 0.8), (0, 1, 0)),
    ((65, 40), (1.0, 0.6), (1, 0)),
    ((33,), (2.0,), (1,)),
]


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    q = rng.random(shape)
    q[q < 0.2] = 0.0
    q[q > 0.85] = 1.0
    return torch.from_numpy(q).cuda()


def _call(fn, shape, L, per, q, m_target, want_mass):
    out = torch.empty_like(q)
    mb = C.c_double(-1.0)
    nd = len(shape)
    st = fn(nd, nat.i64arr(shape), nat.darr(L), (C.c_int * nd)(*per), ptr(q), m_target,
            ptr(out), C.byref(mb) if want_mass else None, stream_handle())
    nat.check(st)
    torch.cuda.synchronize()
    return out, mb.value


@pytest.mark.parametrize("want_mass", [True, False])
def test_noalloc_mass_correction_matches_allocating_entry_point(want_mass):
    lib = nat.lib()
    for i, (shape, L, per) in enumerate(CASES):
        print(f"[gpu-case] mass_correct_noalloc shape={shape} mass_before={want_mass}",
              file=sys.stderr, flush=True)
        q = _field(shape, 40 + i)
        target = float(q.sum()) * 1e-3 + 0.37
        ref, m_ref = _call(lib.tpf_dccd_mass_correct, shape, L, per, q, target, True)
        got, m_got = _call(lib.tpf_dccd_mass_correct_noalloc, shape, L, per, q, target, want_mass)
        assert float((got - ref).abs().max()) <= 1e-15
        if want_mass:
            assert abs(m_got - m_ref) <= 1e-14 * abs(m_ref)


def test_noalloc_mass_correction_rejects_bad_arguments():
    lib = nat.lib()
    q = _field((9, 9), 1)
    out = torch.empty_like(q)
    with pytest.raises(ValueError):
        nat.check(lib.tpf_dccd_mass_correct_noalloc(2, nat.i64arr((9, 9)), nat.darr((1.0, 1.0)),
                                                    (C.c_int * 2)(0, 0), ptr(q), 0.5, ptr(q),
                                                    None, stream_handle()))
    with pytest.raises(ValueError):
        nat.check(lib.tpf_dccd_mass_correct_noalloc(2, nat.i64arr((9, 2)), nat.darr((1.0, 1.0)),
                                                    (C.c_int * 2)(0, 0), ptr(q), 0.5, ptr(out),
                                                    None, stream_handle()))
