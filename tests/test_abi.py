"""CPU checks of the C ABI library: it loads, exports every declared symbol,
maps errors to the reference's exception types, and its SPIKE tables
reproduce the dense-LU oracle through the host model of the kernel."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import ccd as occd
from twophaseflow_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("tpf.h", "tpf_testing.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"TPF_API\s+[\w\s\*]+?\b(tpf_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = nat.lib()
    names = declared_symbols()
    assert len(names) >= 17
    for n in names:
        assert hasattr(lib, n), n
    assert names == set(nat.SIGNATURES), names ^ set(nat.SIGNATURES)
    assert lib.tpf_version() >= 1


def _plan(shape, bcs, L=None):
    L = L or [1.0] * len(shape)
    h = [L[a] / (shape[a] - 1) for a in range(len(shape))]
    plan = C.c_void_p()
    nat.check(nat.lib().tpf_ccd_plan_create(C.byref(plan), len(shape), nat.i64arr(shape),
                                            nat.darr(h), nat.iarr(bcs), 0))
    return plan


def test_wall_axis_with_five_nodes_raises_value_error():
    plan = _plan((5, 9), (0, 0))
    with pytest.raises(ValueError, match="n_pts >= 6"):
        nat.check(nat.lib().tpf_ccd_axis_info(plan, 0, None, None, None, None))
    nat.lib().tpf_ccd_plan_destroy(plan)


def test_bad_arguments_raise_value_error():
    plan = C.c_void_p()
    with pytest.raises(ValueError):
        nat.check(nat.lib().tpf_ccd_plan_create(C.byref(plan), 4, nat.i64arr([9] * 4),
                                                nat.darr([0.1] * 4), nat.iarr([0] * 4), 0))


CASES = [((9, 7), (0, 0)), ((65, 65), (0, 0)), ((65, 65), (1, 1)), ((33, 20, 9), (0, 1, 0)),
         ((257, 3), (1, 0)), ((256, 3), (0, 0)), ((1025, 2), (1, 1)), ((1025, 2), (0, 0)),
         ((6, 6), (0, 0)), ((4, 5), (1, 1)), ((129, 7), (0, 1)), ((100, 8), (1, 0))]


@pytest.mark.parametrize("shape,bcs", CASES)
def test_host_model_matches_dense_lu_oracle(shape, bcs):
    L = [1.0 + 0.3 * a for a in range(len(shape))]
    plan = _plan(shape, bcs, L)
    rng = np.random.default_rng(11)
    f = rng.standard_normal(shape)
    kinds = tuple("periodic" if b else "wall" for b in bcs)
    o = occd.OracleCCD([s - 1 for s in shape], L, kinds)
    for axis in range(len(shape)):
        if kinds[axis] == "wall" and shape[axis] < 6:
            continue
        if kinds[axis] == "periodic" and shape[axis] < 4:
            continue
        d1 = np.full(shape, np.nan)
        d2 = np.full(shape, np.nan)
        nat.check(nat.lib().tpf_test_ccd_host_model(plan, f.ctypes.data, d1.ctypes.data,
                                                    d2.ctypes.data, axis))
        r1, r2 = o.differentiate(f, axis)
        assert np.max(np.abs(d1 - r1)) <= 1e-13 * np.max(np.abs(r1))
        assert np.max(np.abs(d2 - r2)) <= 1e-13 * np.max(np.abs(r2))
    nat.lib().tpf_ccd_plan_destroy(plan)


@pytest.mark.parametrize("counts", [(1033, 1033, 1033), (5, 17, 2), (9,), (0, 4, 4), (64, 1),
                                    (1033, 1033)])
@pytest.mark.parametrize("strided_mask", [0, 0b011])
def test_fused_launch_block_map_is_a_bijection(counts, strided_mask):
    """Every (slot, slot-local workgroup) of a fused launch is visited once,
    with and without the XCD renumbering of strided slots."""
    cnt = np.asarray(counts, dtype=np.uint32)
    total = int(cnt.sum())
    slot = np.full(max(total, 1), -1, dtype=np.int32)
    local = np.zeros(max(total, 1), dtype=np.uint32)
    nat.check(nat.lib().tpf_test_ccd_block_map(len(counts), cnt.ctypes.data, strided_mask,
                                               slot.ctypes.data, local.ctypes.data))
    got = sorted(zip(slot[:total].tolist(), local[:total].tolist()))
    want = [(s, l) for s, c in enumerate(counts) for l in range(c)]
    assert got == want
    if min(counts) > 0:  # interleaved head: slot = b % nslots
        head = len(counts) * min(counts)
        assert (slot[:head] == np.arange(head) % len(counts)).all()


EXPLICIT_CASES = [((257, 9), (1, 0), 16), ((256, 7, 4), (0, 0, 0), 16), ((40, 33), (0, 1), 8),
                  ((65, 30, 7), (1, 0, 1), 8), ((7, 6), (0, 0), 8), ((100, 8), (1, 1), 0)]


@pytest.mark.parametrize("shape,bcs,cap", EXPLICIT_CASES)
def test_explicit_inverse_tables_match_dense_lu_oracle(shape, bcs, cap):
    """The explicit partition-inverse tables (mode 3, strided axes) through
    the scalar model of that kernel, against the dense-LU oracle; the
    contiguous axis refuses the mode."""
    L = [1.0 + 0.2 * a for a in range(len(shape))]
    plan = _plan(shape, bcs, L)
    f = np.random.default_rng(5).standard_normal(shape)
    kinds = tuple("periodic" if b else "wall" for b in bcs)
    o = occd.OracleCCD([s - 1 for s in shape], L, kinds)
    for axis in range(len(shape) - 1):  # strided axes
        if shape[axis] < (4 if kinds[axis] == "periodic" else 6):
            continue
        nat.check(nat.lib().tpf_ccd_plan_set_launch(plan, axis, 3, cap))
        d1 = np.full(shape, np.nan)
        d2 = np.full(shape, np.nan)
        nat.check(nat.lib().tpf_test_ccd_host_model(plan, f.ctypes.data, d1.ctypes.data,
                                                    d2.ctypes.data, axis))
        r1, r2 = o.differentiate(f, axis)
        assert np.max(np.abs(d1 - r1)) <= 1e-13 * np.max(np.abs(r1))
        assert np.max(np.abs(d2 - r2)) <= 1e-13 * np.max(np.abs(r2))
    with pytest.raises(ValueError, match="strided axes only"):
        nat.check(nat.lib().tpf_ccd_plan_set_launch(plan, len(shape) - 1, 3, cap))
    nat.lib().tpf_ccd_plan_destroy(plan)


def test_explicit_inverse_rejects_long_lines():
    plan = _plan((1025, 9), (1, 0))
    with pytest.raises(ValueError, match="256 unknowns"):
        nat.check(nat.lib().tpf_ccd_plan_set_launch(plan, 0, 3, 0))
    nat.lib().tpf_ccd_plan_destroy(plan)


def test_reduced_system_window_truncation_is_negligible():
    plan = _plan((257, 257, 257), (1, 1, 1))
    P, M, W, t = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    nat.check(nat.lib().tpf_ccd_axis_info(plan, 0, C.byref(P), C.byref(M), C.byref(W), C.byref(t)))
    assert P.value == 16 and M.value == 16
    assert W.value < 16 and t.value < 1e-18
    nat.lib().tpf_ccd_plan_destroy(plan)


def test_launch_mode_selection():
    """Few long lines (2-D 1024^2) take the wavefront-per-line kernel (<= 64
    partitions); many lines take the explicit partition inverse on strided
    axes and, on the contiguous axis, the wavefront kernel with four lines of
    16 partitions per wavefront (2); a fused multi-axis call runs every axis
    of one rows cap and partition count in one lane-kernel grid."""
    plan = _plan((1025, 1025), (1, 0))
    P, M, W, t = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    nat.check(nat.lib().tpf_ccd_axis_info(plan, 0, C.byref(P), C.byref(M), C.byref(W), C.byref(t)))
    assert (P.value, M.value) == (64, 16)
    nat.check(nat.lib().tpf_ccd_plan_set_launch(plan, 0, 1, 0))
    nat.check(nat.lib().tpf_ccd_axis_info(plan, 0, C.byref(P), C.byref(M), C.byref(W), C.byref(t)))
    assert (P.value, M.value) == (16, 64)
    assert nat.lib().tpf_ccd_plan_set_launch(plan, 0, 2, 32) == nat.TPF_ERR_VALUE
    nat.lib().tpf_ccd_plan_destroy(plan)
    plan = _plan((257, 257, 257), (1, 1, 1))
    nat.check(nat.lib().tpf_ccd_axis_info(plan, 2, C.byref(P), C.byref(M), C.byref(W), C.byref(t)))
    assert (P.value, M.value) == (16, 16)
    k = C.c_int()
    kernels = []
    for a in range(3):  # strided axes: explicit inverse (3); contiguous: wavefront (2)
        nat.check(nat.lib().tpf_ccd_axis_kernel(plan, a, C.byref(k)))
        kernels.append(k.value)
    assert kernels == [3, 3, 2]
    nat.check(nat.lib().tpf_ccd_axis_info(plan, 2, C.byref(P), C.byref(M), C.byref(W), C.byref(t)))
    assert (P.value, M.value) == (16, 16)
    nat.check(nat.lib().tpf_ccd_plan_set_launch(plan, 0, 1, 0))
    nat.check(nat.lib().tpf_ccd_axis_kernel(plan, 0, C.byref(k)))
    assert k.value == 1
    nat.lib().tpf_ccd_plan_destroy(plan)
    plan = _plan((1025, 1025), (1, 0))
    nat.check(nat.lib().tpf_ccd_axis_kernel(plan, 0, C.byref(k)))
    assert k.value == 2
    nat.lib().tpf_ccd_plan_destroy(plan)
    # 2-D wall: 513 lines of 511 unknowns take a wavefront per line (64 x 8
    # rows); 257 lines of 255 take the lane kernels
    for shape, want in (((513, 513), [2, 2]), ((257, 257), [3, 1])):
        plan = _plan(shape, (0, 0))
        got = []
        for a in range(2):
            nat.check(nat.lib().tpf_ccd_axis_kernel(plan, a, C.byref(k)))
            got.append(k.value)
        assert got == want, shape
        nat.lib().tpf_ccd_plan_destroy(plan)
