"""ige_steepest_descent_noalloc, called through the C ABI, against ige_steepest_descent on
identical states: the same words, tracked energies and report, bit for bit, on two calls
in a row (the second on a smaller state, with a clamp mask and a step budget, so stale
workspace contents from the first call would show), on a dense model whose matrix-core
fields come from the workspace, and the size limit reported as an error."""
import ctypes

import numpy as np
import pytest

from ising_engine_amd import _lib, instances

pytestmark = pytest.mark.gpu


def twin_states(ctx, qb, cells, seed, kernel, sweeps):
    m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    ladder = instances.geometric_betas(3.0, 0.5, cells)
    pair = []
    for _ in range(2):
        st = _lib.State(m, cells, seed, kernel=kernel)
        st.set_ladder(ladder)
        if sweeps:
            st.run(sweeps, exchange=cells > 1)
        pair.append(st)
    np.testing.assert_array_equal(pair[0].spins(), pair[1].spins())
    return m, pair[0], pair[1]


def descent_noalloc(st, max_steps, clamped=None):
    cl = None if clamped is None else np.ascontiguousarray(clamped, dtype=np.uint8)
    rep = _lib.DescentReport()
    ptr = None if cl is None else cl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    rc = st._lib.ige_steepest_descent_noalloc(st.handle, int(max_steps), ptr, ctypes.byref(rep))
    return rc, rep


def assert_same(a, b, ra, rb):
    np.testing.assert_array_equal(b.spins(), a.spins())
    np.testing.assert_array_equal(b.energies(), a.energies())
    np.testing.assert_array_equal(b.energies(), b.recompute_energies())
    assert (rb.steps, rb.converged, rb.flips, rb.proposed) == \
        (ra.steps, ra.converged, ra.flips, ra.proposed)


def test_noalloc_descent_matches_allocating_descent_twice_in_a_row(ctx):
    print("[noalloc] call 1: 3D +-J L=16, 8 cells, multispin, unlimited steps", flush=True)
    _, a, b = twin_states(ctx, instances.hypercubic_pm_j(16, 3, seed=4), 8, 77,
                          _lib.KERNEL_MULTISPIN, 30)
    ra = a.steepest_descent(_lib.STEPS_UNLIMITED)
    rc, rb = descent_noalloc(b, _lib.STEPS_UNLIMITED)
    assert rc == 0, _lib.last_error()
    assert ra.converged == 1 and ra.flips > 0
    assert_same(a, b, ra, rb)
    print("[noalloc] call 2: random QUBO n=300, 2 cells, generic, clamped, 5 steps", flush=True)
    qb = instances.random_qubo(300, 6, seed=9)
    # from the random initial words, so the 5-step budget ends the descent early
    _, a2, b2 = twin_states(ctx, qb, 2, 78, _lib.KERNEL_GENERIC, 0)
    clamped = (np.arange(qb.n) % 7 == 3).astype(np.uint8)
    ra2 = a2.steepest_descent(5, clamped=clamped)
    rc, rb2 = descent_noalloc(b2, 5, clamped)
    assert rc == 0, _lib.last_error()
    assert ra2.steps == 5 and ra2.converged == 0
    assert_same(a2, b2, ra2, rb2)
    # and unlimited on what is left, still equal
    ra3 = a2.steepest_descent(_lib.STEPS_UNLIMITED, clamped=clamped)
    rc, rb3 = descent_noalloc(b2, _lib.STEPS_UNLIMITED, clamped)
    assert rc == 0, _lib.last_error()
    assert_same(a2, b2, ra3, rb3)


def test_noalloc_descent_on_a_dense_model(ctx):
    print("[noalloc] dense SK n=256, 2 cells", flush=True)
    _, a, b = twin_states(ctx, instances.sk_int8(256, seed=3), 2, 5, _lib.KERNEL_DENSE, 6)
    ra = a.steepest_descent(_lib.STEPS_UNLIMITED)
    rc, rb = descent_noalloc(b, _lib.STEPS_UNLIMITED)
    assert rc == 0, _lib.last_error()
    assert ra.flips > 0
    assert_same(a, b, ra, rb)


def test_noalloc_descent_rejects_a_state_beyond_its_workspace(ctx):
    print("[noalloc] size limit", flush=True)
    qb = instances.random_qubo(8, 3, seed=1)
    m = _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)
    big = _lib.State(m, 32768, 1, kernel=_lib.KERNEL_GENERIC)  # 16 MiB of keys alone
    before = big.spins()
    rc, _ = descent_noalloc(big, 1)
    assert rc == _lib.IGE_ERR_INVALID
    assert "no-alloc limit" in _lib.last_error()
    np.testing.assert_array_equal(big.spins(), before)  # refused before any flip
    big.close()
    # the workspace still serves the next call
    _, a, b = twin_states(ctx, qb, 1, 3, _lib.KERNEL_GENERIC, 4)
    ra = a.steepest_descent(_lib.STEPS_UNLIMITED)
    rc, rb = descent_noalloc(b, _lib.STEPS_UNLIMITED)
    assert rc == 0, _lib.last_error()
    assert_same(a, b, ra, rb)
