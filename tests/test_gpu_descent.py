"""Steepest descent, 1-opt polish and incumbent tracking on the device against the oracle,
bit-exact: spins, tracked energies, flip counts, step counts, the incumbent's bits, energy
and the sweep it was found at.
"""
import numpy as np
import pytest

from ising_engine_amd import _lib, instances
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def make(ctx, qb):
    return _lib.Model(ctx, qb.n, qb.offset, qb.linear, qb.pi, qb.pj, qb.q)


def unpack(words):
    cells, n = words.shape
    bits = (words[:, None, :] >> np.arange(64, dtype=np.uint64)[None, :, None]) & np.uint64(1)
    return bits.reshape(cells * 64, n).astype(np.uint8)


DESCENT_CASES = [
    # name, factory, kernel, cells, pre-sweeps, clamp fraction
    ("qubo_n300_generic", lambda: instances.random_qubo(300, 6, 31, offdiag=20, diag=30),
     _lib.KERNEL_GENERIC, 3, 0, 0.0),
    ("qubo_n300_generic_clamped", lambda: instances.random_qubo(300, 6, 32, offdiag=20, diag=30),
     _lib.KERNEL_GENERIC, 2, 0, 0.25),
    ("pmJ_3d_L6_multispin_after_pt", lambda: instances.hypercubic_pm_j(6, 3, 33),
     _lib.KERNEL_MULTISPIN, 4, 6, 0.0),
    ("sk_n300_dense", lambda: instances.sk_int8(300, 34), _lib.KERNEL_DENSE, 2, 0, 0.0),
    ("dense_qubo_n130_clamped", lambda: instances.dense_qubo(130, 35, scale=3),
     _lib.KERNEL_DENSE, 1, 0, 0.2),
]


@pytest.mark.parametrize("case", DESCENT_CASES, ids=[c[0] for c in DESCENT_CASES])
def test_state_descent_matches_oracle(ctx, case):
    name, factory, kernel, cells, pre, clamp_frac = case
    print(f"[descent] {name}", flush=True)
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    seed = 0xDE5C_0000 + len(name)
    st = _lib.State(model, cells, seed, kernel=kernel)
    assert st.stats()[0].kernel == kernel
    ost = ob.OracleState(csr, cells, seed, view=ob.VIEW_WORD if kernel != _lib.KERNEL_MULTISPIN
                         else ob.VIEW_PLANE)
    if pre:
        ladder = instances.geometric_betas(2.0, 0.3, cells)
        st.set_ladder(ladder)
        st.run(pre, exchange=True)
        ost.run(pre, exchange=True, ladder=ladder)
        np.testing.assert_array_equal(st.spins(), ost.words)
    clamped = None
    if clamp_frac:
        clamped = (np.random.default_rng(seed).random(qb.n) < clamp_frac).astype(np.uint8)
    words, energies = ost.words.copy(), ost.energy.copy()
    flips = np.zeros((cells, 64), dtype=np.uint32)
    # a bounded first call, then to the local minimum
    rep = st.steepest_descent(3, clamped=clamped)
    steps, flips = csr.steepest_descent(words, energies, 3, clamped=clamped, flips=flips)
    assert (rep.steps, rep.converged) == (steps, 0) and rep.flips == flips.sum()
    assert rep.proposed == 3 * qb.n * 64 * cells
    np.testing.assert_array_equal(st.spins(), words)
    np.testing.assert_array_equal(st.energies(), energies)
    rep = st.steepest_descent(clamped=clamped)
    steps, flips = csr.steepest_descent(words, energies, 1 << 30, clamped=clamped, flips=flips)
    assert rep.converged == 1 and rep.steps == steps and rep.steps > 0
    assert rep.proposed == (steps + 1) * qb.n * 64 * cells
    np.testing.assert_array_equal(st.spins(), words)
    np.testing.assert_array_equal(st.energies(), energies)
    np.testing.assert_array_equal(st.recompute_energies(), energies)
    # 1-opt: no free single flip lowers any replica's energy
    h = st.fields().astype(np.int64)
    x = unpack(words).reshape(cells, 64, qb.n).transpose(0, 2, 1)
    gain = np.where(x == 1, -h, h)
    free = np.ones(qb.n, dtype=bool) if clamped is None else clamped == 0
    assert (gain[:, free, :] >= 0).all()
    # a converged state stays put
    rep = st.steepest_descent(clamped=clamped)
    assert (rep.steps, rep.converged, rep.flips) == (0, 1, 0)


POLISH_CASES = [
    ("qubo_n200", lambda: instances.random_qubo(200, 5, 41, offdiag=50, diag=80), 100, 0.0),
    ("qubo_n200_clamped", lambda: instances.random_qubo(200, 5, 42, offdiag=50, diag=80), 70,
     0.3),
    ("sk_n257_dense", lambda: instances.sk_int8(257, 43), 65, 0.0),
    ("pmJ_2d_L10", lambda: instances.hypercubic_pm_j(10, 2, 44), 1, 0.0),
]


@pytest.mark.parametrize("case", POLISH_CASES, ids=[c[0] for c in POLISH_CASES])
def test_model_polish_matches_oracle(ctx, case):
    name, factory, count, clamp_frac = case
    print(f"[polish] {name}", flush=True)
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    rng = np.random.default_rng(len(name))
    x0 = rng.integers(0, 2, size=(count, qb.n), dtype=np.uint8)
    clamped = (rng.random(qb.n) < clamp_frac).astype(np.uint8) if clamp_frac else None
    x, e, f = model.polish(x0, clamped=clamped)
    for i in range(count):
        xo = x0[i].copy()
        imp, nflip = csr.polish_1opt(xo, clamped=clamped)
        np.testing.assert_array_equal(x[i], xo)
        assert e[i] == csr.energy(xo) == int(qb.energy(x0[i])) - imp
        assert f[i] == nflip
    np.testing.assert_array_equal(x0.shape, x.shape)


def test_polish_rejects_non_binary(ctx):
    qb = instances.random_qubo(50, 4, 45)
    model = make(ctx, qb)
    x = np.zeros((2, qb.n), dtype=np.uint8)
    x[1, 3] = 2
    with pytest.raises(_lib.IgeError) as ei:
        model.polish(x)
    assert ei.value.code == _lib.IGE_ERR_INVALID


TRACK_CASES = [
    ("qubo_n120_generic_pt", lambda: instances.random_qubo(120, 5, 51, offdiag=6, diag=6),
     _lib.KERNEL_GENERIC, ob.VIEW_WORD, 4),
    ("pmJ_2d_L8_multispin_pt", lambda: instances.hypercubic_pm_j(8, 2, 52),
     _lib.KERNEL_MULTISPIN, ob.VIEW_PLANE, 3),
    ("sk_n100_dense_pt", lambda: instances.sk_int8(100, 53), _lib.KERNEL_DENSE, ob.VIEW_WORD, 2),
]


@pytest.mark.parametrize("case", TRACK_CASES, ids=[c[0] for c in TRACK_CASES])
def test_track_best_matches_oracle(ctx, case):
    name, factory, kernel, view, cells = case
    print(f"[track] {name}", flush=True)
    qb = factory()
    csr = ob.CSR.from_qubo(qb)
    model = make(ctx, qb)
    seed = 0x7EAC_0000 + len(name)
    st = _lib.State(model, cells, seed, kernel=kernel)
    ost = ob.OracleState(csr, cells, seed, view=view)
    with pytest.raises(_lib.IgeError):
        st.best()
    t_max, t_min = instances.suggested_temp_range(qb)
    ladder = instances.geometric_betas(t_max, t_min, cells)
    st.set_ladder(ladder)
    inc = ob.Incumbent(qb.n)
    found = None
    st.update_best()
    if inc.update(ost.words, ost.temp_of, ost.energy):
        found = 0
    sweeps = 25
    st.run(10, exchange=True, track_best=True)
    st.run(sweeps - 10, exchange=True, track_best=True)
    for k in range(sweeps):
        ost.run(1, exchange=True, ladder=ladder)
        if inc.update(ost.words, ost.temp_of, ost.energy):
            found = k + 1
    np.testing.assert_array_equal(st.spins(), ost.words)
    x, e, sw = st.best()
    assert e == inc.energy.value == csr.energy(x)
    assert sw == found
    np.testing.assert_array_equal(x, inc.x)
    # polish of the incumbent (the reference's end-of-run step) agrees as well
    xp, ep, _ = model.polish(x)
    xo = x.copy()
    csr.polish_1opt(xo)
    np.testing.assert_array_equal(xp[0], xo)
    assert ep[0] == csr.energy(xo) <= e


def test_full_size_sk4096_descent(ctx):
    """SK N=4096 at its real shape: descent to the local minimum from the seeded start, on the
    matrix-core fields; tracked energies equal a recompute and every replica is 1-opt."""
    qb = instances.workload_sk_anneal()["qubo"]
    model = make(ctx, qb)
    st = _lib.State(model, 1, 91)
    before = st.energies().copy()
    rep = st.steepest_descent()
    assert rep.converged == 1 and rep.steps > 100
    after = st.energies()
    np.testing.assert_array_equal(st.recompute_energies(), after)
    assert (after < before).all()
    h = st.fields().astype(np.int64)[0]
    x = unpack(st.spins())[:, :].T  # [n][64]
    assert (np.where(x == 1, -h, h) >= 0).all()
