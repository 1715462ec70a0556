"""Pins the oracle: golden fixtures from the reference router, the reference
router itself (when built here), the step-by-step restatement, and
properties the routing must satisfy whatever the implementation."""
import ast
import glob
import os

import numpy as np
import pytest

import oracle
import oracle.ref as ref
from graphs_small import CASES, VARIANTS

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "router_*.npz")))


def _fixture_kwargs(z):
    kw = ast.literal_eval(str(z["kwargs"]))      # repr of a dict of literals written by make_golden.py
    env = ast.literal_eval(str(z["env"]))
    if env.get("LIBRA_SKIP_TC") == "1":
        kw["skip_tc"] = True
    if "LIBRA_SPTC_UNIT_MIN_NNZ" in env:
        kw["unit_min_nnz"] = int(env["LIBRA_SPTC_UNIT_MIN_NNZ"])
    kw.setdefault("skip_tc", False)
    kw.setdefault("unit_min_nnz", 2)
    return kw


def _same(meta, arrays):
    return [n for n, a in zip(oracle.NAMES, arrays)
            if not (meta[n].dtype == a.dtype and meta[n].shape == a.shape and np.array_equal(meta[n], a))]


def test_golden_present():
    assert len(GOLDEN) >= 10


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[7:-4] for p in GOLDEN])
@pytest.mark.parametrize("faithful", [False, True])
def test_oracle_matches_golden(path, faithful):
    z = np.load(path, allow_pickle=False)
    meta = oracle.route(z["rowptr"], z["col"], z["val"], faithful=faithful, **_fixture_kwargs(z))
    bad = _same(meta, [z["out_" + n] for n in oracle.NAMES])
    assert not bad, bad


@pytest.mark.skipif(not ref.available(), reason="reference router not built (needs /root/reference)")
@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_reference_random(seed):
    from trspmm import graphs as g
    rng = np.random.default_rng(100 + seed)
    csr = g.rmat_gcn(int(rng.integers(8, 11)), seed=seed)
    density = int(rng.integers(2, 9))
    arrays, _ = ref.route(csr.rowptr, csr.col, csr.val, density=density)
    meta = oracle.route(csr.rowptr, csr.col, csr.val, density=density)
    assert not _same(meta, arrays)


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("vi", range(len(VARIANTS)))
def test_closed_form_equals_faithful(case, vi):
    csr = CASES[case]()
    a = oracle.route(csr.rowptr, csr.col, csr.val, faithful=False, **VARIANTS[vi])
    b = oracle.route(csr.rowptr, csr.col, csr.val, faithful=True, **VARIANTS[vi])
    assert not _same(a, [b[n] for n in oracle.NAMES])


def _triplets(csr):
    rows = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    return sorted(zip(rows.tolist(), csr.col.tolist(), csr.val.tolist()))


@pytest.mark.parametrize("case", sorted(CASES))
def test_exact_mode_represents_every_nonzero_once(case):
    csr = CASES[case]()
    meta = oracle.route(csr.rowptr, csr.col, csr.val, exact=True)
    r, c, v = oracle.decode(meta)
    assert sorted(zip(r.tolist(), c.tolist(), v.tolist())) == _triplets(csr)


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("floor", [4, 8, 16])
def test_tuned_mode_represents_every_nonzero_once(case, floor):
    """Tuned routing moves sparse 2:4 groups to the residual route: still every
    nonzero exactly once, TC untouched, every kept group at or above the floor."""
    csr = CASES[case]()
    meta = oracle.route(csr.rowptr, csr.col, csr.val, exact=True, sptc_min_nnz=floor)
    r, c, v = oracle.decode(meta)
    assert sorted(zip(r.tolist(), c.tolist(), v.tolist())) == _triplets(csr)
    base = oracle.route(csr.rowptr, csr.col, csr.val, exact=True)
    for n in oracle.NAMES:
        if n.startswith("t_"):
            assert np.array_equal(meta[n], base[n]), n
    kept = (meta["s_pos"] >= 0).reshape(meta["s_pos"].shape[0], 32).sum(axis=1)
    assert np.all(kept >= floor)
    assert meta["s_window"].size <= base["s_window"].size


@pytest.mark.parametrize("case", sorted(CASES))
def test_compat_mode_drops_exactly_the_overflow(case):
    """compat = exact minus the 3rd/4th hits of each 2:4 group row
    (block_online.cpp:690-694, :727); t_* and s_* identical in both modes."""
    csr = CASES[case]()
    a = oracle.route(csr.rowptr, csr.col, csr.val, exact=False)
    b = oracle.route(csr.rowptr, csr.col, csr.val, exact=True)
    for n in oracle.NAMES:
        if n.startswith(("t_", "s_")):
            assert np.array_equal(a[n], b[n]), n
    mask = a["s_mask"].astype(np.int64)
    hits = np.zeros_like(mask)
    for bit in range(4):
        hits += (mask >> bit) & 1
    overflow = int(np.maximum(hits - 2, 0).sum())
    ra, _, _ = oracle.decode(a)
    assert ra.size == csr.nnz - overflow


@pytest.mark.parametrize("case", sorted(CASES))
def test_structural_invariants(case):
    csr = CASES[case]()
    m = oracle.route(csr.rowptr, csr.col, csr.val)
    W = (csr.m_pad + 15) // 16
    so = m["s_offsets"]
    assert so.size == W + 1 and so[0] == 0 and np.all(np.diff(so) >= 0)
    tiles = np.sum((np.diff(so) + 3) // 4)
    assert m["s_tile_window"].size == tiles and m["s_window_tile_offset"][-1] == tiles
    # popcount of each sub-block's occupancy = its value count
    bits = m["t_binary"].view(np.uint64).reshape(-1, 2)
    pc = np.array([bin(int(x)).count("1") + bin(int(y)).count("1") for x, y in bits], dtype=np.int64)
    assert np.array_equal(pc, np.diff(m["t_block_offset"]))
    # every 2:4 row slot holds at most 2 entries with distinct positions < 4
    pos = m["s_pos"].reshape(-1, 2)
    both = (pos[:, 0] >= 0) & (pos[:, 1] >= 0)
    assert np.all(pos[both, 0] != pos[both, 1]) and np.all(pos < 4)
    # groups: sorted columns, -1 padding at the end
    sc = m["s_column"]
    for row in sc:
        valid = row[row >= 0]
        assert np.all(np.diff(valid) > 0) and np.all(row[valid.size:] == -1)


def test_invariant_check_catches_a_planted_mistake():
    """A transposed TC half (lo/hi swapped) must break the exact-mode check."""
    csr = CASES["dense"]()
    meta = oracle.route(csr.rowptr, csr.col, csr.val, exact=True)
    bad = dict(meta.arrays)
    tb = bad["t_binary"].copy().reshape(-1, 2)
    tb = tb[:, ::-1].copy()
    bad["t_binary"] = tb.reshape(-1)
    r, c, v = oracle.decode(bad, 16)
    assert sorted(zip(r.tolist(), c.tolist(), v.tolist())) != _triplets(csr)


def test_f16_conversion_matches_numpy():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-8, 5, 20000),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, 6.0e-8, 5.96e-8, 2.98e-8, 2.99e-8,
                  6.1035156e-05, 6.0975552e-05, 1.0009765625, 1.00048828125, 1.00146484375], np.float32)])
    got = oracle.f32_to_f16_bits(x)
    want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


def test_sort_order_is_a_descending_permutation():
    rng = np.random.default_rng(1)
    for n in (0, 1, 5, 16, 17, 100, 1000):
        k = rng.integers(1, 33, n).astype(np.int32)
        o = oracle.sort_order(k)
        assert sorted(o.tolist()) == list(range(n))
        assert np.all(np.diff(k[o]) <= 0)
    # n <= 16: libstdc++ uses insertion sort only -> already sorted input is unchanged
    k = np.sort(rng.integers(1, 5, 16))[::-1].astype(np.int32)
    assert np.array_equal(oracle.sort_order(k), np.arange(16))


def test_spmm64_against_dense_brute_force():
    csr = CASES["ragged"]()
    meta = oracle.route(csr.rowptr, csr.col, csr.val, exact=True)
    B = np.random.default_rng(2).standard_normal((csr.n, 8))
    A = np.zeros((csr.m, csr.n))
    rows = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    for r, c, v in zip(rows, csr.col, csr.val):
        if r < csr.m:
            A[r, c] += v
    got = oracle.spmm64(meta, B, csr.m, half_values=False)
    assert np.allclose(got, A @ B, rtol=1e-12, atol=1e-12)


def test_oracle_rejects_large_window():
    csr = CASES["ragged"]()
    with pytest.raises(RuntimeError):
        oracle.route(csr.rowptr, csr.col, csr.val, window=20)


def _pair_sequential(order, n):
    """block_online.cpp:365-391 with equal degrees: each unit not yet used, in
    `order`, takes the lowest-index unused unit (other than itself)."""
    used = [False] * n
    pairs = []
    for v in order:
        if used[v]:
            continue
        p = next((k for k in range(n) if not used[k] and k != v), n)
        if p == n:
            continue
        used[v] = used[p] = True
        pairs.append((v, p))
    left = next((k for k in range(n) if not used[k]), -1)
    return pairs, left


def _pair_high_water(order, n):
    """The device router's form of the same pairing: every index up to the
    latest partner hw is used, so a unit is taken at its turn iff v <= hw, and
    an active unit's partner is the first index above hw, other than v, that
    no earlier active unit holds."""
    active = [False] * n
    hw, pairs = -1, []
    for v in order:
        if v <= hw:
            continue
        p = next((k for k in range(hw + 1, n) if k != v and not active[k]), n)
        if p == n:
            continue
        active[v] = True
        pairs.append((v, p))
        hw = p
    left = next((k for k in range(hw + 1, n) if not active[k]), -1) if 2 * len(pairs) < n else -1
    return pairs, left


@pytest.mark.parametrize("seed", range(6))
def test_router_pairing_high_water_form_equals_sequential(seed):
    rng = np.random.default_rng(seed)
    orders = []
    for n in (0, 1, 2, 3, 5, 17, 64, 65, 200):
        orders += [list(range(n)), list(range(n))[::-1], list(rng.permutation(n))]
        # blocks of equal keys permuted among themselves (the tie orders introsort leaves)
        perm = list(range(n))
        for b in range(0, n, 7):
            perm[b:b + 7] = list(rng.permutation(perm[b:b + 7]))
        orders.append(perm)
    for order in orders:
        n = len(order)
        assert _pair_high_water(order, n) == _pair_sequential(order, n), order
