"""Pins the CPU oracle before anything is checked against it.

* decode: every fixture written by the compiled reference compressors must decode to exactly
  the compressor's input (the golden CSR), through the whole-list path and through random
  access at every (v, k);
* offsets: record lengths sum to the .graph bit length minus < 8 bits of padding;
  the n-overcount fixture;
* the hand-derived worked example (header fields, payload widths, two decodes);
* Philox4x32-10: Random123 known-answer vectors and rocRAND's implementation;
* sampling: invariants, determinism and chi-square uniformity (the sampling semantics have
  no reference-held vectors; see the oracle header).
"""
import os

import numpy as np
import pytest

from conftest import CASES, GOLDEN, csr_of, load_case

FMTS = ("le", "gt")


def _graph(oracle_mod, case, fmt):
    c = load_case(case)
    return oracle_mod.Graph(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt), c


# ---------------------------------------------------------------- decode / format

@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", CASES)
def test_roundtrip_golden_csr(oracle_mod, case, fmt):
    g, c = _graph(oracle_mod, case, fmt)
    assert g.n == int(c["n"])
    rowptr, col = g.decode_csr()
    erp, ecol = csr_of(c)
    np.testing.assert_array_equal(rowptr, erp)
    np.testing.assert_array_equal(col, ecol)


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", ["example", "degrees", "small_u", "overcount"])
def test_random_access_brute_force(oracle_mod, case, fmt):
    g, c = _graph(oracle_mod, case, fmt)
    erp, ecol = csr_of(c)
    for v in range(g.n):
        d = int(erp[v + 1] - erp[v])
        assert g.degree(v) == d
        got = [g.get(v, k) for k in range(d)]
        assert got == ecol[erp[v]:erp[v + 1]].astype(np.int64).tolist(), v
        assert g.get(v, d) == -1


@pytest.mark.parametrize("fmt", FMTS)
def test_random_access_hub(oracle_mod, fmt):
    g, c = _graph(oracle_mod, "hub", fmt)
    erp, ecol = csr_of(c)
    hub = int(np.argmax(np.diff(erp.astype(np.int64))))
    d = int(erp[hub + 1] - erp[hub])
    assert d == 30000
    rng = np.random.default_rng(1)
    for k in rng.choice(d, size=500, replace=False).tolist() + [0, d - 1]:
        assert g.get(hub, k) == int(ecol[erp[hub] + k])


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", CASES)
def test_offsets_sum_to_graph_bits(oracle_mod, case, fmt):
    c = load_case(case)
    ubl, nf, lens = oracle_mod.parse_offsets(c[f"{fmt}_offset"].tobytes())
    n = int(c["n"])
    assert ubl == int(c[f"{fmt}_offset"][0])
    assert nf >= n
    assert np.all(lens[n:] == 0) and np.all(lens[:n] > 0)
    total = int(lens.sum())
    gbits = 8 * len(c[f"{fmt}_graph"])
    assert gbits - 8 < total <= gbits
    # ubl is the bit width of the longest record (LE/chunk_compress/compressor.hpp:227-232)
    assert ubl == int(lens.max()).bit_length()


def test_overcount_fixture(oracle_mod):
    c = load_case("overcount")
    for fmt in FMTS:
        ubl, nf, _ = oracle_mod.parse_offsets(c[f"{fmt}_offset"].tobytes())
        assert nf > int(c["n"]) == 3


def test_hub_needs_wide_length_field():
    c = load_case("hub")
    assert c["le_offset"][0] > 16 and c["gt_offset"][0] > 16


def test_worked_example(oracle_mod):
    # v = 5, N = [1, 3, 8, 9, 20], derived by hand from the encoders.
    g, _ = _graph(oracle_mod, "example", "le")
    h = g.header(5)
    assert (h["d"], h["f1"], h["f2"], h["first_pos"], h["C"], h["q"], h["R"]) == (5, 4, 3, 1, 2, 2, 0)
    assert h["header_bits"] == 5 + 5 + 5 + 3
    assert g.get(5, 1) == 3 and g.get(5, 3) == 9
    assert g.payload_bits(5, 0) == 4 and g.payload_bits(5, 1) == 2 * 4 + 3 + 1
    record = "00110" "00101" "00100" "010" "0100" "0011" "1111" "0011" "1011"
    assert g.record_start(6) - g.record_start(5) == len(record)
    gg, _ = _graph(oracle_mod, "example", "gt")
    h = gg.header(5)
    assert (h["d"], h["f1"], h["f2"], h["first_pos"], h["C"], h["q"], h["R"]) == (5, 4, 1, 1, 3, 1, 2)
    assert gg.get(5, 3) == 9 and gg.get(5, 4) == 20
    assert gg.payload_bits(5, 3) == 4 + 1
    assert gg.record_start(6) - gg.record_start(5) == 30
    # a d = 0 GT record is gamma(0) x 4 = "1111"
    assert gg.record_start(1) - gg.record_start(0) == 4


def test_corrupt_offsets_rejected(oracle_mod):
    c = load_case("degrees")
    off = bytearray(c["le_offset"].tobytes())
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.Graph(c["le_graph"].tobytes(), bytes([0]) + bytes(off[1:]), "le")
    with pytest.raises(oracle_mod.OracleError):   # graph truncated by one byte
        oracle_mod.Graph(c["le_graph"].tobytes()[:-1], bytes(off), "le")
    with pytest.raises(oracle_mod.OracleError):   # graph longer than the records
        oracle_mod.Graph(c["le_graph"].tobytes() + b"\0", bytes(off), "le")


# ---------------------------------------------------------------- Philox

R123_KAT = [  # Random123 kat_vectors, philox4x32 10 rounds: ctr[4] key[2] -> out[4]
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,exp", R123_KAT)
def test_philox_random123_kat(oracle_mod, ctr, key, exp):
    assert tuple(int(x) for x in oracle_mod.philox(ctr, key)) == exp


def test_philox_matches_rocrand(oracle_mod):
    rows = [l.split() for l in open(os.path.join(GOLDEN, "philox_rocrand.txt")) if not l.startswith("#")]
    assert len(rows) == 64
    for r in rows:
        v = [int(x, 16) for x in r]
        assert [int(x) for x in oracle_mod.philox(v[0:4], v[4:6])] == v[6:10]


def test_uniform_is_multiply_high(oracle_mod):
    for bound in (1, 2, 3, 7, 1000, 2 ** 31 - 1, 2 ** 40 + 3):
        for w in range(20):
            r = oracle_mod.philox((5, w, 9, 0), (42 & 0xffffffff, 0))
            x = (int(r[1]) << 32) | int(r[0])
            assert oracle_mod.uniform(42, 5, w, 9, 0, bound) == (x * bound) >> 64


# ---------------------------------------------------------------- sampling semantics

def _edge_set(c):
    return set(zip(c["src"].tolist(), c["dst"].tolist()))


@pytest.mark.parametrize("fmt", FMTS)
def test_rw_invariants(oracle_mod, fmt):
    g, c = _graph(oracle_mod, "degrees", fmt)   # directed: has dead ends
    E = _edge_set(c)
    erp, _ = csr_of(c)
    deg = np.diff(erp.astype(np.int64))
    walks, nbytes = g.random_walk(3000, 12, seed=7, instance=3)
    assert walks.shape == (3000, 13) and nbytes > 0
    dead = 0
    for row in walks.tolist():
        for a, b in zip(row, row[1:]):
            if a == g.n:
                assert b == g.n
            elif b == g.n:
                assert deg[a] == 0
                dead += 1
            else:
                assert (a, b) in E
    assert dead > 0
    again, _ = g.random_walk(3000, 12, seed=7, instance=3)
    np.testing.assert_array_equal(walks, again)
    other, _ = g.random_walk(3000, 12, seed=7, instance=4)
    assert not np.array_equal(walks, other)
    # starts are drawn uniformly from [0, n)
    assert walks[:, 0].max() < g.n
    # explicit starts
    st = np.arange(100, dtype=np.uint32) * 17 % g.n
    w2, _ = g.random_walk(100, 3, seed=1, starts=st)
    np.testing.assert_array_equal(w2[:, 0], st)


def test_rw_chi_square_uniform(oracle_mod):
    from scipy.stats import chisquare
    g, c = _graph(oracle_mod, "rmat12", "gt")
    erp, ecol = csr_of(c)
    deg = np.diff(erp.astype(np.int64))
    v = int(np.argsort(deg)[-3])
    W = 200 * int(deg[v])
    walks, _ = g.random_walk(W, 1, seed=11, starts=np.full(W, v, np.uint32))
    nb = ecol[erp[v]:erp[v + 1]]
    pos = np.searchsorted(nb, walks[:, 1])
    counts = np.bincount(pos, minlength=len(nb))
    assert chisquare(counts).pvalue > 1e-4


@pytest.mark.parametrize("fmt", FMTS)
def test_khop_invariants(oracle_mod, fmt):
    g, c = _graph(oracle_mod, "rmat12", fmt)
    E = _edge_set(c)
    erp, _ = csr_of(c)
    deg = np.diff(erp.astype(np.int64))
    rng = np.random.default_rng(3)
    seeds = [rng.integers(0, g.n, size=k).astype(np.uint32) for k in (1, 5, 40, 300)]
    seeds.append(np.array([7, 7, 3, 7], np.uint32))
    fans = [[25, 10], [3], [15, 10], [10, 10, 5], [4, 4]]
    res = g.khop(seeds, fans, seed=5, instances=[11, 12, 13, 14, 15])
    for r in range(len(seeds)):
        e0, e1 = int(res["edge_off"][r]), int(res["edge_off"][r + 1])
        src, dst = res["src"][e0:e1], res["dst"][e0:e1]
        front = list(dict.fromkeys(seeds[r].tolist()))
        pos = 0
        for h, f in enumerate(fans[r]):
            cnt = int(res["hop_count"][r, h])
            hs, hd = src[pos:pos + cnt], dst[pos:pos + cnt]
            pos += cnt
            expect_src = []
            for v in front:
                expect_src += [v] * min(f, int(deg[v]))
            assert hs.tolist() == expect_src
            for v in front:
                sel = hd[hs == v]
                assert len(set(sel.tolist())) == len(sel)
                assert np.all(np.diff(sel.astype(np.int64)) > 0)
                assert all((v, int(x)) in E for x in sel)
            front = list(dict.fromkeys(hd.tolist()))
        assert pos == e1 - e0
    # each request depends only on its own seeds and instance
    alone = g.khop([seeds[2]], [fans[2]], seed=5, instances=[13])
    e0, e1 = int(res["edge_off"][2]), int(res["edge_off"][3])
    np.testing.assert_array_equal(alone["dst"], res["dst"][e0:e1])


def test_khop_subset_uniform(oracle_mod):
    from scipy.stats import chisquare
    g, c = _graph(oracle_mod, "degrees", "le")
    erp, ecol = csr_of(c)
    deg = np.diff(erp.astype(np.int64))
    v = int(np.nonzero(deg == 10)[0][0])
    res = g.khop([np.array([v], np.uint32)] * 4000, [[3]] * 4000, seed=9,
                 instances=np.arange(4000))
    nb = ecol[erp[v]:erp[v + 1]]
    counts = np.bincount(np.searchsorted(nb, res["dst"]), minlength=10)
    assert counts.sum() == 12000
    assert chisquare(counts).pvalue > 1e-4


@pytest.mark.parametrize("fmt", FMTS)
def test_mdrw_invariants(oracle_mod, fmt):
    g, c = _graph(oracle_mod, "rmat12", fmt)
    E = _edge_set(c)
    erp, ecol = csr_of(c)
    res = g.mdrw([20, 50, 7], target=300, seed=3, instances=[1, 2, 3])
    for i, m in enumerate([20, 50, 7]):
        s0, s1 = int(res["step_off"][i]), int(res["step_off"][i + 1])
        assert s1 - s0 == 300 - m
        ss, sd = res["step_src"][s0:s1], res["step_dst"][s0:s1]
        assert all((int(a), int(b)) in E for a, b in zip(ss, sd))
        v0, v1 = int(res["vert_off"][i]), int(res["vert_off"][i + 1])
        verts = res["verts"][v0:v1]
        assert np.all(np.diff(verts.astype(np.int64)) > 0)
        assert set(sd.tolist()) <= set(verts.tolist())
        vs = set(verts.tolist())
        exp = [(int(x), int(y)) for x in verts for y in ecol[erp[x]:erp[x + 1]] if int(y) in vs]
        i0, i1 = int(res["ind_off"][i]), int(res["ind_off"][i + 1])
        got = list(zip(res["ind_src"][i0:i1].tolist(), res["ind_dst"][i0:i1].tolist()))
        assert got == exp


def test_mdrw_degree_proportional(oracle_mod):
    """The first step's source is drawn proportionally to degree among the initial frontier."""
    from scipy.stats import chisquare
    g, c = _graph(oracle_mod, "rmat12", "le")
    erp, _ = csr_of(c)
    deg = np.diff(erp.astype(np.int64))
    I = 3000
    res = g.mdrw([2] * I, target=3, seed=21, instances=np.arange(I))
    # frontier of instance i: two uniform draws with tag 0xFFFFFFFF
    hits = np.zeros(2)
    exp = np.zeros(2)
    for i in range(I):
        f = [oracle_mod.uniform(21, i, j, 0xFFFFFFFF, 0, g.n) for j in range(2)]
        if f[0] == f[1] or deg[f[0]] == 0 or deg[f[1]] == 0:
            continue
        src = int(res["step_src"][res["step_off"][i]])
        hits[0 if src == f[0] else 1] += 1
        p0 = deg[f[0]] / (deg[f[0]] + deg[f[1]])
        exp += [p0, 1 - p0]
    assert chisquare(hits, exp * hits.sum() / exp.sum()).pvalue > 1e-4
