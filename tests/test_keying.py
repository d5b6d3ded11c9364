"""The random-walk keying, recomputed from the golden CSR without the oracle's code.

A pure-Python Philox4x32-10 (Salmon et al., SC'11: two 32x32->64 multiplies per round with
M0 = 0xD2511F53, M1 = 0xCD9E8D57, Weyl key increments 0x9E3779B9 / 0xBB67AE85), itself pinned by
Random123's known-answer vectors, draws
  * the start of walker w of instance i:  floor((r1:r0) * n / 2^64) of Philox(seed, {i, w, 0xFFFFFFFF, 0});
  * step s of walker w at vertex v:        N_v[floor((r1:r0) * d_v / 2^64)] of Philox(seed, {i, w, s, 0}),
with N_v the sorted adjacency of v in the golden CSR (the compressors' input).  The oracle's walks
must equal these; the GPU's walks are checked against the oracle elsewhere, so an off-by-one in
the step counter, a swapped counter word or a wrong key half fails here.
"""
import numpy as np
import pytest

from conftest import csr_of, load_case

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox(ctr, key):
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    for _ in range(10):
        p0, p1 = M0 * c0, M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ c3 ^ k1) & MASK, p0 & MASK
        k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
    return c0, c1, c2, c3


@pytest.mark.parametrize("ctr,key,exp", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((MASK,) * 4, (MASK,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_python_philox_random123_kat(ctr, key, exp):
    assert philox(ctr, key) == exp


def draw(seed, ctr, bound):
    r = philox(ctr, (seed & MASK, seed >> 32))
    return (((r[1] << 32) | r[0]) * bound) >> 64


def walks_from_csr(rowptr, col, n, seed, instance, walkers, length, starts=None):
    out = np.full((walkers, length + 1), n, dtype=np.uint32)
    for w in range(walkers):
        v = int(starts[w]) if starts is not None else draw(seed, (instance, w, 0xFFFFFFFF, 0), n)
        out[w, 0] = v
        for s in range(length):
            d = int(rowptr[v + 1] - rowptr[v])
            if d == 0:
                break
            v = int(col[int(rowptr[v]) + draw(seed, (instance, w, s, 0), d)])
            out[w, s + 1] = v
    return out


@pytest.mark.parametrize("fmt", ["le", "gt"])
@pytest.mark.parametrize("seed,instance", [(42, 0), (7, 12345), ((3 << 32) | 99, 0xFFFFFFF0)])
def test_oracle_walks_follow_the_keying(oracle_mod, fmt, seed, instance):
    c = load_case("rmat12")
    rowptr, col = csr_of(c)
    og = oracle_mod.Graph(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    got, _ = og.random_walk(48, 24, seed, instance=instance)
    np.testing.assert_array_equal(got, walks_from_csr(rowptr, col, og.n, seed, instance, 48, 24))
    starts = np.arange(48, dtype=np.uint32) * 37 % og.n
    got, _ = og.random_walk(48, 24, seed, instance=instance, starts=starts)
    np.testing.assert_array_equal(got, walks_from_csr(rowptr, col, og.n, seed, instance, 48, 24, starts))


def test_dead_end_pads_with_n(oracle_mod):
    """A walk that reaches a vertex without neighbours is padded with n (the directed fixture has
    such vertices)."""
    c = load_case("degrees")
    rowptr, col = csr_of(c)
    og = oracle_mod.Graph(c["le_graph"].tobytes(), c["le_offset"].tobytes(), "le")
    deg = np.diff(rowptr.astype(np.int64))
    assert (deg == 0).any()
    got, _ = og.random_walk(200, 12, 5, instance=3)
    exp = walks_from_csr(rowptr, col, og.n, 5, 3, 200, 12)
    np.testing.assert_array_equal(got, exp)
    assert (exp == og.n).any()
