"""The oracle's two bench drivers, pinned against the pinned scalar drivers and the reference build.

``oracle_ranges_simd`` (targets in AVX2 lanes; the checker behind the bench's accuracy figures and
its CPU baseline) must equal ``oracle_ranges`` bit for bit: same per-pair IEEE operations, and each
target's sum runs over all ranges of its group in the same source order.  ``oracle_direct_chunked``
(FMMData::evaluateKernel's 4000-target chunks, Lib/src/FMMData.cpp:252-283) must equal
``oracle_direct``.  A planted change of summation order (each range summed on its own, then added:
the driver's former behaviour) is shown to break bit equality, so the check can fail.
"""
import numpy as np
import pytest

from conftest import source_values
from oracle import oracle as O

SIMD_UKERS = (O.UK["STOKES_PVEL_FUSED"], O.UK["LAPLACE_PGRAD"], O.UK["RPY_ULAPU"])


def _worklist(rng, n_src, n_trg, sizes, n_ranges_max=5):
    """Groups with the given target counts, each against 1..n_ranges_max random source ranges
    (overlapping, unsorted, possibly empty), as a U-list's range list is shaped."""
    gb, ge, gp, rb, re = [], [], [0], [], []
    t = 0
    for m in sizes:
        gb.append(t)
        ge.append(min(t + m, n_trg))
        t = ge[-1]
        for _ in range(int(rng.integers(1, n_ranges_max + 1))):
            a = int(rng.integers(0, n_src))
            b = int(rng.integers(a, min(n_src, a + 400) + 1))
            rb.append(a)
            re.append(b)
        gp.append(len(rb))
    return [np.asarray(v, dtype=np.int32) for v in (gb, ge, gp, rb, re)]


def _points(rng, n, coincide=None):
    x = rng.random((n, 3))
    if coincide is not None:
        k = min(len(coincide), n // 4)
        x[:k] = coincide[:k]  # exact coincidences: the masked pair must give 0
    return x


@pytest.mark.parametrize("uk", SIMD_UKERS)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ranges_simd_equals_scalar_bitwise(uk, seed):
    rng = np.random.default_rng(100 + seed)
    n_src = 1500
    src = _points(rng, n_src)
    # group sizes off the 8-lane width, single-target groups, an empty group
    sizes = [1, 7, 8, 9, 15, 17, 64, 3, 0, 1, 33, 100]
    n_trg = sum(sizes)
    trg = _points(rng, n_trg, coincide=src[::7])
    v = source_values(uk, n_src, rng)
    wl = _worklist(rng, n_src, n_trg, sizes)
    base = rng.normal(size=(n_trg, O.trg_dim(uk)))  # += semantics on a non-zero output
    a = O.ranges(uk, src, v, trg, *wl, trg_value=base)
    b = O.ranges(uk, src, v, trg, *wl, simd=True, trg_value=base)
    assert np.array_equal(a, b)
    assert np.all(np.isfinite(a))


@pytest.mark.parametrize("uk", SIMD_UKERS)
def test_ranges_simd_equals_scalar_on_selected_groups(uk):
    rng = np.random.default_rng(9)
    n_src = 900
    src = _points(rng, n_src)
    sizes = [5, 13, 1, 40, 8, 27, 2]
    n_trg = sum(sizes)
    trg = _points(rng, n_trg, coincide=src[::3])
    v = source_values(uk, n_src, rng)
    wl = _worklist(rng, n_src, n_trg, sizes)
    sel = np.array([6, 0, 3, 2], dtype=np.int32)
    a = O.ranges(uk, src, v, trg, *wl, sel=sel, trg_value=np.zeros((n_trg, O.trg_dim(uk))))
    b = O.ranges(uk, src, v, trg, *wl, sel=sel, simd=True, trg_value=np.zeros((n_trg, O.trg_dim(uk))))
    assert np.array_equal(a, b)
    # groups outside sel are untouched
    gb, ge = wl[0], wl[1]
    for g in (1, 4, 5):
        assert not b[gb[g]:ge[g]].any()


def test_planted_range_split_breaks_bit_equality():
    """Summing each source range on its own and adding the range sums (the planted mistake) moves
    the result off the scalar driver's bits on multi-range groups: the equality above can fail."""
    uk = O.UK["STOKES_PVEL_FUSED"]
    rng = np.random.default_rng(3)
    n_src = 1200
    src = _points(rng, n_src)
    sizes = [64] * 8
    n_trg = sum(sizes)
    trg = _points(rng, n_trg)
    v = source_values(uk, n_src, rng)
    gb, ge, gp, rb, re = _worklist(rng, n_src, n_trg, sizes, n_ranges_max=6)
    ref = O.ranges(uk, src, v, trg, gb, ge, gp, rb, re, simd=True)
    kt = O.trg_dim(uk)
    split = np.zeros((n_trg, kt))
    for g in range(len(gb)):
        acc = np.zeros((ge[g] - gb[g], kt))
        for r in range(gp[g], gp[g + 1]):
            one = O.ranges(uk, src, v, trg[gb[g]:ge[g]], [0], [ge[g] - gb[g]], [0, 1], [rb[r]], [re[r]])
            acc = acc + one / O.scale(uk)
        split[gb[g]:ge[g]] = O.scale(uk) * acc
    assert not np.array_equal(split, ref)
    np.testing.assert_allclose(split, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("uk", [O.UK["LAPLACE_PGRAD"], O.UK["STOKES_PVEL_FUSED"], O.UK["STOKES_VEL"]])
@pytest.mark.parametrize("n_trg", [1, 3999, 4000, 4001, 8192])
def test_direct_chunked_equals_direct_bitwise(uk, n_trg):
    rng = np.random.default_rng(n_trg)
    n_src = 300
    src = _points(rng, n_src)
    trg = _points(rng, n_trg, coincide=src)
    v = source_values(uk, n_src, rng)
    base = rng.normal(size=(n_trg, O.trg_dim(uk)))
    a = O.direct(uk, src, v, trg, trg_value=base)
    for nth in (1, 3, 0):
        b = O.direct(uk, src, v, trg, trg_value=base, chunked=True, nthreads=nth)
        assert np.array_equal(a, b), nth


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("uk,parts", [(13, ((128, 0, 0, 4), (128, 1, 4, 13))), (11, ((16, 0, 0, 4),)),
                                      (2, ((1, 0, 0, 1),))])
def test_ranges_simd_whole_leaf_matches_live_reference(uk, parts):
    """One whole leaf of an adaptive octree through the SIMD driver against the reference's own
    Test/SimpleKernel.cpp summed over the leaf's U-list (FP64 parity bound of the path, 1e-12)."""
    import stkfmm_amd as S
    from stkfmm_amd import workloads as W

    x = W.pvel_points(8000)
    tr = S.Tree(x, 300)
    xs = x[tr.perm]
    rng = np.random.default_rng(17)
    v = source_values(uk, tr.n, rng)
    leaf = int(np.argmax(tr.leaf_end - tr.leaf_begin))
    kt = O.trg_dim(uk)
    got = O.ranges(uk, xs, v, xs, tr.leaf_begin, tr.leaf_end, tr.ulist_ptr, tr.rng_begin, tr.rng_end,
                   sel=[leaf], simd=True)
    b, e = tr.leaf_begin[leaf], tr.leaf_end[leaf]
    src = np.concatenate([np.arange(tr.rng_begin[r], tr.rng_end[r])
                          for r in range(tr.ulist_ptr[leaf], tr.ulist_ptr[leaf + 1])])
    ref = sum(O.simplekernel_direct(K, layer, c1 - c0, kt, xs[src], v[src, c0:c1], xs[b:e])
              for K, layer, c0, c1 in parts)
    l2, mx = O.max_error(got[b:e], ref)
    assert l2 < 1e-12 and mx < 1e-12, (l2, mx)
