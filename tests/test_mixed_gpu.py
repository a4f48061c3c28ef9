"""SC_MIXED kNN: BF16 MFMA candidates + FP64 replay must equal the oracle bit
for bit (indices in heap-array order and distance bits), i.e. zero index
mismatches against the FP64 brute-force form on every input."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from smartcore_amd import _capi as S
from smartcore_amd import synth

pytestmark = pytest.mark.gpu


def check(ref, qs, k):
    ix = S.Index(ref)
    gi, gd, gc = ix.knn(qs, k, S.SC_EUCLIDIAN, 2, S.SC_MIXED)
    st = S.last_stats()
    oi, od, oc = O.linear_find_batch(ref, qs, k)
    np.testing.assert_array_equal(gc, oc)
    mism = 0
    for r in range(qs.shape[0]):
        c = int(oc[r])
        mism += int(np.sum(gi[r, :c] != oi[r, :c]))
        np.testing.assert_array_equal(gd[r, :c].view(np.uint64), od[r, :c].view(np.uint64))
    assert mism == 0
    return st


def test_mixed_lognormal_128():
    ref = synth.lognormal_rows(61, 30_000, 128)
    qs = synth.lognormal_rows(62, 600, 128)  # two query tiles of 512
    st = check(ref, qs, 10)
    assert st["uncertified"] == 0
    assert st["shortlist_rescored"] >= 600 * 10


@pytest.mark.parametrize("d,k", [(96, 20), (32, 5), (8, 1), (17, 3)])
def test_mixed_shapes(d, k):
    ref = synth.lognormal_rows(70 + d, 12_000, d)
    qs = synth.lognormal_rows(71 + d, 130, d)
    check(ref, qs, k)


def test_mixed_gaussian_96_k20():
    rng = np.random.default_rng(80)
    ref = rng.normal(size=(40_000, 96)).astype(np.float32)
    qs = rng.normal(size=(64, 96)).astype(np.float32)
    check(ref, qs, 20)


def test_mixed_tie_heavy_keeps_heap_order():
    rng = np.random.default_rng(81)
    ref = rng.integers(0, 3, size=(5_000, 16)).astype(np.float32)
    qs = rng.integers(0, 3, size=(100, 16)).astype(np.float32)
    for k in (1, 5, 10):
        check(ref, qs, k)


def test_mixed_falls_back_for_non_finite_rows_and_queries():
    rng = np.random.default_rng(82)
    ref = rng.normal(size=(3_000, 24)).astype(np.float32)
    ref[17, 3] = np.nan
    qs = rng.normal(size=(40, 24)).astype(np.float32)
    check(ref, qs, 6)
    ref2 = rng.normal(size=(3_000, 24)).astype(np.float32)
    qs2 = rng.normal(size=(40, 24)).astype(np.float32)
    qs2[5, 0] = np.inf
    qs2[9, 1] = 1e20
    st = check(ref2, qs2, 6)
    assert st["uncertified"] == 2


def test_mixed_large_k_and_f64_index_use_exact_form():
    rng = np.random.default_rng(83)
    ref = rng.normal(size=(2_000, 16)).astype(np.float32)
    qs = rng.normal(size=(20, 16)).astype(np.float32)
    check(ref, qs, 40)
    check(ref.astype(np.float64), qs.astype(np.float64), 7)


def test_mixed_rejects_minkowski():
    ix = S.Index(np.ones((4, 3), np.float32))
    with pytest.raises(S.Failed):
        ix.knn(np.ones((1, 3), np.float32), 2, S.SC_MINKOWSKI, 3, S.SC_MIXED)


def check_l1(ref, qs, k):
    """SC_MIXED Manhattan (f32 sums with a rigorous bound, candidates rescored
    in the reference arithmetic) against the oracle and the FP64 form."""
    ix = S.Index(ref)
    gi, gd, gc = ix.knn(qs, k, S.SC_MANHATTAN, 1, S.SC_MIXED)
    st = S.last_stats()
    fi, fd, fc = ix.knn(qs, k, S.SC_MANHATTAN, 1, S.SC_PLAIN_FP64)
    np.testing.assert_array_equal(gc, fc)
    np.testing.assert_array_equal(gi, fi)
    np.testing.assert_array_equal(gd.view(np.uint64), fd.view(np.uint64))
    oi, od, oc = O.linear_find_batch(ref, qs[:5], k, O.MANHATTAN)
    for r in range(min(5, qs.shape[0])):
        c = int(oc[r])
        np.testing.assert_array_equal(gi[r, :c], oi[r, :c])
        np.testing.assert_array_equal(gd[r, :c].view(np.uint64), od[r, :c].view(np.uint64))
    return st


@pytest.mark.parametrize("n,d,k", [(40_000, 96, 20), (12_000, 16, 1), (9_000, 33, 64),
                                   (300_000, 96, 20)])
def test_mixed_manhattan_equals_fp64_form(n, d, k):
    ref = synth.lognormal_rows(120 + d, n, d)
    qs = synth.lognormal_rows(121 + d, 150, d)
    st = check_l1(ref, qs, k)
    assert st["shortlist_rescored"] >= 150 * k


def test_mixed_manhattan_ties_non_finite_and_fallbacks():
    rng = np.random.default_rng(125)
    # integer rows: many exact ties, heap order decides
    ref = rng.integers(0, 3, size=(20_000, 12)).astype(np.float32)
    qs = rng.integers(0, 3, size=(200, 12)).astype(np.float32)
    for k in (1, 7, 33):
        check_l1(ref, qs, k)
    # NaN / inf rows and a huge-valued query
    ref = rng.normal(size=(5_000, 20)).astype(np.float32)
    ref[10, 3] = np.nan
    ref[11, 0] = np.inf
    ref[12, :] = 3e38
    qs = rng.normal(size=(30, 20)).astype(np.float32)
    qs[4, :] = -3e38
    check_l1(ref, qs, 9)
    # k above the tiled limit and f64 rows take the exact form
    check_l1(ref, qs, 70)
    check_l1(ref.astype(np.float64), qs.astype(np.float64), 9)


def _coded_scan_ran(fn):
    """Runs fn with the launch timers on; True when the 16-bit coded
    Manhattan scan was launched."""
    lib = S.lib()
    lib.sc_profile_enable.argtypes = [C.c_int]
    lib.sc_profile_read.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    lib.sc_profile_enable(1)
    try:
        out = fn()
        ms, cnt = C.c_double(), C.c_uint64()
        lib.sc_profile_read(b"knn_scan_q16", C.byref(ms), C.byref(cnt))
    finally:
        lib.sc_profile_enable(0)
    return out, cnt.value > 0


@pytest.mark.parametrize("kind,n,d,k", [("gaussian", 30_000, 96, 20), ("gaussian", 300_000, 40, 10),
                                        ("uniform01", 20_000, 128, 5), ("lognormal", 25_000, 33, 64),
                                        ("gaussian", 15_000, 7, 1)])
def test_mixed_manhattan_coded_scan(kind, n, d, k):
    """Indexes whose values span a range small against their mean |v| take
    the 16-bit coded scan (v_sad_u16 over code pairs, bound from the code
    errors); the answers must still equal the FP64 form bit for bit and the
    oracle.  Covers odd widths (padding dimensions), the prefix-seeded plan
    (300,000 rows), k = 1 and k = 64."""
    rng = np.random.default_rng(7 + d)
    if kind == "gaussian":
        ref = rng.normal(size=(n, d)).astype(np.float32)
        qs = rng.normal(size=(120, d)).astype(np.float32)
    elif kind == "uniform01":
        ref = rng.random(size=(n, d)).astype(np.float32)
        qs = rng.random(size=(120, d)).astype(np.float32)
    else:
        ref = synth.lognormal_rows(130 + d, n, d)
        qs = synth.lognormal_rows(131 + d, 120, d)
    _, coded = _coded_scan_ran(lambda: check_l1(ref, qs, k))
    assert coded


def test_mixed_manhattan_coded_scan_query_edges():
    """Queries outside the index's code range are clamped (their code error
    widens the bound), a query with a NaN or a huge value has no threshold
    (its lists overflow and the serial kernel answers it), and exact integer
    data codes without error; every answer equals the FP64 form's."""
    rng = np.random.default_rng(140)
    ref = rng.uniform(-1.0, 1.0, size=(40_000, 24)).astype(np.float32)
    qs = rng.uniform(-3.0, 3.0, size=(100, 24)).astype(np.float32)
    qs[5, 2] = np.nan
    qs[6, :] = 1e35
    qs[7, :] = 0.0
    _, coded = _coded_scan_ran(lambda: check_l1(ref, qs, 8))
    assert coded
    ref = rng.integers(-5, 6, size=(30_000, 16)).astype(np.float32)
    qs = rng.integers(-5, 6, size=(80, 16)).astype(np.float32)
    _, coded = _coded_scan_ran(lambda: check_l1(ref, qs, 12))
    assert coded


def test_mixed_manhattan_coded_scan_declined():
    """Non-finite rows, a value range far wider than the typical value, and
    rows too wide to stage keep the f32 scan; answers are unchanged."""
    rng = np.random.default_rng(150)
    ref = rng.normal(size=(8_000, 20)).astype(np.float32)
    ref[3, 3] = np.inf
    qs = rng.normal(size=(40, 20)).astype(np.float32)
    _, coded = _coded_scan_ran(lambda: check_l1(ref, qs, 5))
    assert not coded
    ref = rng.normal(size=(8_000, 20)).astype(np.float32)
    ref[100, 0] = 1e9
    _, coded = _coded_scan_ran(lambda: check_l1(ref, qs, 5))
    assert not coded
    ref = rng.normal(size=(3_000, 900)).astype(np.float32)
    qs = rng.normal(size=(20, 900)).astype(np.float32)
    _, coded = _coded_scan_ran(lambda: check_l1(ref, qs, 5))
    assert not coded


@pytest.mark.parametrize("d,k,kind,n", [(16, 10, "lognormal", 2_200_000), (48, 32, "ties", 2_200_000),
                                        (16, 1, "lognormal", 2_200_000),
                                        (128, 10, "lognormal", 2_200_000),
                                        (128, 1, "lognormal", 2_200_000),
                                        (128, 10, "ties", 2_200_000),
                                        (96, 20, "gaussian", 2_000_000),
                                        (96, 20, "lognormal", 600_000)])
def test_mixed_seeded_large_index(d, k, kind, n):
    """Indexes of 524,288 rows or more scan a 16,384-row prefix first, then
    row ranges growing 8 times, each tested against the k-th exact distance
    of the earlier candidates.  The result must still equal the FP64 form bit
    for bit (and the oracle on a few queries), also with heavy exact ties and
    with a single neighbour.  The d = 128 cases run the scan the 4,000,000 x 128
    workload times (eight K steps plus the folded bound step) on both the
    middle and the last level.  The 2,000,000 x 96 Gaussian case is the
    served index's shape, and 600,000 rows end in a short last level."""
    nq = 3_000
    if kind == "lognormal":
        ref = synth.lognormal_rows(90 + d, n, d)
        qs = synth.lognormal_rows(91 + d, nq, d)
    elif kind == "gaussian":
        rng = np.random.default_rng(93)
        ref = rng.normal(size=(n, d)).astype(np.float32)
        qs = rng.normal(size=(nq, d)).astype(np.float32)
    else:
        rng = np.random.default_rng(92)
        ref = rng.integers(0, 4, size=(n, d)).astype(np.float32)
        qs = rng.integers(0, 4, size=(nq, d)).astype(np.float32)
    ix = S.Index(ref)
    gi, gd, gc = ix.knn(qs, k, S.SC_EUCLIDIAN, 2, S.SC_MIXED)
    fi, fd, fc = ix.knn(qs, k, S.SC_EUCLIDIAN, 2, S.SC_PLAIN_FP64)
    np.testing.assert_array_equal(gc, fc)
    np.testing.assert_array_equal(gi, fi)
    np.testing.assert_array_equal(gd.view(np.uint64), fd.view(np.uint64))
    oi, od, oc = O.linear_find_batch(ref, qs[:3], k)
    for r in range(3):
        c = int(oc[r])
        np.testing.assert_array_equal(gi[r, :c], oi[r, :c])
        np.testing.assert_array_equal(gd[r, :c].view(np.uint64), od[r, :c].view(np.uint64))
