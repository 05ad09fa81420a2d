"""The work-unit codec and checkpoint text (no GPU needed).

The rank arithmetic is pinned by exact Python integers (sum of d_i (n-1-i)!,
the identity fact_work::VectToBigint implements, distributed/src/fact_work.cpp:76-86)
over n up to 63, where ranks exceed 2^290; the text layout by a checkpoint
written out by hand from the reference's stream operators (solution
operator<<, common/src/solution.cpp:132-142; works::save,
distributed/src/works.cpp:172-191; works operator<<, :558-569;
work::writeHeader / writeIntervals, distributed/src/work.cpp:544-583).  The
reference's own codec needs GMP, which this image lacks, so it is not run:
the identities above pin it instead.
"""
from math import factorial

import numpy as np
import pytest

from pbbperm_amd import factoradic as F


def _rand_digits(rng, n):
    return np.array([rng.integers(0, n - i) for i in range(n)], dtype=np.intc)


@pytest.mark.parametrize("n", [1, 2, 5, 13, 20, 21, 34, 50, 63])
def test_rank_round_trip_matches_exact_integers(n):
    import pbbperm_amd as P
    rng = np.random.default_rng(n)
    cases = [np.zeros(n, np.intc), np.arange(n - 1, -1, -1, dtype=np.intc)]
    cases += [_rand_digits(rng, n) for _ in range(40)]
    for d in cases:
        dec = P.fact_to_dec(d)
        assert int(dec) == F.to_int(d)
        assert np.array_equal(P.dec_to_fact(n, dec), d)
    # the last rank of the full interval is n! - 1
    assert int(P.fact_to_dec(np.arange(n - 1, -1, -1, dtype=np.intc))) == factorial(n) - 1


def test_rank_errors():
    import pbbperm_amd as P
    with pytest.raises(P.PbbError):
        P.fact_to_dec([0, 2, 0])          # digit 1 must be <= 1
    with pytest.raises(P.PbbError):
        P.dec_to_fact(5, str(factorial(5)))   # not below 5!
    with pytest.raises(P.PbbError):
        P.dec_to_fact(5, "12a")
    with pytest.raises(P.PbbError):
        P.dec_to_fact(63, "9" * 100)       # beyond the fixed width


def _reference_text(n, root_cost, root, sb, best_cost, best, units):
    """What works::save writes, spelled out from the reference's operators."""
    def sol(cost, perm):
        return f"{n}\n{cost}\n" + "".join(f"{p} " for p in perm) + "\n"
    out = sol(root_cost, root) + f"{sb} " + sol(best_cost, best) + f"{len(units)}\n"
    for wid, explored, ivs in units:
        size = sum(e - b for _, b, e in ivs if e > b)
        out += f"{wid} 0 {len(ivs)} {explored} {len(ivs)}\n"
        out += "".join(f"{i} {b} {e}\n" for i, b, e in ivs)
        out += f"{size}\n"
    return out


def test_checkpoint_text_has_the_reference_layout():
    import pbbperm_amd as P
    n = 20
    rng = np.random.default_rng(3)
    lo = sorted(int(x) for x in rng.integers(0, factorial(n) - 1, size=6))
    ivs = [(k, lo[2 * k], lo[2 * k + 1]) for k in range(3)]
    ids = [i for i, _, _ in ivs]
    pos = np.array([F.to_digits(b, n) for _, b, _ in ivs])
    end = np.array([F.to_digits(e, n) for _, _, e in ivs])
    root = list(range(n))
    best = list(rng.permutation(n))
    txt = P.checkpoint_format(n, ids, pos, end, root_cost=1300, root_perm=root, simple_bounds=123456789,
                              best_cost=1278, best_perm=best, explored=5000, work_id=0, max_intervals=3)
    assert txt == _reference_text(n, 1300, root, 123456789, 1278, best, [(0, 5000, ivs)])


def test_checkpoint_parse_reads_every_unit_and_skips_empty_intervals():
    import pbbperm_amd as P
    n = 12
    root = list(range(n))
    best = list(range(n - 1, -1, -1))
    u1 = (4, 100, [(0, 10, 20), (1, 5000, 4000), (2, 30, 30)])        # second one is empty
    u2 = (9, 23, [(7, factorial(n) - 50, factorial(n) - 1)])
    txt = _reference_text(n, 999, root, 42, 777, best, [u1, u2])
    hdr, ids, pos, end = P.checkpoint_parse(txt, n)
    assert hdr["root_cost"] == 999 and hdr["best_cost"] == 777 and hdr["simple_bounds"] == 42
    assert hdr["best_perm"] == best and hdr["root_perm"] == root
    assert hdr["explored"] == 123 and hdr["nb"] == 3
    assert list(ids) == [0, 2, 7]
    got = [(F.to_int(p), F.to_int(e)) for p, e in zip(pos, end)]
    assert got == [(10, 20), (30, 30), (factorial(n) - 50, factorial(n) - 1)]
    # round trip through the formatter
    txt2 = P.checkpoint_format(n, ids, pos, end, hdr["root_cost"], root, hdr["simple_bounds"], hdr["best_cost"],
                               best, explored=hdr["explored"])
    hdr2, ids2, pos2, end2 = P.checkpoint_parse(txt2, n)
    assert np.array_equal(pos2, pos) and np.array_equal(end2, end) and list(ids2) == list(ids)


def test_checkpoint_parse_errors():
    import pbbperm_amd as P
    with pytest.raises(P.PbbError):
        P.checkpoint_parse("5\n10\n0 1 2 3\n", 5)                        # truncated root
    good = _reference_text(5, 10, range(5), 1, 10, range(5), [(0, 0, [(0, 0, 200)])])
    with pytest.raises(P.PbbError):
        P.checkpoint_parse(good, 5)                                      # 200 is not below 5!
    with pytest.raises(P.PbbError):
        P.checkpoint_parse(good.replace(" 200\n", " x\n"), 5)
