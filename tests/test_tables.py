"""Coefficient tables, Kronecker composition and block decoding (CPU).

Pins: the Brent equations (exact integer arithmetic) for every shipped table
and for the composed configurations; filename invariants <name>-<R>-<nnz>;
planted mistakes (a flipped sign, a swapped digit order) must be caught.
"""
import os
import re
from fractions import Fraction

import pytest


def test_all_tables_are_exact_algorithms(fmm):
    names = sorted(set(fmm.alg_dic))
    assert len({fmm.alg_dic[n][0] for n in names}) == 40
    for name in names:
        path, dims = fmm.alg_dic[name]
        alg = fmm.load_algorithm(path, dims, name)
        assert fmm.brent_residual(alg) == 0, name


def test_filename_invariants(fmm):
    seen = set()
    for name, (path, dims) in fmm.alg_dic.items():
        base = os.path.basename(path)
        if path in seen:
            continue
        seen.add(path)
        alg = fmm.load_algorithm(path, dims, name)
        mt = re.match(r"^[a-z]+(\d)(\d)(\d)-(\d+)-(\d+)$", base)
        if mt:
            assert tuple(int(x) for x in mt.group(1, 2, 3)) == tuple(dims), base
            assert int(mt.group(4)) == alg.rank, base
            assert sum(alg.nnz()) == int(mt.group(5)), base


@pytest.mark.parametrize("spec,dims,rank,nnz", [
    ((["222"], [1]), (2, 2, 2), 7, (12, 12, 12)),
    ((["222"], [2]), (4, 4, 4), 49, (144, 144, 144)),
    ((["222", "323"], [1, 1]), (6, 4, 6), 105, (384, 384, 360)),
    ((["333"], [1]), (3, 3, 3), 23, (49, 45, 45)),
    ((["232"], [1]), (2, 3, 2), 11, (16, 16, 16)),
    ((["323"], [1]), (3, 2, 3), 15, (32, 32, 30)),
    ((["232", "222"], [1, 1]), (4, 6, 4), 77, (192, 192, 192)),
])
def test_composed_configs(fmm, spec, dims, rank, nnz):
    alg = fmm.build_algorithm(*spec)
    assert alg.dims == dims and alg.rank == rank and alg.nnz() == nnz
    assert fmm.brent_residual(alg) == 0


def test_planted_sign_error_is_caught(fmm):
    alg = fmm.build_algorithm(["222"], [1])
    U = [row[:] for row in alg.U]
    U[0][0] = -U[0][0]
    bad = fmm.Algorithm("bad", alg.levels, U, alg.V, alg.W)
    assert fmm.brent_residual(bad) > 0


def test_planted_digit_order_error_is_caught(fmm):
    """Decoding a composed row with the levels reversed (level 0 taken as the
    least significant digit) must break a heterogeneous 2-level algorithm."""
    alg = fmm.build_algorithm(["222", "323"], [1, 1])
    rev = fmm.Algorithm("rev", list(reversed(alg.levels)), alg.U, alg.V, alg.W)
    assert fmm.brent_residual(rev) > 0


def test_kronecker_rank_and_values(fmm):
    a = fmm.build_algorithm(["222"], [1])
    b = fmm.build_algorithm(["323"], [1])
    c = fmm.compose([a, b])
    # entry (x*rows_b + y, r*R_b + s) == a[x][r] * b[y][s]
    for (x, r, y, s) in [(0, 0, 0, 0), (3, 6, 5, 14), (1, 4, 2, 9)]:
        assert c.U[x * 6 + y][r * 15 + s] == a.U[x][r] * b.U[y][s]


def test_parser_rejects_symbolic(fmm, tmp_path):
    with pytest.raises(fmm.CoeffError):
        fmm.parse_coeff("x")
    with pytest.raises(fmm.CoeffError):
        fmm.parse_coeff("-x2i")
    assert fmm.parse_coeff("-.125") == Fraction(-1, 8)
    assert fmm.parse_coeff(" .5 ") == Fraction(1, 2)
    p = tmp_path / "two_groups"
    p.write_text("1 0\n#\n0 1\n")
    with pytest.raises(fmm.CoeffError):
        fmm.read_coeffs(str(p))


def test_partition_and_jobname(fmm):
    p = fmm.partition(8192, 8192, 8192, (3, 2, 3))
    assert (p.md, p.kd, p.nd) == (8184, 8192, 8184)
    assert (p.ms, p.ks, p.ns) == (2728, 4096, 2728)
    assert p.fringe == (8, 0, 8)
    p = fmm.partition(14400, 14400, 480, (6, 4, 6))
    assert (p.ms, p.ks, p.ns) == (2400, 120, 2400)
    assert fmm.jobname(["222", "323"], [1, 1], "abc") == "222-1_323-1_abc"
