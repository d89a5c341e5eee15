"""Performance model (meta/model_coefficient.py, meta/model_gen.py), CPU.

Pinned against hand-computed values: Strassen's counters read off its table,
and the model formula worked by hand for a plain GEMM at 240^3 and a
one-level Strassen ABC at 480^3 with the reference's constants.
"""
import os

import pytest


def test_strassen_counters(fmm):
    from fmm_gen_amd import model
    alg = fmm.build_algorithm(["222"], [1])
    comp, abc, ab, naive = model.gen_model_coefficient(alg)
    assert comp == [7, 5, 5, 12]
    assert abc == [12, 12, 12, 0, 0, 0]
    assert ab == [12, 12, 7, 0, 0, 36]
    assert naive == [7, 7, 7, 19, 19, 36]


def test_two_level_counters_scale(fmm):
    from fmm_gen_amd import model
    comp, abc, _, _ = model.gen_model_coefficient(fmm.build_algorithm(["222"], [2]))
    assert comp[0] == 49 and abc[:3] == [144, 144, 144]
    assert comp[1] == 144 - 49


def test_model_gflops_hand_values():
    from fmm_gen_amd import model
    # GEMM, 240^3: actual 2*240^3 flops at 28.32 flop/ns, mops 7*240^2 at 0.5
    g = model.get_model_gflops(240, 240, 240, (1, 1, 1), 1, model.GEMM_ROW)
    assert g == pytest.approx(23.47285536676337, rel=1e-12)
    # Strassen L1 ABC, 480^3 (leaf 240^3): actual 196,070,400 flops, mops 48*240^2
    s = model.get_model_gflops(480, 480, 480, (2, 2, 2), 1, [7, 5, 5, 12, 12, 12, 12, 0, 0, 0])
    assert s == pytest.approx(26.630098342671317, rel=1e-12)


def test_coefficient_matrix_layout(fmm):
    from fmm_gen_amd import model
    coef = model.generate_coefficient_mat()
    assert len(coef) == 1 + 6 * 23
    assert coef[0] == model.GEMM_ROW
    i222 = model.alg_set.index("222")
    assert coef[1 + i222] == [7, 5, 5, 12, 12, 12, 12, 0, 0, 0]                 # L1 abc
    assert coef[1 + i222 + 23] == [7, 5, 5, 12, 12, 12, 7, 0, 0, 36]            # L1 ab
    assert coef[1 + i222 + 2 * 23] == [7, 5, 5, 12, 7, 7, 7, 19, 19, 36]        # L1 naive
    assert coef[1 + i222 + 3 * 23][:4] == [49, 95, 95, 144]                      # L2


def test_csv_files(tmp_path):
    from fmm_gen_amd import model
    files = model.csv_files(str(tmp_path))
    assert len(files) == 18
    names = sorted(os.path.basename(f) for f in files)
    assert "model_square_1level_abc_1core.csv" in names
    assert "model_rankk_2level_naive_1core.csv" in names
    lines = open(os.path.join(tmp_path, "model_square_1level_abc_1core.csv")).read().splitlines()
    assert lines[0] == model.HEADER
    assert len(lines) == 1 + len(model.SAMPLES)
    first = [float(x) for x in lines[1].split(",")]
    assert first[0] == 240.0 and first[1] == pytest.approx(23.47, abs=0.006)
