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


def test_mi355x_csvs_written_in_reference_layout(fmm, tmp_path):
    """The 18-file layout of meta/model_gen.py:191-249 for the MI355X FP64
    machine: same names with the machine as suffix, 23 algorithms, GEMM below
    the FP64 matrix peak, and an FMM's modelled edge over GEMM on large
    squares bounded by the flop ratio (8/7 for Strassen)."""
    from fmm_gen_amd import model as M
    files = M.csv_files(str(tmp_path), M.MI355X_F64)
    assert len(files) == 18
    names = sorted(os.path.basename(f) for f in files)
    assert "model_rankk_1level_naive_mi355x-f64.csv" in names
    assert "model_square_2level_abc_mi355x-f64.csv" in names
    import csv
    rows = list(csv.DictReader(open(os.path.join(tmp_path, "model_square_1level_abc_mi355x-f64.csv"))))
    assert len(rows) == len(M.SAMPLES) and len(rows[0]) == 2 + 23
    for r in rows:
        assert 0 < float(r["gemm"]) <= 78.6e3
        assert float(r["222"]) <= float(r["gemm"]) * 8 / 7 * 1.001


def test_model_check_against_sweep_rows(fmm):
    """check_against_sweep compares measured and modelled rates per variant
    and counts winner agreement (rows in sweep.py's CSV layout)."""
    from fmm_gen_amd import model as M
    rows = []
    for s in (2400, 4800, 9600):
        gem = M.get_model_gflops(s, s, s, (1, 1, 1), 1, M.GEMM_ROW, M.MI355X_F64)
        fmm_ = M.get_model_gflops(s, s, s, (2, 2, 2), 1, M.model_row(fmm.build_algorithm(["222"], [1]), "naive"),
                                  M.MI355X_F64)
        rows.append({"m": s, "n": s, "k": s, "gemm": gem, "222-1_naive": fmm_ * 0.9, "222-1_naive/gemm": 0.0})
    out = M.check_against_sweep(rows, M.MI355X_F64)
    assert set(out) == {"222-1_naive"}
    c = out["222-1_naive"]
    assert abs(c["measured_over_model_median"] - 0.9) < 1e-3 and c["points"] == 3


def test_sweep_ranges_match_reference(fmm):
    """The three sweeps of meta/sbatch_util.py:122-157, step 240."""
    from fmm_gen_amd import sweep as S
    sq = S.points("square")
    assert sq[0] == (240, 240, 240) and sq[-1] == (12000, 12000, 12000) and len(sq) == 50
    rk = S.points("rankk")
    assert rk[0] == (14400, 14400, 240) and rk[-1] == (14400, 14400, 14400) and len(rk) == 60
    fk = S.points("fixk")
    assert fk[0] == (240, 240, 1024) and fk[-1] == (19920, 19920, 1024) and len(fk) == 83
    assert S.points("rankk", limit=960) == [(14400, 14400, 240), (14400, 14400, 480), (14400, 14400, 720),
                                            (14400, 14400, 960)]
    v = S.variants(["222", "333"], [1], ["222:2"], ["abc", "naive"])
    assert [x[0] for x in v] == ["222-1_abc", "222-1_naive", "333-1_abc", "333-1_naive", "222-2_abc", "222-2_naive"]


def test_sweep_csv_roundtrip(fmm, tmp_path):
    from fmm_gen_amd import sweep as S
    rows = [(240, 240, 240, {"gemm": 100.0, "222-1_abc": 90.0}), (480, 480, 480, {"gemm": 200.0, "222-1_abc": 210.0})]
    path = str(tmp_path / "s.csv")
    S.write_csv(path, ["222-1_abc"], rows, ratios=True)
    back = S.read_csv(path)
    assert back[1]["m"] == 480 and back[1]["222-1_abc"] == 210.0 and back[1]["222-1_abc/gemm"] == 1.05
