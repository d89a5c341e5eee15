"""The standalone emitter for custom coefficient tables (abc_gen.py), CPU only:
the reference's argument form (meta/abc_gen_mix.py:474-502), composition of
table files in command-line order, exact checking of the tables, and the
emitted job source."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ALG_DIR = os.path.join(ROOT, "fmm-gen_amd", "algorithms")
GEN = os.path.join(ROOT, "fmm-gen_amd", "abc_gen.py")


def test_parse_reference_argument_form(fmm):
    from fmm_gen_amd import abc_gen
    argv = ["abc_gen.py", "2", "algorithms/strassen", "2,2,2", "1", "algorithms/smirnov333-23-139", "3,3,3", "1",
            "my_dgemm_strassen_abc.hip"]
    files, dims, levels, out, lib, pack = abc_gen.parse_args(argv)
    assert files == ["algorithms/strassen", "algorithms/smirnov333-23-139"]
    assert dims == [(2, 2, 2), (3, 3, 3)] and levels == [1, 1]
    assert out == "my_dgemm_strassen_abc.hip" and lib is None and pack == "abc"
    one = ["abc_gen.py", "1", "algorithms/strassen", "2,2,2", "2"]
    files, dims, levels, out, lib, pack = abc_gen.parse_args(one + ["--pack", "naive"])
    assert levels == [2] and out == "a.hip" and pack == "naive"
    assert abc_gen.parse_args(one + ["x.hip", "libx.so"])[3:5] == ("x.hip", "libx.so")
    for bad in (["abc_gen.py"], ["abc_gen.py", "1", "x"], ["abc_gen.py", "1", "x", "2,2", "1"],
                ["abc_gen.py", "1", "x", "2,2,2", "-1"], argv[:5] + ["--pack", "abcd"]):
        with pytest.raises(SystemExit):
            abc_gen.parse_args(bad)


def test_custom_composition_equals_registry(fmm):
    """Table files composed by the emitter equal the registry composition of
    control.py (same rows, same order) and satisfy the Brent equations."""
    from fmm_gen_amd import abc_gen
    alg = abc_gen.build_custom([os.path.join(ALG_DIR, "strassen"), os.path.join(ALG_DIR, "hk323-15-94")],
                               [(2, 2, 2), (3, 2, 3)], [1, 1])
    ref = fmm.build_algorithm(["222", "323"], [1, 1])
    assert alg.dims == ref.dims == (6, 4, 6) and alg.rank == 105
    for a, b in zip(alg.as_arrays(), ref.as_arrays()):
        assert np.array_equal(a, b)
    two = abc_gen.build_custom([os.path.join(ALG_DIR, "strassen")], [(2, 2, 2)], [2])
    assert two.dims == (4, 4, 4) and two.rank == 49


def test_planted_table_mistake_is_rejected(fmm, tmp_path):
    from fmm_gen_amd import abc_gen
    lines = open(os.path.join(ALG_DIR, "strassen")).read().splitlines()
    # flip the sign of the first nonzero coefficient of the first U row
    for i, ln in enumerate(lines):
        toks = ln.split()
        if toks and not ln.startswith("#"):
            j = next(j for j, t in enumerate(toks) if t not in ("0", "0.0"))
            toks[j] = toks[j][1:] if toks[j].startswith("-") else "-" + toks[j]
            lines[i] = " ".join(toks)
            break
    bad = tmp_path / "strassen_bad"
    bad.write_text("\n".join(lines) + "\n")
    with pytest.raises(ValueError, match="Brent"):
        abc_gen.build_custom([str(bad)], [(2, 2, 2)], [1])
    with pytest.raises(Exception):
        abc_gen.build_custom([os.path.join(ALG_DIR, "strassen")], [(3, 3, 3)], [1])   # wrong partition


def test_cli_writes_job_source(fmm, tmp_path):
    out = tmp_path / "my_dgemm_strassen_abc.hip"
    res = subprocess.run([sys.executable, GEN, "2", os.path.join(ALG_DIR, "strassen"), "2,2,2", "1",
                          os.path.join(ALG_DIR, "smirnov333-23-139"), "3,3,3", "1", str(out)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    src = out.read_text()
    assert "bl_dgemm_strassen_abc" in src and "kRank = 161" in src and "<6,6,6>" in src
    assert src.count("M160 =") == 1 and "M161 =" not in src
    res = subprocess.run([sys.executable, GEN, "1", os.path.join(ALG_DIR, "strassen"), "2,2,2", "1", str(out),
                          "--pack", "ab"], capture_output=True, text=True)
    assert res.returncode == 0 and "bl_dgemm_strassen_ab(" in out.read_text()
