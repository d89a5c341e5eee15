"""CPU tests of the autotuner's cost model (fmm_gen_amd/autotune.py)."""
import numpy as np

from fmm_gen_amd import autotune as AT


def test_gemm_features_count_tiles():
    ac = AT.alg_counts(None)
    f = AT.unit_features(300, 200, 50, ac, 256, 128)
    assert f[0] == 2 * 2
    assert f[1] == 4 * 50 * 256 * 128
    assert f[2] == 0                      # one-term operands: no extra loads
    assert f[3] == 4 * 256 * 128


def test_fmm_features_follow_job_split():
    ac = AT.alg_counts((["222"], [1]))
    assert ac.rank == 7 and sum(ac.nu) == 12 and sum(ac.nv) == 12 and sum(ac.nw) == 12
    m, n, k = 2 * 8 * 20 + 5, 2 * 4 * 40 + 3, 2 * 30 + 1   # all three fringes
    bm, bn = 256, 128
    f = AT.unit_features(m, n, k, ac, bm, bn)
    md, nd, kd = m - 5, n - 3, k - 1
    ms, ns, ks = md // 2, nd // 2, kd // 2
    tl = AT._tiles(ms, bm) * AT._tiles(ns, bn)
    units = (AT._tiles(m, bm) * AT._tiles(3, bn) + AT._tiles(5, bm) * AT._tiles(nd, bn)
             + 7 * tl + 4 * tl)
    assert f[0] == units
    flops = (AT._tiles(m, bm) * AT._tiles(3, bn) * k + AT._tiles(5, bm) * AT._tiles(nd, bn) * k
             + 7 * tl * ks + 4 * tl * 1) * bm * bn
    assert f[1] == flops
    assert f[2] == tl * ks * ((12 - 7) * bm + (12 - 7) * bn)
    assert f[3] == (AT._tiles(m, bm) * AT._tiles(3, bn) + AT._tiles(5, bm) * AT._tiles(nd, bn) + 12 * tl
                    + 4 * tl) * bm * bn


def test_choose_prefers_cheapest():
    t = AT.Tuner("f32")
    t.model.theta = np.array([0.0, 1.0, 0.0, 0.0])     # MFMA work only: fewest flops wins
    a, tile, _ = t.choose(8192, 8192, 8192)
    assert a == "222L2"
    t.model.theta = np.array([1.0, 0.0, 0.0, 0.0])     # unit count only: plain GEMM wins
    assert t.choose(8192, 8192, 8192)[0] == "gemm"


def test_calibrate_recovers_constants():
    t = AT.Tuner("f32")
    true = np.array([3e-6, 1.0 / 0.5e12, 4.0 / 120e9, 8.0 / 60e9])

    def device_time(m, n, k, alg, tile, copies):
        return float(t.features(m, n, k, alg, tile) @ true) * copies / AT.CUS
    got = t.calibrate(device_time)
    assert np.allclose(got, true, rtol=1e-6), (got, true)
    assert t.model.calibrated
