"""control.py -- generate and build one FMM job (the reference CLI).

    python control.py N alg1 L1 [alg2 L2 ...] [pack_type] [gen_path]

Same arguments as meta/control.py:120-142: N algorithms, each with a level
count, composed in command-line order; pack_type defaults to 'abc';
gen_path defaults to '../'.  The job directory is <gen_path>/<job> with
job = '<alg1>-<L1>_..._<pack_type>' (meta/control.py:47), holding the
generated HIP source, lib/lib<job>.so and test/test_<job>_{st,mt}.x.

Algorithms are the registry names of alg_dic (e.g. 222, 323, 333,
smirnov333_23_139).  Unlike the reference, the job directory does not need a
copy of the runtime sources next to gen_path, and nothing is built with make:
the HIP compiler is called directly.
"""
from __future__ import annotations

import os
import sys

if __package__ in (None, ""):
    # run as a script: import the package this file belongs to
    _root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, _root)
    from __graft_entry__ import load_package  # noqa: E402
    load_package()
    from fmm_gen_amd import codegen  # noqa: E402
    from fmm_gen_amd.plan import PACK_TYPES, build_algorithm, jobname  # noqa: E402
else:
    from . import codegen
    from .plan import PACK_TYPES, build_algorithm, jobname

USAGE = "USAGE: python control.py 2 m1,n1,p1 L1 m2,n2,p2, L2 pack_type gen_path"


def generateCode(alg_dims_mix, level_mix, pack_type="abc", path_prefix="../", compile=True):
    """Generate (and build) the job; returns the paths written.

    Name and argument order as meta/control.py:42.
    """
    if pack_type not in PACK_TYPES:
        raise ValueError(f"pack_type {pack_type!r} not supported (one of {PACK_TYPES})")
    if not path_prefix.endswith("/"):
        path_prefix += "/"
    job = jobname(alg_dims_mix, level_mix, pack_type)
    alg = build_algorithm(list(alg_dims_mix), [int(l) for l in level_mix])
    job_dir = os.path.join(path_prefix, job)
    out = codegen.generate_job(alg, job, pack_type, job_dir, compile=compile)
    out["job"] = job
    out["job_dir"] = job_dir
    return out


def parse_args(argv):
    try:
        num = int(argv[1])
        if num <= 0:
            raise ValueError
        algs, levels = [], []
        for i in range(num):
            algs.append(argv[2 + 2 * i])
            levels.append(int(argv[3 + 2 * i]))
        pack_type, gen_path = "abc", "../"
        if len(argv) > 2 + 2 * num:
            pack_type = argv[2 + 2 * num]
        if len(argv) > 3 + 2 * num:
            gen_path = argv[3 + 2 * num]
    except (IndexError, ValueError):
        raise SystemExit(USAGE)
    return algs, levels, pack_type, gen_path


def main(argv=None):
    argv = sys.argv if argv is None else argv
    algs, levels, pack_type, gen_path = parse_args(argv)
    out = generateCode(algs, levels, pack_type, gen_path)
    for key in ("source", "library", "driver_st", "driver_mt"):
        if key in out:
            print(f"{key}: {out[key]}")


if __name__ == "__main__":
    main()
