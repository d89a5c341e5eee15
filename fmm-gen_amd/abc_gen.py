"""abc_gen.py -- standalone emitter for custom coefficient tables.

    python abc_gen.py N coeff_file1 mx1,kx1,nx1 L1 [... coeff_fileN dimsN LN] [out_file] [lib_file]
                      [--pack abc|ab|naive]

The coefficient-file form of generation (meta/abc_gen_mix.py:474-502, and
meta/ab_gen_mix.py / meta/naive_gen_mix.py with the same arguments; examples
in meta/exec.sh:2-8): N table files, each with its partition <mx,kx,nx> and
level count, composed in command-line order -- every file's table applied L
times -- into one algorithm.  Unlike control.py, the tables need not be in
the registry (alg_dic).  The tables are checked exactly (the Brent equations)
before anything is written.

Writes the HIP source of the job (out_file, default 'a.hip'): the composed
tables, the product equations and the reference's per-job entry
bl_dgemm_strassen_<pack> over the shared gfx950 kernels.  The reference's
two further arguments name the files its micro-kernel and kernel header go
to; here the kernels are the library's headers, so the next argument is
instead an optional lib_file: when given, the source is compiled into that
shared library (the same C-ABI as control.py's job libraries).
"""
from __future__ import annotations

import os
import sys
from typing import List, Optional, Sequence, Tuple

if __package__ in (None, ""):
    _root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, _root)
    from __graft_entry__ import load_package  # noqa: E402
    load_package()
    from fmm_gen_amd import build as B  # noqa: E402
    from fmm_gen_amd import codegen  # noqa: E402
    from fmm_gen_amd.coeffs import Algorithm, brent_residual, compose, load_algorithm  # noqa: E402
    from fmm_gen_amd.plan import PACK_TYPES  # noqa: E402
else:
    from . import build as B
    from . import codegen
    from .coeffs import Algorithm, brent_residual, compose, load_algorithm
    from .plan import PACK_TYPES

USAGE = ("USAGE: python abc_gen.py 2 coeff_file1 m1,n1,p1 L1 coeff_file2 m2,n2,p2 L2 out_file [lib_file] "
         "[--pack abc|ab|naive]")


def parse_args(argv: Sequence[str]) -> Tuple[List[str], List[Tuple[int, int, int]], List[int], str,
                                             Optional[str], str]:
    """argv (program name first) -> files, dims, levels, out_file, lib_file,
    pack type; positional order as meta/abc_gen_mix.py:474-498."""
    args = list(argv[1:])
    pack = "abc"
    if "--pack" in args:
        i = args.index("--pack")
        if i + 1 >= len(args):
            raise SystemExit(USAGE)
        pack = args[i + 1]
        del args[i:i + 2]
    if pack not in PACK_TYPES:
        raise SystemExit(f"pack type {pack!r} not one of {PACK_TYPES}\n{USAGE}")
    try:
        num = int(args[0])
        if num <= 0:
            raise ValueError
        files, dims, levels = [], [], []
        for i in range(num):
            files.append(args[1 + 3 * i])
            d = tuple(int(x) for x in args[2 + 3 * i].split(","))
            if len(d) != 3 or min(d) <= 0:
                raise ValueError
            dims.append(d)
            levels.append(int(args[3 + 3 * i]))
            if levels[-1] < 0:
                raise ValueError
    except (IndexError, ValueError):
        raise SystemExit(USAGE)
    out_file = args[1 + 3 * num] if len(args) > 1 + 3 * num else "a.hip"
    lib_file = args[2 + 3 * num] if len(args) > 2 + 3 * num else None
    return files, dims, levels, out_file, lib_file, pack


def build_custom(files: Sequence[str], dims: Sequence[Tuple[int, int, int]], levels: Sequence[int]) -> Algorithm:
    """Each file's table applied L times, in order (the composition of
    meta/abc_gen_mix.py:407-431); exact Brent check of every table and of
    the composition."""
    per_level = []
    for f, d, L in zip(files, dims, levels):
        base = load_algorithm(f, d)
        bad = brent_residual(base)
        if bad:
            raise ValueError(f"{f}: not an exact <{d[0]},{d[1]},{d[2]}> algorithm ({bad} Brent equations fail)")
        per_level += [base] * L
    if not per_level:
        raise ValueError("all level counts are zero")
    name = "_".join(f"{os.path.basename(f)}-{L}" for f, L in zip(files, levels))
    alg = compose(per_level, name=name)
    if len(per_level) > 1 and brent_residual(alg):
        raise ValueError(f"{name}: the composition fails the Brent equations")
    return alg


def gen_fmm(files, dims, levels, out_file: str, lib_file: Optional[str] = None, pack: str = "abc") -> dict:
    """Write the job source (and build the library when lib_file is given)."""
    alg = build_custom(files, dims, levels)
    job = alg.name.replace(".", "_") + "_" + pack
    os.makedirs(os.path.dirname(os.path.abspath(out_file)), exist_ok=True)
    with open(out_file, "w") as fh:
        fh.write(codegen.emit_source(alg, job, pack))
    out = {"source": out_file, "job": job, "rank": alg.rank, "dims": alg.dims}
    if lib_file:
        out["library"] = B.build_shared(out_file, lib_file)
    return out


def main(argv=None) -> None:
    argv = sys.argv if argv is None else argv
    files, dims, levels, out_file, lib_file, pack = parse_args(argv)
    out = gen_fmm(files, dims, levels, out_file, lib_file, pack)
    for key in ("source", "library"):
        if key in out:
            print(f"{key}: {out[key]}")


if __name__ == "__main__":
    main()
