"""Algorithm registry: short name -> (table file, (mx, kx, nx)).

Same keys and the same file/partition pairs as meta/alg_dic.py:33-107.  The
tuple is (m-split, k-split, n-split): the CLI's "m,n,p" means <m,k,n>.
Paths are resolved against this package's ``algorithms/`` directory rather
than the reference's job-relative ``../meta/algorithms`` prefix, so the
registry works from any working directory.  The approximate / parameterised
entries the reference lists under ``ignore/`` are not carried (their tables
have symbolic coefficients the generated code never binds).
"""
from __future__ import annotations

import os
from typing import Dict, Tuple

from .coeffs import ALG_DIR

_FULL = {
    "fast323": ("grey323-15-103", (3, 2, 3)),
    "fast424": ("grey424-26-257", (4, 2, 4)),
    "classical222_8_24": ("classical222-8-24", (2, 2, 2)),
    "classical234_24_72": ("classical234-24-72", (2, 3, 4)),
    "classical333_27_81": ("classical333-27-81", (3, 3, 3)),
    "classical423_24_72": ("classical423-24-72", (4, 2, 3)),
    "fast423_130": ("fast423-130", (4, 2, 3)),
    "fast423_134": ("fast423-134", (4, 2, 3)),
    "fast423_138": ("fast423-138", (4, 2, 3)),
    "fast423_156": ("fast423-156", (4, 2, 3)),
    "grey234_20_144": ("grey234-20-144", (2, 3, 4)),
    "grey243_20_144": ("grey243-20-144", (2, 4, 3)),
    "grey252_18_99": ("grey252-18-99", (2, 5, 2)),
    "grey322_11_50": ("grey322-11-50", (3, 2, 2)),
    "grey323_15_103": ("grey323-15-103", (3, 2, 3)),
    "grey324_20_144": ("grey324-20-144", (3, 2, 4)),
    "grey332_15_103": ("grey332-15-103", (3, 3, 2)),
    "grey333_23_142": ("grey333-23-142", (3, 3, 3)),
    "grey333_23_143": ("grey333-23-143", (3, 3, 3)),
    "grey333_23_144": ("grey333-23-144", (3, 3, 3)),
    "grey333_23_152": ("grey333-23-152", (3, 3, 3)),
    "grey333_23_221": ("grey333-23-221", (3, 3, 3)),
    "grey342_20_144": ("grey342-20-144", (3, 4, 2)),
    "grey343_29_234": ("grey343-29-234", (3, 4, 3)),
    "grey422_14_84": ("grey422-14-84", (4, 2, 2)),
    "grey423_20_144": ("grey423-20-144", (4, 2, 3)),
    "grey424_26_257": ("grey424-26-257", (4, 2, 4)),
    "grey432_20_144": ("grey432-20-144", (4, 3, 2)),
    "grey433_29_234": ("grey433-29-234", (4, 3, 3)),
    "grey442_26_257": ("grey442-26-257", (4, 4, 2)),
    "grey522_18_99": ("grey522-18-99", (5, 2, 2)),
    "grey_strassen": ("grey-strassen", (2, 2, 2)),
    "hk323_15_94": ("hk323-15-94", (3, 2, 3)),
    "hk332_15_94": ("hk332-15-94", (3, 3, 2)),
    "madan232_11_48": ("madan232-11-48", (2, 3, 2)),
    "smirnov333_23_139": ("smirnov333-23-139", (3, 3, 3)),
    "smirnov336_40_960": ("smirnov336-40-960", (3, 3, 6)),
    "smirnov343_29_204": ("smirnov343-29-204", (3, 4, 3)),
    "smirnov353_36_280": ("smirnov353-36-280", (3, 5, 3)),
    "smirnov363_40_960": ("smirnov363-40-960", (3, 6, 3)),
    "smirnov633_40_960": ("smirnov633-40-960", (6, 3, 3)),
    "strassen": ("strassen", (2, 2, 2)),
}

# Short "pure partition" names (meta/alg_dic.py:84-106).
_SHORT = {
    "222": "strassen", "232": "madan232-11-48", "234": "grey234-20-144",
    "243": "grey243-20-144", "252": "grey252-18-99", "322": "grey322-11-50",
    "323": "hk323-15-94", "324": "grey324-20-144", "332": "hk332-15-94",
    "333": "smirnov333-23-139", "336": "smirnov336-40-960", "342": "grey342-20-144",
    "343": "smirnov343-29-204", "353": "smirnov353-36-280", "363": "smirnov363-40-960",
    "422": "grey422-14-84", "423": "fast423-138", "424": "grey424-26-257",
    "432": "grey432-20-144", "433": "grey433-29-234", "442": "grey442-26-257",
    "522": "grey522-18-99", "633": "smirnov633-40-960",
}

alg_dic: Dict[str, Tuple[str, Tuple[int, int, int]]] = {
    k: (os.path.join(ALG_DIR, f), d) for k, (f, d) in _FULL.items()
}
for _k, _f in _SHORT.items():
    alg_dic[_k] = (os.path.join(ALG_DIR, _f), (int(_k[0]), int(_k[1]), int(_k[2])))

# The 23 short names the reference's model and sweeps iterate over
# (meta/model_gen.py:37-61).
alg_set = list(_SHORT.keys())


def lookup(name: str) -> Tuple[str, Tuple[int, int, int]]:
    try:
        return alg_dic[name]
    except KeyError:
        raise KeyError(f"unknown algorithm {name!r}; known: {sorted(alg_dic)}") from None
