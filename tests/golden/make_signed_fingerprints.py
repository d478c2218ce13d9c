"""Write tests/golden/signed_fingerprints.json: signed fingerprints of the
oracle's integrals over libcint's test-suite fixture
(testsuite/test_cint.py:46-136, tests/golden/libcint_testsuite.json).

For every quartet of the fixture's (i<=j, k<=l) loop (test_cint.py:235-256)
the block x (Fortran order, components last) contributes sum_k x_k cos(k);
contributions are summed per angular-momentum class (l_i, l_j, l_k, l_l).
Unlike libcint's sum |x| checksums these change when any row of the
spherical transform, any component or any derivative direction changes sign.
The oracle runs with the exact (quad-precision) Rys rule.

    python tests/golden/make_signed_fingerprints.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

INTORS = ("int2e_sph", "int2e_cart", "int2e_ip1_sph", "int3c2e_sph")


def main():
    import oracle
    from conftest import class_fingerprints, fixture_quartets
    oracle.build()
    with open(os.path.join(HERE, "libcint_testsuite.json")) as fh:
        g = json.load(fh)
    atm = np.array(g["atm"], np.int32)
    bas = np.array(g["bas"], np.int32)
    env = np.array(g["env"])
    oracle.set_root_mode(oracle.ROOTS_EXACT)
    res = {"about": "sum_k x_k cos(k) per class over the libcint test-suite fixture, oracle "
                    "with exact roots"}
    for name in INTORS:
        q = fixture_quartets(name)
        out, off, _ = oracle.batch(name, q, oracle.Basis(atm, bas, env))
        res[name] = class_fingerprints(bas, q, out, off, name)[0]
        print(name, len(res[name]), "classes")
    with open(os.path.join(HERE, "signed_fingerprints.json"), "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
