"""Write the golden fixtures of libcint's own test suite as JSON data.

The arrays reproduce the atm/bas/env fixture that testsuite/test_cint.py:46-136
and testsuite/test_3c2e.py:46-158 build (4 atoms, shells l=1..4 plus their
l-1 copies with rescaled first coefficients, a 2-primitive/2-contraction d
shell, and the exponent-0 s shell used by the 3c2e invariant).  The golden
numbers are the ones those scripts assert:
  test_cint.py:478-481      cint2e_sph 56243.88080655417, cint2e_ip1_sph 115489.8647398112
  test_3c2e.py:303-307      cint3c2e_sph 1586.350797347553
  test_rys_roots.py:64-65   Boys F_0..2 at x=1 and x=7
  test_rys_roots.py:72-85   fingerprints fp(u, w) of Rys rules
Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import math
import os

PTR_ENV_START = 20
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF = range(7)


def build():
    natm = 4
    atm = [[0] * 6 for _ in range(natm)]
    bas = [[0] * 8 for _ in range(10)]
    env = [0.0] * 200
    off = PTR_ENV_START
    for i in range(natm):
        atm[i][0] = (i + 1) * 2
        atm[i][1] = off
        env[off + 0] = .2 * (i + 1)
        env[off + 1] = .3 + (i + 1) * .5
        env[off + 2] = .1 - (i + 1) * .5
        off += 3
    off0 = off
    # (atom, l, kappa, nprim, nctr, exps, coefs)
    shells = [
        (0, 1, 1, 1, 1, [1.0], [1.0]),
        (1, 2, 2, 2, 2, [5.0, 3.0], [1.0, 2.0, 4.0, 1.0]),
        (2, 3, 3, 1, 1, [1.0], [1.0]),
        (3, 4, 4, 1, 1, [0.5], [1.0]),
    ]
    nh = 0
    for ia, l, kappa, npr, nct, ex, cf in shells:
        b = bas[nh]
        b[ATOM_OF], b[ANG_OF], b[KAPPA_OF], b[NPRIM_OF], b[NCTR_OF] = ia, l, kappa, npr, nct
        b[PTR_EXP] = off
        env[off:off + npr] = ex
        b[PTR_COEFF] = off + npr
        env[off + npr:off + npr + npr * nct] = cf
        off += npr + npr * nct
        nh += 1
    n = off - off0
    for i in range(n):
        env[off + i] = env[off0 + i]
    for i in range(nh):
        src, dst = bas[i], bas[i + nh]
        dst[ATOM_OF] = src[ATOM_OF]
        dst[ANG_OF] = src[ANG_OF] - 1
        dst[KAPPA_OF] = -src[KAPPA_OF]
        dst[NPRIM_OF] = src[NPRIM_OF]
        dst[NCTR_OF] = src[NCTR_OF]
        dst[PTR_EXP] = src[PTR_EXP] + n
        dst[PTR_COEFF] = src[PTR_COEFF] + n
        env[dst[PTR_COEFF]] /= 2 * env[src[PTR_EXP]]
    c5, c1, e1 = bas[5][PTR_COEFF], bas[1][PTR_COEFF], bas[1][PTR_EXP]
    env[c5 + 0] = env[c1 + 0] / (2 * env[e1 + 0])
    env[c5 + 1] = env[c1 + 1] / (2 * env[e1 + 1])
    env[c5 + 2] = env[c1 + 2] / (2 * env[e1 + 0])
    env[c5 + 3] = env[c1 + 3] / (2 * env[e1 + 1])
    # exponent-0 s shell of test_3c2e.py:128-140, index 8
    off += n
    s0 = bas[2 * nh]
    s0[ATOM_OF], s0[ANG_OF], s0[NPRIM_OF], s0[NCTR_OF] = 0, 0, 1, 1
    s0[PTR_EXP] = off
    env[off] = 0.0
    off += 1
    s0[PTR_COEFF] = off
    env[off] = 2 * math.sqrt(math.pi)
    off += 1
    return atm, bas[: 2 * nh + 1], env[:off]


def main():
    atm, bas, env = build()
    data = {
        "source": "libcint testsuite/test_cint.py:46-136, test_3c2e.py:46-158",
        "atm": atm,
        "bas": bas,
        "env": env,
        "nbas_regular": 8,
        "s0_shell": 8,
        "golden": {
            "cint2e_sph": [56243.88080655417, 8],
            "cint2e_ip1_sph": [115489.8647398112, 8],
            "cint3c2e_sph": [1586.350797347553, 10],
            "boys_x1": [0.746824132812427, 0.18947234582049235, 0.10026879814501737],
            "boys_x7": [0.3349010581765593, 0.023856369729357483, 0.005046944801608424],
            "fp_rys": [[1, 0.3, 0.93475892362307533], [2, 0.8, 0.94983728587499205],
                       [3, 0.5, -2.7165192846520481], [4, 0.5, -11.442155925744938]],
            "fp_rys6": [0.4, 2.898038115730877, 0.1193288112731978],
        },
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcint_testsuite.json")
    with open(out, "w") as fh:
        json.dump(data, fh, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
