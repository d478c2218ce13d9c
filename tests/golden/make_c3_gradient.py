"""Write tests/golden/c3_gradient.json: the oracle's nuclear gradient of
E = 1/2 D J - 1/4 D K for the 20-atom C8H12 cluster (DZ basis with d on H,
libcint_amd.synthetic 'c3') and the density random_density(nao, seed=5),
from int2e_ip1_sph blocks of every (i, j, k >= l) shell tuple contracted
with the density on the CPU (oracle.grad_ip1; tabulated Rys roots).

    python tests/golden/make_c3_gradient.py [--threads N]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SEED = 5


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    import oracle
    from libcint_amd.synthetic import make_molecule, random_density
    oracle.build()
    m = make_molecule("c3")
    dm = random_density(m.nao_sph(), seed=SEED)
    oracle.set_root_mode(oracle.ROOTS_TABLE)
    t = time.time()
    g = oracle.grad_ip1(oracle.Basis(m.atm, m.bas, m.env), dm, a.threads)
    dt = time.time() - t
    res = {"molecule": m.name, "natm": m.natm, "nbas": m.nbas, "nao": m.nao_sph(),
           "density": f"random_density(nao, seed={SEED})", "grad": g.tolist(),
           "translational_residual": float(abs(g.sum(0)).max()), "oracle_seconds": dt}
    with open(os.path.join(HERE, "c3_gradient.json"), "w") as fh:
        json.dump(res, fh, indent=1)
    print(f"wrote c3_gradient.json ({dt:.0f} s), residual {res['translational_residual']:.2e}")


if __name__ == "__main__":
    main()
