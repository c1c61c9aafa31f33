"""Write legacy_ptmc_ising20.txt: columns 1-5 (T, E_av, E_err, C, M) of the
reference's own parallel-tempering Monte Carlo test output
tests/serial/PTMC/EM.dat (data only; the run's input is tests/serial/PTMC/inp:
20x20 square lattice, periodic, `ising` sampling, J_1 = 0.5 meV, c_Ji = -1,
D_ani 0.005d-4 along the easy axis z)."""
import os
import sys

import numpy as np

SRC = "/root/reference/tests/serial/PTMC/EM.dat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "legacy_ptmc_ising20.txt")


def main():
    ref = np.loadtxt(SRC)
    with open(OUT, "w") as f:
        f.write("# columns 1-5 (T, E_av, E_err, C, M) of tests/serial/PTMC/EM.dat: 20x20 square lattice, periodic, "
                "Ising sampling, J_1 = 0.5 meV (c_Ji = -1), on-site term +5e-7 eV for spins along z, "
                "40 temperatures 1..61 K (legacy reference run; data only)\n")
        for r in ref:
            f.write(" ".join("%.10e" % x for x in r[:5]) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
