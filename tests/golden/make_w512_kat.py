"""Writes tests/golden/w512_ewald_kat.npz from the reference's own example
data: the 512-molecule q-TIP4P/F configuration
examples/q-TIP4P-F/W512-initial.xyz (positions in Angstrom, box 24.832 A
from examples/q-TIP4P-F/build.py:11) and the reciprocal-space energies that
examples/q-TIP4P-F/Ewald-log.txt:1-36 records for it (Hartree, alpha*L = 10).

Run in a container that has the reference tree mounted; the .npz it writes
is the committed fixture.
"""
import os
import sys

import numpy as np

REF = "/root/reference/examples/q-TIP4P-F/W512-initial.xyz"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "w512_ewald_kat.npz")


def main():
    with open(REF) as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    names, pos = [], []
    for ln in lines[2:2 + n]:
        p = ln.split()
        names.append(p[0])
        pos.append([float(v) for v in p[1:4]])
    pos = np.array(pos)
    # (label, use_pme, kmax_ewd, kmax_pme, order, energy) -- Ewald-log.txt lines 8-36
    kat = [("ewald_k13", 0, 13, 0, 0, 8.721399599052),
           ("ewald_k26", 0, 26, 0, 0, 8.721400121222),
           ("pme8_k17", 1, 13, 17, 8, 8.721395400508),
           ("pme8_k22", 1, 13, 22, 8, 8.721398904468),
           ("pme8_k25", 1, 13, 25, 8, 8.721399505611)]
    np.savez(OUT, names=np.array(names), pos_A=pos, box_A=24.832, ewald_alpha=10.0,
             labels=np.array([k[0] for k in kat]), use_pme=np.array([k[1] for k in kat]),
             kmax_ewd=np.array([k[2] for k in kat]), kmax_pme=np.array([k[3] for k in kat]),
             order=np.array([k[4] for k in kat]), energy=np.array([k[5] for k in kat]))
    print("wrote", OUT, pos.shape)


if __name__ == "__main__":
    sys.exit(main())
