#!/usr/bin/env python3
"""Write tests/golden/sph_files/in.{rtp,rtm,rlm,rj}: a one-rank set of the
reference's spherical grid / mode files for L=5, Nr=3, Ntheta=9, Nphi=18 on
the standard shell radii, written by the reference's own writers with its
own ordering routines (oracle/ref_sph_files.f90, compiled by
oracle/Makefile.ref; built only where the reference tree is present).

Usage: python tests/golden/make_sph_files.py
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from kemorin_mhd_amd.sht import ShellConfig  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sph_files")
SHAPE = ShellConfig(5, 3, 9, 18)
SPACES = ("rtp", "rtm", "rlm", "rj")


def paths(d=OUT):
    return [os.path.join(d, f"in.{s}") for s in SPACES]


def main():
    if not ref.build():
        sys.exit("oracle/_ref is not built (needs the reference tree)")
    os.makedirs(OUT, exist_ok=True)
    c = SHAPE
    ref.write_sph_files(c.L, c.nr, c.nth, c.nphi, c.radius(), paths())
    for p in paths():
        print(f"wrote {p} ({os.path.getsize(p) / 1e3:.1f} kB)")


if __name__ == "__main__":
    main()
