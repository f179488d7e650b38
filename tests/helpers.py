"""Shared test helpers: fixture loading and error measures."""
import os

import numpy as np

from salmon_amd.synth import System

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_case(name):
    """System and reference outputs of a golden fixture (inputs are stored whole)."""
    d = dict(np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False))
    NL3, aL = d["NL3"], d["aL"]
    s = System(NLx=int(NL3[0]), NLy=int(NL3[1]), NLz=int(NL3[2]), aLx=float(aL[0]), aLy=float(aL[1]),
               aLz=float(aL[2]), lapx=d["lapx"], lapy=d["lapy"], lapz=d["lapz"], nabx=d["nabx"], naby=d["naby"],
               nabz=d["nabz"], vloc=d["vloc"], Rion=d["Rion"], rc=float(d["rc"]), Mps=d["Mps"], Jxyz=d["Jxyz"],
               Jxx=d["Jxx"], Jyy=d["Jyy"], Jzz=d["Jzz"], a_tbl=d["a_tbl"], uV=d["uV"], iuV=d["iuV"],
               kAc0=d["kAc0"], occ=d["occ"], zu=d["zu"], dt=float(d["dt"]), name=name)
    return s, d


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
