"""Writes tests/golden/oracle_golden.npz: states of the CPU oracle on small cases.

Self-generated, not reference-held (the reference ships no expected outputs).  They
freeze the oracle: a CPU test re-runs it against them, and the GPU tests compare the HIP
path with them without running the oracle.  Regenerate only after a deliberate change
to the oracle:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from moses_2pdf_amd import scenarios as S  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402
from util import df004, hsiaolin  # noqa: E402

CASES = {
    "plane24": (lambda: S.landslide_plane(n=24), 20),
    "plane40_cfl04": (lambda: _knobs(S.landslide_plane(n=40)), 25),
    "fractal32_phi": (lambda: S.landslide_fractal(n=32, seed=11), 15),
    "df004": (df004, 30),
    "hsiaolin": (hsiaolin, 4),
}


def _knobs(sc):
    sc.cfl = 0.4
    sc.rho_ratio = 1000.0 / 2700.0
    return sc


def canonical(a: np.ndarray) -> np.ndarray:
    """-0.0 -> +0.0: the sign of a zero never reaches a nonzero value in this scheme, and
    implementations may differ in it."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    return np.where(a == 0.0, 0.0, a)


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(canonical(a).tobytes()).hexdigest()


def run(name):
    make, steps = CASES[name]
    o = Oracle(make())
    o.step(steps)
    u = o.get("u")[:, 3:-3, 3:-3]
    return u, o.scalar("time"), o.scalar("dt")


def main():
    out = {}
    for name in CASES:
        u, t, dt = run(name)
        out[f"{name}/sha256"] = np.array(digest(u))
        out[f"{name}/time"] = np.array(t)
        out[f"{name}/dt"] = np.array(dt)
        out[f"{name}/sum"] = np.array(u.reshape(7, -1).sum(axis=1))
        if u.size <= 40 * 40 * 7:
            out[f"{name}/u"] = u
        print(name, digest(u)[:16], t, dt)
    np.savez_compressed(os.path.join(HERE, "oracle_golden.npz"), **out)


if __name__ == "__main__":
    main()
