"""Writes the golden fixtures under tests/golden/.

eigh3.npz: 200 random symmetric 3x3 matrices (seeded; spread over scales
1e-6..1e6, including near-degenerate and rank-deficient cases) with their
eigenvalues from numpy.linalg.eigh (LAPACK dsyevd), the independent
reference for the FP64 eigen solve of the FTLE stage.

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261203)
    mats = []
    for i in range(200):
        kind = i % 5
        if kind == 0:  # random symmetric, random scale
            a = rng.normal(size=(3, 3)) * 10.0 ** rng.uniform(-6, 6)
            m = a + a.T
        elif kind == 1:  # Cauchy-Green like: J^T J with stretched J
            j = rng.normal(size=(3, 3)) @ np.diag(np.exp(rng.uniform(-8, 8, size=3)))
            m = j.T @ j
        elif kind == 2:  # near-degenerate: rotated diag with two close eigenvalues
            q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            d = np.array([1.0, 1.0 + 1e-9 * rng.uniform(), rng.uniform(2, 5)])
            m = q @ np.diag(d) @ q.T
        elif kind == 3:  # rank deficient PSD
            v = rng.normal(size=(3, 2))
            m = v @ v.T
        else:  # diagonal / rotated diagonal with known spectrum
            q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            d = np.sort(rng.uniform(-3, 3, size=3))
            m = q @ np.diag(d) @ q.T
        m = 0.5 * (m + m.T)
        mats.append(m)
    mats = np.array(mats, np.float64)
    ev = np.linalg.eigh(mats)[0]
    np.savez(os.path.join(HERE, "eigh3.npz"), matrices=mats, eigenvalues=ev)
    print("wrote eigh3.npz", mats.shape)


if __name__ == "__main__":
    main()
