"""Digest of a final state, for runs too large to store as fixtures.

A digest fixes a state bit for bit (SHA-256 of the interior conserved
variables as little-endian float64, C order (4, nj, ni)) and, for a readable
failure, carries per-variable sums and a block-averaged density field.  The
same function digests the reference's dump when the fixture is written and
the GPU's state when it is checked.
"""
import hashlib

import numpy as np

COARSE = 16  # block size of the averaged density field


def sha256_f64(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def coarse_density(interior: np.ndarray, block: int = COARSE) -> np.ndarray:
    """Means of the density over block x block cells (edge blocks truncated)."""
    rho = np.asarray(interior[0], dtype=np.float64)
    nj, ni = rho.shape
    by, bx = -(-nj // block), -(-ni // block)
    out = np.empty((by, bx))
    for j in range(by):
        for i in range(bx):
            out[j, i] = rho[j * block:(j + 1) * block, i * block:(i + 1) * block].mean()
    return out


def state_digest(interior: np.ndarray) -> dict:
    """interior: conserved variables (4, nj, ni)."""
    a = np.asarray(interior, dtype=np.float64)
    return dict(sha256=sha256_f64(a),
                sums=[float(np.sum(a[k])) for k in range(4)],
                abs_sums=[float(np.sum(np.abs(a[k]))) for k in range(4)],
                max_abs=[float(np.max(np.abs(a[k]))) for k in range(4)])
