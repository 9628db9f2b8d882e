"""Julia's isapprox for vectors, as the reference's tests use it:
x == y or norm(x - y) <= max(atol, rtol * max(norm(x), norm(y)))."""
import numpy as np


def isapprox(x, y, rtol=None, atol=0.0):
    x, y = np.asarray(x, dtype=float), np.asarray(y, dtype=float)
    if rtol is None:
        rtol = 0.0 if atol > 0 else np.sqrt(np.finfo(float).eps)
    if np.array_equal(x, y):
        return True
    return np.linalg.norm(x - y) <= max(atol, rtol * max(np.linalg.norm(x), np.linalg.norm(y)))
