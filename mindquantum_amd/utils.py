"""Small helpers (mindquantum/utils/f.py ket_string)."""
import numpy as np


def ket_string(state, tol=1e-7):
    """'amplitude¦bits⟩' strings of the non-negligible amplitudes (little-endian bits)."""
    n = int(np.log2(len(state)))
    out = []
    for i, v in enumerate(state):
        if abs(v) > tol:
            out.append(f"{np.round(v, 8)}¦{format(i, f'0{n}b')}⟩")
    return out
