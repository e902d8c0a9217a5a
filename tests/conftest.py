"""Shared fixtures.  GPU tests carry ``@pytest.mark.gpu`` and run only on an MI355X box."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libmqamd.so")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "reference_values.json")) as f:
        return json.load(f)


def random_circuit(n, depth, seed, kinds=None, ctrl_prob=0.2, param_names=True):
    """Seeded random circuit over every gate kind of the backend (controls included)."""
    import mindquantum_amd as mq
    from mindquantum_amd.core import gates as G

    rng = np.random.default_rng(seed)
    kinds = kinds or ["X", "Y", "Z", "H", "S", "T", "SX", "SWAP", "ISWAP", "RX", "RY", "RZ", "PS", "GP", "Rxx",
                      "Ryy", "Rzz", "MAT1", "MAT2", "MAT3", "Sdag", "CNOT"]
    circ = mq.Circuit()
    pidx = 0
    for _ in range(depth):
        k = kinds[rng.integers(len(kinds))]
        two = k in ("SWAP", "ISWAP", "Rxx", "Ryy", "Rzz", "MAT2", "CNOT")
        three = k == "MAT3"
        need = 3 if three else (2 if two else 1)
        if need > n:
            continue
        qs = [int(q) for q in rng.permutation(n)]
        obj = qs[:need]
        rest = qs[need:]
        ctrl = []
        if rest and rng.random() < ctrl_prob:
            ctrl = rest[: int(rng.integers(1, min(2, len(rest)) + 1))]
        if k in ("RX", "RY", "RZ", "PS", "GP", "Rxx", "Ryy", "Rzz"):
            cls = {"RX": G.RX, "RY": G.RY, "RZ": G.RZ, "PS": G.PhaseShift, "GP": G.GlobalPhase, "Rxx": G.Rxx,
                   "Ryy": G.Ryy, "Rzz": G.Rzz}[k]
            if param_names and rng.random() < 0.8:
                pr = {f"p{pidx}": float(rng.uniform(0.5, 1.5))}
                if rng.random() < 0.3 and pidx > 0:
                    pr[f"p{int(rng.integers(pidx))}"] = float(rng.uniform(-1, 1))
                pidx += 1
                g = cls(mq.ParameterResolver(pr, const=float(rng.uniform(-0.3, 0.3))))
            else:
                g = cls(float(rng.uniform(0, 2 * np.pi)))
        elif k.startswith("MAT"):
            m = need
            a = rng.normal(size=(2**m, 2**m)) + 1j * rng.normal(size=(2**m, 2**m))
            q, _ = np.linalg.qr(a)
            g = G.UnivMathGate(f"U{m}", q)
        elif k == "CNOT":
            g = G.X
            obj, ctrl = obj[:1], [obj[1]] + ctrl
        elif k == "Sdag":
            g = G.S.hermitian()
        else:
            g = getattr(G, k)
        circ += g.on(obj, ctrl)
    return circ


def random_hamiltonian(n, n_terms, seed, max_weight=3):
    import mindquantum_amd as mq

    rng = np.random.default_rng(seed)
    op = mq.QubitOperator("", float(rng.normal()))
    for _ in range(n_terms):
        w = int(rng.integers(1, min(max_weight, n) + 1))
        qs = rng.choice(n, size=w, replace=False)
        s = " ".join(f"{'XYZ'[rng.integers(3)]}{q}" for q in qs)
        op += mq.QubitOperator(s, float(rng.normal()))
    return op


def hea(n, layers, seed=0, encoder=False):
    """Hardware-efficient ansatz: layers of RY, RZ on every qubit and a CNOT ladder."""
    import mindquantum_amd as mq

    circ = mq.Circuit()
    if encoder:
        enc = mq.Circuit()
        for q in range(n):
            enc.rx(f"x{q}", q)
        enc.as_encoder()
        circ += enc
    for l in range(layers):
        for q in range(n):
            circ.ry(f"a{l}_{q}", q)
        for q in range(n):
            circ.rz(f"b{l}_{q}", q)
        for q in range(n - 1):
            circ.x(q + 1, q)
    return circ


def params_for(circ, seed):
    rng = np.random.default_rng(seed)
    names = list(circ.params_name)
    return names, rng.uniform(0, 2 * np.pi, size=len(names))
