"""TEST INFRASTRUCTURE ONLY: dense numpy simulation for small registers (n <= 10).

Independent of both the HIP library and the C++ oracle: every gate is embedded into a full
2^n x 2^n matrix from its closed-form numpy matrix (mindquantum_amd.core.gates, which follow
mindquantum/core/gates/basicgate.py), Hamiltonians are built from Kronecker products of Pauli
matrices, and gradients use the direct product formula

    dE/dtheta_j = sum_k c_kj * 2 Re <psi| U^dag H U_{>k} dG_k U_{<k} |psi>

(no adjoint sweep), so a sign, ordering or jacobian mistake in either simulator shows up here.
"""
import numpy as np

_PAULI = {
    "I": np.eye(2, dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def embed(m, objs, ctrls, n):
    """Full matrix of a gate with target matrix ``m`` (bit b of its index <=> objs[b])."""
    dim = 1 << n
    full = np.zeros((dim, dim), dtype=np.complex128)
    cmask = sum(1 << c for c in ctrls)
    omask = sum(1 << q for q in objs)
    k = len(objs)
    for i in range(dim):
        if (i & cmask) != cmask:
            full[i, i] = 1
            continue
        base = i & ~omask
        col = sum(((i >> q) & 1) << b for b, q in enumerate(objs))
        for row in range(1 << k):
            j = base | sum(((row >> b) & 1) << q for b, q in enumerate(objs))
            full[j, i] = m[row, col]
    return full


def embed_diff(dm, objs, ctrls, n):
    """Derivative of a controlled gate: zero outside the control subspace."""
    dim = 1 << n
    full = np.zeros((dim, dim), dtype=np.complex128)
    cmask = sum(1 << c for c in ctrls)
    omask = sum(1 << q for q in objs)
    k = len(objs)
    for i in range(dim):
        if (i & cmask) != cmask:
            continue
        base = i & ~omask
        col = sum(((i >> q) & 1) << b for b, q in enumerate(objs))
        for row in range(1 << k):
            j = base | sum(((row >> b) & 1) << q for b, q in enumerate(objs))
            full[j, i] = dm[row, col]
    return full


def gate_matrix(g, values):
    """Closed-form matrix of the gate itself (never its lowering into other gates)."""
    if g.param_resolvers():
        return g.matrix(values)
    return g.matrix()


def _acting(circ):
    return [g for g in circ if g.obj_qubits and getattr(g, "name", "") != "BARRIER"]


def unitary(circ, values, n):
    u = np.eye(1 << n, dtype=np.complex128)
    for g in _acting(circ):
        u = embed(gate_matrix(g, values), g.obj_qubits, g.ctrl_qubits, n) @ u
    return u


def state(circ, values, n, init=None):
    psi = np.zeros(1 << n, dtype=np.complex128)
    if init is None:
        psi[0] = 1
    else:
        psi[:] = init
    return unitary(circ, values, n) @ psi


def hamiltonian_matrix(qubit_operator, n):
    """Sum of coeff * (kron of Pauli matrices), qubit 0 as the least significant factor; a matrix
    (numpy or scipy sparse, 2^n x 2^n) is returned as a dense array."""
    if not hasattr(qubit_operator, "terms"):
        m = qubit_operator.toarray() if hasattr(qubit_operator, "toarray") else np.asarray(qubit_operator)
        return np.asarray(m, dtype=np.complex128)
    dim = 1 << n
    h = np.zeros((dim, dim), dtype=np.complex128)
    for key, coeff in qubit_operator.terms.items():
        ops = dict((q, p) for q, p in key)
        m = np.ones((1, 1), dtype=np.complex128)
        for q in range(n - 1, -1, -1):
            m = np.kron(m, _PAULI[ops.get(q, "I")])
        h += coeff * m
    return h


def expectation(circ, values, n, qubit_operator, init=None):
    psi = state(circ, values, n, init)
    return np.vdot(psi, hamiltonian_matrix(qubit_operator, n) @ psi)


def gradient(circ, values, names, n, qubit_operator, init=None):
    """(f, dE/dparam for every name in ``names``) by the direct product formula."""
    dim = 1 << n
    circ = _acting(circ)
    psi0 = np.zeros(dim, dtype=np.complex128)
    if init is None:
        psi0[0] = 1
    else:
        psi0[:] = init
    hm = hamiltonian_matrix(qubit_operator, n)
    mats = [embed(gate_matrix(g, values), g.obj_qubits, g.ctrl_qubits, n) for g in circ]
    # prefix states and suffix unitaries
    prefix = [psi0]
    for m in mats:
        prefix.append(m @ prefix[-1])
    psi = prefix[-1]
    f = np.vdot(psi, hm @ psi)
    suffix = [np.eye(dim, dtype=np.complex128)]
    for m in reversed(mats):
        suffix.append(suffix[-1] @ m)
    suffix = suffix[::-1]  # suffix[k] = U_N ... U_k (product of gates k..N-1), suffix[N] = I
    bra = (hm @ psi).conj()
    grad = np.zeros(len(names))
    idx = {nm: i for i, nm in enumerate(names)}
    for k, g in enumerate(circ):
        prs = g.param_resolvers()
        if not prs or all(pr.is_const() for pr in prs):
            continue
        # one derivative matrix per parameter expression; chain rule through its coefficients
        for pr, dmat in zip(prs, g.diff_matrices(values)):
            if pr.is_const():
                continue
            dm = embed_diff(dmat, g.obj_qubits, g.ctrl_qubits, n)
            val = bra @ (suffix[k + 1] @ (dm @ prefix[k]))
            for name, c in pr.items():
                if name in idx and name in pr.requires_grad_parameters:
                    grad[idx[name]] += 2 * np.real(val) * np.real(c)
    return f, grad


def _state_derivatives(circ, values, names, n, init):
    """Final state and d state / d name for every name (product rule over closed-form gates)."""
    dim = 1 << n
    circ = _acting(circ)
    psi0 = np.zeros(dim, dtype=np.complex128)
    if init is None:
        psi0[0] = 1
    else:
        psi0[:] = init
    mats = [embed(gate_matrix(g, values), g.obj_qubits, g.ctrl_qubits, n) for g in circ]
    prefix = [psi0]
    for m in mats:
        prefix.append(m @ prefix[-1])
    suffix = [np.eye(dim, dtype=np.complex128)]
    for m in reversed(mats):
        suffix.append(suffix[-1] @ m)
    suffix = suffix[::-1]
    idx = {nm: i for i, nm in enumerate(names)}
    d = np.zeros((len(names), dim), dtype=np.complex128)
    for k, g in enumerate(circ):
        prs = g.param_resolvers()
        for pr, dmat in zip(prs, g.diff_matrices(values) if prs else []):
            if pr.is_const():
                continue
            v = suffix[k + 1] @ (embed_diff(dmat, g.obj_qubits, g.ctrl_qubits, n) @ prefix[k])
            for name, c in pr.items():
                if name in idx and name in pr.requires_grad_parameters:
                    d[idx[name]] += np.real(c) * v
    return prefix[-1], d


def nonhermitian_gradient(circ_r, circ_l, values, names, n, qubit_operator, init_r=None, init_l=None):
    """f = <phi_l|H|psi_r> (phi_l = circ_l init_l, psi_r = circ_r init_r) and its complex
    derivative df/dname = <d phi_l|H|psi_r> + <phi_l|H|d psi_r> for every name."""
    hm = hamiltonian_matrix(qubit_operator, n)
    psi_r, d_r = _state_derivatives(circ_r, values, names, n, init_r)
    phi_l, d_l = _state_derivatives(circ_l, values, names, n, init_l)
    f = np.vdot(phi_l, hm @ psi_r)
    g = np.array([np.vdot(d_l[j], hm @ psi_r) + np.vdot(phi_l, hm @ d_r[j]) for j in range(len(names))])
    return f, g
