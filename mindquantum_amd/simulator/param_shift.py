"""Parameter-shift gradients (get_expectation_with_grad(..., pr_shift=True)).

Follows GetExpectationWithGradParameterShiftOneMulti (vector_state.tpp:1033-1243): for every
parameterised gate of the circuit and every non-constant parameter expression k of it, the
expectation is evaluated with that expression shifted by +s and -s, and
d f / d p = sum over (gate, k) of coeff * (f(+s) - f(-s)) * d expr_k / d p.  The shift rule is the
reference's: s = pi/2, coeff = 1/2 for RX, RY, RZ, Rxx, Ryy, Rzz, Rxy, Rxz, Ryz, GlobalPhase,
PhaseShift and U3; s = 1/2, coeff = pi/2 for SWAPalpha; a central difference (s = 1e-3, coeff =
1/(2s)) for every other gate and for every gate with control qubits.  All shifted settings of a
call run as one batch of angle tables on the device (mqa_sim_expectation_angles).
"""
import ctypes

import numpy as np

from .. import _native
from ..core import gates as G
from ..ir import NativeCircuit

_HALF_PI_KINDS = (G.RX, G.RY, G.RZ, G.Rxx, G.Ryy, G.Rzz, G.Rxy, G.Rxz, G.Ryz, G.GlobalPhase, G.PhaseShift, G.U3)


def shift_rule(gate):
    """(shift, coefficient) of the reference for this gate."""
    if gate.ctrl_qubits:
        return 1e-3, 0.5 / 1e-3
    if isinstance(gate, _HALF_PI_KINDS):
        return np.pi / 2, 0.5
    if isinstance(gate, G.SWAPalpha):
        return 0.5, np.pi / 2
    return 1e-3, 0.5 / 1e-3


def slot_table(circ, names):
    """Per angle slot of the lowered circuit: its affine angle (const, [(param index, coeff)]) and
    its origin (gate index, expression index, factor) or None for a constant slot."""
    index = {n: i for i, n in enumerate(names)}
    slots = []
    for gi, g in enumerate(circ):
        for p in g.lowered():
            if not isinstance(p, G.ParameterGate):
                continue
            pr = p.coeff
            terms = [(index[k], float(np.real(v))) for k, v in pr.items()]
            origin = None
            if not pr.is_const():
                k, factor = getattr(p, "_origin", (0, 1.0))
                origin = (gi, k, factor)
            slots.append((float(np.real(pr.const)), terms, origin))
    return slots


def gradient(sim, hams, circ, enc_names, ans_names, enc, ans):
    """f [batch, n_hams] and g [batch, n_hams, n_params] (complex), as the adjoint path returns."""
    names = list(enc_names) + list(ans_names)
    n_enc = len(enc_names)
    nc = NativeCircuit(circ, names)
    if nc.custom:
        raise ValueError("parameter-shift gradients of parameterised matrix gates (custom gates, parameterised Rn) "
                         "are not supported: their matrices do not come from angle settings; use the adjoint gradient")
    slots = slot_table(circ, names)
    # the shifted expressions: (gate, expression, shift, coeff, slot -> factor)
    shifts = []
    for gi, g in enumerate(circ):
        prs = g.param_resolvers()
        if not prs:
            continue
        s, coeff = shift_rule(g)
        for k, pr in enumerate(prs):
            if pr.is_const() or not pr.requires_grad_parameters:
                continue
            slot_f = [(j, o[2]) for j, (_, _, o) in enumerate(slots) if o is not None and o[0] == gi and o[1] == k]
            jac = {index: float(np.real(pr[n])) for index, n in enumerate(names)
                   if n in pr.keys() and n in pr.requires_grad_parameters}
            shifts.append((s, coeff, slot_f, jac))
    batch = enc.shape[0]
    n_slots = len(slots)
    rows = 1 + 2 * len(shifts)
    angles = np.zeros((batch, rows, max(n_slots, 1)))
    for b in range(batch):
        values = np.concatenate([enc[b], ans]) if n_enc else np.asarray(ans, dtype=np.float64)
        base = np.array([c + sum(w * values[i] for i, w in t) for c, t, _ in slots]) if n_slots else np.zeros(0)
        angles[b, :, :n_slots] = base
        for r, (s, _, slot_f, _) in enumerate(shifts):
            for j, fac in slot_f:
                angles[b, 1 + 2 * r, j] += fac * s
                angles[b, 2 + 2 * r, j] -= fac * s
    n_h = len(hams)
    f = np.zeros((batch * rows, n_h, 2))
    handles = (ctypes.c_void_p * n_h)(*[h.native().value for h in hams])
    flat = np.ascontiguousarray(angles[:, :, :n_slots].reshape(batch * rows, n_slots)) if n_slots else \
        np.zeros(batch * rows)
    _native.check(_native.lib().mqa_sim_expectation_angles(sim._h, handles, n_h, nc.handle, _native.ptr(flat),
                                                            batch * rows, _native.ptr(f)))
    fc = (f[..., 0] + 1j * f[..., 1]).reshape(batch, rows, n_h)
    g = np.zeros((batch, n_h, len(names)), dtype=np.complex128)
    for r, (s, coeff, _, jac) in enumerate(shifts):
        d = coeff * (fc[:, 1 + 2 * r, :] - fc[:, 2 + 2 * r, :])
        for index, w in jac.items():
            g[:, :, index] += d * w
    return fc[:, 0, :], g
