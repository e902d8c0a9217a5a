"""Accuracy bars shared by the GPU parity tests.

Every result is compared with the complex128 oracle (C++ restatement of the reference CPU path, or
the dense numpy simulation), whatever the precision of the run:

* complex128 - states: fidelity 1 - |<ref|psi>| <= 1e-12 and max |psi - ref| <= 1e-10;
  expectations: |f - f_ref| <= 1e-10 * max(1, sum |c|); gradients:
  max |g - g_ref| <= 1e-10 * max(1, max |g_ref|).
* complex64 - states: max |psi - ref| <= 1e-5 (unit norm); expectations:
  |f - f_ref| <= 1e-5 * max(1, sum |c|); gradients: max |g - g_ref| <= 1e-5 * max(max |g_ref|, 1e-2 * s)
  with s = max(1, sum |c'|), c' the non-identity coefficients (relative to the largest gradient).

sum |c| is the sum of the absolute Pauli coefficients, a bound on |<H>|; sum |c'| bounds every
derivative (the identity part adds nothing to one).  The floor of the gradient bar: a float32
evaluation of any gradient - the reference's complex64 path included - rounds at about 1e-7 of
sum |c'| (the co-state H psi and the tap sums have that size), so near a stationary point, where the
largest gradient is below a hundredth of that scale, the bar is taken relative to 1e-2 * s instead.
"""
import numpy as np

import mindquantum_amd as mq

C64_REL = 1e-5
C128_REL = 1e-10


def is_c128(dtype):
    return np.dtype(dtype) == np.complex128


def coeff_sum(op, identity=True):
    """sum |c| of a QubitOperator (or a Hamiltonian wrapping one); 1 for other operators.  With
    identity=False the identity term is left out."""
    terms = getattr(op, "terms", None)
    if terms is None and hasattr(op, "operator"):
        terms = getattr(op.operator, "terms", None)
    if not terms:
        return 1.0
    return float(sum(abs(complex(v)) for k, v in terms.items() if identity or len(k) > 0))


def f_tol(dtype, csum=1.0):
    return (C128_REL if is_c128(dtype) else C64_REL) * max(1.0, float(csum))


def g_tol(dtype, g_ref, csum=1.0):
    """csum: sum |c'| of the observable's non-identity coefficients (the floor's scale)."""
    g_ref = np.asarray(g_ref)
    gmax = float(np.max(np.abs(g_ref))) if g_ref.size else 0.0
    if is_c128(dtype):
        return C128_REL * max(1.0, gmax)
    return C64_REL * max(gmax, 1e-2 * max(1.0, float(csum)))


def assert_state(got, ref, dtype):
    got = np.asarray(got, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    err = float(np.max(np.abs(got - ref)))
    if is_c128(dtype):
        fid = 1 - abs(np.vdot(ref, got))
        assert fid <= 1e-12, f"fidelity defect {fid:.3e}"
        assert err <= 1e-10, f"max amplitude error {err:.3e}"
    else:
        assert err <= 1e-5, f"max amplitude error {err:.3e}"


def assert_f(f, f_ref, dtype, csum=1.0, what="f"):
    err = float(np.max(np.abs(np.asarray(f) - np.asarray(f_ref))))
    tol = f_tol(dtype, csum)
    assert err <= tol, f"{what}: error {err:.3e} > bar {tol:.3e}"
    return err


def assert_g(g, g_ref, dtype, what="g", csum=1.0):
    g = np.asarray(g)
    g_ref = np.asarray(g_ref)
    err = float(np.max(np.abs(g - g_ref))) if g_ref.size else 0.0
    tol = g_tol(dtype, g_ref, csum)
    assert err <= tol, f"{what}: error {err:.3e} > bar {tol:.3e}"
    return err


C128 = mq.complex128
