"""rq_expectation_hamiltonian_noalloc (reduction partials in the library's
fixed device array) against rq_expectation_hamiltonian (partials in the
state's own buffer), both called through the C ABI: the same kernels in the
same order, so the energies must be bit-identical.  Two calls in a row, the
second on a smaller state with fewer terms, so stale contents of the shared
array from the first call cannot leak into the second; the oracle's FP64
energy is checked beside them at 1e-12 (c64) / 1e-5 (c32)."""
import ctypes as C

import numpy as np
import pytest

from ruqu_amd import Hamiltonian, QuantumState, workloads
from ruqu_amd._native import lib, ptr
from tests.helpers import orc, random_state

pytestmark = pytest.mark.gpu

CASES = [  # (qubits, precision, terms, seed): the second call is the smaller one
    (18, "c64", 75, 1),
    (11, "c32", 5, 2),
    (9, "c64", 33, 3),
]


def _energy(entry, st, h):
    coeffs, offs, ops = h.to_arrays()
    v = C.c_double()
    rc = getattr(lib(), entry)(st.handle, ptr(coeffs), ptr(offs), ptr(ops) if len(ops) else None, len(coeffs),
                               C.byref(v))
    assert rc == 0, lib().rq_last_error()
    return v.value


def test_noalloc_matches_allocating_entry_point():
    for n, prec, nterms, seed in CASES:
        print("noalloc hamiltonian n=%d %s terms=%d" % (n, prec, nterms), flush=True)
        psi = random_state(n, 100 + seed, np.complex128 if prec == "c64" else np.complex64)
        terms = workloads.random_pauli_hamiltonian(n, nterms, 4, seed)
        h = Hamiltonian([(c, ops) for c, ops in terms], n)
        st = QuantumState.from_amplitudes(psi, n, prec)
        got = _energy("rq_expectation_hamiltonian_noalloc", st, h)
        want = _energy("rq_expectation_hamiltonian", st, h)
        assert got == want
        ref = orc.expectation_hamiltonian(psi.astype(np.complex128), n, terms)
        assert abs(got - ref) < (1e-12 if prec == "c64" else 1e-5)
        again = _energy("rq_expectation_hamiltonian_noalloc", st, h)
        assert again == got
        st.close()


def test_noalloc_errors_and_empty():
    st = QuantumState.from_amplitudes(random_state(4, 5), 4, "c64")
    v = C.c_double(7.0)
    assert lib().rq_expectation_hamiltonian_noalloc(st.handle, None, None, None, 0, C.byref(v)) == 0
    assert v.value == 0.0
    h = Hamiltonian([(1.0, [(9, 1)])], 4)  # X on a qubit past the state: an error, as rq_expectation_hamiltonian
    coeffs, offs, ops = h.to_arrays()
    rc_a = lib().rq_expectation_hamiltonian_noalloc(st.handle, ptr(coeffs), ptr(offs), ptr(ops), 1, C.byref(v))
    rc_b = lib().rq_expectation_hamiltonian(st.handle, ptr(coeffs), ptr(offs), ptr(ops), 1, C.byref(v))
    assert rc_a == rc_b != 0
    st.close()
