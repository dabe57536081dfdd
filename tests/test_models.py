"""Host-side setup (site types, operator lists, MPO automaton) against dense exact
operators built by the oracle. CPU only."""
import numpy as np
import pytest

from oracle import exact


def dense_from_oplist(H):
    return exact.dense_hamiltonian(H.length, [(o, s, c) for o, s, c in zip(H.ops, H.sites, H.coeffs)])


def heisenberg_oplist(N, pauli=True):
    from tnhip.models import OpList

    H = OpList(N)
    for i in range(1, N):
        if pauli:
            for o in ("x", "y", "z"):
                H.add([o, o], [i, i + 1], 0.25)
        else:
            H.add(["s+", "s-"], [i, i + 1], 0.5)
            H.add(["s-", "s+"], [i, i + 1], 0.5)
            H.add(["z", "z"], [i, i + 1], 0.25)
    return H


@pytest.mark.parametrize("pauli", [True, False])
def test_heisenberg_mpo_matches_dense(pauli):
    from tnhip.models import spinhalf, MPO

    st = spinhalf()
    N = 6
    H = heisenberg_oplist(N, pauli)
    M = MPO(st, H, compress=False)
    assert max(w.shape[3] for w in M.tensors[:-1]) == 5  # D = 2 + 3 channels
    assert np.allclose(exact.mpo_to_dense(M.tensors), dense_from_oplist(H))
    # s+/s- form is real; the Pauli-y form is complex
    assert np.isrealobj(M.tensors[2]) == (not pauli)
    # both forms are the same operator
    assert np.allclose(dense_from_oplist(heisenberg_oplist(N, True)), dense_from_oplist(heisenberg_oplist(N, False)))


def test_tfim_with_fields_and_long_range_term():
    from tnhip.models import spinhalf, OpList, MPO

    st = spinhalf()
    N = 5
    H = OpList(N)
    for i in range(1, N + 1):
        H.add("x", i, 1.0)
        H.add("z", i, 0.05)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], 1.2)
    H.add(["x", "z"], [1, 4], 0.3)  # range-4 term with identity gaps
    M = MPO(st, H, compress=False)
    assert np.allclose(exact.mpo_to_dense(M.tensors), dense_from_oplist(H))


def test_oplist_validation():
    from tnhip.models import OpList

    H = OpList(4)
    with pytest.raises(ValueError):
        H.add(["x", "x"], [2, 2])
    with pytest.raises(ValueError):
        H.add(["x"], [5])
    H.add(["z", "x"], [3, 1], 2.0)  # sorted by site
    assert H.sites[-1] == [1, 3] and H.ops[-1] == ["x", "z"]


def test_sitetensor_two_site_gate_layout():
    """sitetensor gives G[o1, i1, o2, i2] = sum of terms starting at the site."""
    from tnhip.models import spinhalf, OpList

    st = spinhalf()
    H = OpList(3)
    H.add(["z", "z"], [1, 2], -1.0)
    H.add("x", 1, -0.5)
    G = H.sitetensor(st, 1)
    Z, X, I = (exact.OPS[k] for k in ("z", "x", "id"))
    ref = -np.einsum("ab,cd->abcd", Z, Z) - 0.5 * np.einsum("ab,cd->abcd", X, I)
    assert np.allclose(G, ref)
    assert H.sitetensor(st, 3) is None


def test_mpo_coefficient_slots_are_the_derivative():
    """W(c) - W(c with c_t -> c_t + 1) is nonzero only at the recorded slot, equal to op:
    the slots are dW/dc_t, which the coupling gradient contracts with dE/dW."""
    from tnhip.models import spinhalf, OpList, mpo_tensors

    st = spinhalf()
    N = 5
    H = OpList(N)
    for i in range(1, N):
        H.add(["z", "z"], [i, i + 1], -1.0 - 0.1 * i)
    for i in range(1, N + 1):
        H.add("x", i, -0.3 * i)
    H.add(["y", "x"], [2, 4], 0.4)
    slots = []
    W0 = mpo_tensors(st, H, slots)
    assert len(slots) == len(H.coeffs)
    for t, (site, row, col, op) in enumerate(slots):
        Hp = H * 1.0
        Hp.coeffs[t] += 1.0
        W1 = mpo_tensors(st, Hp)
        for i in range(N):
            diff = np.asarray(W1[i], dtype=complex) - np.asarray(W0[i], dtype=complex)
            if i == site - 1:
                expect = np.zeros_like(diff)
                expect[row, :, :, col] = op
                assert np.allclose(diff, expect), t
            else:
                assert np.abs(diff).max() == 0.0, (t, i)
