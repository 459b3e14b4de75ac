"""The oracle's copies between field-major rj spectra and the interleaved
send / receive buffers (copy_mhd_spectr_to_send / _from_recv), pinned bit
for bit against the reference's own set_to_send_buf_* and
set_from_recv_buf_rev_* (SOLVER_src/send_recv/field_to_send_buffer.f90),
compiled from source into oracle/_ref; and against what the copies must
satisfy: to-send then from-receive through inverse tables gives the fields
back, and the dummy record reads zero."""
import numpy as np
import pytest

import oracle as O
from oracle import ref

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")

# the MHD step's backward field set: velo, vort, magne, current (vectors),
# temp, light (scalars), in a field array that also holds other fields, and
# a tensor
FIELDS_SEND = [(3, 0, 3), (9, 3, 3), (0, 6, 3), (15, 9, 3), (6, 12, 1), (21, 13, 1), (22, 14, 6)]
NTOT = 28
NB = 20


def _tables(n, seed):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n).astype(np.int32) + 1          # 1-based
    return perm


@needs_ref
def test_to_send_bitexact_vs_reference():
    n = 517
    rng = np.random.default_rng(1)
    fld = rng.standard_normal((NTOT, n))
    ws0 = rng.standard_normal(NB * n)          # components no field names must stay
    inod = _tables(n, 2)
    want = ref.to_send(FIELDS_SEND, NTOT, n, fld, NB, ws0, inod)
    got = O.rj_to_send(FIELDS_SEND, NTOT, n, fld, NB, ws0, inod_export=inod)
    np.testing.assert_array_equal(got, want)
    ident = O.rj_to_send(FIELDS_SEND, NTOT, n, fld, NB, ws0)
    np.testing.assert_array_equal(ident, ref.to_send(FIELDS_SEND, NTOT, n, fld, NB, ws0,
                                                     np.arange(1, n + 1, dtype=np.int32)))


@needs_ref
def test_from_recv_bitexact_vs_reference():
    n, nrecv = 433, 400
    rng = np.random.default_rng(3)
    wr = rng.standard_normal(NB * nrecv)
    irev = rng.integers(1, nrecv + 2, size=n).astype(np.int32)   # nrecv + 1: the dummy record
    irev[:5] = nrecv + 1
    d0 = rng.standard_normal(NTOT * n)
    want = ref.from_recv_rev(FIELDS_SEND, NB, nrecv, wr, NTOT, n, d0, irev)
    got = O.rj_from_recv(FIELDS_SEND, NB, nrecv, wr, NTOT, n, d0, irev=irev)
    np.testing.assert_array_equal(got, want)


def test_send_then_receive_is_the_identity():
    n = 301
    rng = np.random.default_rng(5)
    fld = rng.standard_normal((NTOT, n))
    perm = _tables(n, 6)
    inv = np.empty_like(perm)
    inv[perm - 1] = np.arange(1, n + 1, dtype=np.int32)
    ws = O.rj_to_send(FIELDS_SEND, NTOT, n, fld, NB, np.zeros(NB * n), inod_export=perm)
    back = O.rj_from_recv(FIELDS_SEND, NB, n, ws, NTOT, n, np.zeros(NTOT * n),
                          irev=inv).reshape(NTOT, n)
    for i_fld, _, nc in FIELDS_SEND:
        np.testing.assert_array_equal(back[i_fld:i_fld + nc], fld[i_fld:i_fld + nc])
    # untouched components stay zero; the dummy record reads zero
    named = {i_fld + c for i_fld, _, nc in FIELDS_SEND for c in range(nc)}
    assert all((back[c] == 0).all() for c in range(NTOT) if c not in named)
    z = O.rj_from_recv(FIELDS_SEND, NB, n, ws, NTOT, 4, np.ones(NTOT * 4),
                       irev=np.full(4, n + 1, dtype=np.int32)).reshape(NTOT, 4)
    assert all((z[c] == 0).all() for c in named)
