"""The reference's own verifier (whir_verifier.hpp, through the CLI's
main.cpp:434-446 flow, compiled from its sources) as the acceptance check of
GPU proofs.  Pinned here on the CPU: it accepts the reference prover's proofs
and rejects every tampering (a flipped narg or hint byte anywhere, a
truncated or extended proof, a wrong claimed value, the other hash engine)."""
import numpy as np
import pytest

import oracle
import whirplonk_amd as W

pytestmark = pytest.mark.skipif("not __import__('oracle').ref()")


def _case(nv, nforms, seed, batch=1):
    coeffs, st = W.lcg_uniform(seed, batch << nv)
    pts, _ = W.lcg_uniform(st, 3 * nv * nforms)
    return coeffs, pts.reshape(nforms, nv, 3)


@pytest.mark.parametrize("nv,kw", [(10, dict(nforms=2)), (12, dict(nforms=1, rate=2, hash_id=2)),
                                   (11, dict(nforms=2, batch=3)), (12, dict(nforms=1, unique=1))])
def test_reference_verifier_accepts_and_rejects(nv, kw):
    kw = dict(kw)
    nforms = kw.pop("nforms")
    hid = kw.pop("hash_id", 1)
    coeffs, pts = _case(nv, nforms, 0x7E51F00 + nv, kw.get("batch", 1))
    ev, narg, hints = oracle.ref_prove(nv, coeffs, pts, hash_id=hid, **kw)
    assert oracle.ref_verify(nv, pts, ev, narg, hints, hash_id=hid, **kw)
    for pos in np.linspace(0, len(narg) - 1, 9).astype(int):
        b = bytearray(narg)
        b[pos] ^= 0x10
        assert not oracle.ref_verify(nv, pts, ev, bytes(b), hints, hash_id=hid, **kw), f"narg byte {pos}"
    # hint bytes past the first vector's u64 length prefix: a changed length
    # trips the reference's assert (matrix_commit.hpp:176), which ends the
    # process instead of rejecting
    for pos in np.linspace(8, len(hints) - 1, 9).astype(int):
        b = bytearray(hints)
        b[pos] ^= 0x01
        assert not oracle.ref_verify(nv, pts, ev, narg, bytes(b), hash_id=hid, **kw), f"hint byte {pos}"
    assert not oracle.ref_verify(nv, pts, ev, narg[:-1], hints, hash_id=hid, **kw)
    assert not oracle.ref_verify(nv, pts, ev, narg + b"\0", hints, hash_id=hid, **kw)
    assert not oracle.ref_verify(nv, pts, ev, narg, hints[:-8], hash_id=hid, **kw)
    bad = ev.copy()
    bad[-1, 2] = (int(bad[-1, 2]) + 1) % oracle.P
    assert not oracle.ref_verify(nv, pts, bad, narg, hints, hash_id=hid, **kw)
    assert not oracle.ref_verify(nv, pts, ev, narg, hints, hash_id=3 - hid, **kw)
