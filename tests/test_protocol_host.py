"""Host-side protocol layer, no GPU: the transcript, challenge derivation,
proof-of-work scan, configuration derivation and domain separator.

The oracle's restatements are pinned against the reference's goldens
(transcript, pow, indices, utils) and, where it is built, against the
reference's own C++ (oracle/_ref).  The product's host code (the transcript
and config of libwhir_gpu.so, callable without a device) is then checked
against the oracle and the reference build.
"""
import hashlib
import random

import numpy as np
import pytest

from oracle import transcript as OT

P = 0xFFFFFFFF00000001


def _ref_available():
    import oracle
    return oracle.ref() is not None


# ------------------------------------------------------------ oracle pins

def test_oracle_transcript_golden(fixtures):
    t = fixtures["transcript"]
    s = OT.Transcript()
    s.absorb(b"hello")
    assert s.squeeze(32).hex() == t["sponge"]["squeeze0"]
    s.absorb(b"world")
    assert s.squeeze(32).hex() == t["sponge"]["squeeze1"]
    assert s.squeeze(16).hex() == t["sponge"]["squeeze2_16b"]
    x = OT.Transcript()
    x.absorb(t["xof"]["absorb"].encode())
    a = x.squeeze(20) + x.squeeze(20)
    y = OT.Transcript()
    y.absorb(t["xof"]["absorb"].encode())
    assert a.hex() == y.squeeze(40).hex() == t["xof"]["once_40"]
    assert hashlib.sha3_512(bytes.fromhex(t["sha3"]["protocol_cbor"])).hexdigest() == t["sha3"]["protocol_id"]
    assert hashlib.sha3_256(bytes.fromhex(t["sha3"]["session_cbor"])).hexdigest() == t["sha3"]["session_id"]


def _round_trip(make, rt):
    """dump_transcript.cpp CASE 2 driven through a transcript factory."""
    pid = hashlib.sha3_512(bytes.fromhex(rt["protocol_cbor"])).digest()
    sess = rt["session"].encode()
    sid = hashlib.sha3_256(bytes([0x60 + len(sess)]) + sess).digest()
    t = make(pid, sid, rt["instance_u64"].to_bytes(8, "little"))
    t.message((0xCAFEBABEDEADBEEF).to_bytes(8, "little"))
    t.message((0x0123456789ABCDEF).to_bytes(8, "little"))
    ch = [int.from_bytes(t.squeeze(8), "little"), int.from_bytes(t.squeeze(8), "little")]
    t.hint(bytes([1, 2, 3, 4, 5]))
    t.message((0xFF00FF00FF00FF00).to_bytes(8, "little"))
    ch.append(int.from_bytes(t.squeeze(8), "little"))
    return ["%016x" % c for c in ch], t


def test_oracle_round_trip_golden(fixtures):
    rt = fixtures["transcript"]["round_trip"]
    ch, t = _round_trip(OT.Transcript, rt)
    assert ch == rt["challenges"]
    assert t.narg.hex() == rt["narg"] and t.hints.hex() == rt["hints"]


def test_oracle_indices_golden(fixtures):
    for c in fixtures["pow_indices"]["indices"]:
        got = OT.indices_from_entropy(bytes.fromhex(c["entropy"]), c["num_leaves"], c["count"], c["dedup"])
        assert got == c["indices"], c


def test_oracle_pow_golden(fixtures, orc):
    g = fixtures["pow_indices"]["pow"]
    ch = bytes([g["fill"]]) * 32
    n = orc.pow_find(ch, g["threshold"])
    assert n == g["nonce"]
    h = orc.blake3(ch + n.to_bytes(8, "little") + bytes(24))
    assert h.hex() == g["hash"]
    assert int.from_bytes(h[:8], "little") <= g["threshold"]
    # minimality: every smaller nonce misses the threshold
    for m in range(n):
        hm = orc.blake3(ch + m.to_bytes(8, "little") + bytes(24))
        assert int.from_bytes(hm[:8], "little") > g["threshold"]
    assert orc.pow_find(ch, 2**64 - 1) == 0


def test_oracle_challenge_decoding():
    """40 squeezed bytes as a little-endian integer mod p, and geometric powers."""
    t = OT.Transcript()
    t.absorb(b"decode")
    raw = OT.Transcript()
    raw.absorb(b"decode")
    b = raw.squeeze(40)
    assert t.challenge_gl() == int.from_bytes(b, "little") % P
    g = t.geometric(4)
    x = (int.from_bytes(raw.squeeze(40), "little") % P, int.from_bytes(raw.squeeze(40), "little") % P,
         int.from_bytes(raw.squeeze(40), "little") % P)
    assert g[0] == (1, 0, 0) and g[1] == x and g[2] == OT.e3_mul(x, x)
    assert t.geometric(1) == [(1, 0, 0)] and t.geometric(0) == []


def _lift(v):
    return np.array([[int(x), 0, 0] for x in v], dtype=np.uint64).reshape(-1)


def test_oracle_eq_geometric_mle_golden(fixtures, orc):
    u = fixtures["utils"]
    for c in u["eval_eq"]:
        k = len(c["point"])
        acc = orc.eval_eq_ext3(np.zeros(3 << k, dtype=np.uint64), _lift(c["point"]), [c["scalar"], 0, 0])
        assert np.array_equal(acc, _lift(c["result"]))
    ga = u["geometric_accumulate"]
    acc = orc.geometric_accumulate_ext3(_lift(ga["acc"]), _lift(ga["points"]), _lift(ga["scalars"]))
    assert np.array_equal(acc, _lift(ga["result"]))
    for c in u["multilinear_extend"]:
        out = orc.mle_evaluate(np.array(c["evals"], dtype=np.uint64), _lift(c["point"]), True)
        assert list(out) == [c["result"], 0, 0]


def _e3(rnd):
    return [rnd.randrange(P) for _ in range(3)]


def test_oracle_eq_geometric_mle_closed_form(orc):
    """Ext3 inputs: eq(x, i) = prod_k (bit ? x_k : 1 - x_k) with x_0 the top bit,
    sum_j s_j z_j^i, and the MLE as sum_i f_i eq(x, i), in Python big ints."""
    rnd = random.Random(9)
    for k in (0, 1, 3, 5):
        x = [_e3(rnd) for _ in range(k)]
        s = _e3(rnd)
        acc = orc.eval_eq_ext3(np.zeros(3 << k, dtype=np.uint64), np.array(x, dtype=np.uint64), s)
        f = [_e3(rnd) for _ in range(1 << k)]
        mle = (0, 0, 0)
        for i in range(1 << k):
            w = (1, 0, 0)
            for j in range(k):
                bit = (i >> (k - 1 - j)) & 1
                w = OT.e3_mul(w, x[j] if bit else [(1 - x[j][0]) % P, -x[j][1] % P, -x[j][2] % P])
            assert list(acc[3 * i:3 * i + 3]) == list(OT.e3_mul(w, s))
            t = OT.e3_mul(w, f[i])
            mle = tuple((a + b) % P for a, b in zip(mle, t))
        got = orc.mle_evaluate(np.array(f, dtype=np.uint64).reshape(-1), np.array(x, dtype=np.uint64), False)
        assert tuple(int(v) for v in got) == mle
    n, pts, scs = 16, [_e3(rnd) for _ in range(3)] + [[rnd.randrange(P), 0, 0]], [_e3(rnd) for _ in range(4)]
    base = [_e3(rnd) for _ in range(n)]
    acc = orc.geometric_accumulate_ext3(np.array(base, dtype=np.uint64).reshape(-1), np.array(pts, dtype=np.uint64),
                                        np.array(scs, dtype=np.uint64))
    for i in range(n):
        want = tuple(base[i])
        for z, s in zip(pts, scs):
            zi = (1, 0, 0)
            for _ in range(i):
                zi = OT.e3_mul(zi, z)
            want = tuple((a + b) % P for a, b in zip(want, OT.e3_mul(s, zi)))
        assert tuple(int(v) for v in acc[3 * i:3 * i + 3]) == want


# ------------------------------------------------------------ product host code

def test_product_transcript_golden(fixtures, W):
    t = fixtures["transcript"]
    s = W.Transcript()
    s.absorb(b"hello")
    assert s.squeeze(32).hex() == t["sponge"]["squeeze0"]
    s.absorb(b"world")
    assert s.squeeze(32).hex() == t["sponge"]["squeeze1"]
    assert s.squeeze(16).hex() == t["sponge"]["squeeze2_16b"]
    rt = t["round_trip"]
    ch, tr = _round_trip(lambda a, b, c: W.Transcript(a, b, c), rt)
    assert ch == rt["challenges"]
    assert tr.proof() == (bytes.fromhex(rt["narg"]), bytes.fromhex(rt["hints"]))


def test_product_transcript_vs_oracle(W):
    """Random interleavings of messages, hints, raw squeezes (crossing the
    168-byte rate), field challenges and index challenges."""
    rnd = random.Random(5)
    for trial in range(20):
        pid, sid = rnd.randbytes(64), rnd.randbytes(32)
        inst = rnd.randbytes(rnd.choice([0, 8, 100]))
        a, b = OT.Transcript(pid, sid, inst), W.Transcript(pid, sid, inst)
        for _ in range(40):
            op = rnd.randrange(6)
            if op == 0:
                m = rnd.randbytes(rnd.choice([8, 24, 32, 200, 500]))
                a.message(m)
                b.message(m)
            elif op == 1:
                h = rnd.randbytes(rnd.choice([5, 32]))
                a.hint(h)
                b.hint(h)
            elif op == 2:
                cnt = rnd.randrange(4)
                v = rnd.randbytes(24 * cnt)
                a.hint_vector(v, cnt)
                b.hint_vector(v, cnt, 24)
            elif op == 3:
                n = rnd.choice([1, 32, 40, 167, 168, 169, 400])
                assert a.squeeze(n) == b.squeeze(n)
            elif op == 4:
                assert a.challenge_gl() == b.challenge_gl()
            else:
                n = 1 << rnd.choice([0, 1, 7, 8, 9, 16, 20, 26])
                cnt, dd = rnd.randrange(0, 60), rnd.random() < 0.5
                assert a.indices(n, cnt, dd) == b.indices(n, cnt, dd)
        assert b.proof() == (bytes(a.narg), bytes(a.hints))


# derived shapes of the CLI defaults (security 128, PoW 20, list decoding)
CLI_DEFAULT_SHAPES = {
    # nv, log_inv_rate: in-domain samples of the initial commit and each round
    (22, 1): [252, 56, 32, 22, 17],
    (16, 2): [117, 45, 28, 20],
    (24, 2): [117, 45, 28, 20, 16, 13],
}


def _irs(flat, k):
    """(in_domain, codeword_length) of the initial committer (k = 0) or round k - 1."""
    off = 1 if k == 0 else 1 + 7 + 3 + 1 + (k - 1) * 11
    return int(flat[off + 4]), int(flat[off + 2])


@pytest.mark.parametrize("shape", sorted(CLI_DEFAULT_SHAPES))
def test_product_config_shapes(W, shape):
    nv, rate = shape
    flat = W.config_flat(1 << nv, W.Params.default(starting_log_inv_rate=rate))
    rounds = int(flat[0])
    assert rounds == len(CLI_DEFAULT_SHAPES[shape]) - 1
    assert [_irs(flat, k)[0] for k in range(rounds + 1)] == CLI_DEFAULT_SHAPES[shape]


@pytest.mark.skipif("not __import__('oracle').ref()")
def test_product_config_vs_reference(W, orc):
    bad = []
    for nv in range(4, 31):
        for rate in (1, 2, 3):
            for ff in (1, 2, 3, 4, 5, 6):
                if ff > nv:
                    continue
                for sec, pw in ((128, 20), (100, 16), (80, 0), (128, 0)):
                    a = orc.ref_config(nv, rate, ff, sec, pw)
                    b = W.config_flat(1 << nv, W.Params.default(starting_log_inv_rate=rate, initial_folding_factor=ff,
                                                                folding_factor=ff, security_level=sec, pow_bits=pw))
                    if not np.array_equal(a, b):
                        bad.append((nv, rate, ff, sec, pw))
    assert not bad


@pytest.mark.skipif("not __import__('oracle').ref()")
def test_product_domain_separator_vs_reference(W, orc):
    for kw in ({}, {"rate": 2}, {"ff": 3, "iff": 3}, {"sec": 100, "pow_bits": 16}, {"unique": 1}, {"batch": 4}):
        pid, sid = orc.ref_cli_ds(**kw)
        p = W.Params.default(unique_decoding=kw.get("unique", 0), starting_log_inv_rate=kw.get("rate", 1),
                             initial_folding_factor=kw.get("iff", 4), folding_factor=kw.get("ff", 4),
                             security_level=kw.get("sec", 128), pow_bits=kw.get("pow_bits", 20),
                             batch_size=kw.get("batch", 1))
        assert W.domain_separator_cli(p) == (pid, sid)


def test_product_domain_separator_sha2_vs_reference(W, orc):
    """--hash Sha2 changes the engine id hashed into the protocol id."""
    for kw in ({}, {"rate": 2, "pow_bits": 16}):
        pid, sid = orc.ref_cli_ds(hash_id=2, **kw)
        assert pid != orc.ref_cli_ds(hash_id=1, **kw)[0]
        p = W.Params.default(starting_log_inv_rate=kw.get("rate", 1), pow_bits=kw.get("pow_bits", 20), hash_id=2)
        assert W.domain_separator_cli(p) == (pid, sid)
