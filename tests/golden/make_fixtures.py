#!/usr/bin/env python3
"""Regenerate tests/golden/fixtures.json from the reference's committed goldens.

Reads the reference's golden files (cpp/dump_output/**, data written by its
dumpers cpp/tests/dump_*.cpp) and re-derives each case's *inputs* by replaying
the dumpers' documented input generation (the shared LCG
x <- x*6364136223846793005 + 1442695040888963407 with each dumper's seed,
drawn in the dumper's order).  The result is a small JSON file of
inputs + expected outputs; the reference itself is not needed afterwards and
no reference code is run or copied.

Usage:  python tests/golden/make_fixtures.py [/path/to/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys

MASK = (1 << 64) - 1
P = 0xFFFFFFFF00000001
HERE = os.path.dirname(os.path.abspath(__file__))


class Lcg:
    def __init__(self, seed: int):
        self.s = seed & MASK

    def next(self) -> int:
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & MASK
        return self.s

    def gl(self) -> int:
        v = self.next()
        return v - P if v >= P else v


def read(ref: str, rel: str, utf16: bool = False) -> str:
    path = os.path.join(ref, "cpp", "dump_output", rel)
    raw = open(path, "rb").read()
    txt = raw.decode("utf-16-le") if utf16 else raw.decode("utf-8")
    return txt.replace("\r", "").lstrip("﻿")


def cases(text: str, section: str) -> list[list[str]]:
    """Split one '# SECTION name' block into CASE blocks (list of lines)."""
    m = re.search(r"# SECTION " + section + r"\n(.*?)(?=\n\s*# SECTION|\Z)", text, re.S)
    body = m.group(1)
    out, cur = [], None
    for line in body.split("\n"):
        if line.startswith("CASE"):
            cur = [line]
            out.append(cur)
        elif cur is not None and line.strip():
            cur.append(line.strip())
    return out


def kv(lines: list[str]) -> dict[str, list[str]]:
    d = {}
    for ln in lines[1:]:
        parts = ln.split()
        d[parts[0]] = parts[1:]
    return d


# ------------------------------------------------------------------ fields
def fields(ref: str) -> dict:
    """cpp/tests/dump_algebra_fields.cpp: seed 0xCAFEBABEDEADBEEF; base draws
    (a, b) x3, ext2 draws 4 x3, ext3 draws 6 x3; the CASE line prints the raw
    u64 draws, mapped with from_u64."""
    txt = read(ref, "field_ouput/golden_fields_cpp.txt")
    base = []
    for c in cases(txt, "base"):
        raw = [int(x) for x in c[0].split()[2:]]
        d = kv(c)
        base.append({"a_raw": raw[0], "b_raw": raw[1],
                     **{k: int(v[0]) for k, v in d.items()}})
    ext3 = []
    for c in cases(txt, "ext3"):
        raw = [int(x) for x in c[0].split()[2:]]
        d = kv(c)
        ext3.append({"a_raw": raw[:3], "b_raw": raw[3:],
                     **{k: [int(x) for x in v] for k, v in d.items()}})
    return {"base": base, "ext3": ext3}


# -------------------------------------------------------------------- hash
def hashes(ref: str) -> list:
    """cpp/tests/dump_hash.cpp: one LCG(0xAAAA...) byte stream shared by all
    cases in order; each case draws size*count bytes, whole u64 words LE,
    discarding the unused tail of the last word."""
    txt = read(ref, "hash_output/golden_hash_cpp.txt")
    rng = Lcg(0xAAAAAAAAAAAAAAAA)
    out = []
    for c in cases(txt, "hash"):
        m = re.match(r"CASE (\d+) (\w+) size=(\d+) count=(\d+)", c[0])
        engine, size, count = m.group(2), int(m.group(3)), int(m.group(4))
        n = size * count
        buf = bytearray()
        while len(buf) < n:
            w = rng.next()
            for _ in range(8):
                if len(buf) < n:
                    buf.append(w & 0xFF)
                w >>= 8
        digests = [c[1 + i].split()[1] for i in range(count)]
        out.append({"engine": engine, "size": size, "count": count,
                    "input": bytes(buf).hex(), "digests": digests})
    return out


# ----------------------------------------------------------- matrix commit
def matrix_commit(ref: str) -> list:
    """cpp/tests/dump_matrix_commit.cpp: LCG(0xBBBB...); case 0 base 4x8 BLAKE3,
    case 1 Ext2 2x8 BLAKE3 (c0,c1 per element), later cases use SHA-256/Copy."""
    txt = read(ref, "matrix_commit_output/golden_matrix_commit_cpp.txt")
    cs = cases(txt, "matrix_commit")
    rng = Lcg(0xBBBBBBBBBBBBBBBB)
    out = []
    m0 = [rng.gl() for _ in range(4 * 8)]
    out.append({"field": "base", "hash": "blake3", "rows": 4, "cols": 8, "matrix": m0,
                "leaves": [ln.split()[1] for ln in cs[0][1:]]})
    m1 = [rng.gl() for _ in range(2 * 8 * 2)]  # (c0, c1) per element
    out.append({"field": "ext2", "hash": "blake3", "rows": 2, "cols": 8, "matrix": m1,
                "leaves": [ln.split()[1] for ln in cs[1][1:]]})
    return out


# ------------------------------------------------------------- merkle tree
def merkle(ref: str) -> list:
    """cpp/tests/dump_merkle_tree.cpp: leaf i byte j = (i*31 + j) & 0xFF;
    open indices per case as listed in the dumper."""
    txt = read(ref, "merkle_tree_output/golden_merkle_tree_cpp.txt")
    idx = {"blake3-16-2": [3, 5, 11], "blake3-32-3": [0, 1, 2, 31], "sha2-8-2": [2, 6]}
    out = []
    for c in cases(txt, "merkle_tree"):
        m = re.match(r"CASE (\d+) (\S+) num_leaves=(\d+)", c[0])
        label, nl = m.group(2), int(m.group(3))
        d = kv(c)
        leaves = bytes(((i * 31 + j) & 0xFF) for i in range(nl) for j in range(32))
        nh = int(d["num_hints"][0])
        out.append({"label": label, "hash": label.split("-")[0], "num_leaves": nl,
                    "leaves": leaves.hex(), "indices": idx[label],
                    "root": d["root"][0], "hints": [d[f"hint{i}"][0] for i in range(nh)]})
    return out


# ------------------------------------------------------------------- utils
def utils(ref: str) -> dict:
    """cpp/tests/dump_algebra_utils.cpp, SECTION utilities (LCG 0x1111...) case 3
    univariate_evaluate and SECTION sumcheck (LCG 0x3333...) cases 0-2."""
    txt = read(ref, "utils_output/golden_utils_cpp.txt")
    rng = Lcg(0x1111111111111111)
    rng.gl()                                   # geometric_sequence base
    [rng.gl() for _ in range(16)]              # dot a, b
    [rng.gl() for _ in range(5)]               # tensor_product aa, bb
    coeffs = [rng.gl() for _ in range(8)]
    x = rng.gl()
    uni = kv(cases(txt, "utilities")[3])
    out = {"univariate_evaluate": {"coeffs": coeffs, "x": x, "result": int(uni["result"][0])}}
    [rng.gl() for _ in range(17)]              # scalar_mul_add acc, w, vec
    acc2 = [rng.gl() for _ in range(8)]
    scal = [rng.gl() for _ in range(3)]
    pts = [rng.gl() for _ in range(3)]
    ga = cases(txt, "utilities")[5]
    out["geometric_accumulate"] = {"acc": acc2, "scalars": scal, "points": pts,
                                   "result": [int(v) for v in ga[2:]]}

    # SECTION multilinear (LCG 0x2222...): multilinear_extend k=3,4,5, eval_eq k=3,4
    rng = Lcg(0x2222222222222222)
    ml = cases(txt, "multilinear")
    ext = []
    for i, k in enumerate((3, 4, 5)):
        ev = [rng.gl() for _ in range(1 << k)]
        pt = [rng.gl() for _ in range(k)]
        ext.append({"evals": ev, "point": pt, "result": int(kv(ml[i])["result"][0])})
    eqs = []
    for i, k in ((3, 3), (4, 4)):
        pt = [rng.gl() for _ in range(k)]
        sc = rng.gl()
        eqs.append({"point": pt, "scalar": sc, "result": [int(v) for v in ml[i][2:]]})
    out["multilinear_extend"] = ext
    out["eval_eq"] = eqs

    rng = Lcg(0x3333333333333333)
    sc = cases(txt, "sumcheck")
    comp = []
    for i, n in enumerate((8, 16)):
        a = [rng.gl() for _ in range(n)]
        b = [rng.gl() for _ in range(n)]
        d = kv(sc[i])
        comp.append({"a": a, "b": b, "acc0": int(d["acc0"][0]), "acc2": int(d["acc2"][0])})
    vals = [rng.gl() for _ in range(16)]
    wt = rng.gl()
    folded = [int(x) for x in sc[2][2:]]
    out["sumcheck_compute"] = comp
    out["fold"] = {"values": vals, "weight": wt, "folded": folded}
    return out


# ---------------------------------------------------------------- sumcheck
def sumcheck(ref: str) -> dict:
    """golden_sumcheck (UTF-16LE): a, b printed in full; replay each round's
    compute and fold with the golden's own r (the squeezed r is stale against
    the current transcript, so only the algebra is pinned)."""
    txt = read(ref, "sumcheck_output/golden_sumcheck_cpp.txt", utf16=True)
    d = {}
    for line in txt.split("\n"):
        p = line.split()
        if len(p) >= 2 and not p[0].startswith(("CASE", "#")):
            d.setdefault(p[0], [x for x in p[1:]])
    rounds = []
    for r in range(3):
        rounds.append({k: int(d[f"round{r}_{k}"][0]) for k in ("c0", "c2", "r", "sum")})
    return {"a": [int(x) for x in d["vector_a"]], "b": [int(x) for x in d["vector_b"]],
            "initial_sum": int(d["initial_sum"][0]), "rounds": rounds,
            "final_a": int(d["final_a"][0]), "final_b": int(d["final_b"][0])}


# -------------------------------------------------------- transcript, PoW
def _cbor_text(t: str) -> bytes:
    assert len(t) < 24
    return bytes([0x60 + len(t)]) + t.encode()


def _cbor_test_config() -> bytes:
    """The CBOR map the dumper hashes in CASE 1: {version: 1, field_size: p,
    security_bits: 100}."""
    return (b"\xa3" + _cbor_text("version") + b"\x01" + _cbor_text("field_size") + b"\x1b"
            + P.to_bytes(8, "big") + _cbor_text("security_bits") + b"\x18\x64")


def transcript(ref: str) -> dict:
    """golden_transcript (UTF-16LE; cpp/tests/dump_transcript.cpp).  CASE 0 and 4:
    sponge absorb/squeeze of the dumper's ASCII inputs; CASE 1: SHA3-512/256 of the
    dumper's CBOR bytes; CASE 2: a prover round trip under DomainSeparator
    protocol(uint32 0xDEADBEEF).session("round_trip") with instance u64 1024.
    CASE 3 (vector framing) is stale against the current serializer, which writes
    an 8-byte element count (transcript.hpp:300-307), and is not used."""
    txt = read(ref, "transcript_output/golden_transcript_cpp.txt", utf16=True)
    cs = [kv(c) for c in cases(txt, "transcript")]
    one = lambda d, k: d[k][0]
    return {
        "sponge": {"absorb": ["hello", "world"], "squeeze0": one(cs[0], "squeeze0"),
                   "squeeze1": one(cs[0], "squeeze1"), "squeeze2_16b": one(cs[0], "squeeze2_16b")},
        "sha3": {"protocol_cbor": _cbor_test_config().hex(),
                 "protocol_id": one(cs[1], "protocol_id_h1") + one(cs[1], "protocol_id_h2"),
                 "session_cbor": _cbor_text("test_session_42").hex(), "session_id": one(cs[1], "session_id")},
        "round_trip": {"protocol_cbor": "1adeadbeef", "session": "round_trip", "instance_u64": 1024,
                       "challenges": [one(cs[2], f"challenge{i}") for i in range(3)],
                       "narg": one(cs[2], "proof_narg"), "hints": one(cs[2], "proof_hints")},
        "xof": {"absorb": "test_xof_streaming", "once_40": one(cs[4], "squeeze_40b_once"),
                "first_20": one(cs[4], "squeeze_20b_first"), "second_20": one(cs[4], "squeeze_20b_second")},
    }


def pow_and_indices(ref: str) -> dict:
    """golden_pow (cpp/tests/dump_pow.cpp: challenge = 32 bytes of the case's fill
    byte, threshold 2^60) and golden_indices (cpp/tests/dump_challenge_indices.cpp:
    the dumper's entropy bytes per case)."""
    txt = read(ref, "pow_output/golden_pow_cpp.txt")
    pc = cases(txt, "pow")
    b3 = kv(pc[0])
    pw = {"fill": 0xAA, "threshold": 1 << 60, "nonce": int(b3["nonce"][0]), "hash": b3["hash"][0]}
    txt = read(ref, "indices_output/golden_indices_cpp.txt")
    ent = [(128, 5, True, "0123456789"), (128, 5, False, "0123456789"),
           (8192, 5, True, "0123456789abcdef1234"), (1 << 20, 4, True, "123456789abcdef0112233 44".replace(" ", "")),
           (128, 5, True, "2040206040"), (1, 3, True, ""), (1, 3, False, ""), (8, 0, True, "")]
    idx = []
    for (n, k, d, e), c in zip(ent, cases(txt, "challenge_indices")):
        got = c[1].split()[1:] if len(c) > 1 else []
        idx.append({"num_leaves": n, "count": k, "dedup": d, "entropy": e, "indices": [int(v) for v in got]})
    return {"pow": pw, "indices": idx}


# -------------------------------------------------------------------- whir
def whir(ref: str) -> list:
    """golden_whir (cpp/tests/dump_whir.cpp): LCG(0xCAFE) shared across cases
    (0: 1x2, 1: 1x4, 2: 1x8, 3: 3x4, 4: 4x8); depth 1, codeword = 2*size,
    1 OOD point; the witness matrix and OOD values pin RS encode and Horner.
    proof_narg / proof_hints are kept for the transcript checks."""
    txt = read(ref, "whir_output/golden_whir_cpp.txt")
    shapes = [(1, 2), (1, 4), (1, 8), (3, 4), (4, 8)]
    out = []
    for (nv, vs), c in zip(shapes, cases(txt, "whir")):
        d = kv(c)
        vecs = [[int(x) for x in d[f"input_vector{i}"]] for i in range(nv)]
        out.append({"num_vectors": nv, "vector_size": vs, "vectors": vecs,
                    "witness_matrix": [int(x) for x in d["witness_matrix"]],
                    "ood_points": [int(x) for x in d["ood_points"]],
                    "ood_matrix": [int(x) for x in d["ood_matrix"]],
                    "proof_narg": d["proof_narg"][0], "proof_hints": d["proof_hints"][0]})
    return out


def main() -> None:
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    fx = {
        "source": "reference cpp/dump_output/* (dumpers cpp/tests/dump_*.cpp)",
        "fields": fields(ref),
        "hash": hashes(ref),
        "matrix_commit": matrix_commit(ref),
        "merkle_tree": merkle(ref),
        "utils": utils(ref),
        "sumcheck": sumcheck(ref),
        "whir": whir(ref),
        "transcript": transcript(ref),
        "pow_indices": pow_and_indices(ref),
    }
    path = os.path.join(HERE, "fixtures.json")
    with open(path, "w") as f:
        json.dump(fx, f, indent=1)
        f.write("\n")
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
