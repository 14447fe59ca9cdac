"""CPU restatement of the prover transcript — TEST INFRASTRUCTURE ONLY (the
checker; never imported by the product).

Follows cpp/include/whir/transcript/transcript.hpp and the challenge helpers:
  * Shake128DuplexSponge (:80-139): absorb appends to one SHAKE-128 input and
    drops the output reader; a squeeze reads SHAKE-128(all absorbed bytes)
    from the start after an absorb, and continues the stream otherwise.
  * ProverState (:595-662): prover_message absorbs the value's raw bytes and
    appends them to the narg string; prover_hint appends to the hints only;
    vectors are framed with an 8-byte little-endian element count
    (narg_serialize, :263-330).
  * verifier_message<Goldilocks> (:401-453): 40 bytes as a little-endian
    integer mod p; Ext3 = three such values.
  * DomainSeparator::absorb_into (:531-538): protocol_id, session_id, instance.
  * challenge_indices / indices_from_entropy
    (protocols/challenge_indices.hpp:54-119) and geometric_challenge
    (protocols/geometric_challenge.hpp).
SHAKE-128 and SHA3 come from Python's hashlib (FIPS 202), pinned separately
against the reference's transcript golden.
"""
from __future__ import annotations

import hashlib

P = 0xFFFFFFFF00000001


class Transcript:
    def __init__(self, protocol_id: bytes | None = None, session_id: bytes | None = None,
                 instance: bytes = b""):
        self.absorbed = bytearray()
        self.pos = None  # read position of the current output stream, None = no reader
        self.narg = bytearray()
        self.hints = bytearray()
        if protocol_id is not None:
            self.absorb(protocol_id)
            self.absorb(session_id)
            self.absorb(instance)

    def absorb(self, data: bytes) -> None:
        self.pos = None
        self.absorbed += data

    def squeeze(self, n: int) -> bytes:
        if n == 0:
            return b""
        if self.pos is None:
            self.pos = 0
        out = hashlib.shake_128(bytes(self.absorbed)).digest(self.pos + n)[self.pos:]
        self.pos += n
        return out

    def message(self, data: bytes) -> None:
        self.absorb(data)
        self.narg += data

    def message_u64(self, v: int) -> None:
        self.message(v.to_bytes(8, "little"))

    def message_e3(self, v) -> None:
        self.message(b"".join(int(c).to_bytes(8, "little") for c in v))

    def hint(self, data: bytes) -> None:
        self.hints += data

    def hint_vector(self, data: bytes, count: int) -> None:
        self.hints += count.to_bytes(8, "little") + data

    def challenge_gl(self) -> int:
        return int.from_bytes(self.squeeze(40), "little") % P

    def challenge_e3(self) -> tuple[int, int, int]:
        return (self.challenge_gl(), self.challenge_gl(), self.challenge_gl())

    def indices(self, num_leaves: int, count: int, dedup: bool) -> list[int]:
        if count == 0:
            return []
        if num_leaves == 1:
            return [0] if dedup else [0] * count
        nb = ((num_leaves.bit_length() - 1) + 7) // 8
        return indices_from_entropy(self.squeeze(count * nb), num_leaves, count, dedup)

    def geometric(self, count: int) -> list[tuple[int, int, int]]:
        if count == 0:
            return []
        if count == 1:
            return [(1, 0, 0)]
        x = self.challenge_e3()
        out, cur = [], (1, 0, 0)
        for _ in range(count):
            out.append(cur)
            cur = e3_mul(cur, x)
        return out


def indices_from_entropy(entropy: bytes, num_leaves: int, count: int, dedup: bool) -> list[int]:
    """challenge_indices.hpp:54-85: big-endian groups of ceil(log2(n)/8) bytes, mod n."""
    if count == 0:
        return []
    if num_leaves == 1:
        return [0] if dedup else [0] * count
    nb = ((num_leaves.bit_length() - 1) + 7) // 8
    assert len(entropy) == count * nb
    out = [int.from_bytes(entropy[i * nb:(i + 1) * nb], "big") % num_leaves for i in range(count)]
    return sorted(set(out)) if dedup else out


def e3_mul(a, b) -> tuple[int, int, int]:
    """(a0 + a1 X + a2 X^2)(b0 + b1 X + b2 X^2) mod X^3 - 2 (goldilocks_ext3.hpp:83-89)."""
    c = [0] * 5
    for i in range(3):
        for j in range(3):
            c[i + j] += a[i] * b[j]
    return ((c[0] + 2 * c[3]) % P, (c[1] + 2 * c[4]) % P, c[2] % P)
