#!/usr/bin/env python3
"""Write tests/golden/proofs.json: digests of whole proofs made by the
reference's own prover (built from its sources by `make -C oracle ref` into
oracle/_ref/), for seeded inputs.  Each case records the parameters, the seed
(coefficients and constraint points drawn with whirplonk_amd.lcg_uniform), the
constraint evaluations and the SHA-256 of the narg string and of the hints.

Usage:  python tests/golden/make_proof_fixtures.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
import whirplonk_amd as W  # noqa: E402

CASES = [
    # the full-prove workload: 2^22 coefficients, rate 1/2, folding factor 4, 3 MLE constraints
    (22, dict(nforms=3), 0x5748495200000004),
    (16, dict(nforms=3), 0x5748495200000016),
    (12, dict(nforms=2, rate=2), 0x5748495200000012),
    (11, dict(nforms=2, batch=3), 0x5748495200000011),
]


def inputs(nv, nforms, batch, seed):
    coeffs, st = W.lcg_uniform(seed, batch << nv)
    pts, _ = W.lcg_uniform(st, 3 * nv * nforms)
    return coeffs, pts.reshape(nforms, nv, 3) if nforms else np.zeros((0, nv, 3), dtype=np.uint64)


def main() -> None:
    if oracle.ref() is None:
        raise SystemExit("build the reference first: make -C oracle ref")
    out = []
    for nv, prm, seed in CASES:
        kw = dict(prm)
        nforms = kw.pop("nforms")
        coeffs, pts = inputs(nv, nforms, kw.get("batch", 1), seed)
        ev, narg, hints = oracle.ref_prove(nv, coeffs, pts, **kw)
        out.append({"nv": nv, "params": prm, "seed": seed,
                    "evaluations": [int(v) for v in ev.reshape(-1)],
                    "narg_len": len(narg), "hints_len": len(hints),
                    "narg_sha256": hashlib.sha256(narg).hexdigest(),
                    "hints_sha256": hashlib.sha256(hints).hexdigest()})
        print(nv, prm, len(narg), len(hints))
    path = os.path.join(HERE, "proofs.json")
    with open(path, "w") as f:
        json.dump({"source": "reference prover built from cpp/ sources (oracle/Makefile ref)",
                   "cases": out}, f, indent=1)
        f.write("\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
