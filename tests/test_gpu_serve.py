"""The serving loop: every proof produced in a mix of concurrent requests of
different sizes, admitted under an HBM budget, is byte-identical to the same
request proved alone on the plain path (and to the reference build)."""
import random

import numpy as np
import pytest

import oracle
import whirplonk_amd as W

pytestmark = pytest.mark.gpu


def _request(nv, nforms, seed):
    coeffs, st = W.lcg_uniform(seed, 1 << nv)
    pts, _ = W.lcg_uniform(st, 3 * nv * nforms)
    return coeffs, pts.reshape(nforms, nv, 3)


def _alone(ctx, params, nv, coeffs, pts):
    pid, sid = W.domain_separator_cli(params)
    d = ctx.upload(coeffs)
    ev = np.array([ctx.mle_evaluate(d, True, pts[f]) for f in range(len(pts))], dtype=np.uint64)
    with W.Prover(ctx, params, 1 << nv, pid, sid) as pr:
        pr.commit(d, 1)
        pr.prove(pts, ev)
        out = pr.proof()
    ctx.free(d)
    return out, ev


@pytest.mark.parametrize("slots", [1, 4])
def test_mix_equals_alone(slots):
    params = W.Params.default(starting_log_inv_rate=2)
    rnd = random.Random(slots)
    sizes = [rnd.choice([8, 10, 12, 14]) for _ in range(24)]
    reqs = [(nv, rnd.randrange(4)) for nv in sizes]
    inputs = []
    for i, (nv, nf) in enumerate(reqs):
        c, p = _request(nv, nf, 0xC0FFEE00 + i)
        inputs.append((nv, c, p))
    ctx = W.Context()
    alone = [_alone(ctx, params, nv, c, p) for nv, c, p in inputs]
    ctx.close()
    # a budget of about two of the largest requests: forces queueing
    with W.Server(params, num_slots=slots, hbm_budget=1 << 20) as probe:
        big = probe.request_bytes(1 << 14)
    budget = 2 * big + (1 << 20)
    print("serve", slots, "budget", budget, flush=True)
    with W.Server(params, num_slots=slots, hbm_budget=budget) as srv:
        ids = []
        for i, (nv, c, p) in enumerate(inputs):
            ev = alone[i][1] if i % 2 else None  # half computed on the device
            ids.append(srv.submit(1 << nv, c, p, ev, borrow=(i % 3 == 0)))
        for i, rid in enumerate(ids):
            narg, hints, ms = srv.wait(rid)
            assert (narg, hints) == alone[i][0], i
            assert ms[2] >= ms[1] >= 0
        st = srv.stats()
        assert st["completed"] == len(inputs)
        assert st["peak"] <= budget and st["in_use"] == 0
        # every proof-of-work search went through the batcher, and the
        # initial commits' encodes of the 2^14 requests through the other
        assert st["pow_searches"] >= st["pow_launches"] > 0
        assert st["encodes"] == sizes.count(14) and st["encodes"] >= st["encode_launches"]
        assert (st["encode_launches"] > 0) == (14 in sizes)


def test_served_vs_reference():
    if oracle.ref() is None:
        pytest.skip("reference build not available")
    params = W.Params.default()
    with W.Server(params, num_slots=3, hbm_budget=1 << 30) as srv:
        cases = []
        for i, nv in enumerate([12, 16, 13, 15]):
            c, p = _request(nv, 3, 0xABCD00 + i)
            cases.append((nv, c, p, srv.submit(1 << nv, c, p)))
        for nv, c, p, rid in cases:
            _, narg, hints = oracle.ref_prove(nv, c, p)
            assert srv.wait(rid)[:2] == (narg, hints)


def test_served_proofs_verify():
    """Every proof of a mixed served batch is accepted by the reference
    verifier (the claimed values from the GPU's MLE evaluation)."""
    if oracle.ref() is None:
        pytest.skip("reference build not available")
    params = W.Params.default()
    rnd = random.Random(0x5E7)
    sizes = [rnd.choice([12, 14, 16, 18, 20, 22]) for _ in range(12)]
    with W.Server(params, num_slots=4, hbm_budget=4 << 30) as srv, W.Context() as ctx:
        reqs = []
        for i, nv in enumerate(sizes):
            c, p = _request(nv, 2, 0x5E7000 + i)
            reqs.append((nv, c, p, srv.submit(1 << nv, c, p)))
        for nv, c, p, rid in reqs:
            narg, hints, _ = srv.wait(rid)
            d = ctx.upload(c)
            ev = np.array([ctx.mle_evaluate(d, True, p[f]) for f in range(len(p))], dtype=np.uint64)
            ctx.free(d)
            print("served verify", nv, flush=True)
            assert oracle.ref_verify(nv, p, ev, narg, hints)


def test_request_larger_than_budget():
    params = W.Params.default()
    with W.Server(params, num_slots=1, hbm_budget=1 << 20) as srv:
        c, p = _request(16, 1, 1)
        with pytest.raises(W.WhirGpuError):
            srv.submit(1 << 16, c, p)


SERVE_SIZES, SERVE_WEIGHTS, SERVE_SEED = (16, 18, 20, 22, 24), (0.3, 0.3, 0.2, 0.15, 0.05), 0x5748495200000005


def test_served_workload_mix():
    """The served workload's mix at rate 1/4: 40 requests with num_variables
    drawn from {16, 18, 20, 22, 24} (weights .3/.3/.2/.15/.05, seeded, at
    least one 2^24 request), 16 slots, 3 MLE constraints each with the
    evaluations computed on the device.  Every served proof equals the plain
    one-at-a-time path's; the first request of each size equals the
    reference prover's proof (whir_prover.hpp:42-400); every proof is
    accepted by the reference verifier."""
    rnd = random.Random(SERVE_SEED)
    sizes = rnd.choices(SERVE_SIZES, weights=SERVE_WEIGHTS, k=40)
    if 24 not in sizes:
        sizes[-1] = 24
    params = W.Params.default(starting_log_inv_rate=2)
    inputs = [_request(nv, 3, SERVE_SEED + 97 * i) for i, nv in enumerate(sizes)]
    print("served mix", sizes, flush=True)
    ctx = W.Context()
    alone = [_alone(ctx, params, nv, c, p) for nv, (c, p) in zip(sizes, inputs)]
    ctx.close()
    with W.Server(params, num_slots=16, hbm_budget=24 << 30) as srv:
        ids = [srv.submit(1 << nv, c, p, borrow=(i % 2 == 0)) for i, (nv, (c, p)) in enumerate(zip(sizes, inputs))]
        got = [srv.wait(i)[:2] for i in ids]
        st = srv.stats()
    assert st["completed"] == len(sizes) and st["in_use"] == 0
    # the initial commits of every request up to 2^22 coefficients went
    # through the encode batcher
    print("encodes", st["encodes"], "in", st["encode_launches"], "launches", flush=True)
    assert st["encodes"] == sum(1 for nv in sizes if nv <= 22) and st["encode_launches"] > 0
    for i in range(len(sizes)):
        assert got[i] == alone[i][0], (i, sizes[i])
    if oracle.ref() is None:
        pytest.skip("reference build not available")
    for nv in sorted(set(sizes)):
        i = sizes.index(nv)
        print("served vs reference prover", nv, flush=True)
        _, narg, hints = oracle.ref_prove(nv, inputs[i][0], inputs[i][1], rate=2)
        assert got[i] == (narg, hints), nv
    for i, nv in enumerate(sizes):
        assert oracle.ref_verify(nv, inputs[i][1], alone[i][1], *got[i], rate=2), (i, nv)


def test_failed_requests_are_isolated():
    """Requests made to fail part-way (at their 1st and 5th device waits,
    with copies and kernels still queued; one reads a borrowed buffer that
    is a converted temporary) report the error; their HBM blocks go back to
    the arena and are reused by later requests, and every other proof, before
    and after, is byte-identical to the same request proved alone."""
    params = W.Params.default(starting_log_inv_rate=2)
    rnd = random.Random(0xFA11)
    sizes = [rnd.choice([10, 12, 14]) for _ in range(14)]
    inputs = [_request(nv, 2, 0xFA1100 + i) for i, nv in enumerate(sizes)]
    ctx = W.Context()
    alone = [_alone(ctx, params, nv, c, p) for nv, (c, p) in zip(sizes, inputs)]
    ctx.close()
    with W.Server(params, num_slots=1, hbm_budget=1 << 20) as probe:
        big = probe.request_bytes(1 << 14)
    with W.Server(params, num_slots=3, hbm_budget=2 * big + (1 << 20)) as srv:
        fail = {3: 1, 6: 5, 11: 2}  # request id (1-based submit order) -> failing wait
        for rid, at in fail.items():
            srv.inject_failure(rid, at)
        ids = []
        for i, (nv, (c, p)) in enumerate(zip(sizes, inputs)):
            src = [int(x) for x in c] if i + 1 == 6 else c  # id 6 borrows a list: a converted copy
            ids.append(srv.submit(1 << nv, src, p, borrow=(i % 2 == 1 or i + 1 == 6)))
        assert ids == list(range(1, len(sizes) + 1))
        for i, rid in enumerate(ids):
            if rid in fail:
                with pytest.raises(W.WhirGpuError, match="injected failure"):
                    srv.wait(rid)
            else:
                assert srv.wait(rid)[:2] == alone[i][0], i
        st = srv.stats()
        assert st["in_use"] == 0 and st["running"] == 0 and st["completed"] == len(sizes)
        # the released blocks serve new requests; proofs unchanged
        again = [srv.submit(1 << nv, c, p) for nv, (c, p) in zip(sizes[:6], inputs[:6])]
        for i, rid in enumerate(again):
            assert srv.wait(rid)[:2] == alone[i][0], ("again", i)
