"""Serving: each result is the same alone and in the mix (bit-identical), and equals the
single-request entry points; admission under a small memory budget still completes every request."""
import numpy as np
import pytest

import decontamination_amd as dam
from decontamination_amd.serving import SOMServer, pinned_empty

pytestmark = pytest.mark.gpu


def make_soms():
    rng = np.random.default_rng(7)
    soms = []
    for (m, n, d, topo) in [(16, 16, 16, "hexagonal"), (10, 12, 32, "square"), (32, 32, 16, "hexagonal")]:
        s = dam.SOM_Batch(m, n, d, topology=topo, precision="bf16_rerank")
        s._weights[...] = rng.standard_normal((m * n, d)).astype(np.float32)
        soms.append(s)
    return soms


def make_requests(soms):
    rng = np.random.default_rng(11)
    sizes = [0, 1, 255, 256, 257, 1000, 4096, 50_000, 200_001]
    reqs = []
    for i, rows in enumerate(sizes * 2):
        mi = i % len(soms)
        d = soms[mi].dim
        x = pinned_empty((rows, d), np.float32)
        x[...] = rng.standard_normal((rows, d)).astype(np.float32)
        if rows > 10:
            x[3] = np.nan
        kind = "winners" if i % 2 == 0 else "activation"
        rho = None
        if kind == "activation" and i % 4 == 1 and rows > 0:
            rho = rng.random(rows) * 2.0
            rho[::9] = 0.0
        reqs.append((mi, kind, x, rho))
    # uploads of several 16 MB pieces (row counts not multiples of the 512-row tile): their first tiles
    # are computed while the later pieces are still on the way
    for mi, kind, rows in ((1, "activation", 600_001), (0, "winners", 700_003)):
        d = soms[mi].dim
        x = pinned_empty((rows, d), np.float32)
        x[...] = rng.standard_normal((rows, d)).astype(np.float32)
        rho = rng.random(rows) * 2.0 if kind == "activation" else None
        reqs.append((mi, kind, x, rho))
    return reqs


def run_alone(server, reqs):
    out = []
    for mi, kind, x, rho in reqs:
        t = server.submit(mi, kind, x, rho)
        out.append(server.wait(t)[0].copy())
    return out


def run_mixed(server, reqs):
    tickets = [server.submit(mi, kind, x, rho) for mi, kind, x, rho in reqs]
    return [server.wait(t)[0].copy() for t in tickets]


def test_alone_equals_mixed_equals_single_entry_points():
    soms = make_soms()
    reqs = make_requests(soms)
    with SOMServer(soms, memory_budget=2 << 30) as server:
        a = run_alone(server, reqs)
        b = run_mixed(server, reqs)
        st = server.stats()
    assert st["completed"] == 2 * len(reqs)
    for (mi, kind, x, rho), ra, rb in zip(reqs, a, b):
        assert np.array_equal(ra, rb, equal_nan=True)
        som = soms[mi]
        if kind == "winners":
            assert ra.dtype == np.int32 and ra.shape == (x.shape[0],)
            if x.shape[0]:
                assert np.array_equal(ra, som.get_winners(np.asarray(x)))
        else:
            assert ra.shape == (som.m, som.n) and ra.dtype == np.float32
            assert np.array_equal(ra, som.get_activation_map(np.asarray(x), rho))


def test_small_budget_admission():
    soms = make_soms()[:1]
    rng = np.random.default_rng(3)
    xs = [np.ascontiguousarray(rng.standard_normal((30_000, 16)).astype(np.float32)) for _ in range(12)]
    ref = [soms[0].get_winners(x) for x in xs]
    # one request needs about 30000*(64+4) bytes = 2 MB: a 3 MB budget admits one at a time
    with SOMServer(soms, memory_budget=3 << 20) as server:
        tickets = [server.submit(0, "winners", x) for x in xs]
        res = [server.wait(t) for t in tickets]
        with pytest.raises(ValueError):
            server.submit(0, "winners", np.zeros((100_000, 16), np.float32))
    for (r, lat), e in zip(res, ref):
        assert np.array_equal(r, e)
        assert lat > 0.0


def test_serving_config_models_vs_oracle():
    """The serving configuration's three resident models (32x32/D16, 64x64/D16, 100x100/D32, trained on
    catalogue-like rows so that near-ties are common) under a mixed load that includes 4M-row requests:
    served winners equal the oracle's on a row sample of every request (REF32 for the default server, FP64
    for a server with FP64 rerank arithmetic); activation maps equal the oracle's maps on whole requests
    up to 100k rows, and on the 4M-row requests equal the counts of the same rows' served winners."""
    from decontamination_amd import synth
    from oracle import FP64, REF32
    from oracle import som_ref as R
    x16 = synth.catalogue_rows(4_000_000, 8, 6, seed=31)
    x32 = synth.catalogue_rows(4_000_000, 16, 14, seed=32)
    soms = []
    for m, data in ((32, x16), (64, x16), (100, x32)):
        s = dam.SOM_Batch(m, m, data.shape[1])
        s.init_rand(seed=0)
        s.train(data[: 1 << 19], n_epochs=3)
        soms.append(s)
    pools = {16: pinned_empty(x16.shape, np.float32), 32: pinned_empty(x32.shape, np.float32)}
    pools[16][...] = x16
    pools[32][...] = x32
    rng = np.random.default_rng(5)
    plan = [(0, "winners", 1_000), (1, "winners", 4_000_000), (2, "activation", 100_000), (1, "activation", 77_777),
            (2, "winners", 4_000_000), (0, "activation", 100_000), (1, "winners", 333_333), (2, "winners", 2_048),
            (1, "activation", 4_000_000), (0, "winners", 1_234_567)]
    for arith, mode in (("ref32", REF32), ("fp64", FP64)):
        print(f"serving config models, {arith} rerank", flush=True)
        reqs = []
        for mi, kind, rows in plan:
            pool = pools[soms[mi].dim]
            off = int(rng.integers(0, pool.shape[0] - rows + 1))
            reqs.append((mi, kind, pool[off:off + rows]))
        with SOMServer(soms, memory_budget=64 << 30, arithmetic=arith) as server:
            res = run_mixed(server, [(mi, kind, x, None) for mi, kind, x in reqs])
            wins = {}
            for i, (mi, kind, x) in enumerate(reqs):  # winners of the rows of the 4M-row activation request
                if kind == "activation" and x.shape[0] > 100_000:
                    wins[i] = server.wait(server.submit(mi, "winners", x))[0].copy()
        for i, ((mi, kind, x), r) in enumerate(zip(reqs, res)):
            W = soms[mi].weights
            if kind == "winners":
                sel = np.sort(rng.choice(x.shape[0], size=min(4096, x.shape[0]), replace=False))
                exp, _, _ = R.top2_omp(W, np.ascontiguousarray(x[sel]), mode)
                assert np.array_equal(r[sel], exp), (arith, mi, x.shape[0], int(np.sum(r[sel] != exp)))
            elif x.shape[0] <= 100_000:
                exp = R.activation_map_omp(W, x, None, mode).reshape(soms[mi].m, soms[mi].n)
                assert np.array_equal(r, exp.astype(np.float32)), (arith, mi, x.shape[0])
            else:
                w = wins[i]
                sel = np.sort(rng.choice(x.shape[0], size=4096, replace=False))
                exp, _, _ = R.top2_omp(W, np.ascontiguousarray(x[sel]), mode)
                assert np.array_equal(w[sel], exp)
                counts = np.bincount(w, minlength=W.shape[0]).astype(np.float32).reshape(soms[mi].m, soms[mi].n)
                assert np.array_equal(r, counts) and r.sum() == x.shape[0]
