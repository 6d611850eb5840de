"""GPU tests of grouped launches, captured-graph replay, the autotuner and the
serving engine. Bars as in test_gpu_parity.py; in addition every path here must
agree bit for bit with the single-chain launch of the same request (each tile's
arithmetic does not depend on which launch or batch carries it)."""
import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O
from test_gpu_parity import assert_rel, TOL_BF16, TOL_F32, dev, empty, rand_chain

pytestmark = pytest.mark.gpu


def make_requests(rng, count, fitting=False):
    """Mixed dtypes, chain lengths 1..4, uneven N, M and K (multiples of 32 and not).
    fitting=True keeps every request on the grouped on-chip path (N % 4 == 0 and
    widths in whole 16-byte groups for every dtype)."""
    import torch
    reqs = []
    widths = [32, 48, 64, 96, 128, 256] if fitting else [32, 48, 64, 96, 128, 200, 256]
    for r in range(count):
        dtype = [fg.DTYPE_BF16, fg.DTYPE_INT8, fg.DTYPE_F32][r % 3]
        n = 1 + r % 4
        dims = [int(rng.choice(widths)) for _ in range(n + 1)]
        N = int(rng.integers(1, 300))
        if fitting:
            N = 4 * (N // 4 + 1)
        B, As = rand_chain(rng, dtype, dims[0], N, dims[1:])
        C = empty((dims[-1], N), torch.int32 if dtype == fg.DTYPE_INT8 else torch.float32)
        dB, dA = dev(B), [dev(a) for a in As]
        reqs.append(dict(dtype=dtype, B=B, As=As, C=C, dB=dB, dA=dA, N=N,
                         req=fg.chain_req(dtype, C, dB, dA, N=N, Ms=[a.shape[0] for a in As],
                                          Ks=[a.shape[1] for a in As])))
    return reqs


def check_against_oracle(q, got):
    if q["dtype"] == fg.DTYPE_INT8:
        assert np.array_equal(got, O.chain_int8(q["B"], q["As"], O.LAYOUT_MATH))
    elif q["dtype"] == fg.DTYPE_BF16:
        assert_rel(got, O.chain_bf16(q["B"], q["As"], O.LAYOUT_MATH), TOL_BF16)
    else:
        assert_rel(got, O.chain_f32(q["B"], q["As"]), TOL_F32)


def single_launch(q):
    """The same chain alone through fg_chain with default options: the planner picks
    the fused path exactly when the request fits a grouped tile (same k order), and
    otherwise the same unfused plan the grouped launch falls back to."""
    import torch
    C = torch.empty_like(q["C"])
    fg.chain(q["dtype"], C, q["dB"], q["dA"], N=q["N"], Ms=[a.shape[0] for a in q["As"]],
             Ks=[a.shape[1] for a in q["As"]])
    torch.cuda.synchronize()
    return C.cpu().numpy()


def test_grouped_mixed_batch_matches_oracle_and_single_launch(gpu):
    import torch
    rng = np.random.default_rng(42)
    reqs = make_requests(rng, 24)
    print("grouped batch of", len(reqs), flush=True)
    fg.grouped_chain([q["req"] for q in reqs])
    torch.cuda.synchronize()
    for q in reqs:
        got = q["C"].cpu().numpy()
        check_against_oracle(q, got)
        assert np.array_equal(got.view(np.uint32), single_launch(q).view(np.uint32))


def test_grouped_prepacked_requests_same_bits(gpu):
    """Requests whose weights were packed beforehand (a_packed=1) ride in the same
    grouped launch as ones the launch packs itself, with the same bits as before."""
    import torch
    rng = np.random.default_rng(17)
    reqs = make_requests(rng, 12, fitting=True)
    fg.grouped_chain([q["req"] for q in reqs])
    torch.cuda.synchronize()
    want = [q["C"].cpu().numpy() for q in reqs]
    keep, mixed = [], []
    for j, q in enumerate(reqs):
        q["C"].zero_()
        if j % 2:
            mixed.append(q["req"])
            continue
        Ms, Ks = [a.shape[0] for a in q["As"]], [a.shape[1] for a in q["As"]]
        packed = [empty((fg.packed_weight_bytes(q["dtype"], m, k),), torch.uint8) for m, k in zip(Ms, Ks)]
        for a, m, k, p in zip(q["dA"], Ms, Ks, packed):
            fg.pack_weight(q["dtype"], a, m, k, p)
        keep.append(packed)
        mixed.append(fg.chain_req(q["dtype"], q["C"], q["dB"], packed, N=q["N"], Ms=Ms, Ks=Ks, a_packed=1))
    print("grouped batch, half pre-packed:", len(mixed), flush=True)
    fg.grouped_chain(mixed)
    torch.cuda.synchronize()
    for q, w in zip(reqs, want):
        assert np.array_equal(q["C"].cpu().numpy().view(np.uint32), w.view(np.uint32))


def test_grouped_oversized_request_runs_unfused(gpu):
    import torch
    rng = np.random.default_rng(3)
    B, As = rand_chain(rng, fg.DTYPE_F32, 2048, 40, [2048, 64])  # 4096-row arena: too big at tile 32
    C = empty((64, 40), torch.float32)
    req = fg.chain_req(fg.DTYPE_F32, C, dev(B), [dev(a) for a in As], N=40)
    assert not fg.grouped_fits(req)
    fg.grouped_chain([req])
    torch.cuda.synchronize()
    assert_rel(C.cpu().numpy(), O.chain_f32(B, As), TOL_F32)


def test_graph_replays_match_eager(gpu):
    """Replay the captured graph for several batches of different sizes, with eager
    work and synchronisations in between; every replay must equal the eager result."""
    import torch
    rng = np.random.default_rng(9)
    g = fg.GroupedGraph(max_reqs=64, max_tiles=4096)
    try:
        for rep, count in enumerate([5, 17, 3, 17]):
            reqs = make_requests(rng, count, fitting=True)
            assert all(fg.grouped_fits(q["req"]) for q in reqs)
            print(f"replay {rep}: {count} requests", flush=True)
            g.launch([q["req"] for q in reqs])
            torch.cuda.synchronize()
            # eager work between replays
            x = torch.randn(512, 512, device="cuda")
            (x @ x).sum().item()
            for q in reqs:
                got = q["C"].cpu().numpy()
                check_against_oracle(q, got)
                assert np.array_equal(got.view(np.uint32), single_launch(q).view(np.uint32))
    finally:
        g.close()


def test_graph_rejects_oversized(gpu):
    import torch
    rng = np.random.default_rng(4)
    B, As = rand_chain(rng, fg.DTYPE_BF16, 2048, 8, [2048, 2048])
    C = empty((2048, 8), torch.float32)
    req = fg.chain_req(fg.DTYPE_BF16, C, dev(B), [dev(a) for a in As], N=8)
    g = fg.GroupedGraph(max_reqs=4, max_tiles=64)
    try:
        with pytest.raises(fg.FlashgemmError) as e:
            g.launch([req])
        assert e.value.status == -7
    finally:
        g.close()


def test_autotune_table_roundtrip_and_planner(gpu, tmp_path):
    fg.autotune_clear()
    Ms, Ks, N = [256, 256], [256, 256], 8192
    best = fg.autotune_chain(fg.DTYPE_F32, N, Ms, Ks)
    print("autotuned", best, flush=True)
    assert fg.autotune_size() == 1
    fused, tn = fg.chain_plan(fg.DTYPE_F32, N, Ms, Ks)
    assert fused == best["fused"] and (not fused or tn == best["tile_n"])
    path = str(tmp_path / "table.txt")
    fg.autotune_save(path)
    fg.autotune_clear()
    assert fg.autotune_size() == 0
    fg.autotune_load(path)
    assert fg.autotune_size() == 1
    assert fg.chain_plan(fg.DTYPE_F32, N, Ms, Ks) == (fused, tn)
    (tmp_path / "bad.txt").write_text("flashgemm-autotune v0 gfx950\n")
    with pytest.raises(fg.FlashgemmError):
        fg.autotune_load(str(tmp_path / "bad.txt"))
    fg.autotune_clear()


def _server_setup(rng, srv):
    """Three resident weight sets (bf16, int8, f32) and device inputs for requests."""
    sets = []
    for dtype, dims in ((fg.DTYPE_BF16, [128, 256, 64]), (fg.DTYPE_INT8, [64, 128, 96, 64]),
                        (fg.DTYPE_F32, [96, 64, 128])):
        _, As = rand_chain(rng, dtype, dims[0], 1, dims[1:])
        dA = [dev(a) for a in As]
        sid = srv.add_weights(dtype, dA)
        sets.append(dict(dtype=dtype, As=As, dA=dA, sid=sid, K1=dims[0], M=dims[-1]))
    return sets


def _request(rng, s, N):
    """Host B, device B and C. The server's streams do not wait for torch's stream,
    so the buffers are made ready (synchronize) before they are submitted."""
    import torch
    B, _ = rand_chain(rng, s["dtype"], s["K1"], N, [1])
    C = torch.zeros((s["M"], N), dtype=torch.int32 if s["dtype"] == fg.DTYPE_INT8 else torch.float32,
                    device="cuda")
    dB = dev(B)
    torch.cuda.synchronize()
    return B, dB, C


def test_server_result_same_alone_and_in_the_mix(gpu):
    import torch
    rng = np.random.default_rng(77)
    srv = fg.Server(num_streams=2, max_batch=64)
    try:
        sets = _server_setup(rng, srv)
        probe_B, probe_dB, probe_C_alone = _request(rng, sets[0], 257)
        srv.submit(sets[0]["sid"], probe_dB, probe_C_alone)
        srv.drain()
        alone = probe_C_alone.cpu().numpy()
        # the same request again inside a mixed trace
        probe_C_mix = torch.zeros_like(probe_C_alone)
        ids, Ns, Bs, Cs, hosts = [], [], [], [], []
        for i in range(60):
            s = sets[i % 3]
            N = int(rng.integers(1, 400))
            hB, dB, C = _request(rng, s, N)
            ids.append(s["sid"]); Ns.append(N); Bs.append(dB); Cs.append(C); hosts.append((s, hB))
            if i == 30:
                ids.append(sets[0]["sid"]); Ns.append(257); Bs.append(probe_dB); Cs.append(probe_C_mix)
                hosts.append((sets[0], probe_B))
        arrivals = sorted(rng.uniform(0, 0.02, len(ids)).tolist())
        torch.cuda.synchronize()  # probe_C_mix's fill is done before the engine writes it
        srv.run_trace(ids, Ns, Bs, Cs, arrivals)
        mix = probe_C_mix.cpu().numpy()
        assert np.array_equal(alone.view(np.uint32), mix.view(np.uint32))
        lat = srv.latencies()
        assert len(lat) == len(ids) + 1 and min(lat) > 0
        for (s, hB), C in zip(hosts, Cs):
            got = C.cpu().numpy()
            check_against_oracle(dict(dtype=s["dtype"], B=hB, As=s["As"]), got)
    finally:
        srv.close()


def test_server_budget_and_slots(gpu):
    rng = np.random.default_rng(5)
    srv = fg.Server(budget_bytes=4 << 20, max_inflight=1, num_streams=1, use_graph=False)
    try:
        sets = _server_setup(rng, srv)
        _, dB, C = _request(rng, sets[1], 64)
        for _ in range(5):  # one slot: requests run one after another, all complete
            srv.submit(sets[1]["sid"], dB, C)
        srv.drain()
        _, dBbig, Cbig = _request(rng, sets[0], 16000)  # > 4 MiB of B + C
        with pytest.raises(fg.FlashgemmError) as e:
            srv.submit(sets[0]["sid"], dBbig, Cbig)
        assert e.value.status == -5
    finally:
        srv.close()


def test_server_per_gemm_baseline_matches(gpu):
    rng = np.random.default_rng(6)
    srv = fg.Server(per_gemm_baseline=True, num_streams=1, max_batch=1)
    try:
        sets = _server_setup(rng, srv)
        reqs = []
        for i in range(6):
            s = sets[i % 3]
            hB, dB, C = _request(rng, s, 100 + i)
            srv.submit(s["sid"], dB, C)
            reqs.append((s, hB, dB, C))  # dB stays referenced until the request completes
        srv.drain()
        for s, hB, _, C in reqs:
            check_against_oracle(dict(dtype=s["dtype"], B=hB, As=s["As"]), C.cpu().numpy())
    finally:
        srv.close()
