"""Batched uneven SPD factor + solve (mi_dposv_vbatched / mi_sposv_vbatched).

Each system is checked against the oracle's restatement of llt_inplace::blocked / unblocked
and LLT::_solve_impl (oracle/llt_oracle.c) on the same input: the factor within 1e-12
relative to ||L||, solve backward error <= 1e-14 (FP64); FP32 against the FP64 oracle within
FP32 bounds.  Failing pivots, argument errors, the upper triangle left untouched, the
longest-first schedule, graph replay and the autotuner are covered too."""
import numpy as np
import pytest
import torch

import mi_llt
import oracle
from mi_llt import batched, workloads

pytestmark = pytest.mark.gpu


def fill(vb, seed=0, shift=0.1, upper_junk=True):
    rng = np.random.default_rng(seed)
    mats, rhss = [], []
    for i in range(vb.count):
        n, r = int(vb.ns[i]), int(vb.rs[i])
        a = workloads.gram_spd(n, seed=seed + i, shift=shift) if n else np.zeros((0, 0))
        if upper_junk and n > 1:
            a = a.copy()
            a[np.triu_indices(n, 1)] = 1e30  # must never be read
        b = rng.standard_normal((n, r))
        if n:
            vb.A(i).copy_(torch.from_numpy(a))
        if n * r:
            vb.B(i).copy_(torch.from_numpy(b))
        mats.append(a)
        rhss.append(b)
    return mats, rhss


def check_system(vb, i, a, b, ftol, stol):
    n, r = int(vb.ns[i]), int(vb.rs[i])
    if n == 0:
        return
    asym = np.tril(a) + np.tril(a, -1).T
    l = np.tril(vb.A(i).cpu().double().numpy())
    up = np.triu(vb.A(i).cpu().double().numpy(), 1)
    assert np.all(up == np.triu(a, 1).astype(up.dtype)), f"system {i}: upper triangle modified"
    m_ref, info = oracle.llt_blocked(np.asfortranarray(asym))
    assert info == -1
    l_ref = np.tril(m_ref)
    assert np.linalg.norm(l - l_ref) <= ftol * np.linalg.norm(l_ref), f"system {i} n={n}: factor"
    if r:
        x = vb.B(i).cpu().double().numpy()
        res = asym @ x - b
        be = (np.linalg.norm(res, axis=0) / (np.linalg.norm(asym) * np.linalg.norm(x, axis=0))).max()
        assert be <= stol, f"system {i} n={n} r={r}: solve berr {be:.2e}"


SIZES = [0, 1, 2, 3, 7, 8, 15, 16, 17, 31, 32, 33, 47, 64, 100, 127, 128, 129, 191, 200, 255, 256]


@pytest.mark.parametrize("group,lpt", [(4, True), (2, False)])
def test_vbatched_fp64_matches_oracle(dev, group, lpt):
    ns = SIZES + [5, 250, 96]
    rs = [(i * 5) % 9 for i in range(len(ns))]
    rs[-1] = 17  # two groups of right-hand sides
    print(f"[gpu] vbatched fp64 group={group} lpt={lpt} count={len(ns)}", flush=True)
    vb = batched.VBatch(ns, rs, device=dev)
    mats, rhss = fill(vb, seed=10)
    h = mi_llt.Handle()
    batched.set_config(h, 2, group, lpt)
    info = batched.posv_vbatched_(vb, h)
    assert torch.all(info == 0), info
    for i in range(vb.count):
        check_system(vb, i, mats[i], rhss[i], 1e-12, 1e-14)


def test_vbatched_fp32_within_fp32_bounds(dev):
    ns = SIZES
    rs = [1 + i % 8 for i in range(len(ns))]
    vb = batched.VBatch(ns, rs, dtype=torch.float32, device=dev)
    mats, rhss = fill(vb, seed=20)
    info = batched.posv_vbatched_(vb)
    assert torch.all(info == 0)
    for i in range(vb.count):
        n = int(vb.ns[i])
        # u_fp32 * n-ish bounds against the FP64 oracle
        check_system(vb, i, mats[i].astype(np.float32).astype(np.float64), rhss[i].astype(np.float32).astype(np.float64),
                     2e-5 * max(1.0, n / 16), 2e-6 * max(1.0, n / 16))


def test_vbatched_failing_pivots_and_arguments(dev):
    ns = [40, 100, 300, 50, 129, 64]
    rs = [2, 2, 1, 3, 1, 1]
    vb = batched.VBatch(ns, rs, device=dev)
    mats, rhss = fill(vb, seed=30)
    # system 0: indefinite at column 17; system 4: at column 128; system 2: n = 300 > 256
    for i, k in ((0, 17), (4, 128)):
        a = workloads.indefinite_at(int(ns[i]), k, seed=40 + i)
        vb.A(i).copy_(torch.from_numpy(a))
        mats[i] = a
    lda_bad = vb.d_lda.clone()
    lda_bad[3] = 10  # system 3: lda < n
    vb.d_lda.copy_(lda_bad)
    b_before = [vb.B(i).clone() for i in range(vb.count)]
    info = batched.posv_vbatched_(vb).cpu().tolist()
    assert info == [18, 0, -1, -2, 129, 0]
    for i in (0, 2, 3, 4):
        assert torch.equal(vb.B(i), b_before[i])  # B unchanged for failed / skipped systems
    for i in (1, 5):
        check_system(vb, i, mats[i], rhss[i], 1e-12, 1e-14)


def test_vbatched_graph_replay(dev):
    vb = batched.gram_batch(300, seed=3, device=dev)
    a0, b0 = vb.a_flat.clone(), vb.b_flat.clone()
    h = mi_llt.Handle()
    batched.posv_vbatched_(vb, h)  # eager: sizes the workspace
    torch.cuda.synchronize()
    ref_a, ref_b = vb.a_flat.clone(), vb.b_flat.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    vb.a_flat.copy_(a0)
    vb.b_flat.copy_(b0)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        batched.posv_vbatched_(vb, h)
    for rep in range(3):
        vb.a_flat.copy_(a0)
        vb.b_flat.copy_(b0)
        _ = (vb.a_flat * 2.0).sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(vb.a_flat, ref_a) and torch.equal(vb.b_flat, ref_b), f"replay {rep}"
        assert torch.all(vb.d_info == 0)


def test_vbatched_autotune_keeps_results(dev, tmp_path):
    vb = batched.gram_batch(200, seed=4, device=dev)
    a0, b0 = vb.a_flat.clone(), vb.b_flat.clone()
    h = mi_llt.Handle()
    batched.posv_vbatched_(vb, h)
    ref_a, ref_b = vb.a_flat.clone(), vb.b_flat.clone()
    vb.a_flat.copy_(a0)
    vb.b_flat.copy_(b0)
    out = batched.autotune(vb, h, candidates=[(1, 2, True), (2, 4, True), (2, 4, False)], reps=1,
                           cache=str(tmp_path / "tune.json"))
    assert tuple(out["config"]) == batched.get_config(h)
    batched.posv_vbatched_(vb, h)
    torch.cuda.synchronize()
    assert torch.equal(vb.a_flat, ref_a) and torch.equal(vb.b_flat, ref_b)  # deterministic per system
    again = batched.autotune(vb, h, cache=str(tmp_path / "tune.json"))  # from the cache
    assert again["config"] == out["config"]


def test_vbatched_full_workload_backward_error(dev):
    """20,000 systems, n ~ U{8..256}, 1..8 right-hand sides: every system's factor and solve
    backward errors (FP64, computed here with batched torch products per size class)."""
    vb = batched.gram_batch(20000, seed=0, device=dev)
    a0, b0 = vb.a_flat.clone(), vb.b_flat.clone()
    info = batched.posv_vbatched_(vb)
    assert int((info != 0).sum()) == 0
    worst_f, worst_s = 0.0, 0.0
    for n in np.unique(vb.ns):
        idx = np.nonzero(vb.ns == n)[0]
        n = int(n)
        ao = torch.from_numpy(vb.a_off[idx]).to(dev)
        pos = ao[:, None] + torch.arange(n * n, device=dev)[None, :]
        A = a0[pos].view(-1, n, n).transpose(1, 2)
        Lf = torch.tril(vb.a_flat[pos].view(-1, n, n).transpose(1, 2))
        rf = torch.linalg.matrix_norm(A - Lf @ Lf.transpose(1, 2).contiguous()) / torch.linalg.matrix_norm(A)
        worst_f = max(worst_f, float(rf.max()))
        for k in idx[:8]:  # solves: a sample of each size class
            r = int(vb.rs[k])
            o = int(vb.b_off[k])
            Ak = a0[int(vb.a_off[k]):int(vb.a_off[k]) + n * n].view(n, n).t()
            Bk = b0[o:o + n * r].view(r, n).t()
            Xk = vb.b_flat[o:o + n * r].view(r, n).t()
            be = float(((Ak @ Xk - Bk).norm(dim=0) / (torch.linalg.matrix_norm(Ak) * Xk.norm(dim=0))).max())
            worst_s = max(worst_s, be)
    assert worst_f <= 1e-14, worst_f
    assert worst_s <= 1e-14, worst_s
