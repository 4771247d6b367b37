"""Learned warm start kernels vs the CPU oracle (oracle/fno.py).

The oracle rounds both operands of every matrix product to bfloat16 as the
MFMA GEMM does, so forward outputs and gradients agree to FP32-accumulation
level (rel-L2 <= 2e-3, a few operands may round across a BF16 boundary
differently); against the unrounded float64 oracle the bar is 3e-2."""
import numpy as np
import pytest

from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a, np.float32), device="cuda")


@pytest.mark.parametrize("M,N,K,transA,transB,nb", [
    (64, 64, 32, False, False, (1, 1, 1)),
    (37, 91, 45, False, False, (1, 1, 1)),
    (100, 33, 70, True, False, (2, 1, 1)),
    (32, 130, 16, False, True, (2, 3, 1)),
    (5, 7, 300, True, True, (1, 2, 3)),
])
def test_gemm_matches_numpy(gpu, M, N, K, transA, transB, nb):
    from oracle.fno import round_bf16
    from wavesim_amd.learned import kernels as Kn
    rng = np.random.default_rng(0)
    nbt = int(np.prod(nb))
    A = rng.standard_normal((nbt, K, M) if transA else (nbt, M, K)).astype(np.float32)
    B = rng.standard_normal((nbt, N, K) if transB else (nbt, K, N)).astype(np.float32)
    Cm = rng.standard_normal((nbt, M, N)).astype(np.float32)
    sA = (1, M) if transA else (K, 1)
    sB = (1, K) if transB else (N, 1)
    def bstr(per):
        return (per * nb[1] * nb[2], per * nb[2], per)
    dA, dB, dC = _t(A), _t(B), _t(Cm)
    Kn.gemm(M, N, K, dA, sA + bstr(M * K), dB, sB + bstr(K * N), dC, (N, 1) + bstr(M * N),
            alpha=0.5, beta=-1.0, nb=nb)
    An = A.transpose(0, 2, 1) if transA else A
    Bn = B.transpose(0, 2, 1) if transB else B
    ref = 0.5 * np.einsum("bmk,bkn->bmn", round_bf16(An), round_bf16(Bn)) - Cm
    out = dC.cpu().numpy()
    assert np.allclose(out, ref, rtol=1e-5, atol=1e-4 * np.abs(ref).max())


def test_gemm_without_unit_strides_matches_numpy(gpu):
    """Operands whose rows and columns are both strided (every second
    element): the staging reads them element by element."""
    from oracle.fno import round_bf16
    from wavesim_amd.learned import kernels as Kn
    rng = np.random.default_rng(3)
    M, N, K = 70, 45, 50
    A = rng.standard_normal((M, 2 * K)).astype(np.float32)
    B = rng.standard_normal((2 * K, 2 * N)).astype(np.float32)
    Cm = np.zeros((M, N), np.float32)
    dA, dB, dC = _t(A), _t(B), _t(Cm)
    Kn.gemm(M, N, K, dA, (2 * K, 2), dB, (2 * 2 * N, 2), dC, (N, 1))
    ref = round_bf16(A[:, ::2]) @ round_bf16(B[::2, ::2])
    assert np.allclose(dC.cpu().numpy(), ref, rtol=1e-5, atol=1e-4 * np.abs(ref).max())


def _pair(seed=0):
    from oracle import fno as O
    from wavesim_amd.learned import fno as G
    cfgo = O.Config(L=16, cin=4, width=8, modes=4, layers=2, proj=16)
    cfgg = G.Config(L=16, cin=4, width=8, modes=4, layers=2, proj=16)
    p = O.init_params(cfgo, seed)
    rng = np.random.default_rng(seed + 1)
    for k in p:
        p[k] = p[k] + 0.2 * rng.standard_normal(p[k].shape)
    x = rng.standard_normal((2, 4, 16, 16, 16))
    t = rng.standard_normal((2, 2, 16, 16, 16))
    return O, cfgo, G.FNO(cfgg, p), p, x, t


def test_fno_forward_matches_oracle(gpu):
    O, cfgo, net, p, x, t = _pair()
    y, _ = net.forward(_t(x.reshape(2, 4, -1)))
    yo, _ = O.Net(cfgo, bf16=True).forward(p, x)
    y64, _ = O.Net(cfgo, bf16=False).forward(p, x)
    y = y.cpu().numpy().reshape(yo.shape)
    assert rel_l2(y, yo) <= 2e-3
    assert rel_l2(y, y64) <= 3e-2


def test_fno_gradients_match_oracle(gpu):
    O, cfgo, net, p, x, t = _pair(3)
    loss, _ = net.loss_and_grad(_t(x.reshape(2, 4, -1)), _t(t.reshape(2, 2, -1)))
    Lo, go, _ = O.Net(cfgo, bf16=True).loss(p, x, t)
    _, g64, _ = O.Net(cfgo, bf16=False).loss(p, x, t)
    assert abs(loss - Lo) <= 1e-3 * abs(Lo)
    gg = net.grads()
    for k in go:
        e = rel_l2(gg[k], go[k])
        e64 = rel_l2(gg[k], g64[k])
        print(k, f"{e:.2e} {e64:.2e}", flush=True)
        assert e <= 2e-3, k
        assert e64 <= 3e-2, k


def test_adam_step_matches_formula(gpu):
    import torch
    O, cfgo, net, p, x, t = _pair(5)
    net.loss_and_grad(_t(x.reshape(2, 4, -1)), _t(t.reshape(2, 2, -1)))
    p0 = net.flat.cpu().numpy().copy()
    g = net.grad.cpu().numpy().copy()
    net.adam_step(lr=1e-2)
    p1 = net.flat.cpu().numpy()
    m = 0.1 * g
    v = 0.001 * g * g
    ref = p0 - 1e-2 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.allclose(p1, ref, rtol=1e-5, atol=1e-6)


def test_warmstart_preprocessing_kernels(gpu):
    """wfno_wkb_prior, wfno_features and wfno_compose_prior vs numpy."""
    import ctypes as C
    import torch
    from oracle import warmstart as O
    from wavesim_amd.learned import kernels as Kn
    rng = np.random.default_rng(7)
    nz = ny = nx = 16
    n = (1.33 + 0.3 * rng.random((nz, ny, nx))).astype(np.float32)
    u = (rng.standard_normal(nz) + 1j * rng.standard_normal(nz)).astype(np.complex64)
    k0h = 2 * np.pi * 0.25
    phase_o, P_o = O.wkb_prior(n, u, k0h, 1.33)
    dn = torch.from_numpy(n.ravel()).cuda()
    du = torch.from_numpy(u.view(np.float32).copy()).cuda()
    ph = torch.empty(n.size, device="cuda")
    P = torch.empty(2 * n.size, device="cuda")
    shape = (C.c_int * 3)(nz, ny, nx)
    Kn.check(Kn.lib().wfno_wkb_prior(Kn.ptr(dn), Kn.ptr(ph), Kn.ptr(P), Kn.ptr(du), shape, k0h, 1.33,
                                     Kn.stream_ptr()))
    assert np.allclose(ph.cpu().numpy().reshape(n.shape), phase_o, atol=1e-5)
    assert rel_l2(P.cpu().numpy().view(np.complex64).reshape(n.shape), P_o) <= 1e-5
    X = torch.empty(6 * 8 ** 3, device="cuda")
    Kn.check(Kn.lib().wfno_features(Kn.ptr(dn), Kn.ptr(ph), Kn.ptr(X), shape, 2, 1.33, Kn.stream_ptr()))
    assert rel_l2(X.cpu().numpy().reshape(6, 8, 8, 8), O.features(n, phase_o, 2, 1.33)) <= 1e-5
    grid, lo, roi, foot = (24, 22, 20), (3, 4, 2), (16, 16, 16), (3, 19, 4, 20)
    A = rng.standard_normal((2, 8, 8, 8)).astype(np.float32)
    uf = (rng.standard_normal(grid[2]) + 1j * rng.standard_normal(grid[2])).astype(np.complex64)
    E0 = torch.empty(2 * int(np.prod(grid)), device="cuda")
    dA = torch.from_numpy(A.ravel()).cuda()      # kept alive: raw pointers below
    duf = torch.from_numpy(uf.view(np.float32).copy()).cuda()
    Kn.check(Kn.lib().wfno_compose_prior(
        Kn.ptr(E0), Kn.ptr(dA), Kn.ptr(P), Kn.ptr(duf), (C.c_int * 3)(*grid), (C.c_int * 3)(*lo),
        (C.c_int * 3)(*roi), 2, (C.c_int * 4)(*foot), Kn.stream_ptr()))
    ref = O.compose(grid, lo, roi, 2, P.cpu().numpy().view(np.complex64).reshape(n.shape), A, uf, foot)
    got = E0.cpu().numpy().view(np.complex64).reshape(grid, order="F")
    assert rel_l2(got, ref) <= 1e-6


def test_warmstart_forward_graph_replays_match_eager(gpu):
    """The captured forward graph gives the eager forward's output on every
    replay, for changing inputs, with eager work in between."""
    import torch
    from wavesim_amd.learned.fno import Config
    from wavesim_amd.learned.warmstart import CIN, WarmStart
    ws = WarmStart(Config(L=16, cin=CIN, width=8, modes=4, layers=2, proj=16), factor=2, seed=3)
    rng = np.random.default_rng(0)
    for k in range(4):
        X = torch.from_numpy(rng.standard_normal(CIN * 16 ** 3).astype(np.float32)).cuda()
        ref = ws.forward(X, graph=False).clone()
        torch.cuda.synchronize()
        got = ws.forward(X, graph=True).clone()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), k


def test_learned_warm_start_gains_over_its_prior_and_zero_start(gpu):
    """What the trained network gains over the method without it, on held-out
    media (seeds 0-2; the training set used seeds 1000-1599): every start runs
    to the zero start's absolute residual.  The straight-ray prior alone saves
    iterations over the zero start, the network saves more on top of it, and
    every warm start ends as close to a tight solve as the zero start does
    (within a factor of three)."""
    import numpy as np
    from wavesim_amd.learned.warmstart import DEFAULT_WEIGHTS, WarmStart, evaluate
    ws = WarmStart.load(DEFAULT_WEIGHTS)
    rows, _ = evaluate(ws, range(3))
    for r in rows:
        print({k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}, flush=True)
    its = {k: sum(r[f"{k}_its"] for r in rows) for k in ("zero", "prior", "learned")}
    print(its, flush=True)
    assert its["learned"] < its["prior"] < its["zero"]
    zero_err = np.mean([r["zero_err"] for r in rows])
    for k in ("prior", "learned"):
        assert np.mean([r[f"{k}_err"] for r in rows]) <= 3 * zero_err
