"""HIP MLP kernels (bf16 MFMA, fp32 accumulate) against the float64 oracle, and the learned
warm start of the condensation Newton solve against the CPU oracle given the same guesses.

Tolerances: the forward pass is compared with an exact emulation of the bf16 rounding the
kernels apply (inputs, W1, W2, first hidden layer) to 1e-4 of the output scale; the loss
gradient is compared with the unrounded float64 gradient, which the bf16 operands perturb by
about 1e-2 relative, to a relative L2 error of 5e-2 and cosine similarity 0.998."""
import ctypes as C

import numpy as np
import pytest

from cleo_amd import abi
from cleo_amd import populations as P
from oracle import mlp_oracle as M
from tests.helpers import SEED, assert_state_match, assert_supers_match, cond_params, make_domain

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(m):
    print(m, flush=True)


def _blob(rng):
    b = np.zeros(abi.MLP_NBLOB, np.float32)
    b[:M.NPARAM] = rng.normal(0, 0.25, M.NPARAM)
    b[M.NPARAM:M.NPARAM + 6] = rng.normal(0, 1, 6)
    b[M.NPARAM + 6:M.NPARAM + 12] = rng.uniform(0.5, 2, 6)
    b[-2], b[-1] = 0.05, 0.3
    return b


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dom(gpu):
    pop = P.condensation(SEED, ndims=(2, 2, 2), mean=10)
    return make_domain(gpu, pop)


def test_mlp_forward_layout(gpu):
    _say("case: MLP forward vs bf16 emulation, n=1000")
    rng = np.random.default_rng(1)
    d = _dom(gpu)
    blob = _blob(rng)
    assert gpu.lib.cleo_mlp_set_weights(d.h, _fp(blob)) == 0
    n = 1000
    x = np.zeros((n, 8), np.float32)
    x[:, :6] = rng.normal(0, 1.5, (n, 6))
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(n, dtype=torch.float32, device="cuda")
    assert gpu.lib.cleo_mlp_forward(d.h, xd.data_ptr(), yd.data_ptr(), n) == 0
    got = yd.cpu().numpy().astype(np.float64)
    yq, _ = M.forward(blob, x, quantize=True)
    want = yq * blob[-1] + blob[-2]
    scale = np.abs(want).max()
    assert np.max(np.abs(got - want)) <= 1e-4 * scale, np.max(np.abs(got - want)) / scale


def test_mlp_gradient_vs_oracle(gpu):
    _say("case: MLP loss gradient vs float64 oracle, n=4096")
    rng = np.random.default_rng(2)
    d = _dom(gpu)
    blob = _blob(rng)
    assert gpu.lib.cleo_mlp_set_weights(d.h, _fp(blob)) == 0
    n = 4096
    x = np.zeros((n, 8), np.float32)
    x[:, :6] = rng.normal(0, 1.5, (n, 6))
    y = rng.normal(0.05, 0.3, n).astype(np.float32)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    loss = C.c_float()
    g = np.zeros(M.NPARAM, np.float32)
    assert gpu.lib.cleo_mlp_loss_grad(d.h, xd.data_ptr(), yd.data_ptr(), n, C.byref(loss), _fp(g)) == 0
    l_ref, g_ref = M.loss_grad(blob, x, y)
    assert loss.value == pytest.approx(l_ref, rel=2e-2)
    g = g.astype(np.float64)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    cos = g @ g_ref / (np.linalg.norm(g) * np.linalg.norm(g_ref))
    _say(f"  gradient rel L2 error {rel:.3e}, cosine {cos:.6f}")
    assert rel < 5e-2 and cos > 0.998


@pytest.mark.parametrize("variant,defer", [(0, 0), (1, 0), (1, 3), (2, 0)])
def test_training_and_warm_start(gpu, oracle, variant, defer):
    _say(f"case: train on solver samples, warm-started condensation vs oracle with the same guesses, "
         f"condensation variant {variant} (1: guess evaluated in the drop kernel), defer {defer}")
    pop = P.condensation(SEED, ndims=(8, 8, 8), mean=40)
    pop.state["qvap"] *= 1.005
    d = make_domain(gpu, pop)
    t = abi.Tiles(1, 1, 1, variant, 0, 0, 0, 0, 0, defer)
    assert gpu.lib.cleo_set_tiles(d.h, abi.ctypes_byref(t)) == 0
    cp = cond_params()
    n = pop.n
    xd = torch.zeros((n, 8), dtype=torch.float32, device="cuda")
    yd = torch.zeros(n, dtype=torch.float32, device="cuda")
    nw = C.c_uint64()
    assert gpu.lib.cleo_mlp_make_samples(d.h, C.byref(cp), xd.data_ptr(), yd.data_ptr(), n, C.byref(nw)) == 0
    ns = int(nw.value)
    assert ns > n // 2
    tp = abi.MLPTrainParams(2, 1024, 1e-3, 0.9, 0.999, 1e-8, 42)
    l1 = C.c_float()
    assert gpu.lib.cleo_mlp_train(d.h, xd.data_ptr(), yd.data_ptr(), ns, C.byref(tp), C.byref(l1)) == 0
    tp2 = abi.MLPTrainParams(40, 1024, 1e-3, 0.9, 0.999, 1e-8, 0)  # continue from current weights
    l2 = C.c_float()
    assert gpu.lib.cleo_mlp_train(d.h, xd.data_ptr(), yd.data_ptr(), ns, C.byref(tp2), C.byref(l2)) == 0
    _say(f"  loss after 2 epochs {l1.value:.4e}, after 42 {l2.value:.4e}")
    assert l2.value < l1.value
    # warm-started condensation on the GPU, recording its guesses
    d.record_niters(True)
    assert gpu.lib.cleo_record_guess(d.h, 1) == 0
    wp = cond_params(warm=1, warm_iters=10)
    assert d.condensation(wp) == 0
    assert d.error()[0] == 0
    guess = np.zeros(n)
    assert gpu.lib.cleo_download_guess(d.h, guess.ctypes.data_as(C.POINTER(C.c_double)), n) == 0
    sw = d.stats()
    # the oracle runs the same learned-guess Newton from the same guesses
    o = make_domain(oracle, pop)
    o.record_niters(True)
    oracle.lib.orc_set_cond_guess(o.h, guess.ctypes.data_as(C.POINTER(C.c_double)), n)
    assert o.condensation(wp) == 0
    assert_supers_match(d.download(), o.download(), what="warm")
    assert_state_match(d.state(), o.state(), what="warm")
    same = np.mean(d.niters() == o.niters())
    assert same >= 0.999
    # plain solve of the same population for the iteration comparison
    p = make_domain(gpu, pop)
    assert p.condensation(cp) == 0
    sp = p.stats()
    it_w = sw["newton_iters"] / sw["cond_drops"]
    it_p = sp["newton_iters"] / sp["cond_drops"]
    _say(f"  Newton iterations per drop: plain {it_p:.3f}, learned {it_w:.3f}, restarts {sw['warm_restarts']}")
    assert it_w < it_p


def test_fused_warm_start_matches_unfused(gpu):
    _say("case: learned guesses in the sort's gather (motion + condensation fused) vs the separate passes")
    pop = P.coupled(SEED, ndims=(8, 6, 4), mean=64)
    rng = np.random.default_rng(5)
    blob = _blob(rng)
    outs, stats = [], []
    # motion variant 1 fuses the condensation into the sort's gather; variant 0 keeps them apart
    for motion_flat, defer in ((1, 8), (0, 8), (0, 0)):
        d = make_domain(gpu, pop)
        t = abi.Tiles(1, 1, 1, 1, motion_flat, 64, 0, 64, 1, defer)
        assert gpu.lib.cleo_set_tiles(d.h, abi.ctypes_byref(t)) == 0
        assert gpu.lib.cleo_mlp_set_weights(d.h, _fp(blob)) == 0
        s = abi.SDMParams()
        s.enable_cond, s.enable_colls, s.enable_motion = 1, 1, 1
        s.cond = cond_params(warm=1, warm_iters=10)
        s.coll = abi.CollParams(100, abi.PROB_LONG, 1.0, abi.ENACT_COALBURE, 5.0, 0.85)
        s.motion = abi.MotionParams(100, abi.TV_SIMMEL)
        d.set_sdm(s)
        d.rng_counter(13)
        for step in range(3):
            assert d.run_step(step * 100, (step + 1) * 100) == 0
        assert d.error()[0] == 0, d.error()
        outs.append(d.download())
        stats.append(d.stats())
    for o in outs[1:]:
        assert_supers_match(outs[0], o, exact_fp=True, what="fused warm start")
    for st in stats[1:]:
        assert st["newton_iters"] == stats[0]["newton_iters"] and st["warm_restarts"] == stats[0]["warm_restarts"]
    _say(f"  Newton iterations {stats[0]['newton_iters']}, warm restarts {stats[0]['warm_restarts']}")
