"""The learned pair-cost model: the CPU oracle pinned by finite differences and
crafted inputs, then the device kernels (features, BF16 forward, backward,
training) checked against it, and what the trained network gains over a
length-only predictor."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

ACGT = np.array(list("ACGT"))


def _rand_seq(rng, n):
    return "".join(rng.choice(ACGT, n))


def _mutate(rng, s, rate):
    out = []
    for ch in s:
        r = rng.random()
        if r < rate / 3:
            out.append(str(rng.choice([c for c in "ACGT" if c != ch])))
        elif r < 2 * rate / 3:
            out.append(ch + str(rng.choice(ACGT)))
        elif r < rate:
            continue
        else:
            out.append(ch)
    return "".join(out)


def _rand_params(rng):
    w = np.zeros(mo.N_PARAMS, np.float32)
    w[mo.W1:mo.B1] = rng.uniform(-1.2, 1.2, mo.B1 - mo.W1)
    w[mo.B1:mo.W2] = rng.uniform(-0.1, 0.3, mo.K_H)
    w[mo.W2:mo.B2] = rng.uniform(-0.3, 0.3, mo.B2 - mo.W2)
    w[mo.B2:mo.W3] = rng.uniform(-0.1, 0.3, mo.K_H)
    w[mo.W3:mo.B3] = rng.uniform(-0.3, 0.3, mo.K_H)
    w[mo.B3] = 0.1
    return mo.bf16_round(w)


# ---------------------------------------------------------------------------
# oracle pins (CPU)

def test_bf16_rounding_known_values():
    def f(bits):
        return np.array([bits], np.uint32).view(np.float32)
    # exact, tie to even (down), tie to even (up), above half, sign, inf
    cases = [(0x3F800000, 0x3F800000), (0x3F808000, 0x3F800000), (0x3F818000, 0x3F820000),
             (0x3F808001, 0x3F810000), (0xBF808001, 0xBF810000), (0x7F800000, 0x7F800000)]
    for src, want in cases:
        assert mo.bf16_round(f(src)).view(np.uint32)[0] == want, hex(src)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = _rand_params(rng).astype(np.float64)
    X = rng.uniform(0, 1, (9, 4))
    y = rng.uniform(0, 1, 9)
    g = mo.gradient(w, X, y, dtype=np.float64)
    idx = np.concatenate([rng.choice(mo.N_PARAMS, 120, replace=False),
                          np.arange(mo.B1, mo.W2), [mo.B3]])
    eps = 1e-6
    for q in idx:
        wp, wm = w.copy(), w.copy()
        wp[q] += eps
        wm[q] -= eps
        fd = (mo.loss(wp, X, y, np.float64) - mo.loss(wm, X, y, np.float64)) / (2 * eps)
        assert abs(fd - g[q]) <= 1e-6 + 1e-5 * abs(fd), (q, fd, g[q])


def test_finite_differences_catch_a_planted_mistake():
    """A transposed W2 in the hidden-layer backprop must fail the same check."""
    rng = np.random.default_rng(4)
    w = _rand_params(rng).astype(np.float64)
    X = rng.uniform(0, 1, (9, 4))
    y = rng.uniform(0, 1, 9)

    def wrong_gradient(w, X, y):
        w1, b1, w2, b2, w3, b3 = mo.unpack(w)
        _, (X, z1, h1, z2, h2) = mo.forward(w, X, np.float64)
        g = (2.0 / len(X)) * (mo.forward(w, X, np.float64)[0] - y)
        d2 = np.outer(g, w3) * (z2 > 0)
        d1 = (d2 @ w2.T) * (z1 > 0)           # planted: should be d2 @ w2
        return (d1.T @ X).reshape(-1)
    wrong = wrong_gradient(w, X, y)
    bad = 0
    for q in range(0, mo.B1 - mo.W1, 7):
        wp, wm = w.copy(), w.copy()
        wp[mo.W1 + q] += 1e-6
        wm[mo.W1 + q] -= 1e-6
        fd = (mo.loss(wp, X, y, np.float64) - mo.loss(wm, X, y, np.float64)) / 2e-6
        bad += abs(fd - wrong[q]) > 1e-6 + 1e-5 * abs(fd)
    assert bad > 0


def test_oracle_features_crafted():
    rng = np.random.default_rng(5)
    p = _rand_seq(rng, 1000)
    x = mo.features_one(p, p)
    assert x[3] == 1.0 and x[2] == 0.0
    assert np.isclose(x[0], np.log2(1001) / 16, rtol=1e-6) and x[0] == x[1]
    # unrelated sequences: random 12-mers almost never occur in a ~200-base window
    x = mo.features_one(p, _rand_seq(rng, 900))
    assert x[3] <= 2 / 16
    assert np.isclose(x[2], np.log2(101) / 16, rtol=1e-6)
    # shorter than a k-mer: no overlap feature
    assert mo.features_one("ACGTACGTAC", "ACGTACGTAC")[3] == 0.0
    # deleting the text copy of the first sampled 12-mer loses exactly that one
    key_pos = ((2 * 0 + 1) * (400 - 12)) // 32
    q = _rand_seq(rng, 400)
    t_cut = q[:key_pos] + q[key_pos + 12:] + _rand_seq(rng, 12)
    assert mo.features_one(q, q)[3] == 1.0
    assert mo.features_one(q, t_cut)[3] == 15 / 16


# ---------------------------------------------------------------------------
# device kernels vs the oracle

bw = pytest.importorskip("biwfa_gpu_pii2025_amd")


@pytest.fixture(scope="module")
def cm():
    from biwfa_gpu_pii2025_amd import costmodel
    return costmodel


def _say(msg):
    print("[costmodel] " + msg, flush=True)


@pytest.mark.gpu
def test_device_features_match_oracle(cm):
    _say("features")
    rng = np.random.default_rng(6)
    ds = bw.generate_dataset(bw.CONFIGS["C4"], num_pairs=150)
    pairs = ds.pairs() + [("", ""), ("ACGT", "ACGTTT"), ("A" * 12, "A" * 12), ("ACGTACGTACGTA", "C" * 40)]
    for _ in range(20):
        p = _rand_seq(rng, int(rng.integers(12, 3000)))
        pairs.append((p, _mutate(rng, p, float(rng.uniform(0, 0.3)))))
    X = cm.batch_features(bw.pack_pairs(pairs))
    ref = mo.features(pairs)
    assert np.array_equal(X[:, 3], ref[:, 3])
    assert np.max(np.abs(X[:, :3] - ref[:, :3])) <= 2e-7


@pytest.mark.gpu
def test_device_forward_and_bf16_weights_match_oracle(cm):
    _say("forward")
    model = cm.CostModel(seed=7)
    master, w = model.params()
    assert np.array_equal(w, mo.bf16_round(master))
    # the forward kernel runs 32-sample tiles per wave (hidden layer on FP32 MFMA):
    # whole and partial tiles, FP32 against the float64 restatement
    rng = np.random.default_rng(8)
    for n in (1, 31, 33, 3000):
        X = rng.uniform(0, 1, (n, 4)).astype(np.float32)
        y = model.predict(X)
        ref, _ = mo.forward(w.astype(np.float64), X.astype(np.float64), np.float64)
        assert np.max(np.abs(y - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref))), n
    model.close()


@pytest.mark.gpu
def test_device_gradient_matches_oracle(cm):
    """Backward kernel (per-block partials, fixed-order reduction) vs the float64
    restatement on the same BF16 weights."""
    _say("gradient")
    rng = np.random.default_rng(9)
    model = cm.CostModel(seed=11)
    master = _rand_params(rng)
    model.set_params(master)
    _, w = model.params()
    for n in (1, 63, 64, 1000, 4099):
        X = rng.uniform(0, 1, (n, 4)).astype(np.float32)
        y = rng.uniform(0, 1, n).astype(np.float32)
        g = model.gradient(X, y)
        ref = mo.gradient(w.astype(np.float64), X.astype(np.float64), y.astype(np.float64), np.float64)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(g - ref)) <= 2e-5 * scale + 1e-7, n
        # deterministic: same inputs, bit-identical gradient
        assert np.array_equal(g, model.gradient(X, y))
    model.close()


@pytest.mark.gpu
def test_training_learns_and_beats_length_only(cm):
    """Trained on aligner-labelled mixed-length pairs, the network predicts the log
    score of held-out pairs much better than the best length-only fit, because the
    k-mer overlap feature sees divergence that lengths cannot."""
    import dataclasses
    _say("training")
    train_cfg = dataclasses.replace(bw.CONFIGS["C4_TRAIN"], len_max=8000)
    model, info = cm.train_cost_model(train_cfg, num_pairs=6000, epochs=40, seed=1)
    assert info["final_mse"] < 0.5 * info["initial_mse"]
    test = bw.generate_dataset(dataclasses.replace(bw.CONFIGS["C4"], len_max=8000), num_pairs=1500)
    packed = test.pack()
    y_true = cm.score_to_target(cm.aligner_labels(packed))
    X = cm.batch_features(packed)
    err_mlp = np.mean(np.abs(model.predict(X) - y_true))
    # length-only baseline: least squares on [1, x0, x1, x2] fitted on the training set
    A = np.c_[np.ones(len(info["X"])), info["X"][:, :3]]
    coef, *_ = np.linalg.lstsq(A, info["y"], rcond=None)
    err_len = np.mean(np.abs(np.c_[np.ones(len(X)), X[:, :3]] @ coef - y_true))
    _say("held-out MAE of log2(1+score)/16: mlp %.4f, length-only %.4f" % (err_mlp, err_len))
    assert err_mlp < 0.6 * err_len
    model.close()
