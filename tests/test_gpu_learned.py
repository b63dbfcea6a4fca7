"""GPU tests of the learned SolidLiquid path: features against the oracle, the
BF16-MFMA network and its gradients against the CPU oracle (oracle/mlp.py),
training on exact SolidLiquid labels, and the accuracy the trained network
reaches with and without the exact fallback."""

import numpy as np
import numpy.testing as npt
import pytest

import freud_amd
from freud_amd.learned import LearnedSolidLiquid
from oracle import mlp

pytestmark = pytest.mark.gpu


def _bt(box):
    return (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)


def _nl(box, pts):
    return freud_amd.AABBQuery(box, pts).query(pts, dict(r_max=1.5, exclude_ii=True)).toNeighborList()


@pytest.mark.parametrize("lmax", [8, 12])
def test_features_match_oracle(oracle, lmax):
    box, pts, _ = freud_amd.data.two_phase(6, seed=2)
    nl = _nl(box, pts)
    ls = LearnedSolidLiquid(feature_lmax=lmax)
    f = ls.features((box, pts), neighbors=nl)
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=1.5, exclude_ii=True), nthreads=8)
    ref = mlp.features(_bt(box), pts, ref_nl, nthreads=8)
    q_cols = [0, 1, 2] if lmax == 8 else [0, 1, 2, 3, 4]
    w_cols = [c + 5 for c in q_cols]
    npt.assert_array_equal(f[:, 10], ref[:, 10])
    npt.assert_allclose(f[:, q_cols], ref[:, q_cols], atol=1e-5)  # q_l
    npt.assert_allclose(f[:, w_cols], ref[:, w_cols], atol=1e-4)  # normalised w_l (other summation order)
    unused = [c for c in range(10) if c not in q_cols + w_cols]
    assert np.all(f[:, unused] == 0.0)
    with pytest.raises(ValueError):
        LearnedSolidLiquid(feature_lmax=10)


def test_fallback_everywhere_is_exact_solidliquid():
    box, pts, _ = freud_amd.data.two_phase(6, seed=4)
    nl = _nl(box, pts)
    exact = LearnedSolidLiquid().exact_labels((box, pts), neighbors=nl)
    ls = LearnedSolidLiquid(margin=np.inf).compute((box, pts), neighbors=nl)
    assert ls.num_fallback == len(pts)
    npt.assert_array_equal(ls.solid, exact)


def _random_params(rng):
    p = np.zeros(mlp.N_PARAMS, np.float32)
    p[mlp.OFF_W1:mlp.OFF_B1] = rng.normal(0, 0.5, mlp.N_H * mlp.N_IN)
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.normal(0, 0.1, mlp.N_H)
    p[mlp.OFF_W2:mlp.OFF_B2] = rng.normal(0, 0.2, mlp.N_H * mlp.N_H)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.normal(0, 0.1, mlp.N_H)
    p[mlp.OFF_W3:mlp.OFF_B3] = rng.normal(0, 0.3, mlp.N_H)
    p[mlp.OFF_B3] = 0.1
    return p


def test_network_forward_matches_oracle():
    rng = np.random.default_rng(1)
    ls = LearnedSolidLiquid()
    p = _random_params(rng)
    ls.params = p
    mu = rng.normal(0, 0.2, 11).astype(np.float32)
    sigma = rng.uniform(0.5, 2.0, 11).astype(np.float32)
    ls.set_normalization(mu, sigma)
    feat = rng.normal(0, 1, (1000, 12)).astype(np.float32)  # not a multiple of 16
    y = ls.logits_of(feat)
    ref, _ = mlp.forward(p, feat, mu, sigma)
    # FP32 accumulation in another order can move a layer-1 activation across a
    # BF16 rounding boundary: a few 2^-8 relative steps at most
    err = np.abs(y - ref) / (1 + np.abs(ref))
    assert np.median(err) < 1e-5
    assert err.max() < 2e-2


def test_network_gradient_matches_oracle():
    rng = np.random.default_rng(2)
    ls = LearnedSolidLiquid()
    p = _random_params(rng)
    ls.params = p
    mu, sigma = np.zeros(11, np.float32), np.ones(11, np.float32)
    ls.set_normalization(mu, sigma)
    feat = rng.normal(0, 1, (777, 12)).astype(np.float32)
    t = rng.integers(0, 2, 777).astype(np.uint8)
    rows = rng.permutation(777)[:500].astype(np.uint32)
    g, loss = ls.gradient(feat, t, rows)
    gr, lr = mlp.gradients(p, feat[rows], t[rows], mu, sigma)
    npt.assert_allclose(loss, lr, rtol=1e-4)
    for lo, hi in ((mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
                   (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.N_PARAMS)):
        rel = np.linalg.norm(g[lo:hi] - gr[lo:hi]) / max(np.linalg.norm(gr[lo:hi]), 1e-12)
        assert rel < 1e-2, (lo, hi, rel)


def _frames(seeds, cells=9):
    out = []
    for s in seeds:
        box, pts, _ = freud_amd.data.two_phase(cells, sigma=0.03 + 0.01 * (s % 3), seed=s)
        out.append((box, pts))
    return out


def test_training_reaches_exact_labels():
    # 8 training frames + 2 validation frames that set the fallback margin
    train = _frames(range(100, 110))
    ls = LearnedSolidLiquid(margin=0.5, seed=3)
    losses = ls.fit(train, epochs=8, batch=2048, lr=3e-3, validation=2, target_agreement=0.997)
    assert losses[-1] < 0.5 * losses[0], losses
    assert ls.margin >= 0.5
    box, pts = _frames([7], cells=10)[0]
    nl = _nl(box, pts)
    exact = ls.exact_labels((box, pts), neighbors=nl)
    ls.compute((box, pts), neighbors=nl)
    agree = np.mean(ls.solid == exact)
    frac_fb = ls.num_fallback / len(pts)
    assert agree >= 0.995, (agree, frac_fb, ls.margin)
    assert frac_fb < 0.1, (frac_fb, ls.margin)
    # the network alone (no fallback) versus an untrained network: what training buys
    net_only = np.mean((ls.logits > 0) == exact)
    untrained = LearnedSolidLiquid(margin=0.5, seed=3)
    untrained.set_normalization(*ls.normalization)
    untrained.compute((box, pts), neighbors=nl)
    assert net_only > 0.97, net_only
    assert net_only > np.mean((untrained.logits > 0) == exact) + 0.1


def test_learned_full_size_agreement():
    """The learned classifier at its benchmark size: trained on 32 generated
    two-phase frames (64,000 particles each, sigma 0.03-0.05 a) plus 2
    validation frames that calibrate the fallback margin, then applied to the
    2,097,152-particle two-phase box (FCC 64^3 x 4 at density 1.05, sigma
    0.04 a, beside as many liquid particles at density 0.95), r_max = 1.5,
    l = 6, q_threshold 0.7, solid_threshold 6.  Its solid-like labels must
    agree with the exact SolidLiquid (SolidLiquid.cc:51-96) on >= 99.5 % of
    the particles."""
    frames = []
    for s in range(34):
        b, p, _ = freud_amd.data.two_phase(20, sigma=0.03 + 0.005 * (s % 5), seed=1000 + s)
        frames.append((b, p))
    ls = LearnedSolidLiquid(q_threshold=0.7, solid_threshold=6, margin=0.5, seed=11)
    losses = ls.fit(frames, epochs=6, batch=8192, lr=3e-3, validation=2, target_agreement=0.997)
    assert losses[-1] < losses[0], losses
    box, pts, _ = freud_amd.data.two_phase(64, rho_solid=1.05, rho_liquid=0.95, sigma=0.04, seed=5)
    assert len(pts) == 2_097_152
    nl = _nl(box, pts)
    exact = ls.exact_labels((box, pts), neighbors=nl)
    ls.compute((box, pts), neighbors=nl)
    agree = float(np.mean(ls.solid == exact))
    print(f"agreement {agree:.5f}, fallback {ls.num_fallback / len(pts):.4f}, margin {ls.margin:.3f}")
    assert agree >= 0.995, (agree, ls.margin)
    # particles that fell back carry the exact label
    fb = ls.fallback
    npt.assert_array_equal(ls.solid[fb], exact[fb])
