"""Learned SolidLiquid: a classifier that stands in for SolidLiquid's bond dot
products (freud/order.py:645-813, SolidLiquid.cc:51-96) on most particles.

SolidLiquid labels a particle solid-like when at least ``solid_threshold`` of
its bonds have a normalised q_6 dot product above ``q_threshold``; that needs a
second pass over every bond gathering the neighbour's 13 q_6m.  Here one fused
pass over the bonds computes per-particle rotation invariants (q_l and
normalised w_l for l = 4..12 even, and the bond count; the q_6m come out of the
same pass; ``feature_lmax=8`` keeps l = 4, 6, 8 only, which halves the
per-bond spherical-harmonic work), an 11 -> 64 -> 64 -> 1 network on the
matrix cores turns them into
a logit, and only particles whose logit lies within ``margin`` of 0 take the
exact bond-dot path.  The network is trained on the GPU on labels that the
exact SolidLiquid produces.
"""

import ctypes

import numpy as np

from freud_amd import _lib
from freud_amd._lib import DeviceArray, check, lib
from freud_amd.locality import AABBQuery, _PairCompute
from freud_amd.util import _Compute, computed_property


class LearnedSolidLiquid(_PairCompute):
    """Solid-like particles of SolidLiquid(l=6) predicted by a small network.

    Args:
        q_threshold (float): bond dot-product cutoff of the exact path.
        solid_threshold (int): solid-like bonds needed to be solid.
        normalize_q (bool): normalise the bond dot products (the exact path).
        margin (float): particles with |logit| < margin take the exact path.
        seed (int): initial weights.
        feature_lmax (int): largest l of the invariants, 8 (q_l, w_l for
            l = 4, 6, 8) or 12 (l = 4 .. 12); the others are zero.
    """

    def __init__(self, q_threshold=0.7, solid_threshold=6, normalize_q=True, margin=0.5, seed=0, feature_lmax=8):
        if feature_lmax not in (8, 12):
            raise ValueError("feature_lmax must be 8 or 12")
        self._feature_lmax = int(feature_lmax)
        self._q_threshold = float(np.float32(q_threshold))
        self._solid_threshold = int(solid_threshold)
        self._normalize_q = bool(normalize_q)
        self._margin = float(margin)
        h = ctypes.c_void_p()
        check(lib().fa_mlp_create(int(seed), ctypes.byref(h)))
        self._h = h
        self._mu = np.zeros(11, np.float32)
        self._sigma = np.ones(11, np.float32)
        self._trained = False
        self._called_compute = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.fa_mlp_destroy(h)

    # ------------------------------------------------------------ parameters
    @property
    def params(self):
        out = np.zeros(lib().fa_mlp_num_params(), np.float32)
        check(lib().fa_mlp_get_params(self._h, out.ctypes.data))
        return out

    @params.setter
    def params(self, value):
        v = np.ascontiguousarray(value, np.float32)
        if v.size != lib().fa_mlp_num_params():
            raise ValueError("wrong number of network parameters")
        check(lib().fa_mlp_set_params(self._h, v.ctypes.data))

    @property
    def normalization(self):
        return self._mu.copy(), self._sigma.copy()

    def set_normalization(self, mu, sigma):
        self._mu = np.ascontiguousarray(mu, np.float32)
        self._sigma = np.ascontiguousarray(sigma, np.float32)
        check(lib().fa_mlp_set_norm(self._h, self._mu.ctypes.data, self._sigma.ctypes.data))

    # ------------------------------------------------------------ pieces
    def _features(self, nq, nlist):
        n = nq._n
        feat = _lib.reuse(self, "_d_feat", (n, 12), np.float32)
        qlm6 = _lib.reuse(self, "_d_qlm6", (n, 13, 2), np.float32)
        ql6 = _lib.reuse(self, "_d_ql6", (n,), np.float32)
        check(lib().fa_sl_features(ctypes.byref(nq._cbox), nq._d_points.ptr, n, ctypes.byref(nlist._c),
                                   self._feature_lmax, feat.ptr, qlm6.ptr, ql6.ptr))
        return feat, qlm6, ql6

    def exact_labels(self, system, neighbors=None):
        """Solid-like flags of the exact SolidLiquid(6) (num_connections >= solid_threshold)."""
        from freud_amd.order import SolidLiquid

        sl = SolidLiquid(6, self._q_threshold, self._solid_threshold, self._normalize_q)
        sl.compute(system, neighbors=neighbors)
        return sl.num_connections >= self._solid_threshold

    def features(self, system, neighbors=None):
        """(N, 12) features of the particles of ``system`` (host array)."""
        nq, nlist, qargs, d_q, nqp = self._preprocess_arguments(system, neighbors=neighbors)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, nqp)
        return self._features(nq, nlist)[0].numpy()

    # ------------------------------------------------------------ training
    def fit(self, frames, r_max=1.5, epochs=6, batch=4096, lr=2e-3, seed=1, validation=0, target_agreement=0.997):
        """Train on the exact SolidLiquid labels of ``frames`` (list of (box, points)).

        With ``validation`` > 0 the last ``validation`` frames are held out and
        the fallback margin is raised (never lowered below the constructor's
        margin) to the smallest value at which the labels of the held-out
        frames agree with the exact ones at ``target_agreement``.
        Returns the per-epoch training loss."""
        feats, labels = [], []
        n_train = len(frames) - int(validation)
        if n_train < 1:
            raise ValueError("no training frames left after the validation split")
        for box, pts in frames:
            aq = AABBQuery(box, pts)
            nl = aq.query(pts, dict(r_max=r_max, exclude_ii=True)).toNeighborList()
            labels.append(self.exact_labels(aq, neighbors=nl).astype(np.uint8))
            feats.append(self._features(aq, nl)[0].numpy())
        self.validation_margin = None
        if validation:
            vf = np.concatenate(feats[n_train:])
            vt = np.concatenate(labels[n_train:]).astype(bool)
        feats, labels = feats[:n_train], labels[:n_train]
        f = np.concatenate(feats)
        t = np.concatenate(labels)
        keep = f[:, 10] > 0  # particles without bonds have no invariants (and are never solid)
        f, t = np.ascontiguousarray(f[keep]), np.ascontiguousarray(t[keep])
        mu = f[:, :11].astype(np.float64).mean(0).astype(np.float32)
        sigma = f[:, :11].astype(np.float64).std(0).astype(np.float32)
        sigma[~(sigma > 0)] = 1.0
        self.set_normalization(mu, sigma)
        d_f = DeviceArray.from_numpy(f)
        d_t = DeviceArray.from_numpy(t)
        losses = np.zeros(epochs, np.float32)
        check(lib().fa_mlp_train(self._h, d_f.ptr, d_t.ptr, len(t), int(epochs), int(batch), float(lr), int(seed),
                                 losses.ctypes.data))
        self._trained = True
        self._train_size = len(t)
        if validation:
            y = self.logits_of(vf)
            # particles the exact path would label (no bonds: NaN logit) are always right
            wrong = np.isfinite(y) & ((y > 0) != vt)
            allowed = int(np.floor((1.0 - target_agreement) * len(vt)))
            mags = np.sort(np.abs(y[wrong]))[::-1]
            need = float(np.nextafter(np.float32(mags[allowed]), np.float32(np.inf))) if len(mags) > allowed else 0.0
            self.validation_margin = need
            self._margin = max(self._margin, need)
        return losses

    @property
    def margin(self):
        return self._margin

    def gradient(self, features, labels, rows=None):
        """Gradient and loss of the mean BCE loss over feature rows (GPU kernels)."""
        f = DeviceArray.from_numpy(np.ascontiguousarray(features, np.float32).reshape(-1, 12))
        t = DeviceArray.from_numpy(np.ascontiguousarray(labels, np.uint8))
        idx = np.arange(len(labels), dtype=np.uint32) if rows is None else np.ascontiguousarray(rows, np.uint32)
        d_idx = DeviceArray.from_numpy(idx)
        g = np.zeros(lib().fa_mlp_num_params(), np.float32)
        loss = ctypes.c_float()
        check(lib().fa_mlp_gradient(self._h, f.ptr, t.ptr, d_idx.ptr, len(idx), g.ctypes.data, ctypes.byref(loss)))
        return g, float(loss.value)

    def logits_of(self, features):
        """Network output for host feature rows (N, 12)."""
        host = np.ascontiguousarray(features, np.float32).reshape(-1, 12)
        f = DeviceArray.from_numpy(host)
        out = DeviceArray((len(host),), np.float32)
        check(lib().fa_mlp_forward(self._h, f.ptr, len(host), out.ptr))
        return out.numpy()

    # ------------------------------------------------------------ inference
    def compute(self, system, neighbors=None):
        nq, nlist, qargs, d_q, nqp = self._preprocess_arguments(system, neighbors=neighbors)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, nqp)
        n = nq._n
        feat, qlm6, ql6 = self._features(nq, nlist)
        logits = _lib.reuse(self, "_d_logits", (n,), np.float32)
        check(lib().fa_mlp_forward(self._h, feat.ptr, n, logits.ptr))
        solid = _lib.reuse(self, "_d_solid", (max(n, 1),), np.uint8)
        fb = _lib.reuse(self, "_d_fb", (max(n, 1),), np.uint8)
        nfb = ctypes.c_uint32()
        check(lib().fa_sl_decide(n, ctypes.byref(nlist._c), logits.ptr, self._margin, qlm6.ptr, ql6.ptr,
                                 int(self._normalize_q), self._q_threshold, self._solid_threshold, solid.ptr, fb.ptr,
                                 ctypes.byref(nfb)))
        self._n = n
        self._d = dict(feat=feat, logits=logits, solid=solid, fallback=fb)
        self._n_fallback = int(nfb.value)
        self._nlist = nlist
        return self

    @computed_property
    def solid(self):
        """(N,) bool: solid-like (classifier, or exact where it fell back)."""
        return self._d["solid"].numpy()[: self._n].astype(bool)

    @computed_property
    def fallback(self):
        """(N,) bool: particles labelled by the exact bond-dot path."""
        return self._d["fallback"].numpy()[: self._n].astype(bool)

    @computed_property
    def num_fallback(self):
        return self._n_fallback

    @computed_property
    def logits(self):
        return self._d["logits"].numpy()

    @computed_property
    def particle_features(self):
        return self._d["feat"].numpy()

    @computed_property
    def nlist(self):
        return self._nlist

    @property
    def feature_lmax(self):
        return self._feature_lmax

    def __repr__(self):
        return (f"freud_amd.learned.LearnedSolidLiquid(q_threshold={self._q_threshold}, "
                f"solid_threshold={self._solid_threshold}, normalize_q={self._normalize_q}, margin={self._margin}, "
                f"feature_lmax={self._feature_lmax})")
