"""The warm start's neighbour search against an FP64 k-d tree.

Each row's list must hold its K = 32 nearest neighbours under the
covariance metric, nearest first, restricted to the table cutoff (argument
8), with -1 past the last neighbour.  The device computes distances in
FP32 in tile frames, so distances are compared to 1e-4 relative and
neighbours at equal distance may trade places; the count may differ only by
points within FP32 rounding of the cutoff.

The geometries cover the search's paths: skewer surveys of three densities
(the first ring holds the K nearest for most rows), uneven subsets of the
300,000-pixel survey (rows that widen their ring, rows with fewer than K
neighbours inside the cutoff, tiles with more candidates than the ring
search keeps) and a tight cluster of 200 points inside a sparse survey
(more candidates than the ring search keeps, so its tiles take the general
search)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = 32
CUT = 199.0           # table argument of the cutoff (lib/signal_covar.h)
DX_INV = 199.0 / 8.0


@pytest.fixture(scope="module")
def D():
    import dachshund_amd as D
    D.default_context()
    return D


def _scaled(pix, w):
    s = np.array([np.sqrt(DX_INV / (2 * w.l_perp ** 2))] * 2 +
                 [np.sqrt(DX_INV / (2 * w.l_para ** 2))])
    return pix[:, :3] * s


def _device_lists(D, pix, w):
    from dachshund_amd.warmstart import WarmStart
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
    WarmStart(seed=0).predict_into(wf)
    nbr, perm = wf.geometry("knn"), wf.geometry("perm")
    wf.close()
    n = len(pix)
    slot = np.empty(n, dtype=np.int64)
    real = perm >= 0
    slot[perm[real]] = np.nonzero(real)[0]
    lists = nbr[slot]
    return np.where(lists >= 0, perm[np.maximum(lists, 0)], -1)


def _check(P, lists):
    from scipy.spatial import cKDTree
    n = len(P)
    dd, ii = cKDTree(P).query(P, k=K + 1)
    # drop each row's own entry (the first one if it is absent: more than K
    # coincident points never occur here)
    own = ii == np.arange(n)[:, None]
    drop = np.where(own.any(1), own.argmax(1), K)
    keep = np.ones_like(own)
    keep[np.arange(n), drop] = False
    d_ref = (dd[keep].reshape(n, K)) ** 2
    d_ref[d_ref >= CUT] = np.inf
    have = lists >= 0
    d_dev = np.full((n, K), np.inf)
    r, c = np.nonzero(have)
    d_dev[r, c] = np.sum((P[lists[r, c]] - P[r]) ** 2, axis=1)
    # every listed neighbour is a distinct other point, nearest first
    assert not np.any(lists == np.arange(n)[:, None])
    srt = np.sort(np.where(have, lists, -1 - np.arange(K)[None, :]), axis=1)
    assert not np.any(srt[:, 1:] == srt[:, :-1])
    fin = np.isfinite(d_dev)
    assert np.all(np.diff(np.where(fin, d_dev, 1e30), axis=1)
                  >= -1e-4 * (1 + d_dev[:, 1:].clip(0, 1e30)))
    # lists end with -1 only
    assert np.all(np.diff(have.astype(int), axis=1) <= 0)
    both = fin & np.isfinite(d_ref)
    err = np.abs(d_dev[both] - d_ref[both]) / (1 + d_ref[both])
    assert err.max(initial=0.0) <= 1e-4, err.max()
    # count mismatches only where the reference's extra (or missing)
    # neighbour sits at the cutoff
    diff = fin != np.isfinite(d_ref)
    edge = np.where(np.isfinite(d_ref), d_ref, np.where(fin, d_dev, CUT))
    assert np.all(np.abs(edge[diff] - CUT) <= 1e-3 * CUT)
    return int(fin.sum())


@pytest.mark.parametrize("name", ["c1", "c2", "small"])
def test_knn_skewer_surveys(D, name):
    from dachshund_amd import synth
    w = synth.workload(name)
    pix = synth.make_pixels(w)
    found = _check(_scaled(pix, w), _device_lists(D, pix, w))
    assert found > 0.9 * K * len(pix)


def _survey_subset(kind):
    """Skewers of the 300,000-pixel survey: `strip` keeps those with x < 16
    (the first ring, sized from the mean density over the bounding box,
    leaves about 8 % of the rows short of K candidates, so they widen their
    ring), `sparse` a random 5 % (more than half of the rows have fewer than
    K neighbours inside the cutoff), `mixed` the strip plus a random 3 %
    elsewhere (a first ring too wide for the strip's rows: more candidates
    than the ring search keeps, so those tiles take the general search)."""
    from dachshund_amd import synth
    w = synth.workload("c5")
    pix = synth.make_pixels(w)
    sk = np.unique(pix[:, :2], axis=0)
    rng = np.random.default_rng(3)
    pick = {"strip": sk[:, 0] < 16,
            "sparse": rng.random(len(sk)) < 0.05,
            "mixed": (sk[:, 0] < 16) | (rng.random(len(sk)) < 0.03)}[kind]
    key = pix[:, 0] * 1e6 + pix[:, 1]
    sel = sk[pick]
    return pix[np.isin(key, sel[:, 0] * 1e6 + sel[:, 1])], w


@pytest.mark.parametrize("kind", ["strip", "sparse", "mixed"])
def test_knn_uneven_surveys(D, kind):
    pix, w = _survey_subset(kind)
    lists = _device_lists(D, pix, w)
    _check(_scaled(pix, w), lists)
    if kind != "strip":
        assert np.mean(lists[:, -1] < 0) > 0.01   # short lists do occur


def test_knn_dense_cluster_general_search(D):
    """200 points within 0.4 scaled units of one spot (every ring radius is
    at least 1, so each of them has more than 128 candidates in its first
    ring) inside a sparse survey."""
    from dachshund_amd import synth
    w = synth.workload("c5")
    pix = synth.make_pixels(w)
    sk = np.unique(pix[:, :2], axis=0)[:40]
    keep = (pix[:, None, 0] == sk[None, :, 0]).any(1) & \
        (pix[:, None, 1] == sk[None, :, 1]).any(1)
    pix = pix[keep]
    rng = np.random.default_rng(7)
    s = np.array([np.sqrt(DX_INV / (2 * w.l_perp ** 2))] * 2 +
                 [np.sqrt(DX_INV / (2 * w.l_para ** 2))])
    centre = pix[len(pix) // 2, :3]
    v = rng.normal(size=(200, 3))
    v *= (0.4 * rng.random(200) ** (1 / 3) / np.linalg.norm(v, axis=1))[:, None]
    cl = np.empty((200, 5))
    cl[:, :3] = centre + v / s
    cl[:, 3] = rng.uniform(0.1, 0.3, 200)
    cl[:, 4] = rng.normal(0, 0.2, 200)
    pix = np.concatenate([pix, cl])
    _check(_scaled(pix, w), _device_lists(D, pix, w))
