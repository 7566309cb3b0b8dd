"""The device geometry builder against the host builder: the ordering, the
culled tile lists with their sub-block masks, the per-slot arrays, the FP32
frames, the symmetric kernel's half lists and column-sum slots and the
launch orders of a problem built on the device equal, byte for byte, those
the host threads build for the same pixels (the host builder's lists are
checked against brute force in test_host.py).  The cases cover the one-
wavefront path (nodes of at most 1024 points), the level-by-level path
(multi-chunk nodes up to 300,000 points), coincident points, far outliers,
ragged last blocks and batches of subvolumes of different sizes."""
import numpy as np
import pytest

from conftest import fixture_pixels

pytestmark = pytest.mark.gpu

ARRAYS = ("perm", "pos", "rhs", "diag", "wvec", "loc32", "tgt32", "tcount",
          "tbox_lo", "tbox_hi", "tptr", "tidx", "tmask", "toff32", "hptr",
          "hrec", "gptr", "gslot", "blk_order", "ord32")
STATS = ("pairs_per_matvec", "sub_pairs_per_matvec", "sub_pairs_evaluated",
         "f64_pairs_per_matvec")


@pytest.fixture(scope="module")
def D():
    import dachshund_amd as D
    D.default_context()
    return D


@pytest.fixture(scope="module")
def host_ctx(D):
    ctx = D.Context(0)
    ctx.set_host_geometry(True)
    return ctx


def _pixels(name):
    from dachshund_amd import synth
    rng = np.random.default_rng(31)
    if name == "fixture":
        return fixture_pixels()
    if name in ("tiny", "small", "c1", "c5"):
        return synth.make_pixels(synth.workload(name))
    if name.startswith("n"):
        n = int(name[1:])
        p = np.empty((n, 5))
        p[:, :3] = rng.uniform(0, 6 + n ** (1 / 3), (n, 3))
        p[:, 3] = rng.uniform(0.1, 0.3, n)
        p[:, 4] = rng.standard_normal(n)
        return p
    if name == "dups":
        # coincident points (ties in every split key), a far loner
        base = _pixels("n65")
        p = np.repeat(base[:40], 30, axis=0)
        p[-1, :3] = [1e4, -3e3, 5e5]
        return p
    if name == "outliers":
        p = synth.make_pixels(synth.workload("small")).copy()
        p[5, :3] = [1e4, -3e3, 5e5]
        p[700, :3] = [-2e3, 40.0, 7.0]
        p[1300:1310, :3] = p[1300, :3] + np.array([0.0, 0.0, 1e3])
        return p
    if name == "sub40k":
        return synth.make_pixels(synth.batch_workloads(seed=4, count=3,
                                                       n_min=30000,
                                                       n_max=40000)[0])
    raise KeyError(name)


def _same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def _compare(dev, host):
    for name in ARRAYS:
        a, b = dev.geometry(name), host.geometry(name)
        assert _same(a, b), name
    sd, sh = dev.stats(), host.stats()
    for f in STATS:
        assert getattr(sd, f) == getattr(sh, f), f


@pytest.mark.parametrize("form", ["w", "sn"])
@pytest.mark.parametrize("case", ["n1", "n2", "n17", "n63", "n64", "n65",
                                  "n129", "n1500", "tiny", "small",
                                  "fixture", "dups", "outliers", "c1"])
def test_device_geometry_equals_host(D, host_ctx, case, form):
    pix = _pixels(case)
    sp = D.SignalCovarParams(0.05, 2.5, 2.1)
    cls = D.WFVectorW if form == "w" else D.WFVectorSN
    conv = D.pixel_to_wpixel if form == "w" else D.pixel_to_npixel
    recs = conv(pix)
    dev = cls(recs, sp)
    host = cls(recs, sp, ctx=host_ctx)
    _compare(dev, host)


@pytest.mark.parametrize("case", ["sub40k", "c5"])
def test_device_geometry_equals_host_large(D, host_ctx, case):
    from dachshund_amd import synth
    pix = _pixels(case)
    if case == "c5":
        w = synth.workload("c5")
        sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    else:
        sp = D.SignalCovarParams(0.05, 2.0, 2.5)
    recs = D.pixel_to_wpixel(pix)
    dev = D.WFVectorW(recs, sp)
    host = D.WFVectorW(recs, sp, ctx=host_ctx)
    _compare(dev, host)
    # the device build is the faster one
    assert dev.stats().setup_seconds < host.stats().setup_seconds


def test_device_geometry_equals_host_batch(D, host_ctx):
    """Subvolumes of 1 to 9,000 points with their own correlation lengths;
    a batch's arrays are its subvolumes' arrays side by side."""
    rng = np.random.default_rng(5)
    parts, sps = [], []
    for k, n in enumerate([1, 70, 9000, 2048, 1025, 3, 4100, 640]):
        p = np.empty((n, 5))
        side = 4.0 + n ** (1 / 3)
        p[:, :3] = rng.uniform(0, side, (n, 3))
        p[:, 3] = rng.uniform(0.1, 0.3, n)
        p[:, 4] = rng.standard_normal(n)
        parts.append(D.pixel_to_wpixel(p))
        sps.append(D.SignalCovarParams(0.05, 1.5 + 0.25 * k, 3.0 - 0.2 * k))
    dev = D.BatchWF(parts, sps)
    host = D.BatchWF(parts, sps, ctx=host_ctx)
    _compare(dev, host)
    p64d, p32d = dev.sub_pairs()
    p64h, p32h = host.sub_pairs()
    assert np.array_equal(p64d, p64h) and np.array_equal(p32d, p32h)


def test_device_geometry_solves_like_host(D, host_ctx):
    """Same arrays, same kernels: the two problems' solves are identical."""
    pix = _pixels("small")
    sp = D.SignalCovarParams(0.05, 2.5, 2.1)
    recs = D.pixel_to_wpixel(pix)
    prm = D.PCGParams(len(pix), 200, 50, 1e-8)
    xd, xh = np.zeros(len(pix)), np.zeros(len(pix))
    rd = D.WFVectorW(recs, sp).solve_pcg(prm, xd)
    rh = D.WFVectorW(recs, sp, ctx=host_ctx).solve_pcg(prm, xh)
    assert rd.success and rd.num_iters == rh.num_iters and _same(xd, xh)


@pytest.mark.parametrize("ties", [False, True])
def test_device_tile_order_every_block_size(D, host_ctx, ties):
    """One subvolume per block size from 17 to 64 points (each a single
    block whose tiles the block kernel orders in registers), and two of 65
    and 200 points (blocks cut from a larger node), on spread-out points and
    on points with only three distinct values per axis (ties in every
    split key, settled by the point index)."""
    rng = np.random.default_rng(11 if ties else 12)
    parts, sps = [], []
    for n in list(range(17, 65)) + [65, 200]:
        p = np.empty((n, 5))
        if ties:
            p[:, :3] = rng.integers(0, 3, (n, 3)) * 1.5
        else:
            p[:, :3] = rng.uniform(0, 5.0, (n, 3))
        p[:, 3] = rng.uniform(0.1, 0.3, n)
        p[:, 4] = rng.standard_normal(n)
        parts.append(D.pixel_to_wpixel(p))
        sps.append(D.SignalCovarParams(0.05, 2.0, 2.5))
    dev = D.BatchWF(parts, sps)
    host = D.BatchWF(parts, sps, ctx=host_ctx)
    _compare(dev, host)
