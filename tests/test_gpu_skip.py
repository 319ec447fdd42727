"""Empty-space skipping in the raycast: results with and without it are bit-identical.

The raycast leaves out samples that the volume's brick map proves far (unobserved, or
tsdf >= 0.999, so the march takes its far step there and neither stops nor hits). The
map is kept by integrate (both kernels), reset and upload, and rebuilt from the voxels
before every raycast once the raw storage has been handed out. Each test compares the
raycast with skipping against the raycast of the same volume without it (every sample
taken), and the FP32 exact path also against the oracle.
"""
import numpy as np
import pytest

import oracle
from azu_amd import native as N
from azu_amd import scene as sc
from tests.gpu_util import frames, gpu_volume, oracle_volume, say, small_params

pytestmark = pytest.mark.gpu


def _raycast_both(g, T):
    g.set_skip(True)
    a = g.raycast(T, sc.K640, with_t=True)
    g.set_skip(False)
    b = g.raycast(T, sc.K640, with_t=True)
    g.set_skip(True)
    return a, b


def _same(a, b):
    return all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))


@pytest.mark.parametrize("fmt,mode", [(N.FMT_F32, N.MODE_EXACT), (N.FMT_F16_U16, N.MODE_EXACT),
                                      (N.FMT_F32, N.MODE_FAST), (N.FMT_F16_U16, N.MODE_FAST)])
def test_skip_equivalence_over_sequence(fmt, mode):
    """Integrate a jittered orbit; raycast after every frame from that frame's pose and
    from the first pose, with and without skipping."""
    p = small_params(256, 0.004)
    g = gpu_volume(p, fmt=fmt, mode=mode)
    fr = frames(6)
    hits = []
    for mm, T in fr:
        g.integrate(oracle.depth_mm(mm), sc.K640, T)
        for Tq in (T, fr[0][1]):
            a, b = _raycast_both(g, Tq)
            assert _same(a, b), (fmt, mode)
            hits.append(float((a[0][..., 2] != 0).mean()))
    say("skip equivalence", fmt, mode, "hit fraction %.3f-%.3f" % (min(hits), max(hits)))
    assert min(hits) > 0.3


def test_skip_after_storage_writes_and_upload():
    """A new surface written through the raw storage pointer into space the map holds
    clear is seen by the next raycast (the map is rebuilt), bit-identical to the oracle;
    the same after an upload."""
    p = small_params(128, 0.008)
    g = gpu_volume(p, fmt=N.FMT_F32, mode=N.MODE_EXACT)
    o = oracle_volume(p)
    fr = frames(3)
    for mm, T in fr:
        d = oracle.depth_mm(mm)
        g.integrate(d, sc.K640, T)
        o.integrate(d, sc.K640, T)
    T0 = fr[0][1]
    before = g.raycast(T0, sc.K640)
    # A slab across the view, in front of the sphere: observed, with a + -> - crossing
    # along z, in voxels that were observed free (tsdf 1) or unobserved.
    t, w = o.get()
    t = t.copy()
    w = w.copy()
    r = p["resolution"]
    z0 = int(round((0.40 - p["origin"][2]) / p["voxel_size"]))
    ys, xs = slice(r // 4, 3 * r // 4), slice(r // 4, 3 * r // 4)
    for k, val in enumerate((0.8, 0.3, -0.2, -0.7)):
        t[z0 + k, ys, xs] = val
        w[z0 + k, ys, xs] = 4.0
    o.set(t, w)
    words = np.stack([t.reshape(-1), w.reshape(-1)], axis=1).astype(np.float32)
    ptr, nbytes = g.storage()
    assert nbytes == words.nbytes
    N.check(N.lib.azu_memcpy_h2d(ptr, words.ctypes.data, words.nbytes, None))
    N.synchronize()
    vg, ng = g.raycast(T0, sc.K640)
    vo, no = o.raycast(T0, sc.K640, 640, 480)
    changed = float((vg.view(np.uint32) != before[0].view(np.uint32)).any(axis=-1).mean())
    say("pixels whose hit moved to the written slab: %.3f" % changed)
    assert changed > 0.05
    assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32))
    assert np.array_equal(ng.view(np.uint32), no.view(np.uint32))
    a, b = _raycast_both(g, T0)
    assert _same(a, b)
    # upload into a fresh volume (map built by a scan of the uploaded voxels)
    g2 = gpu_volume(p, fmt=N.FMT_F32, mode=N.MODE_EXACT)
    g2.upload(t, w)
    v2, n2 = g2.raycast(T0, sc.K640)
    assert np.array_equal(v2.view(np.uint32), vo.view(np.uint32))
    assert np.array_equal(n2.view(np.uint32), no.view(np.uint32))
