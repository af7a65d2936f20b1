"""Host wrappers keep alive what their native objects point at.

rzk_lightslice keeps its scene and VPL set by pointer, and rzk_clusters its
scene: a Python wrapper that dropped its reference to either would let the
native object read freed memory (a freed rzk_vpls hands the next gather a stale
device pointer).  A stand-in library runs the wrappers without a GPU."""
import gc
import weakref

from renzoku_amd import _lib


class _StandInLib:
    def __getattr__(self, name):
        return lambda *args: 0              # every entry point succeeds


class _StandInCtx:
    lib = _StandInLib()
    h = None

    def check(self, rc, what):
        assert rc == 0, what


class _Handle:
    h = None


def test_lightslice_and_clusters_keep_their_native_dependencies():
    ctx = _StandInCtx()
    scene, vpls = _Handle(), _Handle()
    w_scene, w_vpls = weakref.ref(scene), weakref.ref(vpls)
    ls = _lib.LightSlice(ctx, scene, vpls, width=16, height=16, tile=8, num_clusters=2, num_neighbor_slices=1)
    del scene, vpls
    gc.collect()
    assert w_scene() is not None and w_vpls() is not None
    cl = _lib.Clusters(ls)
    del ls
    gc.collect()
    assert w_scene() is not None                    # the clusters still refer to the scene
    assert w_vpls() is None                         # nothing holds the VPLs once the LightSlice is gone
    del cl
    gc.collect()
    assert w_scene() is None
