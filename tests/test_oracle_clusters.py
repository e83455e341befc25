"""Pins the oracle's two-level cluster cull (CPU only).

Object pass (PreClusterDrawCull.comp.glsl:13-47): its expansion must be exactly the PLAIN (TEMPORAL with a pyramid)
pass's emitted objects and LODs, each followed by the chosen LOD's cluster range in order. Cluster tests (this
library's definition): the backface cone against a float64 restatement of meshoptimizer's sphere-corrected test,
the cluster frustum against the object frustum when the cluster sphere equals the surface sphere.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from blitzen_amd import _capi as A
from blitzen_amd import cull as bz
from oracle import pyoracle as orc
from tests import helpers as H


@pytest.fixture(scope="module")
def cscene():
    sc = bz.synth_scene(H.SEED_C4, 20000, 128, 8, clusters=(16, 128))
    return sc


def test_synth_clusters_follow_the_config(cscene):
    cl, lods = cscene.clusters, cscene.lods
    l0 = lods[::8]
    assert (l0["clusterCount"] >= 16).all() and (l0["clusterCount"] <= 128).all()
    # contiguous, non-overlapping ranges in LOD order
    np.testing.assert_array_equal(lods["clusterOffset"][1:], (lods["clusterOffset"] + lods["clusterCount"])[:-1])
    assert lods["clusterOffset"][-1] + lods["clusterCount"][-1] == len(cl)
    axis = np.stack([cl["coneAxis"][:, k] for k in range(3)], -1).astype(np.float64) / 127
    assert np.allclose(np.linalg.norm(axis, axis=1), 1, atol=0.02)
    cut = cl["coneCutoff"].astype(np.float64) / 127
    assert cut.min() >= -0.51 and cut.max() <= 0.91
    assert (cl["triangleCount"] == 64).all()


@pytest.mark.parametrize("with_pyr", [False, True])
def test_expansion_is_object_pass_times_cluster_ranges(cscene, with_pyr):
    scene = H.as_tuple(cscene)
    v = H.view(pyr=(1024, 512))
    v.lodTarget = 2e-5
    vb = H.view_bytes(v)
    pyr = None
    if with_pyr:
        pyr = orc.pyramid(orc.pyramid_build(H.c1_depth(seed=H.SEED_C4, occluders=256), 1024, 512, 11), 1024, 512, 11)
    recs, n = orc.cluster_expand(scene, vb, pyr)
    cmds, k, lod = orc.cull(scene, vb, A.PASS_TEMPORAL if with_pyr else A.PASS_PLAIN, pyr, want_lod=True, threads=4)
    lods = cscene.lods
    exp_obj = np.repeat(cmds[:, 0], lods["clusterCount"][lod])
    exp_lod = np.repeat(lod, lods["clusterCount"][lod])
    starts = np.repeat(lods["clusterOffset"][lod], lods["clusterCount"][lod])
    within = np.arange(len(exp_obj)) - np.repeat(np.cumsum(lods["clusterCount"][lod]) - lods["clusterCount"][lod],
                                                 lods["clusterCount"][lod])
    assert n == len(exp_obj) and n > 1000
    np.testing.assert_array_equal(recs[:, 0], exp_obj)
    np.testing.assert_array_equal(recs[:, 1], exp_lod)
    np.testing.assert_array_equal(recs[:, 2], starts + within)
    # capacity: exact count, prefix written
    part, n2 = orc.cluster_expand(scene, vb, pyr, capacity=77)
    assert n2 == n
    np.testing.assert_array_equal(part, recs[:77])


def visible(cl, t, vb, tests, pyr=None):
    vbuf = np.frombuffer(vb, np.uint8).copy()
    cl, t = np.array(cl), np.array(t)  # records -> 0-d arrays with their own buffers
    return bool(orc.load().orc_cluster_visible(cl.ctypes.data, t.ctypes.data, vbuf.ctypes.data,
                                               C.byref(pyr) if pyr is not None else None, tests))


def quat_mat(q):
    x, y, z, w = (float(a) for a in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_cone_test_matches_float64(cscene):
    cl_all, xf = cscene.clusters, cscene.transforms
    v = H.view(pos=(3.0, -2.0, 1.0), yaw=0.4, pitch=-0.2)
    vb = H.view_bytes(v)
    f = orc.view_floats(vb).astype(np.float64)
    M = f[:16].reshape(4, 4).T
    rng = np.random.default_rng(12)
    culled = kept = 0
    for _ in range(4000):
        cl = cl_all[int(rng.integers(len(cl_all)))].copy()
        t = xf[int(rng.integers(len(xf)))].copy()
        # place the object near the camera so both outcomes occur
        t["pos"] = rng.uniform(-30, 30, 3).astype(np.float32)
        R = quat_mat(t["orientation"])
        c = M[:3, :3] @ (R @ cl["center"].astype(np.float64) * float(t["scale"]) + t["pos"]) + M[:3, 3]
        r = float(cl["radius"]) * float(t["scale"])
        a = M[:3, :3] @ (R @ (cl["coneAxis"].astype(np.float64) / 127))
        cut = float(cl["coneCutoff"]) / 127
        lhs, rhs = float(c @ a), cut * float(np.linalg.norm(c)) + r
        if abs(lhs - rhs) < 1e-3 * (1 + abs(lhs)):
            continue
        got = visible(cl, t, vb, A.CLUSTER_CONE)
        assert got == (lhs < rhs), (lhs, rhs)
        culled += not got
        kept += got
    assert culled > 500 and kept > 500


def test_cluster_frustum_equals_object_frustum(cscene):
    """A cluster whose sphere is the surface's sphere gets the object's frustum decision."""
    objects, xf, surfaces = cscene.objects, cscene.transforms, cscene.surfaces
    vb = H.view_bytes(H.view())
    plain, _ = orc.cull(H.as_tuple(cscene), vb, A.PASS_PLAIN)
    vis_obj = np.zeros(len(objects), bool)
    vis_obj[plain[:, 0]] = True
    cl = np.zeros(1, A.CLUSTER)
    for i in range(0, len(objects), 7):
        s = surfaces[objects["surfaceId"][i]]
        cl["center"] = s["center"]
        cl["radius"] = s["radius"]
        t = xf[objects["transformId"][i]].copy()
        assert visible(cl[0].copy(), t, vb, A.CLUSTER_FRUSTUM) == vis_obj[i]
        assert visible(cl[0].copy(), t, vb, 0)  # no tests: every cluster is drawn (InitialClusterCull.comp.glsl)


def test_cluster_cull_command_fields(cscene):
    scene = H.as_tuple(cscene)
    vb = H.view_bytes(H.view())
    recs, n = orc.cluster_expand(scene, vb)
    cmds, k = orc.cluster_cull(recs, cscene.objects, cscene.transforms, cscene.clusters, vb, None, 0)
    assert k == n  # no tests: one draw per record
    cl = cscene.clusters[recs[:, 2]]
    np.testing.assert_array_equal(cmds[:, 0], recs[:, 0])
    np.testing.assert_array_equal(cmds[:, 1], cl["triangleCount"].astype(np.uint32) * 3)
    np.testing.assert_array_equal(cmds[:, 2], 1)
    np.testing.assert_array_equal(cmds[:, 3], cl["dataOffset"])
    np.testing.assert_array_equal(cmds[:, 4:], 0)
    sub, k2 = orc.cluster_cull(recs, cscene.objects, cscene.transforms, cscene.clusters, vb, None,
                               A.CLUSTER_FRUSTUM | A.CLUSTER_CONE)
    assert 0 < k2 < k
    # survivors are a subsequence of the records, in record order
    key = recs[:, 0].astype(np.uint64) << 32 | recs[:, 2]
    skey = sub[:, 0].astype(np.uint64) << 32 | (sub[:, 3] // 192)
    assert np.isin(skey, key).all() and (np.diff(skey.astype(np.int64)) > 0).all()
