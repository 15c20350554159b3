"""The centerline graph on the device against the oracle, bit for bit:
linking (the edge list in the 2D pyramid's order), component labels and
sizes, and the rasterisation of the large trees (removeSmallTrees)."""
import numpy as np
import pytest
import scipy.ndimage as ndi

pytestmark = pytest.mark.gpu


def polyline_vertices(seed, n_lines, steps, box):
    g = np.random.default_rng(seed)
    pts = []
    for _ in range(n_lines):
        p = g.uniform(4, box - 4, 3)
        d = g.normal(size=3)
        d /= np.linalg.norm(d)
        for _ in range(steps):
            p = p + g.uniform(1.5, 4.0) * d
            d = d + 0.25 * g.normal(size=3)
            d /= np.linalg.norm(d)
            pts.append(np.clip(np.round(p), 0, box - 1))
    return np.unique(np.array(pts, np.int32), axis=0)


def smooth_tdf(shape, seed):
    g = np.random.default_rng(seed)
    t = ndi.gaussian_filter(g.standard_normal(shape), 2.0)
    return ((t - t.min()) / (t.max() - t.min())).astype(np.float32)


def run_graph(tsf, v, T, maxd, min_avg, min_tree):
    dv = tsf.DeviceArray.from_host(v)
    dT = tsf.DeviceArray.from_host(T)
    n = tsf.link_centerpoints(dv, len(v), dT, maxd, min_avg)
    edges = tsf.DeviceArray((max(n, 1), 2), np.int32)
    assert tsf.link_centerpoints(dv, len(v), dT, maxd, min_avg, edges) == n
    labels = tsf.DeviceArray((len(v),), np.int32)
    sizes = tsf.DeviceArray((len(v),), np.int32)
    tsf.graph_components(edges, n, len(v), labels, sizes)
    out = tsf.DeviceArray(T.shape, np.uint8)
    out.zero()
    if n:
        tsf.rasterize_tree_edges(edges, n, dv, labels, sizes, min_tree, out)
    return edges.download()[:n], labels.download(), sizes.download(), out.download()


@pytest.mark.parametrize("seed,min_avg,box", [(1, 0.5, 48), (2, 0.3, 64), (3, 0.0, 40), (4, 0.45, 96)])
def test_link_graph_vs_oracle(tsf, orc, seed, min_avg, box):
    v = polyline_vertices(seed, 10, 14, box)
    T = smooth_tdf((box, box, box), seed)
    e, L, S, C = run_graph(tsf, v, T, 25.0, min_avg, 5)
    oe = orc.link_centerpoints(v, T, 25.0, min_avg)
    print(f"{len(v)} vertices, {len(e)} edges (oracle {len(oe)})")
    assert np.array_equal(e, oe)
    oL, oS = orc.graph_components(oe, len(v))
    assert np.array_equal(L, oL) and np.array_equal(S, oS)
    assert np.array_equal(C, orc.rasterize_tree_edges(oe, v, oL, oS, 5, T.shape))


def test_link_chain_and_limits(tsf, orc):
    T = np.ones((12, 12, 20), np.float32)
    v = np.array([[2 + 3 * i, 5, 6] for i in range(5)], np.int32)
    e, L, S, _ = run_graph(tsf, v, T, 25.0, 0.5, 1)
    assert {tuple(map(int, r)) for r in e} == {(1, 0), (1, 2), (2, 3), (3, 4)}
    assert np.array_equal(L, np.zeros(5)) and S[0] == 5 and not S[1:].any()
    # nothing above the minimum mean TDF: no edges; labels are the identity
    e, L, S, C = run_graph(tsf, v, np.zeros_like(T), 25.0, 0.5, 1)
    assert len(e) == 0 and np.array_equal(L, np.arange(5)) and np.array_equal(S, np.ones(5))
    assert not C.any()
    # a list too short for the edges: TSF_ERR_CAPACITY, the count still reported
    dv, dT = tsf.DeviceArray.from_host(v), tsf.DeviceArray.from_host(T)
    small = tsf.DeviceArray((2, 2), np.int32)
    with pytest.raises(tsf.TSFError) as ex:
        tsf.link_centerpoints(dv, len(v), dT, 25.0, 0.5, small)
    assert ex.value.status == 5


def test_link_graph_on_tube_volume(tsf, orc):
    # vertices from the candidate pipeline of a rendered tube tree (the
    # oracle's TDF, threshold, compaction, suppression and cube maximum)
    from tsf import synth
    shape = (64, 64, 64)
    vol, _ = synth.tube_volume(shape, 8, 1.5, 3.5, 21, noise=0.05)
    F = orc.vector_field(orc.blur(vol, 0.5, "f32"), 0.2, 1, "f32")
    V = orc.gvf_euler(F, 0.05, 60, "f32")
    t, r = orc.tdf_circle(V, 1.0, 4.0, 0.5, "f32")
    pos = orc.hp_compact(orc.candidates(t, 0.5, "f32"), "sort")
    keep = orc.candidate_nms(t, r, V, pos, 5, "f32")
    verts = orc.hp_compact(orc.cube_max(t, keep, 4, "f32"), "sort")
    assert len(verts) >= 8
    e, L, S, C = run_graph(tsf, verts, t, 25.0, 0.5, 5)
    oe = orc.link_centerpoints(verts, t, 25.0, 0.5)
    oL, oS = orc.graph_components(oe, len(verts))
    print(f"tube volume: {len(verts)} vertices, {len(e)} edges, {int((oS >= 5).sum())} trees of >= 5")
    assert len(e) > 0 and np.array_equal(e, oe)
    assert np.array_equal(L, oL) and np.array_equal(S, oS)
    assert np.array_equal(C, orc.rasterize_tree_edges(oe, verts, oL, oS, 5, shape))
