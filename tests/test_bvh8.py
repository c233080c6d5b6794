"""CPU checks of the library's BVH8 builder (the structure the HIP traversal
walks): every triangle stored once with the exact Möller–Trumbore operands,
meta/child encoding consistent, and every quantised child box, decoded the way
the kernel decodes it, containing all geometry below it."""
import numpy as np
import pytest

import nori_gpu as ng


def _decode(node, axis, q):
    scale = np.float32(2.0) ** (int(node["e"][axis]) - 127)
    # q * scale is exact (power-of-two scale); the one rounding is the add, as in fmaf
    return np.float32(node["p"][axis]) + np.float32(q) * np.float32(scale)


def _check(scene):
    info, nodes, tris = ng.build_bvh8_host(scene.positions, scene.indices)
    nt = len(scene.indices)
    assert info.triangles == nt and info.leaf_tris_max <= 3
    assert np.array_equal(np.sort(tris["prim"]), np.arange(nt))
    p = scene.positions[scene.indices[tris["prim"]]]
    assert np.array_equal(tris["p0"], p[:, 0])
    assert np.array_equal(tris["e1"], (p[:, 1] - p[:, 0]).astype(np.float32))
    assert np.array_equal(tris["e2"], (p[:, 2] - p[:, 0]).astype(np.float32))
    # per-node geometry bounds, children before parents (reverse breadth-first order)
    lo = np.full((len(nodes), 3), np.inf, np.float32)
    hi = np.full((len(nodes), 3), -np.inf, np.float32)
    seen_nodes = np.zeros(len(nodes), int)
    seen_tris = np.zeros(nt, int)
    for k in range(len(nodes) - 1, -1, -1):
        nd = nodes[k]
        inner = 0
        for s in range(8):
            meta = int(nd["meta"][s])
            if meta == 0:
                assert not (nd["imask"] >> s) & 1
                continue
            if (meta & 0x18) == 0x18:
                assert (nd["imask"] >> s) & 1 and meta == (1 << 5) | (24 + s)
                c = int(nd["child_base"]) + inner
                inner += 1
                assert c > k
                seen_nodes[c] += 1
                clo, chi = lo[c], hi[c]
            else:
                cnt = bin(meta >> 5).count("1")
                off = meta & 0x1F
                assert meta >> 5 == (1 << cnt) - 1 and off + cnt <= 24
                idx = int(nd["tri_base"]) + off + np.arange(cnt)
                seen_tris[idx] += 1
                v = p[idx].reshape(-1, 3)
                clo, chi = v.min(0), v.max(0)
            for a in range(3):
                assert _decode(nd, a, nd["qlo"][a][s]) <= clo[a], (k, s, a)
                assert _decode(nd, a, nd["qhi"][a][s]) >= chi[a], (k, s, a)
            lo[k] = np.minimum(lo[k], clo)
            hi[k] = np.maximum(hi[k], chi)
        assert inner == bin(int(nd["imask"])).count("1")
    assert seen_nodes[0] == 0 and np.all(seen_nodes[1:] == 1)
    assert np.all(seen_tris == 1)
    return info


def test_bvh8_cornell(cornell):
    info = _check(cornell)
    assert info.max_depth <= 12


def test_bvh8_disney_grid(disney_small):
    _check(disney_small)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 9, 25])
def test_bvh8_tiny_and_degenerate(n):
    # n coincident triangles (identical centroids) still form a valid tree
    pos = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (n, 1))
    idx = np.arange(3 * n).reshape(n, 3)
    sc = ng.HostScene(pos, idx, [0], [n], ng.default_materials(1))
    _check(sc)
