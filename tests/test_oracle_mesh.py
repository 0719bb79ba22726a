"""The oracle's Kuhn mesher, P2 promotion and boundary sets, pinned by closed forms and invariants."""
import numpy as np
import pytest

from oracle import fem_oracle as O


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16])
def test_counts_and_volume(n):
    p, t = O.kuhn_box(n)
    assert p.shape == ((n + 1) ** 3, 3)
    assert t.shape == (6 * n**3, 4)
    _, vol = O.p1_geometry(p, t)
    # six tets of equal volume tile each voxel: total volume 1, each h^3/6
    assert np.allclose(vol, 1.0 / (6 * n**3), rtol=1e-12)
    assert abs(vol.sum() - 1.0) < 1e-12


def test_node_numbering_and_kuhn_order():
    # node id (i*(ny+1)+j)*(nz+1)+k, voxel-major tets (jno/domain/geometries.py:226-257)
    nx, ny, nz = 2, 3, 4
    p, t = O.kuhn_box(nx, ny, nz)
    i, j, k = 1, 2, 3
    nid = (i * (ny + 1) + j) * (nz + 1) + k
    assert np.allclose(p[nid], [i / nx, j / ny, k / nz])
    v = (1 * ny + 2) * nz + 3  # voxel (1,2,3)
    corner = lambda c: ((1 + (c >> 2 & 1)) * (ny + 1) + (2 + (c >> 1 & 1))) * (nz + 1) + 3 + (c & 1)  # noqa: E731
    for s, tet in enumerate(O.KUHN):
        assert list(t[6 * v + s]) == [corner(c) for c in tet]


def test_kuhn_tets_share_main_diagonal_and_are_conforming():
    p, t = O.kuhn_box(3)
    # every tet contains its voxel's corners 0 and 7
    assert np.all(t[:, 0] < t[:, 3])
    # conformity: each interior facet is shared by exactly two tets
    allf = np.sort(np.concatenate([t[:, list(c)] for c in O._TET_FACES]), axis=1)
    _, counts = np.unique(allf, axis=0, return_counts=True)
    assert set(np.unique(counts)) <= {1, 2}
    assert (counts == 1).sum() == 12 * 3**2  # 6 faces x n^2 quads x 2 triangles


@pytest.mark.parametrize("n", [2, 4, 7])
def test_boundary_nodes(n):
    p, t = O.kuhn_box(n)
    bn = O.boundary_nodes(p, t)
    assert len(bn) == 6 * n**2 + 2
    on = np.any((p == 0.0) | (p == 1.0), axis=1)
    assert np.array_equal(bn, np.flatnonzero(on))
    left = O.boundary_nodes(p, t, "left")
    assert len(left) == (n + 1) ** 2 and np.all(p[left, 0] == 0.0)


def test_face_triangles_cover_the_boundary_facets():
    n = 3
    p, t = O.kuhn_box(n)
    faces = O.kuhn_box_faces(n)
    tri = np.sort(np.concatenate(list(faces.values())), axis=1)
    bf = np.sort(O.boundary_facets(t), axis=1)
    assert {tuple(r) for r in tri} == {tuple(r) for r in bf}


@pytest.mark.parametrize("n", [1, 2, 3])
def test_p2_promotion_vectorised_equals_loop(n):
    p, t = O.kuhn_box(n)
    p2, c2 = O.promote_p2(p, t)
    p2l, c2l = O.promote_p2_loop(p, t)
    assert np.array_equal(c2, c2l) and np.array_equal(p2, p2l)
    assert p2.shape[0] == (2 * n + 1) ** 3  # vertices + edge midpoints of the lattice
    # vertices keep ids, edge nodes are the midpoints of their edge
    assert np.array_equal(c2[:, :4], t)
    for e, (a, b) in enumerate(O.P2_EDGES):
        assert np.allclose(p2[c2[:, 4 + e]], 0.5 * (p[t[:, a]] + p[t[:, b]]))


def test_p2_first_appearance_order():
    p, t = O.kuhn_box(2)
    _, c2 = O.promote_p2(p, t)
    new = c2[:, 4:].reshape(-1)
    nv = p.shape[0]
    first = []
    seen = set()
    for g in new:
        if g not in seen:
            seen.add(g)
            first.append(g)
    assert first == list(range(nv, nv + len(first)))


def test_p2_boundary_includes_facet_midpoints():
    n = 2
    p, t = O.kuhn_box(n)
    p2, c2 = O.promote_p2(p, t)
    bn = O.boundary_nodes(p2, c2)
    on = np.any((p2 == 0.0) | (p2 == 1.0), axis=1)
    assert np.array_equal(bn, np.flatnonzero(on))
    left = O.boundary_nodes(p2, c2, "left")
    assert len(left) == (2 * n + 1) ** 2
