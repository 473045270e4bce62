"""The load-step checkpoint's embedded CheckpointFile image, checked on the
CPU against the bytes the compiled reference's CheckpointFile::save wrote
(tests/golden/checkpoint.npz, FEM/adaptive_fem_engine.hpp:1347-1398): record
layout, and the closed forms the library writes for the element records and
the states (volume and quality of FEM/mesh_adaptation.hpp:110-146, F, and the
Mandel stress C S = mu (C - I) + lam ln J I of neo_hookean_response, :79-129)."""
import numpy as np

from conftest import golden


def parse_image(b, nn, ne):
    """Split CheckpointFile::save bytes into (header, nodes, elements, states, u)."""
    hdr = b[:20].view(np.uint32)
    o = 20
    nodes = b[o:o + 48 * nn].view(np.float64).reshape(nn, 6)
    o += 48 * nn
    er = b[o:o + 40 * ne].reshape(ne, 40)
    conn = er[:, :16].copy().view(np.uint32).reshape(ne, 4)
    geo = er[:, 16:].copy().view(np.float64).reshape(ne, 3)
    o += 40 * ne
    st = b[o:o + 184 * ne].view(np.float64).reshape(ne, 23)
    o += 184 * ne
    u = b[o:o + 8 * 3 * nn].view(np.float64)
    assert o + 24 * nn == b.size
    return hdr, nodes, conn, geo, st, u


def mandel(F, mu, lam):
    C = F.T @ F
    J = abs(np.linalg.det(F))
    M = mu * (C - np.eye(3)) + lam * np.log(max(J, 0.01)) * np.eye(3)
    s2 = np.sqrt(2.0)
    return np.array([M[0, 0], M[1, 1], M[2, 2], s2 * M[0, 1], s2 * M[0, 2], s2 * M[1, 2]])


def test_reference_checkpoint_image_layout_and_closed_forms(O):
    g = golden("checkpoint.npz")
    X, conn, u = g["X"], g["conn"], g["u"]
    nn, ne = X.size // 3, conn.size // 4
    hdr, nodes, c, geo, st, uu = parse_image(g["ref_bytes"], nn, ne)
    assert list(hdr) == [7, 0, nn, ne, 3 * nn]
    assert np.array_equal(nodes[:, :3], X.reshape(3, nn).T) and np.array_equal(nodes[:, 3:], u.reshape(nn, 3))
    assert np.array_equal(c, conn.reshape(4, ne).T) and np.array_equal(uu, u)
    E, nu = float(g["E"][0]), float(g["nu"][0])
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    P = X.reshape(3, nn).T
    ofl = O.element_fields(X, conn, np.full(ne, mu), np.full(ne, lam), u)
    for e in range(ne):
        v = P[c[e]]
        v6 = abs(np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], axis=1)))
        sq = sum(np.sum((v[a] - v[b]) ** 2) for a in range(4) for b in range(a + 1, 4))
        # tet_quality's sqrt 2 is 1.41421356237, 2.2e-12 below it
        assert np.allclose(geo[e, :2], [v6 / 6, v6], rtol=1e-14, atol=0)
        assert np.allclose(geo[e, 2], 12 * np.sqrt(2) * v6 / sq ** 1.5, rtol=5e-12, atol=0)
        F = ofl["F"][:, e].reshape(3, 3)
        assert np.allclose(st[e, :9], F.ravel(), rtol=0, atol=1e-14)
        # the reference scales the shear terms by 1.41421356237 (2.2e-12 below sqrt 2)
        assert np.allclose(st[e, 9:15], mandel(F, mu, lam), rtol=0, atol=5e-12 * np.abs(st[e, 9:15]).max())
        assert not st[e, 15:].any()
