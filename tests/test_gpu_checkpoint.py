"""Load-step checkpoint through the C ABI: the embedded CheckpointFile image
against the bytes the compiled reference's CheckpointFile::save wrote for the
same mesh, displacement and material (tests/golden/checkpoint.npz,
FEM/adaptive_fem_engine.hpp:1347-1398); round trip, restart of an empty
context, and rejection of corrupted files and of files of another mesh or
material."""
import numpy as np
import pytest

from conftest import golden
from test_checkpoint_format import parse_image

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fm():
    import flexmesh

    flexmesh.lib()
    return flexmesh


@pytest.fixture
def S(fm):
    s = fm.Solver(0)
    yield s
    s.close()


def fnv1a(b):
    h = 1469598103934665603
    for x in b.tobytes():
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def golden_context(fm, S):
    g = golden("checkpoint.npz")
    X, conn, u = g["X"], g["conn"], g["u"]
    ne = conn.size // 4
    E, nu = float(g["E"][0]), float(g["nu"][0])
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    S.set_mesh(X, conn)
    S.set_material(np.full(ne, mu), np.full(ne, lam))
    S.set_u(u)
    return g, mu, lam


def test_checkpoint_embeds_the_reference_image(fm, S, tmp_path):
    g, mu, lam = golden_context(fm, S)
    X, conn = g["X"], g["conn"]
    nn, ne = X.size // 3, conn.size // 4
    p = str(tmp_path / "ck.bin")
    S.checkpoint_save(p, int(g["load_step"][0]))
    b = np.fromfile(p, dtype=np.uint8)
    ref = g["ref_bytes"]
    assert b[:8].tobytes() == b"FMCKPT02" and list(b[8:16].view(np.uint32)) == [2, 0]
    L = int(b[16:24].view(np.uint64)[0])
    assert L == ref.size and b.size == 24 + L + 16 * ne + 8
    img = b[24:24 + L]
    h1, n1, c1, g1, s1, u1 = parse_image(img, nn, ne)
    h0, n0, c0, g0, s0, u0 = parse_image(ref, nn, ne)
    assert np.array_equal(h1, h0) and np.array_equal(n1, n0) and np.array_equal(c1, c0) and np.array_equal(u1, u0)
    # volume and jacobian_det; the reference's tet_quality also scales by
    # 1.41421356237 (2.2e-12 below sqrt 2)
    assert np.allclose(g1[:, :2], g0[:, :2], rtol=1e-14, atol=0)
    assert np.allclose(g1[:, 2], g0[:, 2], rtol=5e-12, atol=0)
    assert np.abs(s1[:, :9] - s0[:, :9]).max() <= 1e-14
    # the reference's sqrt 2 constant (1.41421356237) is 2.2e-12 below sqrt 2
    assert np.abs(s1[:, 9:15] - s0[:, 9:15]).max() <= 5e-12 * np.abs(s0[:, 9:15]).max()
    assert not s1[:, 15:].any() and not s0[:, 15:].any()
    ext = b[24 + L:24 + L + 16 * ne].view(np.float64)
    assert np.array_equal(ext[:ne], np.full(ne, mu)) and np.array_equal(ext[ne:], np.full(ne, lam))
    assert int(b[-8:].view(np.uint64)[0]) == fnv1a(b[:-8])


def test_checkpoint_round_trip_and_restart(fm, S, tmp_path):
    cfg = fm.configs.make(3, dims=(3, 4, 3))
    cfg["n_steps"] = 3
    fm.configs.load(S, cfg)
    c = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64)
    S.newton_step(1, c)
    u1 = S.get_u()
    p = str(tmp_path / "ck.bin")
    S.checkpoint_save(p, 1)
    S.newton_step(2, c)
    u2 = S.get_u()
    assert S.checkpoint_load(p) == 1 and np.array_equal(S.get_u(), u1)
    # an empty context takes the mesh, the material and u from the file
    R = fm.Solver(0)
    try:
        assert R.checkpoint_load(p) == 1 and np.array_equal(R.get_u(), u1)
        # the binding checks nodal arrays against the context's node count
        with pytest.raises(ValueError):
            R.set_u(np.zeros(7))
        R.set_bcs(cfg["dir_dofs"], cfg["dir_vals"])
        R.newton_step(2, c)
        assert np.array_equal(R.get_u(), u2)
    finally:
        R.close()


def test_checkpoint_rejects_corruption_and_other_meshes(fm, S, tmp_path):
    cfg = fm.configs.make(3, dims=(3, 3, 3))
    fm.configs.load(S, cfg)
    S.set_u(np.random.default_rng(6).standard_normal(cfg["X"].size) * 1e-3)
    p = str(tmp_path / "ck.bin")
    S.checkpoint_save(p, 4)
    data = bytearray(open(p, "rb").read())
    bad = bytearray(data)
    bad[100] ^= 1
    q = str(tmp_path / "bad.bin")
    open(q, "wb").write(bytes(bad))
    with pytest.raises(fm.FlexMeshError, match="checksum"):
        S.checkpoint_load(q)
    open(q, "wb").write(bytes(data[:-100]))
    with pytest.raises(fm.FlexMeshError):
        S.checkpoint_load(q)
    # the same sizes, another material
    S.set_material(cfg["mu"] * 1.01, cfg["lam"])
    with pytest.raises(fm.FlexMeshError, match="material"):
        S.checkpoint_load(p)
    S.set_material(cfg["mu"], cfg["lam"])
    # the same sizes, moved nodes
    X2 = cfg["X"].copy()
    X2[5] += 1e-3
    S.set_coords(X2)
    with pytest.raises(fm.FlexMeshError, match="coordinates"):
        S.checkpoint_load(p)
    S.set_coords(cfg["X"])
    assert S.checkpoint_load(p) == 4
