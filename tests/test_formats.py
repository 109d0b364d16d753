"""On-disk formats (rgp.formats) against the byte layouts the reference
documents and writes: convert_npy_ply/palette_npy_to_pply.py:13-39 (.gswp
header and payload, .pply conventions) and 176-270 (the writers),
gsplat/gau_io.py:10-15 (vanilla record type), finetune.py:440-468 (the
checkpoint dict).  Expected bytes are assembled here with struct from those
layouts, independently of the writers."""
import os
import struct

import numpy as np
import pytest

from rgp import formats as F
from rgp.scene import make_scene


def dense_ckpt(n=3, K=4, nsh=2, seed=0):
    rng = np.random.default_rng(seed)
    f = lambda *s: rng.standard_normal(s).astype(np.float32)
    return {"pw": f(n, 3), "rot": f(n, 4), "scale": np.exp(f(n, 3)), "alpha": rng.uniform(0.01, 0.99, n).astype(
        np.float32), "palette": rng.uniform(0, 1, (K, 3)).astype(np.float32), "shs_w": f(n, K * nsh),
            "low_rank": False}


def lowrank_ckpt(n=3, K=5, P=5, Q=15, seed=1):
    c = dense_ckpt(n, K, 1, seed)
    del c["shs_w"]
    rng = np.random.default_rng(seed + 10)
    c.update(low_shs_w=rng.standard_normal((n, K)).astype(np.float32),
             high_shs_A=rng.standard_normal((n, P)).astype(np.float32),
             high_shs_B=rng.standard_normal((n, Q)).astype(np.float32), low_rank=True, reshape_P=P, reshape_Q=Q)
    return c


def test_gswp_dense_bytes(tmp_path):
    c = dense_ckpt(n=2, K=3, nsh=2)
    p = tmp_path / "s.gswp"
    F.write_gswp(p, c)
    sh = c["shs_w"].reshape(2, 2, 3)
    want = struct.pack("<4sIIIIIII", b"GSWP", 1, 2, 3, 2, 0, 0, 0)
    want += c["palette"].astype("<f4").tobytes() + sh[:, 0, :].astype("<f4").tobytes()
    want += sh[:, 1:, :].reshape(2, 3).astype("<f4").tobytes()
    assert p.read_bytes() == want
    assert len(want) == 32 + 4 * (3 * 3 + 2 * 3 + 2 * 3)


def test_gswp_lowrank_bytes(tmp_path):
    c = lowrank_ckpt(n=2)
    p = tmp_path / "s.gswp"
    F.write_gswp(p, c)
    want = struct.pack("<4sIIIIIII", b"GSWP", 1, 2, 5, 16, 1, 5, 15)
    for k in ("palette", "low_shs_w", "high_shs_A", "high_shs_B"):
        want += c[k].astype("<f4").tobytes()
    assert p.read_bytes() == want


def test_pply_bytes(tmp_path):
    c = dense_ckpt(n=2)
    c["scale"][:] = 1.0       # log 1 = 0 exactly
    c["alpha"][:] = 0.5       # logit 0.5 = 0 exactly
    p = tmp_path / "s.pply"
    F.write_pply(p, c)
    props = ["x", "y", "z", "scale_0", "scale_1", "scale_2", "opacity", "rot_0", "rot_1", "rot_2", "rot_3"]
    head = "\n".join(["ply", "format binary_little_endian 1.0",
                      "comment palette-based 3DGS - geometry only; pair with .gswp sidecar", "element vertex 2"]
                     + [f"property float {q}" for q in props] + ["end_header", ""]).encode("ascii")
    body = b"".join(struct.pack("<11f", *c["pw"][i], 0.0, 0.0, 0.0, 0.0, *c["rot"][i]) for i in range(2))
    assert p.read_bytes() == head + body


def test_pply_log_and_logit(tmp_path):
    c = dense_ckpt(n=64, seed=3)
    c["alpha"][:3] = [0.0, 1.0, 1e-9]  # clipped to [1e-6, 1 - 1e-6]
    p = tmp_path / "s.pply"
    F.write_pply(p, c)
    cols, comments = F.read_pply(p)
    assert comments == ["palette-based 3DGS - geometry only; pair with .gswp sidecar"]
    ls = np.stack([cols[f"scale_{i}"] for i in range(3)], 1).astype(np.float64)
    assert np.allclose(ls, np.log(c["scale"].astype(np.float64)), rtol=0, atol=2e-7)
    # the clip runs in float32, as in the reference (alpha is float32 there too)
    a = np.clip(c["alpha"], 1e-6, 1 - 1e-6).astype(np.float64)
    assert np.allclose(cols["opacity"], np.log(a / (1 - a)), rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize("kind", ["dense", "lowrank"])
def test_scene_round_trip(tmp_path, kind):
    if kind == "dense":
        sc = make_scene(500, seed=4, K=5)
        c = F.checkpoint(sc.pws, sc.rots, sc.scales, sc.alphas, sc.palette, shs=sc.shs)
    else:
        sc = make_scene(500, seed=4, K=5, lowrank=True, rank=1, P=5, Q=15)
        c = F.checkpoint(sc.pws, sc.rots, sc.scales, sc.alphas, sc.palette, dc=sc.dc, A=sc.A, B=sc.B)
    pply, gswp = F.export(c, tmp_path / "scene.pply")
    got = F.load_scene_files(pply)
    assert np.array_equal(got["pws"], sc.pws) and np.array_equal(got["rots"], sc.rots)
    assert np.array_equal(got["palette"], sc.palette) and got["K"] == 5 and got["num_sh"] == 16
    assert np.allclose(got["scales"], sc.scales, rtol=4e-7, atol=0)
    assert np.allclose(got["alphas"], sc.alphas, rtol=0, atol=1e-6)
    if kind == "dense":
        assert np.array_equal(got["shs"], sc.shs)
    else:
        assert got["low_rank"] and (got["P"], got["Q"]) == (5, 15)
        assert np.array_equal(got["dc"], sc.dc) and np.array_equal(got["A"], sc.A) and np.array_equal(got["B"], sc.B)


def test_rank2_needs_dense():
    sc = make_scene(10, seed=1, K=5, lowrank=True, rank=2, P=5, Q=15)
    with pytest.raises(ValueError, match="rank 1"):
        F.checkpoint(sc.pws, sc.rots, sc.scales, sc.alphas, sc.palette, dc=sc.dc, A=sc.A, B=sc.B)


def test_checkpoint_npy_needs_trust(tmp_path):
    c = lowrank_ckpt()
    p = str(tmp_path / "ck.npy")
    F.save_palette_npy(p, c)
    with pytest.raises(PermissionError):
        F.load_palette_npy(p)
    back = F.load_palette_npy(p, trusted=True)
    assert sorted(back) == sorted(c)
    assert all(np.array_equal(np.asarray(back[k]), np.asarray(c[k])) for k in c)
    z = str(tmp_path / "ck.npz")
    F.save_palette_npz(z, c)
    back = F.load_palette_npz(z)
    assert back["low_rank"] is True and back["reshape_Q"] == 15
    assert all(np.array_equal(np.asarray(back[k]), np.asarray(c[k])) for k in c)


def test_convert_cli_layout(tmp_path):
    c = dense_ckpt(n=5, K=5, nsh=16)
    z = tmp_path / "truck_palette.npz"
    F.save_palette_npz(str(z), c)
    assert F.main([str(z), "--quiet"]) == 0
    out = tmp_path / "truck_palette"
    assert sorted(os.listdir(out)) == ["truck_palette.gswp", "truck_palette.pply"]
    h = F.read_gswp(out / "truck_palette.gswp")
    assert (h["N"], h["splat_dim"], h["num_sh"], h["low_rank"]) == (5, 5, 16, False)
    assert F.main([str(z), str(tmp_path / "x" / "a.pply"), "--sidecar", str(tmp_path / "b.gswp"), "--quiet"]) == 0
    assert (tmp_path / "x" / "a.pply").exists() and (tmp_path / "b.gswp").exists()
    info = F.inspect(str(z))
    assert info["num_sh"] == 16


def test_validation_errors():
    c = dense_ckpt()
    for key in ("pw", "palette", "shs_w"):
        bad = dict(c)
        del bad[key]
        with pytest.raises(KeyError):
            F.palette_info(bad)
    bad = dict(c, shs_w=c["shs_w"][:, :-1])
    with pytest.raises(ValueError):
        F.palette_info(bad)
    bad = dict(c, rot=c["rot"][:, :3])
    with pytest.raises(ValueError):
        F.palette_info(bad)
    lr = lowrank_ckpt(P=4, Q=14)  # 56 is not a multiple of K = 5
    with pytest.raises(ValueError, match="multiple"):
        F.palette_info(lr)


def test_bad_sidecars(tmp_path):
    c = dense_ckpt()
    p = tmp_path / "s.gswp"
    F.write_gswp(p, c)
    data = p.read_bytes()
    (tmp_path / "short.gswp").write_bytes(data[:-4])
    with pytest.raises(ValueError, match="implies"):
        F.read_gswp(tmp_path / "short.gswp")
    (tmp_path / "magic.gswp").write_bytes(b"GSWQ" + data[4:])
    with pytest.raises(ValueError, match="magic"):
        F.read_gswp(tmp_path / "magic.gswp")
    (tmp_path / "ver.gswp").write_bytes(data[:4] + struct.pack("<I", 2) + data[8:])
    with pytest.raises(ValueError, match="version"):
        F.read_gswp(tmp_path / "ver.gswp")


def test_vanilla_npy(tmp_path):
    rng = np.random.default_rng(5)
    n = 7
    arrs = [rng.standard_normal((n, d)).astype(np.float32) for d in (3, 4, 3)]
    alphas, shs = rng.uniform(0, 1, n).astype(np.float32), rng.standard_normal((n, 48)).astype(np.float32)
    p = str(tmp_path / "v.npy")
    F.save_vanilla_npy(p, *arrs, alphas, shs)
    rec = np.load(p, allow_pickle=False)
    assert rec.dtype == F.vanilla_dtype(48) and rec.dtype.names == ("pw", "rot", "scale", "alpha", "sh")
    got = F.load_vanilla_npy(p)
    for k, v in zip(("pws", "rots", "scales", "alphas", "shs"), (*arrs, alphas, shs)):
        assert np.array_equal(got[k], v)
    np.save(str(tmp_path / "plain.npy"), np.zeros((3, 3), np.float32))
    with pytest.raises(ValueError):
        F.load_vanilla_npy(str(tmp_path / "plain.npy"))
