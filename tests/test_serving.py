"""Serving loop: admission by memory budget and request generation on the
CPU; on the GPU, every result of a continuously batched run (and of the
one-at-a-time run) equals the same request rendered alone with the
exact-size pipeline, bit for bit."""
import numpy as np
import pytest

from rgp.serving import AdmissionError, SceneSpec, SceneStore, edited_palette, make_requests

SMALL = [SceneSpec("a", 20_000, 5, False, seed=1), SceneSpec("b", 30_000, 7, True, seed=2),
         SceneSpec("c", 40_000, 5, True, seed=3)]
SMALL_RES = ((320, 192), (480, 272))


def test_admission_budget():
    specs = [SceneSpec("x", 1_000_000, 5, False), SceneSpec("y", 3_000_000, 7, True)]
    need = [s.footprint() for s in specs]
    store = SceneStore(need[0] + need[1] - 1)
    store.admit(specs[0])
    with pytest.raises(AdmissionError):
        store.admit(specs[1])
    store.release("x")
    store.admit(specs[1])
    assert store.used == need[1]


def test_default_workload_fits_200gb():
    from rgp.serving import default_specs
    store = SceneStore(200 * 2**30)
    for s in default_specs():
        store.admit(s)
    assert store.used < 200 * 2**30


def test_request_stream():
    specs = SMALL
    reqs = make_requests(specs, 4000, rate=500.0, seed=3)
    first = {}
    for r in reqs:
        first.setdefault(r.scene, r)
    assert all(r.kind == "pose" for r in first.values())
    frac_pose = np.mean([r.kind == "pose" for r in reqs])
    assert 0.35 < frac_pose < 0.45
    gaps = np.diff([r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 500.0) < 2e-4


def test_edited_palette():
    P = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.2, 0.9]], np.float32)
    out = edited_palette(P, np.full((3, 3), 0.6, np.float32), 1.2)
    assert (out[:2] == P[:2]).all()
    assert np.allclose(out[2], np.clip((P[2] + 0.6) * 1.2, 0, 1))


def expected_outputs(engine, reqs):
    """Each request alone: exact-size pipeline of rgp.gsmodel + mix."""
    import torch
    from rgp import gsmodel, ops
    view = {}
    out = {}
    for r in sorted(reqs, key=lambda q: q.arrival):
        sc = engine.scenes[r.scene]
        if r.kind == "pose":
            view[r.scene] = (r.pose, r.res)
        pose, res = view[r.scene]

        class Cam:
            pass
        cam = Cam()
        host = engine.camera(pose, *res).numpy()
        cam.Rcw, cam.tcw, cam.twc = host[0:9].reshape(3, 3), host[9:12], host[12:15]
        cam.fx, cam.fy, cam.cx, cam.cy = (float(v) for v in host[15:19])
        cam.width, cam.height = res
        kw = dict(shs=sc.shs) if sc.mode == 0 else dict(dc=sc.dc, A=sc.A, B=sc.B, Q=sc.Q)
        layer, _ = gsmodel.render_weights(sc.pws, sc.alphas, sc.scales, sc.rots, cam, sc.K, **kw)
        P = sc.palette if r.kind == "pose" else edited_palette(sc.palette, r.dpal, r.lum)
        out[r.rid] = ops.mix(layer, torch.from_numpy(P.astype(np.float32)).to(layer.device)).cpu().numpy()
    return out


def test_default_workload_with_batch_slots_fits_200gb():
    from rgp.serving import default_specs
    store = SceneStore(200 * 2**30, batch_views=8)
    for s in default_specs():
        store.admit(s)
    assert store.used < 200 * 2**30
    assert store.used > SceneStore(200 * 2**30).admit(default_specs()[0])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["batched", "batched_oneview", "batched_smallcap", "serial", "serial_rerender"])
def test_results_alone_equal_in_mix(device, mode, monkeypatch):
    """batched: poses of a scene and resolution pending together render as
    multi-view batches; _oneview: one pose per launch; _smallcap: pair buffers
    planned far too small, so batches overflow, grow and re-render the
    requests that overflowed."""
    import rgp.serving as S
    monkeypatch.setattr(S, "RESOLUTIONS", SMALL_RES)
    print(f"case: serving {mode}", flush=True)
    kw = dict(batch_views=0) if mode == "batched_oneview" else {}
    if mode == "batched_smallcap":
        kw = dict(cap_margin=0.05)
    eng = S.Engine(SMALL, device, n_poses=8, **kw)
    reqs = S.make_requests(SMALL, 120, rate=0.0, seed=5, n_poses=8)
    eng.prepare(reqs)
    # 32 requests admitted per loop iteration: the run spans several iterations
    rep = eng.run(reqs, mode=mode.split("_")[0], realtime=False, rerender_edits=mode.endswith("rerender"),
                  max_batch=32)
    assert rep["requests"] == 120 and rep["edits"] > 40
    assert all(r.done is not None for r in reqs)
    exp = expected_outputs(eng, reqs)
    for r in reqs:
        got = r.out.cpu().numpy()
        assert np.array_equal(got, exp[r.rid]), (r.rid, r.kind, r.scene)
    if mode == "batched":
        assert any(v > 1 for bs in eng.bslots.values() for v in bs.graphs)  # some batches held several views
        # the same engine then serves one request at a time, with the same results
        eng.run(reqs, mode="serial", realtime=False, max_batch=32)
        for r in reqs:
            assert np.array_equal(r.out.cpu().numpy(), exp[r.rid]), ("serial after batched", r.rid)


@pytest.mark.gpu
def test_scenes_from_files(device, monkeypatch, tmp_path):
    """Scenes exported as .pply + .gswp (dense and rank-1) are admitted from
    their sidecar headers, load the arrays the files hold, and serve results
    equal to the same requests rendered alone."""
    import rgp.serving as S
    from rgp import formats
    from rgp.scene import make_scene
    monkeypatch.setattr(S, "RESOLUTIONS", SMALL_RES)
    print("case: serving scenes from .pply + .gswp", flush=True)
    specs = []
    for name, lowrank in (("file_dense", False), ("file_rank1", True)):
        sc = make_scene(20_000, seed=7, extent=3.0, K=5, lowrank=lowrank, rank=1, P=5 if lowrank else None, Q=15)
        kw = dict(dc=sc.dc, A=sc.A, B=sc.B) if lowrank else dict(shs=sc.shs)
        pply, _ = formats.export(formats.checkpoint(sc.pws, sc.rots, sc.scales, sc.alphas, sc.palette, **kw),
                                 tmp_path / f"{name}.pply")
        spec = S.SceneSpec.from_files(name, pply)
        assert (spec.n, spec.K, spec.lowrank) == (20_000, 5, lowrank)
        specs.append(spec)
    eng = S.Engine(specs, device, n_poses=8)
    for spec in specs:
        want = formats.load_scene_files(*spec.files)
        got = eng.scenes[spec.name]
        assert np.array_equal(got.scales.cpu().numpy(), want["scales"])
        assert np.array_equal(got.alphas.cpu().numpy(), want["alphas"])
        assert np.array_equal(got.palette, want["palette"])
    reqs = S.make_requests(specs, 60, rate=0.0, seed=6, n_poses=8)
    eng.prepare(reqs)
    rep = eng.run(reqs, mode="batched", realtime=False)
    assert rep["requests"] == 60
    exp = expected_outputs(eng, reqs)
    for r in reqs:
        assert np.array_equal(r.out.cpu().numpy(), exp[r.rid]), (r.rid, r.kind, r.scene)
