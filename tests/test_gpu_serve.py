"""Serving (pt_server_*) on the GPU: convergence detection, independence of
how samples are split and interleaved, the relative-RMSE estimate against a
converged render, scheduling order and HBM admission."""
import json

import numpy as np
import pytest

import pathtracer_amd as pt
from pathtracer_amd import scenes

pytestmark = pytest.mark.gpu


def no_caustic_cornell(seed=0):
    s = scenes.cornell_dict(4, seed)
    s["material"][4] = {"type": "lambert", "gaussians": [[0.6, 600.0, 80.0]]}
    return json.dumps(s)


def run_until_done(srv, ids, max_steps=10000):
    done = set()
    for _ in range(max_steps):
        sts = srv.step()
        if sts is None:
            break
        done |= {st.id for st in sts if st.state == pt.REQ_DONE}
        if done >= set(ids):
            break
    return {i: srv.status(i) for i in ids}


@pytest.mark.parametrize("kind", ["spheres", "sdf", "shapes"])
def test_result_is_the_same_alone_and_in_the_mix(kind):
    """Continuous batching: a request's paths share waves with other requests
    (other scenes, SDF and analytic, with and without lenses and cyclides,
    other depths and resolutions), in quanta of whatever size the wave leaves;
    its per-pixel sums are added in sample order, so the image is
    bit-identical to the request traced alone."""
    from tests.test_gpu_ext_shapes import shapes_cornell
    text, depth = {"spheres": (scenes.random_spheres(n=256, scene_seed=3, wavelengths=4, seed=11), 8),
                   "sdf": (scenes.sdf_scene(jitter=0.05, scene_seed=4, seed=12), 6),
                   "shapes": (shapes_cornell(emissive=True), 8)}[kind]
    others = [(scenes.cornell(4, seed=1), 24, 24, 6), (scenes.sdf_scene(seed=2), 32, 16, 5),
              (scenes.random_spheres(n=64, scene_seed=9, wavelengths=1, seed=3), 40, 20, 12),
              (shapes_cornell(glass_lens=True), 16, 16, 7)]
    imgs = []
    # graph 0: each wave shape captured once into a hipGraph and replayed every
    # step with new tables; graph 1: kernels launched one by one
    for wave, mix, graph in ((1 << 22, False, 1), (20000, False, 0), (1 << 22, True, 0), (30000, True, 0),
                             (30000, True, 1)):
        with pt.Context(0, wave_paths=wave, graph_mode=graph) as ctx, pt.Server(ctx) as srv:
            rid = srv.submit(text, 48, 32, spp_max=48, path_length=depth, target=1e-6)  # runs to spp_max
            extra = [srv.submit(t, w, h, spp_max=24, path_length=d, target=1e-6) for t, w, h, d in others] if mix else []
            st = run_until_done(srv, [rid] + extra)
            assert st[rid].state == pt.REQ_DONE and st[rid].spp == 48 and not st[rid].converged
            assert all(st[i].spp == 24 for i in extra)
            imgs.append(srv.read(rid))
    for im in imgs[1:]:
        assert np.array_equal(imgs[0], im)
    assert np.isfinite(imgs[0]).all() and imgs[0][..., 1].mean() > 0


def test_mean_matches_the_renderer(gpu_ctx):
    """Served mean == the running-mean texels of pt_dispatch with the same samples."""
    text = no_caustic_cornell(seed=4)
    with pt.Server(gpu_ctx) as srv:
        rid = srv.submit(text, 40, 40, spp_max=32, path_length=8, target=1e-6)
        run_until_done(srv, [rid])
        served = srv.read(rid)[..., :3].astype(np.float64)
    sc = gpu_ctx.scene_from_json(text)
    r = pt.Renderer(gpu_ctx, sc, 40, 40, spf=32, path_length=8, mode=pt.MODE_EXT)
    r.render_frame()
    ref = r.acc.read()[..., :3].astype(np.float64)
    # the batched kernels do the single-scene kernels' arithmetic, path for path
    np.testing.assert_array_equal(served, ref)


def test_relative_error_estimate_tracks_the_converged_render(gpu_ctx):
    """Stopping at a 3 % estimate gives an image whose actual relative RMSE
    against a 64x longer render is close to 3 %."""
    text = no_caustic_cornell(seed=6)
    W = H = 48
    with pt.Server(gpu_ctx) as srv:
        a = srv.submit(text, W, H, spp_max=4096, path_length=8, target=0.03)
        st = run_until_done(srv, [a])[a]
        # converged without overshooting far past the target
        assert st.converged and 0.025 <= st.rel_error <= 0.03
        img = srv.read(a)[..., :3].astype(np.float64)
        b = srv.submit(text, W, H, spp_max=64 * st.spp, path_length=8, target=1e-6)
        run_until_done(srv, [b])
        ref = srv.read(b)[..., :3].astype(np.float64)
    # the reference shares the first st.spp samples: compare against its other samples only
    n = 64 * st.spp
    ref_other = (ref * n - img * st.spp) / (n - st.spp)
    actual = np.sqrt(((img - ref_other) ** 2).sum()) / np.sqrt((ref_other ** 2).sum())
    print(f"\nspp {st.spp}: estimate {st.rel_error:.4f}, actual {actual:.4f}")
    assert 0.6 * st.rel_error <= actual <= 1.5 * st.rel_error


def test_policies_order_requests(gpu_ctx):
    big = scenes.random_spheres(n=1024, scene_seed=2, wavelengths=4, seed=1)
    small = no_caustic_cornell(seed=2)
    order = {}
    # "aged": the big request arrived long before the small ones
    for policy, aged in ((pt.SERVE_FIFO, False), (pt.SERVE_SRPT, False), (pt.SERVE_HRRN, False),
                         (pt.SERVE_HRRN, True), (pt.SERVE_SRPT, True)):
        with pt.Server(gpu_ctx, policy=policy) as srv:
            t = srv.now()
            ids = [srv.submit(big, 256, 256, spp_max=256, path_length=8, target=1e-6,
                              t_arrive=t - 1000.0 if aged else t)]
            ids += [srv.submit(small, 32, 32, spp_max=64, path_length=8, target=1e-6, t_arrive=t) for _ in range(3)]
            st = run_until_done(srv, ids)
            order[policy, aged] = sorted(ids, key=lambda i: st[i].t_done)
            big_id = ids[0]
    assert order[pt.SERVE_FIFO, False][0] == big_id  # one request at a time in arrival order
    assert order[pt.SERVE_SRPT, False][-1] == big_id  # the small ones overtake it
    assert order[pt.SERVE_SRPT, True][-1] == big_id  # whatever its wait
    # highest response ratio first: with equal waits the shorter remaining time
    # leads; a request that has waited long enough leads whatever its size
    assert order[pt.SERVE_HRRN, False][-1] == big_id
    assert order[pt.SERVE_HRRN, True][0] == big_id


def test_hbm_budget_limits_admission_and_release_frees(gpu_ctx):
    text = scenes.random_spheres(n=64, scene_seed=5, wavelengths=4, seed=2)
    # a 4096-path wave traces one sample per pixel of a 64 x 64 request per step
    with pt.Context(0, wave_paths=4096) as ctx, pt.Server(ctx, hbm_budget=1) as srv:  # one request fits at a time
        ids = [srv.submit(text, 64, 64, spp_max=8, path_length=6, target=1e-6) for _ in range(3)]
        sts = srv.step()
        assert len(sts) == 1 and sts[0].id == ids[0] and sts[0].spp == 1
        waiting, active, used = srv.counts()
        assert active == 1 and waiting == 2 and used > 0
        run_until_done(srv, ids)
        for i in ids:
            assert srv.status(i).spp == 8
            srv.release(i)
        assert srv.counts() == (0, 0, 0)


def test_bad_requests_are_rejected(gpu_ctx):
    with pt.Server(gpu_ctx) as srv:
        with pytest.raises(pt.PTError):
            srv.submit(scenes.cornell(4), 0, 16, spp_max=4, path_length=4)
        rid = srv.submit('{"camera": 1}', 16, 16, spp_max=4, path_length=4)
        with pytest.raises(pt.PTError):
            srv.step()
        assert srv.status(rid).state == pt.REQ_FAILED
        assert srv.step() is None


def test_request_larger_than_a_wave_is_rejected_at_submit():
    """A batched wave traces at least one sample of every pixel of a request,
    so a request with more pixels than the wave holds paths is refused when it
    is submitted (PT_ERR_CAPACITY), and one that fits is still served."""
    with pt.Context(0, wave_paths=4096) as ctx, pt.Server(ctx) as srv:
        with pytest.raises(pt.PTError) as e:
            srv.submit(no_caustic_cornell(), 65, 64, spp_max=4, path_length=4)
        assert e.value.status == pt.PT_ERR_CAPACITY
        rid = srv.submit(no_caustic_cornell(), 64, 64, spp_max=4, path_length=4, target=1e-6)
        st = run_until_done(srv, [rid])
        assert st[rid].state == pt.REQ_DONE and st[rid].spp == 4


def test_scene_with_overflowing_bounds_is_rejected(gpu_ctx):
    """A sphere whose box overflows the float range (centre 3e38, radius 1e38)
    is refused when the extended-mode scene is built (its first dispatch, or a
    request's admission) instead of reaching the BVH builder."""
    d = json.loads(scenes.random_spheres(n=8, scene_seed=2, wavelengths=4))
    d["sphere"][3]["position"] = [3e38, 3e38, 3e38]
    d["sphere"][3]["radius"] = 1e38
    text = json.dumps(d)
    scene = gpu_ctx.scene_from_json(text)
    r = pt.Renderer(gpu_ctx, scene, 16, 16, spf=1, path_length=4, mode=pt.MODE_EXT)
    with pytest.raises(pt.PTError) as e:
        r.render_frame()
    assert e.value.status == pt.PT_ERR_INVALID_ARG and "float range" in str(e.value)
    with pt.Server(gpu_ctx) as srv:
        rid = srv.submit(text, 16, 16, spp_max=4, path_length=4)
        with pytest.raises(pt.PTError):
            srv.step()
        assert srv.status(rid).state == pt.REQ_FAILED


def attach_shipped_denoiser(ctx, srv, size=128, batch=1):
    from pathtracer_amd import denoise_train as T
    dn = pt.Denoiser(ctx, batch, size, size)
    T.load_weights(dn)
    srv.set_denoiser(dn)
    return dn


def test_denoised_result_is_the_same_alone_and_in_the_mix():
    """A request that may finish denoised, run to its sample budget: its image
    (denoised or, as fallback, plain) depends only on its own sums. A denoiser
    with room for a batch of 2 runs both half images of the error estimate in
    one pass, which must not change the estimate either."""
    text = scenes.sdf_scene(jitter=0.05, scene_seed=4, seed=12)
    others = [(scenes.cornell(4, seed=1), 24, 24, 6), (scenes.random_spheres(n=64, scene_seed=9, seed=3), 40, 20, 12)]
    outs = []
    for wave, mix, graph, batch in ((1 << 22, False, 0, 1), (20000, True, 0, 2), (30000, True, 1, 2)):
        with pt.Context(0, wave_paths=wave, graph_mode=graph) as ctx, pt.Server(ctx) as srv:
            dn = attach_shipped_denoiser(ctx, srv, 64, batch)
            rid = srv.submit(text, 48, 32, spp_max=40, path_length=6, target=1e-6, denoise=True)
            extra = [srv.submit(t, w, h, spp_max=24, path_length=d, target=1e-6) for t, w, h, d in others] if mix else []
            st = run_until_done(srv, [rid] + extra)[rid]
            assert st.state == pt.REQ_DONE and st.spp == 40
            outs.append((srv.read(rid), st.denoised, st.rel_error, st.rel_plain))
            dn.close()
    for img, den, rel, relp in outs[1:]:
        assert np.array_equal(outs[0][0], img) and den == outs[0][1] and rel == outs[0][2] and relp == outs[0][3]
    img, den, rel, relp = outs[0]
    # at the budget the image with the lower estimate is returned
    assert rel == (min(rel, relp) if den else relp)
    assert np.isfinite(img).all()


def test_denoised_request_meets_its_target_with_fewer_samples(gpu_ctx):
    """A 2 % target on the mandelbulb + menger scene: the denoised request
    stops several times earlier than the plain one; the dual-buffer estimate
    tracks the actual relative RMSE against an 8192-sample render, which meets
    the target within the estimate's spread."""
    from pathtracer_amd import denoise_train as T
    text, depth = T.training_scene(101)
    W = H = 128
    with pt.Server(gpu_ctx) as srv:
        dn = attach_shipped_denoiser(gpu_ctx, srv, 128)
        a = srv.submit(text, W, H, spp_max=4096, path_length=depth, target=0.02, denoise=True)
        b = srv.submit(text, W, H, spp_max=4096, path_length=depth, target=0.02)
        st = run_until_done(srv, [a, b])
        img_dn = srv.read(a)[..., :3]
        img_plain = srv.read(b)[..., :3]
        dn.close()
    sa, sb = st[a], st[b]
    print(f"\ndenoised: spp {sa.spp} est {sa.rel_error:.4f} (plain est {sa.rel_plain:.4f}); "
          f"plain: spp {sb.spp} est {sb.rel_error:.4f}")
    assert sa.converged and sa.denoised and sb.converged and not sb.denoised
    assert sa.spp * 3 <= sb.spp
    sc = gpu_ctx.scene_from_json(text)
    ref = T.render_mean(gpu_ctx, sc, W, H, 8192, depth, first=10 ** 6)
    sc.release()
    act_dn, act_plain = T.rel_rmse(img_dn, ref), T.rel_rmse(img_plain, ref)
    print(f"actual: denoised {act_dn:.4f}, plain {act_plain:.4f}")
    assert act_dn <= 1.2 * 0.02 and act_plain <= 1.5 * 0.02
    assert 0.7 * act_dn <= sa.rel_error <= 1.4 * act_dn


def test_denoising_falls_back_to_plain(gpu_ctx):
    text = scenes.sdf_scene(seed=3)
    with pt.Server(gpu_ctx) as srv:
        with pytest.raises(pt.PTError):  # no denoiser attached
            srv.submit(text, 32, 32, spp_max=8, path_length=4, denoise=True)
        dn = attach_shipped_denoiser(gpu_ctx, srv, 64)
        # 30 x 30 is not a multiple of 4, 96 x 64 is wider than the network's 64: served plain
        ids = [srv.submit(text, w, h, spp_max=16, path_length=4, target=1e-6, denoise=True)
               for w, h in ((30, 30), (96, 64))]
        st = run_until_done(srv, ids)
        assert all(st[i].state == pt.REQ_DONE and not st[i].denoised for i in ids)
        assert all(st[i].rel_error == st[i].rel_plain for i in ids)
        dn.close()
