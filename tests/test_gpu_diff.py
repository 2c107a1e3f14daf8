"""GPU parity of the differentiable form (primal and path-replay adjoint) through
the C-ABI and the torch autograd extension.

  * constant textures equal to the material values give the plain megakernel
    render (per-pixel rtol 1e-5: FP32 mean vs FP64 film sums of the same samples);
  * random textures: primal images against the FP32 oracle's (same samples),
    per pixel |gpu - oracle| <= 1e-4 |oracle| + 1e-5 on >= 98% of pixels;
  * adjoint against FP64 central differences (h = 1e-3) of the oracle's loss
    sum(dimg * mean L) with sampling and Russian roulette frozen at the texture
    values, on 64 random texture entries: relative error <= 1e-2 where the
    gradient is above 1% of the largest, absolute <= 1e-4 * max elsewhere;
  * the torch Function's forward equals the C-ABI call bit for bit, its backward
    equals orf_diff_adjoint up to FP32 atomic summation order;
  * every adjoint launch reports zero samples whose replay did not end on the
    radiance of its first pass (orf_diff_counters).
"""
import sys

import numpy as np
import pytest

from open_rendering_framework_amd import scenes

pytestmark = pytest.mark.gpu

W = 32
TEX = 8
SHIFT = 0x40000000


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _scene():
    return scenes.path_scene(width=W, height=W, level=3, env_size=(64, 32))


def _desc(rt):
    return rt.DiffDesc.make(scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, (TEX, TEX),
                            scenes.DIFF_ALBEDO_UV, scenes.DIFF_ROUGHNESS_UV)


def test_constant_textures_equal_plain_render():
    import torch

    from open_rendering_framework_amd import rt

    _say("diff: constant textures vs plain render")
    d = _scene()
    ctx = scenes.to_context(d)
    views = scenes.ring_views(n=2, size=W, spp=4)
    alb = torch.full((TEX, TEX, 3), 0.5, device="cuda")
    rough = torch.full((TEX, TEX), 0.3, device="cuda")
    img = ctx.diff_render(_desc(rt), views, alb, rough).cpu().numpy()
    for i, v in enumerate(views):
        for k in ("eye", "U", "V", "W"):
            ctx.set_float(k, *v[k])
        ctx.film_clear()
        ctx.launch(W, W, v["frame0"], v["spp"])
        s = ctx.film_sums(W, W)
        want = s[..., :3] / np.maximum(s[..., 3:], 1)
        np.testing.assert_allclose(img[i], want, rtol=1e-5, atol=1e-6)


def test_primal_matches_oracle():
    import torch

    from open_rendering_framework_amd import rt
    from tests.oracle_scene import OracleScene

    _say("diff: primal vs FP32 oracle")
    d = _scene()
    ctx = scenes.to_context(d)
    views = scenes.ring_views(n=2, size=W, spp=8)
    a, r = scenes.diff_textures(TEX, seed=5)
    img = ctx.diff_render(_desc(rt), views, torch.from_numpy(a).cuda(), torch.from_numpy(r).cuda()).cpu().numpy()
    orc = OracleScene(d)
    orc.diff_set((TEX, TEX), scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, a, r)
    _, want = orc.diff_eval(views)
    ok = (np.abs(img - want) <= 1e-4 * np.abs(want) + 1e-5).all(axis=-1)
    _say("  pixels within tolerance: %.5f" % ok.mean())
    assert ok.mean() >= 0.98


def _fd_setup():
    import torch

    from open_rendering_framework_amd import rt

    d = _scene()
    ctx = scenes.to_context(d)
    views = scenes.ring_views(n=2, size=W, spp=8)
    a, r = scenes.diff_textures(TEX, seed=5)
    rng = np.random.default_rng(3)
    dimg = rng.standard_normal((2, W, W, 3)).astype(np.float32)
    ga, gr = ctx.diff_adjoint(_desc(rt), views, torch.from_numpy(a).cuda(), torch.from_numpy(r).cuda(),
                              torch.from_numpy(dimg).cuda(), frame_shift=SHIFT)
    _assert_replays_match(ctx)
    return d, views, a, r, dimg, ga.cpu().numpy(), gr.cpu().numpy()


def _assert_replays_match(ctx):
    st = ctx.diff_counters()
    _say("  adjoint counters %s" % st)
    assert st["replay_mismatch"] == 0, st


def test_adjoint_matches_fp64_central_differences():
    from tests.oracle_scene import OracleScene

    _say("diff: adjoint vs FP64 central differences")
    d, views, a, r, dimg, ga, gr = _fd_setup()
    o64 = OracleScene(d, fp64=True)
    a0, r0, dd = a.astype(np.float64), r.astype(np.float64), dimg.astype(np.float64)
    h = 1e-3

    def loss(aa, rr):
        o64.diff_set((TEX, TEX), scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, aa, rr, a0, r0)
        return o64.diff_eval(views, shift=SHIFT, dimg=dd, images=False)[0]

    rng = np.random.default_rng(11)
    # 64 distinct entries of the 8x8x3 albedo + 8x8 roughness parameter vector
    picks = rng.choice(TEX * TEX * 4, size=64, replace=False)
    got, fd = [], []
    for p in picks:
        aa, rr = a0.copy(), r0.copy()
        if p < TEX * TEX * 3:
            idx = np.unravel_index(p, a0.shape)
            aa[idx] += h
            lp = loss(aa, r0)
            aa[idx] -= 2 * h
            lm = loss(aa, r0)
            got.append(ga[idx])
        else:
            idx = np.unravel_index(p - TEX * TEX * 3, r0.shape)
            rr[idx] += h
            lp = loss(a0, rr)
            rr[idx] -= 2 * h
            lm = loss(a0, rr)
            got.append(gr[idx])
        fd.append((lp - lm) / (2 * h))
    got, fd = np.array(got, np.float64), np.array(fd)
    big = np.abs(fd) > 1e-2 * np.abs(fd).max()
    rel = np.abs(got - fd) / np.maximum(np.abs(fd), 1e-30)
    _say("  entries %d (large %d): max rel %.3e, median rel %.3e" % (len(fd), big.sum(), rel[big].max(),
                                                                       np.median(rel[big])))
    assert big.sum() >= 32
    assert np.all(rel[big] <= 1e-2), list(zip(got[big][rel[big] > 1e-2], fd[big][rel[big] > 1e-2]))
    assert np.all(np.abs(got - fd)[~big] <= 1e-4 * np.abs(fd).max())


def test_lds_and_global_gradient_accumulation_agree():
    """The adjoint adds texel gradients into a per-workgroup LDS copy when the
    textures fit (flushed once per workgroup) and straight into global memory
    otherwise; both give the same gradients up to FP32 summation order."""
    import os

    import torch

    from open_rendering_framework_amd import rt

    ctx = scenes.to_context(_scene())
    views = scenes.ring_views(n=2, size=W, spp=8)
    a, r = scenes.diff_textures(TEX, seed=5)
    dimg = np.random.default_rng(3).standard_normal((2, W, W, 3)).astype(np.float32)
    args = (_desc(rt), views, torch.from_numpy(a).cuda(), torch.from_numpy(r).cuda(), torch.from_numpy(dimg).cuda())
    ga0, gr0 = ctx.diff_adjoint(*args, frame_shift=SHIFT)
    os.environ["ORF_DIFF_GLOBAL_GRAD_ATOMICS"] = "1"
    try:
        ga1, gr1 = ctx.diff_adjoint(*args, frame_shift=SHIFT)
    finally:
        del os.environ["ORF_DIFF_GLOBAL_GRAD_ATOMICS"]
    _assert_replays_match(ctx)
    for x, y in ((ga0, ga1), (gr0, gr1)):
        x, y = x.cpu().numpy().astype(np.float64), y.cpu().numpy().astype(np.float64)
        assert np.abs(x).max() > 0
        np.testing.assert_allclose(x, y, rtol=1e-4, atol=1e-6 * np.abs(y).max())


def test_replayed_paths_give_the_same_gradients():
    """An adjoint sample keeps its gradient vertices' terms until its radiance is
    known and replays the path when it holds more of them than the record takes;
    with the record cut to 1 vertex (ORF_DIFF_RECORDS) most paths are replayed,
    every replay reproduces its first pass, and the gradients are the same up to
    FP32 summation order."""
    import os

    import torch

    from open_rendering_framework_amd import rt

    ctx = scenes.to_context(_scene())
    views = scenes.ring_views(n=2, size=W, spp=8)
    a, r = scenes.diff_textures(TEX, seed=5)
    dimg = np.random.default_rng(3).standard_normal((2, W, W, 3)).astype(np.float32)
    args = (_desc(rt), views, torch.from_numpy(a).cuda(), torch.from_numpy(r).cuda(), torch.from_numpy(dimg).cuda())
    ga0, gr0 = ctx.diff_adjoint(*args, frame_shift=SHIFT)
    os.environ["ORF_DIFF_RECORDS"] = "1"
    try:
        ga1, gr1 = ctx.diff_adjoint(*args, frame_shift=SHIFT)
    finally:
        del os.environ["ORF_DIFF_RECORDS"]
    _assert_replays_match(ctx)
    for x, y in ((ga0, ga1), (gr0, gr1)):
        x, y = x.cpu().numpy().astype(np.float64), y.cpu().numpy().astype(np.float64)
        assert np.abs(x).max() > 0
        np.testing.assert_allclose(x, y, rtol=1e-4, atol=1e-6 * np.abs(y).max())


def test_torch_function_matches_c_abi():
    import torch

    from open_rendering_framework_amd import diff, rt

    _say("diff: torch autograd Function vs C-ABI")
    d = _scene()
    ctx = scenes.to_context(d)
    views = scenes.ring_views(n=2, size=W, spp=4)
    a, r = scenes.diff_textures(TEX, seed=7)
    alb = torch.from_numpy(a).cuda().requires_grad_(True)
    rough = torch.from_numpy(r).cuda().requires_grad_(True)
    R = diff.DiffRenderer(ctx, views, scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, (TEX, TEX),
                          scenes.DIFF_ALBEDO_UV, scenes.DIFF_ROUGHNESS_UV, frame_shift=SHIFT)
    img = R(alb, rough)
    target = torch.zeros_like(img)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    want_img = ctx.diff_render(_desc(rt), views, alb.detach(), rough.detach())
    assert torch.equal(img.detach(), want_img)
    ga, gr = ctx.diff_adjoint(_desc(rt), views, alb.detach(), rough.detach(), (2.0 * img.detach()).contiguous(),
                              frame_shift=SHIFT)
    # same samples; FP32 atomics add in a different order, so compare against the scale
    torch.testing.assert_close(alb.grad, ga, rtol=1e-4, atol=1e-4 * float(ga.abs().max()))
    torch.testing.assert_close(rough.grad, gr, rtol=1e-4, atol=1e-4 * float(gr.abs().max()))
    R.close()


def test_roughness_texture_refuses_the_microsurface_walk():
    import torch

    from open_rendering_framework_amd import rt

    _say("diff: roughness texture on a multiple-scattering conductor is refused")
    d = _scene()
    d.materials[scenes.DIFF_ROUGHNESS_MATERIAL].uints["microfacet_model"] = (2,)
    ctx = scenes.to_context(d)
    views = scenes.ring_views(n=1, size=W, spp=1)
    alb = torch.full((TEX, TEX, 3), 0.5, device="cuda")
    rough = torch.full((TEX, TEX), 0.3, device="cuda")
    with pytest.raises(rt.OrfError, match="multiple-scattering"):
        ctx.diff_render(_desc(rt), views, alb, rough)


SSS_MILK = 1
SSS_W = 20
SSS0 = np.array([0.0011, 0.0024, 0.014, 2.55, 3.21, 3.77])


def _sss_views():
    views = []
    for i, eye in enumerate(((-1.0, 1.2, 2.6), (-2.8, 0.9, 1.2))):
        e, U, V, Wv = scenes.pinhole(eye, (-1.0, -0.1, 0.0), (0.0, 1.0, 0.0), 35.0, SSS_W, SSS_W)
        views.append({"width": SSS_W, "height": SSS_W, "spp": 8, "frame0": 0, "eye": e, "U": U, "V": V, "W": Wv})
    return views


@pytest.mark.parametrize("dipole_model", [0, 1])
def test_scattering_adjoint_matches_fp64_central_differences(dipole_model):
    """The translucent material's absorption and scattering (6 values) as
    parameters: the adjoint against FP64 central differences of the oracle's
    loss with the gather's exit points, the extinction roulette and the probes
    frozen at the parameters' values (relative step 1e-3), to 1e-2 relative
    (1e-4 of the largest for gradients below 1% of it)."""
    import torch

    from open_rendering_framework_amd import rt
    from tests.oracle_scene import OracleScene

    _say("diff: scattering adjoint vs FP64 central differences, dipole %d" % dipole_model)
    d = scenes.subsurface_scene(width=SSS_W, height=SSS_W, level=4, glass_level=2, env_size=(64, 32),
                                probes=4096, probe_samples=4, dipole_model=dipole_model)
    ctx = scenes.to_context(d)
    ctx.accel_build()
    views = _sss_views()
    desc = rt.DiffDesc.make(-1, -1, (1, 1), scattering_material=SSS_MILK)
    theta = torch.tensor(SSS0, dtype=torch.float32, device="cuda")
    img = ctx.diff_render(desc, views, None, None, scattering=theta).cpu().numpy()
    rng = np.random.default_rng(17)
    dimg = rng.standard_normal(img.shape).astype(np.float32)
    _, _, g = ctx.diff_adjoint_sss(desc, views, None, None, theta, torch.from_numpy(dimg).cuda(), frame_shift=SHIFT)
    _assert_replays_match(ctx)
    g = g.cpu().numpy().astype(np.float64)
    # the probes are frozen sampling inputs: the oracle gathers from the GPU's probe
    # values (its own probe pass differs from them in the rare probes whose paths
    # branch apart, and single probes carry large values)
    pr = ctx.probes()
    o64 = OracleScene(d, fp64=True)
    o64.probes_set(pr, 4, SSS_MILK)
    o64.probes_values(pr["L"], pr["w12"], pr["sn"])
    center = SSS0.astype(np.float32).astype(np.float64)
    fd = np.zeros(6)
    for k in range(6):
        h = 1e-3 * center[k]
        lp, lm = [], []
        for sgn, out in ((1, lp), (-1, lm)):
            e = center.copy()
            e[k] += sgn * h
            o64.diff_set_sss(SSS_MILK, e, center)
            out.append(o64.diff_eval(views, shift=SHIFT, dimg=dimg.astype(np.float64), images=False)[0])
        fd[k] = (lp[0] - lm[0]) / (2 * h)
    o64.diff_set_sss(-1, None)
    rel = np.abs(g - fd) / np.abs(fd)
    _say("  adjoint %s\n  central differences %s\n  rel %s" % (g, fd, rel))
    # as for the textures: relative 1e-2 where the gradient is above 1% of the
    # largest, absolute 1e-4 of the largest elsewhere
    big = np.abs(fd) > 1e-2 * np.abs(fd).max()
    assert big.sum() >= 4
    assert np.all(rel[big] <= 1e-2), rel
    assert np.all(np.abs(g - fd)[~big] <= 1e-4 * np.abs(fd).max()), (g, fd)


def test_scattering_parameters_equal_the_material_render():
    """With the material's own coefficients as parameters the differentiable
    primal is the plain render's mean (per pixel rtol 1e-5)."""
    import torch

    from open_rendering_framework_amd import rt

    d = scenes.subsurface_scene(width=SSS_W, height=SSS_W, level=4, glass_level=2, env_size=(64, 32),
                                probes=4096, probe_samples=4)
    ctx = scenes.to_context(d)
    views = _sss_views()
    desc = rt.DiffDesc.make(-1, -1, (1, 1), scattering_material=SSS_MILK)
    theta = torch.tensor(SSS0, dtype=torch.float32, device="cuda")
    img = ctx.diff_render(desc, views, None, None, scattering=theta).cpu().numpy()
    for i, v in enumerate(views):
        for k in ("eye", "U", "V", "W"):
            ctx.set_float(k, *v[k])
        ctx.film_clear()
        ctx.launch(SSS_W, SSS_W, v["frame0"], v["spp"])
        s = ctx.film_sums(SSS_W, SSS_W)
        want = s[..., :3] / np.maximum(s[..., 3:], 1)
        np.testing.assert_allclose(img[i], want, rtol=1e-5, atol=1e-6)


def test_torch_function_scattering_gradient_matches_c_abi():
    """DiffRenderer with the scattering coefficients as a third parameter: its
    backward equals orf_diff_adjoint_ex on the same samples."""
    import torch

    from open_rendering_framework_amd import diff, rt

    d = scenes.subsurface_scene(width=SSS_W, height=SSS_W, level=4, glass_level=2, env_size=(64, 32),
                                probes=4096, probe_samples=4)
    ctx = scenes.to_context(d)
    views = _sss_views()
    alb = torch.full((1, 1, 3), 0.5, device="cuda")
    rough = torch.full((1, 1), 0.3, device="cuda")
    theta = torch.tensor(SSS0, dtype=torch.float32, device="cuda", requires_grad=True)
    R = diff.DiffRenderer(ctx, views, -1, -1, (1, 1), frame_shift=SHIFT, scattering_material=SSS_MILK)
    img = R(alb, rough, theta)
    (img ** 2).sum().backward()
    desc = rt.DiffDesc.make(-1, -1, (1, 1), scattering_material=SSS_MILK)
    _, _, gs = ctx.diff_adjoint_sss(desc, views, alb, rough, theta.detach(), (2.0 * img.detach()).contiguous(),
                                    frame_shift=SHIFT)
    torch.testing.assert_close(theta.grad, gs, rtol=1e-4, atol=1e-4 * float(gs.abs().max()))
    assert float(theta.grad.abs().max()) > 0
    R.close()


def test_inverse_rendering_workload_adjoint_matches_fp64_central_differences():
    """The inverse-rendering workload's shapes: the benchmark scene at 256x256,
    64x64 albedo and roughness textures, 1 spp x 2 ring views; the adjoint
    against FP64 central differences (h = 1e-3) on 64 texture entries (48 among
    those the views see, 16 anywhere), same bars as above."""
    import torch

    from open_rendering_framework_amd import rt
    from tests.oracle_scene import OracleScene

    size, tex = 256, 64
    d = scenes.path_scene(width=size, height=size)
    ctx = scenes.to_context(d)
    views = scenes.ring_views(n=2, size=size, spp=1)
    a, r = scenes.diff_textures(tex, seed=5, rough_range=(0.05, 0.8))
    desc = rt.DiffDesc.make(scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, (tex, tex),
                            scenes.DIFF_ALBEDO_UV, scenes.DIFF_ROUGHNESS_UV)
    dimg = np.random.default_rng(3).standard_normal((2, size, size, 3)).astype(np.float32)
    ga, gr = ctx.diff_adjoint(desc, views, torch.from_numpy(a).cuda(), torch.from_numpy(r).cuda(),
                              torch.from_numpy(dimg).cuda(), frame_shift=SHIFT)
    _assert_replays_match(ctx)
    g = np.concatenate([ga.cpu().numpy().ravel(), gr.cpu().numpy().ravel()]).astype(np.float64)
    o64 = OracleScene(d, fp64=True)
    a0, r0, dd = a.astype(np.float64), r.astype(np.float64), dimg.astype(np.float64)
    h = 1e-3

    def loss(aa, rr):
        o64.diff_set((tex, tex), scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, aa, rr, a0, r0)
        return o64.diff_eval(views, shift=SHIFT, dimg=dd, images=False, threads=16)[0]

    rng = np.random.default_rng(11)
    seen = np.flatnonzero(g != 0.0)
    assert len(seen) >= 1000
    picks = np.concatenate([rng.choice(seen, size=48, replace=False), rng.choice(len(g), size=16, replace=False)])
    na = a0.size
    fd = []
    for p in picks:
        aa, rr = a0.copy(), r0.copy()
        v = aa.reshape(-1) if p < na else rr.reshape(-1)
        k = p if p < na else p - na
        v[k] += h
        lp = loss(aa, rr)
        v[k] -= 2 * h
        lm = loss(aa, rr)
        fd.append((lp - lm) / (2 * h))
    got, fd = g[picks], np.array(fd)
    big = np.abs(fd) > 1e-2 * np.abs(fd).max()
    rel = np.abs(got - fd) / np.maximum(np.abs(fd), 1e-30)
    _say("  entries %d (large %d, seen %d): max rel %.3e, median rel %.3e" % (len(fd), big.sum(), len(seen),
                                                                             rel[big].max(), np.median(rel[big])))
    assert big.sum() >= 24
    assert np.all(rel[big] <= 1e-2), list(zip(got[big][rel[big] > 1e-2], fd[big][rel[big] > 1e-2]))
    assert np.all(np.abs(got - fd)[~big] <= 1e-4 * np.abs(fd).max())


def test_inverse_rendering_workload_primal_image_rmse():
    """The inverse-rendering workload's primal at its real shape (the benchmark
    scene at 256x256, 8 ring views, 16 spp, 64x64 albedo and roughness
    textures) against the oracle's images of the same samples: relative RMSE
    (RMSE / mean) <= 1e-5 against the FP32 oracle (the same arithmetic;
    measured 4.3e-6) and <= 2.5e-3 against the FP64 oracle (measured 1.7e-3).
    The FP64 figure is set by the sun: its 0.5-degree lobe is narrower than one
    texel of the 1024x512 map, so the FP32 texel coordinate (acos, atan2 and the
    u*w product, a few 1e-5 texel) moves the bilinear value of sun samples
    (radiance ~2e4 against an image mean ~1) by ~1e-4 relative; the error falls
    as 1/sqrt(spp) (the 512x512 image at 64 spp measures 9.2e-4)."""
    import torch

    from open_rendering_framework_amd import rt
    from tests.oracle_scene import OracleScene

    size, tex = 256, 64
    d = scenes.path_scene(width=size, height=size)
    ctx = scenes.to_context(d)
    views = scenes.ring_views(n=8, size=size, spp=16)
    a, r = scenes.diff_textures(tex, seed=5, rough_range=(0.05, 0.8))
    desc = rt.DiffDesc.make(scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, (tex, tex),
                            scenes.DIFF_ALBEDO_UV, scenes.DIFF_ROUGHNESS_UV)
    img = ctx.diff_render(desc, views, torch.from_numpy(a).cuda(), torch.from_numpy(r).cuda()).cpu().numpy()
    assert np.isfinite(img).all()
    for fp64, bar in ((False, 1e-5), (True, 2.5e-3)):
        o = OracleScene(d, fp64=fp64)
        o.diff_set((tex, tex), scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, a.astype(np.float64),
                   r.astype(np.float64))
        _, want = o.diff_eval(views, threads=16)
        rel = np.sqrt(((img - want) ** 2).mean()) / want.mean()
        _say("diff workload primal: rel rmse vs fp%d oracle %.3g over %d views" % (64 if fp64 else 32, rel,
                                                                                     len(views)))
        assert rel <= bar
