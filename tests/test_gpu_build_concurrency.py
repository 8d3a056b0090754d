"""GPU regression test: LBVH builds that run while another stream renders.

A build starts by describing the root level (one node holding every
triangle).  When other kernels occupy the CUs, the waves of one block start
at different times, and that description must not depend on their order: a
build that loses it leaves the previous tree in place.  Here every build runs
on its own stream while frames of a second scene render on another stream,
the geometry changes from build to build, and every tree must be the
oracle's tree of its own geometry.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_builds_beside_rendering_frames(ctx):
    import rtx
    from rtx import scenes
    sd = scenes.config_c3(n=65, res=(32, 32))
    s = scenes.load(rtx.Scene(ctx), sd)
    bg = scenes.config_c2(n_spheres=60, res=(640, 480), slices=24, stacks=16)
    sb = scenes.load(rtx.Scene(ctx), bg)
    sb.build(30, 1)
    cam = scenes.camera_of(bg)
    R = rtx.Renderer(ctx, bg.width, bg.height, bg.max_bounces)
    R.frame_async(sb, cam, rebuild=False)
    ctx.sync()
    render_stream, build_stream = rtx.Stream(), rtx.Stream()
    idx, mm = sd.idx, sd.matmap
    for i in range(32):
        R.frame_async(sb, cam, rebuild=False, stream=render_stream.handle)
        R.frame_async(sb, cam, rebuild=False, stream=render_stream.handle)
        s.displace_waves(sd.waves, 0.05 * i, stream=build_stream.handle)
        s.build(30, 1, stream=build_stream.handle)
        ctx.device_sync()
        pos, _ = s.vertices()
        b = O.Build(pos, idx, mm, 30, 1)
        n, _ = s.info()
        assert n == b.node_count, f"build {i}: node count {n} != {b.node_count}"
        g = s.nodes()
        assert np.array_equal(g["a"], b.nodes["a"]) and np.array_equal(g["b"], b.nodes["b"]), \
            f"build {i}: topology differs"
        assert np.array_equal(g["lo"][:, :3], b.nodes["lo"][:, :3]) and \
            np.array_equal(g["hi"][:, :3], b.nodes["hi"][:, :3]), f"build {i}: node AABBs differ"
        idx, mm = b.idx.reshape(-1, 3), b.matmap
    render_stream.destroy()
    build_stream.destroy()
    R.destroy()
    sb.destroy()
    s.destroy()
