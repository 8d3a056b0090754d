"""GPU parity of textured diffuse and the cubemap background (shade_sample,
src/kernel/ray-shader.cl:187-240, 275-307): whole frames, primary and
secondary rays, against the oracle bit for bit.

The scene is C1's heightfield (textured: an image of odd size, uvs running
over [-1, 2] so that the sampler wraps) and its mirror sphere, under a
cubemap of six random faces of different sizes; the camera sees the sky, so
primary misses take the cubemap and the sphere's reflections take both the
texture and the cubemap.  Also: an index without an image (black diffuse),
the cubemap switched off (misses add nothing, as before), and a captured
frame replayed.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _images(seed):
    rng = np.random.default_rng(seed)

    def img(h, w):
        im = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        im[..., 3] = 255
        return im
    textures = [img(23, 37), img(16, 16)]
    faces = [img(32, 32), img(17, 31), img(8, 8), img(64, 48), img(5, 7), img(32, 32)]
    return textures, faces


def _scene(ctx, tex_index=0, sphere_tex=-1):
    import rtx
    from rtx import scenes, make_materials
    sd = scenes.config_c1()
    sd.uv = (sd.uv * 3.0 - 1.0).astype(np.float32)
    mats = make_materials([(0.6, 0.5, 0.4, 0.3, 0.0, 0.0), (1.0, 1.0, 1.0, 1.0, 0.8, 0.0)]).view(np.int32)
    mats[7] = tex_index
    mats[8 + 7] = sphere_tex
    sd.materials = mats.view(np.uint8)
    s = scenes.load(rtx.Scene(ctx), sd)
    return sd, s


def _oracle(sd, s, cam, w, h, bounces, tex):
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    return O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h, bounces,
                    tex=tex)


def _camera(sd, w, h):
    import rtx
    # lifted and looking past the sphere: sky in the upper part of the frame
    return rtx.camera_set((0.0, 2.5, -6.0), (0.0, 1.2, 0.0), (0.0, 1.0, 0.0), sd.camera["fov"], w / h)


@pytest.mark.parametrize("bounces", [0, 2])
def test_textured_frame_with_cubemap(ctx, bounces):
    import rtx
    textures, faces = _images(1)
    sd, s = _scene(ctx)
    s.set_textures(textures)
    s.set_cubemap(faces)
    w, h = 200, 150
    cam = _camera(sd, w, h)
    R = rtx.Renderer(ctx, w, h, max(bounces, 1))
    R.set_bounces(bounces)
    R.frame(s, cam, rebuild=True)
    acc = R.fb.accum()
    oacc, st, _ = _oracle(sd, s, cam, w, h, bounces, O.TexSet(textures, faces))
    plain, _, _ = _oracle(sd, s, cam, w, h, bounces, None)
    assert (oacc != plain).mean() > 0.3, "texture and sky should change much of the frame"
    assert np.array_equal(acc, oacc), f"accumulators differ in {(acc != oacc).any(1).sum()} pixels"
    # a captured frame replays the same image
    R.capture(s, cam, rebuild=False)
    R.replay()
    ctx.sync()
    assert np.array_equal(R.fb.accum(), oacc)
    R.destroy()
    s.destroy()


def test_texture_index_without_image_and_cubemap_off(ctx):
    import rtx
    textures, faces = _images(2)
    # the heightfield's texture 5 does not exist: black diffuse; the sphere uses image 1
    sd, s = _scene(ctx, tex_index=5, sphere_tex=1)
    s.set_textures(textures)
    s.set_cubemap(faces)
    w, h = 160, 120
    cam = _camera(sd, w, h)
    R = rtx.Renderer(ctx, w, h, 2)
    R.frame(s, cam, rebuild=True)
    oacc, _, _ = _oracle(sd, s, cam, w, h, 2, O.TexSet(textures, faces))
    assert np.array_equal(R.fb.accum(), oacc)
    # cubemap off: misses add nothing
    s.set_cubemap(None)
    R.frame(s, cam, rebuild=True)
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    b = O.Build(sd.pos, b.idx.reshape(-1, 3), b.matmap, 30, 1)  # the second build starts from the first's order
    oacc2, _, _ = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h, 2,
                           tex=O.TexSet(textures))
    assert np.array_equal(R.fb.accum(), oacc2)
    # textures removed: the textured materials sample black
    s.set_textures([])
    R.frame(s, cam, rebuild=False)
    oacc3, _, _ = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h, 2)
    assert np.array_equal(R.fb.accum(), oacc3)
    R.destroy()
    s.destroy()


def test_image_errors(ctx):
    import rtx
    sd, s = _scene(ctx)
    with pytest.raises(rtx.RtxError):
        s.set_cubemap([np.zeros((4, 4, 4), np.uint8)] * 5 + [np.zeros((0, 4, 4), np.uint8)])
    with pytest.raises(rtx.RtxError):
        s.set_textures([np.zeros((4, 4, 4), np.uint8)] * 65)
    s.destroy()
