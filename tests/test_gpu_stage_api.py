"""The C++ stage classes (include/rt/stages.hpp) drive a frame the way the
reference's apps drive Renderer; the framebuffer must equal the oracle's.
Runs tests/cpp/stage_render.cpp (built by build()) against librtx.so."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tesina-rt_amd" / "bin" / "stage_render"


def write_scene(d: Path, sd, w, h, bounces, depth, min_leaf, frames):
    sd.pos.astype(np.float32).tofile(d / "pos.f32")
    sd.nrm.astype(np.float32).tofile(d / "nrm.f32")
    sd.uv.astype(np.float32).tofile(d / "uv.f32")
    sd.idx.astype(np.uint32).tofile(d / "idx.u32")
    sd.matmap.astype(np.int32).tofile(d / "matmap.i32")
    sd.materials.tofile(d / "mats.bin")
    (d / "lights.bin").write_bytes(bytes(sd.lights))
    c = sd.camera
    (d / "view.txt").write_text(" ".join(repr(float(x)) for x in (*c["pos"], *c["look_at"], *c["up"],
                                                                     np.float32(c["fov"]))))
    (d / "params.txt").write_text(f"{w} {h} {bounces} {depth} {min_leaf} {frames}")


@pytest.mark.parametrize("frames", [1, 2])
def test_cpp_renderer_matches_oracle(tmp_path, frames):
    import rtx
    from rtx import scenes
    assert BIN.exists(), "stage_render not built (run __graft_entry__.build())"
    sd = scenes.config_c1()
    w, h, bounces, depth = 160, 120, 2, 30
    write_scene(tmp_path, sd, w, h, bounces, depth, 1, frames)
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(BIN), str(tmp_path), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    acc = np.fromfile(out / "accum.u32", np.uint32).reshape(-1, 3)
    rays, sec, nodes, _ = (float(x) for x in (out / "stats.txt").read_text().split())
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], sd.camera["fov"], w / h)
    b = O.Build(sd.pos, sd.idx, sd.matmap, depth, 1)
    for _ in range(frames - 1):  # every frame rebuilds from the permuted buffers
        b = O.Build(sd.pos, b.idx.reshape(-1, 3), b.matmap, depth, 1)
    # the reference Renderer's bundles hold 3 tiles (src/rt/renderer.cpp:376)
    oacc, st, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h,
                               bounces, cap=3 * w * h)
    assert int(rays) == orays and int(sec) == st["secondary"] and int(nodes) == b.node_count
    assert np.array_equal(acc, oacc)
    rgba = np.fromfile(out / "rgba.u8", np.uint8).reshape(-1, 4)
    assert np.array_equal(rgba, O.resolve(oacc))


def test_cpp_renderer_textures_and_cubemap(tmp_path):
    """TextureAtlas and Cubemap of the C++ Scene, shaded by RayShader through
    the Renderer: the framebuffer equals the oracle's with the same images."""
    import rtx
    from test_gpu_textures import _images
    from rtx import scenes, make_materials
    sd = scenes.config_c1()
    sd.uv = (sd.uv * 3.0 - 1.0).astype(np.float32)
    mats = make_materials([(0.6, 0.5, 0.4, 0.3, 0.0, 0.0), (1.0, 1.0, 1.0, 1.0, 0.8, 0.0)]).view(np.int32)
    mats[7] = 0
    sd.materials = mats.view(np.uint8)
    textures, faces = _images(3)
    w, h, bounces = 160, 120, 2
    sd.camera = dict(pos=(0.0, 2.5, -6.0), look_at=(0.0, 1.2, 0.0), up=(0.0, 1.0, 0.0), fov=sd.camera["fov"])
    write_scene(tmp_path, sd, w, h, bounces, 30, 1, 1)
    lines = [str(len(textures))] + [f"{t.shape[1]} {t.shape[0]}" for t in textures] + ["1"] + \
        [f"{f.shape[1]} {f.shape[0]}" for f in faces]
    (tmp_path / "images.txt").write_text("\n".join(lines) + "\n")
    for k, t in enumerate(textures):
        t.tofile(tmp_path / f"tex{k}.rgba")
    for k, f in enumerate(faces):
        f.tofile(tmp_path / f"face{k}.rgba")
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(BIN), str(tmp_path), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    acc = np.fromfile(out / "accum.u32", np.uint32).reshape(-1, 3)
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], np.float32(sd.camera["fov"]), w / h)
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    oacc, _, _ = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h,
                          bounces, cap=3 * w * h, tex=O.TexSet(textures, faces))
    assert np.array_equal(acc, oacc)


OBJ_BIN = ROOT / "tesina-rt_amd" / "bin" / "stage_objects"


def _write_objects(d: Path, meshes, objects, moved, ori4, dir4, lights):
    (d / "meshes.txt").write_text(f"{len(meshes)}\n" + "".join(f"{len(m[0])} {len(m[3])}\n" for m in meshes))
    for k, (p, n, uv, idx) in enumerate(meshes):
        p.astype(np.float32).tofile(d / f"mesh{k}.pos.f32")
        n.astype(np.float32).tofile(d / f"mesh{k}.nrm.f32")
        uv.astype(np.float32).tofile(d / f"mesh{k}.uv.f32")
        idx.astype(np.uint32).tofile(d / f"mesh{k}.idx.u32")

    def mats(tr, tri):
        return " ".join(repr(float(x)) for x in (*np.asarray(tr, np.float32).ravel(),
                                                  *np.asarray(tri, np.float32).ravel()))
    (d / "objects.txt").write_text(f"{len(objects)}\n" + "".join(f"{m} {mats(tr, tri)}\n" for m, tr, tri in objects))
    if moved is not None:
        (d / "moved.txt").write_text("".join(f"{mats(tr, tri)}\n" for _, tr, tri in moved))
    np.concatenate([ori4, dir4], axis=1).astype(np.float32).tofile(d / "rays.f32")
    (d / "lights.bin").write_bytes(bytes(lights))


def _read_pass(out: Path, p: int, n: int):
    b = out / f"pass{p}"
    return {"flags": np.fromfile(f"{b}.flags.u32", np.uint32), "t": np.fromfile(f"{b}.t.f32", np.float32),
            "id": np.fromfile(f"{b}.id.i32", np.int32), "bary": np.fromfile(f"{b}.bary.f32", np.float32).reshape(n, 2),
            "n4": np.fromfile(f"{b}.n4.f32", np.float32).reshape(n, 4),
            "p4": np.fromfile(f"{b}.p4.f32", np.float32).reshape(n, 4),
            "uv2": np.fromfile(f"{b}.uv2.f32", np.float32).reshape(n, 2)}


def test_cpp_objects_match_oracle(tmp_path):
    """Objects through the C++ Scene / Tracer: create_bvhs, trace, shadow,
    then new transforms through update_bvh_roots; both passes equal the
    oracle's multi-root traversal."""
    from inst_scene import rot, xform, two_meshes, rays_at
    from test_gpu_instances import _lights
    assert OBJ_BIN.exists(), "stage_objects not built (run __graft_entry__.build())"
    meshes = two_meshes()
    objects = [(0, *xform(rot((0, 1, 0), 0.7), 1.0, (-2.0, 0.3, 0.0))),
               (1, *xform(rot((1, 0, 1), 1.9), 2.0, (2.0, -0.5, 1.0))),
               (0, *xform(rot((0, 0, 1), -0.4), 0.5, (0.0, 2.0, -1.0)))]
    moved = [(m, *xform(rot((0, 1, 0), 0.3 * i + 0.2), 1.1, (1.5 * i - 1.5, 0.2, 0.0))) for i, (m, _, _) in
             enumerate(objects)]
    n = 3000
    ori4, dir4 = rays_at(21, n)
    lights = _lights()
    _write_objects(tmp_path, meshes, objects, moved, ori4, dir4, lights)
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(OBJ_BIN), str(tmp_path), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    for p, objs in enumerate((objects, moved)):
        ms = O.MultiScene(builds, [m[1] for m in meshes], [m[2] for m in meshes], objs)
        o = O.trace_multi(ms, ori4, dir4)
        o["flags"] = O.trace_shadow_multi(ms, O.lights_from(lights), o["flags"], o["p4"])
        assert ((o["flags"] & 1) != 0).sum() > 300
        g = _read_pass(out, p, n)
        for k in ("flags", "t", "id", "bary", "n4", "p4", "uv2"):
            assert np.array_equal(g[k], o[k]), f"pass {p}: {k} differs"


def test_cpp_single_object_is_traced_untransformed(tmp_path):
    """With one object and no aggregate mesh, root_count() is 1 and the
    Tracer takes the single-root kernels, which trace the mesh as stored
    (src/rt/tracer.cpp:97: the transform of a lone object is not applied)."""
    from inst_scene import rot, xform, sphere_mesh, rays_at
    from test_gpu_instances import _lights
    mesh = sphere_mesh()
    objects = [(0, *xform(rot((0, 1, 0), 0.7), 2.0, (3.0, 0.0, 0.0)))]
    n = 2000
    ori4, dir4 = rays_at(22, n)
    lights = _lights()
    _write_objects(tmp_path, [mesh], objects, None, ori4, dir4, lights)
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(OBJ_BIN), str(tmp_path), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    b = O.Build(mesh[0], mesh[3], None, 30, 1)
    o = O.trace(b, mesh[1], mesh[2], ori4, dir4)
    o["flags"] = O.trace_shadow(b, O.lights_from(lights), o["flags"], o["p4"])
    assert ((o["flags"] & 1) != 0).sum() > 100
    g = _read_pass(out, 0, n)
    for k in ("flags", "t", "id", "bary", "n4", "p4", "uv2"):
        assert np.array_equal(g[k], o[k]), f"{k} differs"


@pytest.mark.parametrize("count", [3, 1])
def test_cpp_renderer_over_objects(tmp_path, count):
    """Renderer over a scene of objects with use_lbvh = 0 (the reference's
    Renderer with Scene::create_bvhs): the multi-root trace and shadow rays,
    secondary rays and shading through the objects' material map (a mesh takes
    its first object's material), textures and the cubemap; then the objects
    move (update_bvh_roots) and the next frame follows.  With one object the
    single-root kernels trace its mesh as stored, shaded with its material."""
    import rtx
    from rtx import make_materials
    from inst_scene import rot, xform, two_meshes, rays_at
    from test_gpu_instances import _lights
    from test_gpu_textures import _images
    assert OBJ_BIN.exists(), "stage_objects not built (run __graft_entry__.build())"
    meshes = two_meshes()
    objects = [(0, *xform(rot((0, 1, 0), 0.7), 1.0, (-1.6, 0.3, 0.0))),
               (1, *xform(rot((1, 0, 1), 1.9), 1.5, (1.6, -0.5, 1.0))),
               (0, *xform(rot((0, 0, 1), -0.4), 0.5, (0.0, 1.6, -1.0)))][:count]
    moved = [(m, *xform(rot((0, 1, 0), 0.3 * i + 0.2), 1.1, (1.5 * i - 1.5, 0.2, 0.0))) for i, (m, _, _) in
             enumerate(objects)]
    n = 500
    ori4, dir4 = rays_at(23, n)
    lights = _lights()
    _write_objects(tmp_path, meshes, objects, moved, ori4, dir4, lights)
    rows = [(0.9, 0.2, 0.2, 0.5, 0.6, 1.5), (0.6, 0.8, 0.3, 1.0, 0.7, 0.0), (0.3, 0.3, 0.9, 0.2, 0.0, 0.0)]
    mats = make_materials(rows[:count])
    mi = mats.view(np.int32).reshape(-1, 8)
    mi[0, 7] = 0                      # the first object's mesh is textured
    if count > 1:
        mi[1, 7] = 1
    mats.tofile(tmp_path / "objmats.bin")
    textures, faces = _images(3)
    lines = [str(len(textures))] + [f"{t.shape[1]} {t.shape[0]}" for t in textures] + ["1"] + \
        [f"{f.shape[1]} {f.shape[0]}" for f in faces]
    (tmp_path / "images.txt").write_text("\n".join(lines) + "\n")
    for k, t in enumerate(textures):
        t.tofile(tmp_path / f"tex{k}.rgba")
    for k, f in enumerate(faces):
        f.tofile(tmp_path / f"face{k}.rgba")
    w, h, bounces = 160, 120, 3
    pos, look, up, fov = (0.0, 1.0, -7.0), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), np.float32(np.pi / 3)
    (tmp_path / "render.txt").write_text(f"{w} {h} {bounces}\n" +
                                         " ".join(repr(float(x)) for x in (*pos, *look, *up, fov)) + "\n")
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(OBJ_BIN), str(tmp_path), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    cam = O.camera_from(rtx.camera_set(pos, look, up, fov, w / h))
    tex = O.TexSet(textures, faces)
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    for p, objs in enumerate((objects, moved)):
        if count > 1:
            ms = O.MultiScene(builds, [m[1] for m in meshes], [m[2] for m in meshes], objs)
            oacc, ost, orays = O.render_multi(ms, mats, O.lights_from(lights), cam, w, h, bounces, cap=3 * w * h,
                                              tex=tex)
        else:
            b, m0 = builds[0], meshes[0]
            oacc, ost, orays = O.render(b, m0[1], m0[2], mats, O.lights_from(lights), cam, w, h, bounces,
                                        cap=3 * w * h, tex=tex)
        acc = np.fromfile(out / f"frame{p}.accum.u32", np.uint32).reshape(-1, 3)
        rays, sec = (int(x) for x in (out / f"frame{p}.stats.txt").read_text().split())
        assert sec > 0 and rays == orays, (rays, sec, orays)
        assert np.array_equal(acc, oacc), f"frame {p}: accumulators differ in {(acc != oacc).any(1).sum()} pixels"
