"""GPU parity of the light model on the single-root path: SPOT, DIR and POINT
shadow rays (shadow_trace_single, src/kernel/shadow-trace-bvh.cl:375-418, with
the cone rule at :399-405), whole frames lit by each light type, and a second
shadow pass over the same hits after the lights move (the reference assigns
shadow_hit on every call, so a pass must clear the bits of pairs that became
lit).  Everything is compared with the oracle bit for bit.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _scene(ctx):
    import rtx
    from rtx import scenes
    sd = scenes.config_c1()
    s = scenes.load(rtx.Scene(ctx), sd)
    s.build(30, 1)
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    return sd, s, b, scenes.camera_of(sd)


LIGHT_SETS = {
    # a spot over the sphere whose cone leaves much of the ground outside
    "spot": [dict(type=0, pos=(1.0, 6.0, -2.0), dir=(-0.1690309, -0.9860133, 0.0), angle=0.93,
                  color=(1.0, 0.9, 0.8))],
    "dir": [dict(type=1, dir=(0.3713907, -0.9284767, 0.0), color=(0.9, 0.9, 1.0))],
    "point": [dict(type=2, pos=(4.0, 8.0, -4.0), color=(1.0, 1.0, 1.0))],
    # one of each, so the shadow bits of three lights are set and cleared per ray
    "mixed": [dict(type=0, pos=(1.0, 6.0, -2.0), dir=(-0.1690309, -0.9860133, 0.0), angle=0.93),
              dict(type=1, dir=(0.3713907, -0.9284767, 0.0), color=(0.5, 0.5, 0.5)),
              dict(type=2, pos=(-3.0, 5.0, -6.0), color=(0.3, 0.6, 0.3))],
}


@pytest.fixture(scope="module")
def c1s(ctx):
    sd, s, b, cam = _scene(ctx)
    yield sd, s, b, cam
    s.destroy()


def _primary_hits(rtx, ctx, s, cam, w, h):
    n = w * h
    r = rtx.Rays(ctx, n)
    hb = rtx.Hits(ctx, n)
    r.primary(cam, w, h)
    rtx.trace(s, r, hb, n)
    return r, hb


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("which", ["spot", "dir", "point", "mixed"])
def test_shadow_flags_per_light_type(ctx, c1s, which, variant):
    import rtx
    sd, s, b, cam = c1s
    w, h = 256, 256
    L = rtx.make_lights((0.1, 0.1, 0.1), LIGHT_SETS[which])
    s.set_lights(L)
    old = ctx.tuned("trace_variant")
    ctx.tune("trace_variant", variant)
    try:
        r, hb = _primary_hits(rtx, ctx, s, cam, w, h)
        rtx.trace_shadow(s, r, hb, w * h)
        g = hb.read(w * h)
    finally:
        ctx.tune("trace_variant", old)
        s.set_lights(sd.lights)
    oo, od, _, _ = O.raygen_primary(O.camera_from(cam), w, h)
    o = O.trace(b, sd.nrm, sd.uv, oo, od)
    of = O.trace_shadow(b, O.lights_from(L), o["flags"], o["p4"])
    assert np.array_equal(g["flags"], of), f"shadow flags differ ({which}, variant {variant})"
    hit = (of & 1) != 0
    for li in range(len(LIGHT_SETS[which])):
        bit = (of[hit] >> (8 + li)) & 1
        assert bit.any() and not bit.all(), f"light {li} of {which} should shadow some hits and light others"
    r.destroy()
    hb.destroy()


@pytest.mark.parametrize("variant", [1, 3])
def test_second_shadow_pass_after_light_moves(ctx, c1s, variant):
    # shadow with one light set, move the lights, shadow the same hits again:
    # the flags must equal a fresh trace + shadow under the new lights
    import rtx
    sd, s, b, cam = c1s
    w, h = 256, 256
    old = ctx.tuned("trace_variant")
    ctx.tune("trace_variant", variant)
    try:
        s.set_lights(rtx.make_lights((0.1, 0.1, 0.1), LIGHT_SETS["mixed"]))
        r, hb = _primary_hits(rtx, ctx, s, cam, w, h)
        rtx.trace_shadow(s, r, hb, w * h)
        first = hb.read(w * h)["flags"]
        moved = rtx.make_lights((0.1, 0.1, 0.1), LIGHT_SETS["mixed"][::-1])
        s.set_lights(moved)
        rtx.trace_shadow(s, r, hb, w * h)
        again = hb.read(w * h)["flags"]
    finally:
        ctx.tune("trace_variant", old)
        s.set_lights(sd.lights)
    oo, od, _, _ = O.raygen_primary(O.camera_from(cam), w, h)
    o = O.trace(b, sd.nrm, sd.uv, oo, od)
    of = O.trace_shadow(b, O.lights_from(moved), o["flags"], o["p4"])
    assert not np.array_equal(first, of), "the light move should change some shadow bits"
    assert np.array_equal(again, of), "a second shadow pass kept stale bits"
    r.destroy()
    hb.destroy()


@pytest.mark.parametrize("which", ["spot", "dir", "mixed"])
def test_frame_per_light_type(ctx, c1s, which):
    # a whole 1-bounce frame (mirror sphere) lit by each light type
    import rtx
    sd, s, b, cam = c1s
    w, h = 192, 160
    L = rtx.make_lights((0.1, 0.1, 0.1), LIGHT_SETS[which])
    s.set_lights(L)
    R = rtx.Renderer(ctx, w, h, 1)
    try:
        st = R.frame(s, cam, rebuild=False)
        acc = R.fb.accum()
    finally:
        R.destroy()
        s.set_lights(sd.lights)
    oacc, ost, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(L), O.camera_from(cam), w, h, 1)
    assert st.primary_rays + st.secondary_rays == orays
    assert np.array_equal(acc, oacc), f"frame differs ({which})"
