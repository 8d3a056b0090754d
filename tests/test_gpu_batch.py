"""GPU tests of batches of frames of different scenes rendered side by side
(several streams, one capturable graph): every frame of the batch must equal
the same frame rendered alone, eagerly and in graph replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_frames_equal_single_frames(ctx):
    import rtx
    from rtx import scenes
    sds = [scenes.config_c1(), scenes.config_c2(n_spheres=30, res=(96, 64), slices=20, stacks=12),
           scenes.config_c3(n=65, res=(80, 60))]
    scs = []
    for sd in sds:
        sc = scenes.load(rtx.Scene(ctx), sd)
        sc.build(30, 1)
        scs.append(sc)
    # 7 frames over 3 scenes, mixed sizes and depths
    plan = [(0, 64, 48, 1), (1, 96, 64, 3), (2, 80, 60, 2), (0, 120, 90, 2), (1, 64, 64, 1), (2, 50, 40, 3),
            (0, 32, 32, 4)]
    cams, rs, ss = [], [], []
    for k, w, h, d in plan:
        c = sds[k].camera
        cams.append(rtx.camera_set(c["pos"], c["look_at"], c["up"], c["fov"], w / h))
        rs.append(rtx.Renderer(ctx, w, h, d))
        ss.append(scs[k])
    ref = []
    for (k, w, h, d), cam in zip(plan, cams):
        R = rtx.Renderer(ctx, w, h, d)
        R.frame_async(scs[k], cam, rebuild=False)
        ctx.sync()
        ref.append((R.fb.accum(), R.last_stats()))
        R.destroy()
    B = rtx.Batch(ctx, rs, ss, cams, streams=3)
    B.render()
    ctx.sync()
    tot = 0
    for i, R in enumerate(rs):
        assert np.array_equal(R.fb.accum(), ref[i][0]), f"batch frame {i} differs (eager)"
        tot += ref[i][1].primary_rays + ref[i][1].secondary_rays
    assert B.stats()[0] == tot
    B.capture()
    for rep in range(2):
        B.replay()
        ctx.sync()
        for i, R in enumerate(rs):
            assert np.array_equal(R.fb.accum(), ref[i][0]), f"batch frame {i} differs (replay {rep})"
    B.destroy()
    for R in rs:
        R.destroy()
    for sc in scs:
        sc.destroy()
