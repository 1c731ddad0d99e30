"""The renderer-facing host arithmetic: the LOD part of
TressFXHairObject::UpdateRenderingParameters (TressFXHairObject.cpp:461-481)
and the draw size of DrawStrands (:260-277), pinned by closed forms and
checked against a float32/uint32 restatement over a grid of inputs."""
import numpy as np
import pytest

import tressfx_amd as tfx

f32 = np.float32


def lod_restated(s, distance, shadow):
    """UpdateRenderingParameters' LOD arithmetic in float32."""
    radius, density = f32(s.fiber_radius), f32(1.0)
    if s.enable_hair_lod:
        a = f32(s.shadow_lod_start_distance if shadow else s.lod_start_distance)
        b = f32(s.shadow_lod_end_distance if shadow else s.lod_end_distance)
        lo, hi = min(a, b), max(a, b)
        d = f32(distance)
        if d > lo:
            ratio = min(f32((d - lo) / max(f32(hi - lo), f32(0.00001))), f32(1.0))
            mult = f32(s.shadow_lod_width_multiplier if shadow else s.lod_width_multiplier)
            max_r = f32(radius * mult)
            radius = f32(radius + f32(ratio * f32(max_r - radius)))
            pct = f32(s.shadow_lod_percent if shadow else s.lod_percent)
            density = f32(f32(1.0) + f32(ratio * f32(pct - f32(1.0))))
    return radius, density


def draw_restated(total, V, F, density):
    prims = np.uint32(total // 3)
    if f32(density) != f32(1.0):
        prims = np.uint32(int(f32(f32(int(prims)) * f32(density))))
        per = np.uint32((V - 1) * 2)
        rem = prims % per
        prims = np.uint32(prims + per - rem) if rem > 0 else prims
        prims = np.uint32(prims - prims % np.uint32(per * (F + 1)))
    return int(prims) * 3


def test_lod_closed_forms():
    s = tfx.RenderingLodSettings.defaults()
    assert (s.lod_start_distance, s.lod_end_distance, s.lod_percent, s.lod_width_multiplier) == (1.0, 5.0, 0.5, 2.0)
    assert s.enable_hair_lod == 0 and s.fiber_radius == f32(0.0021)
    r0 = float(f32(0.0021))
    assert tfx.rendering_lod(s, 100.0) == (r0, 1.0)                  # LOD off: nothing changes
    s.enable_hair_lod = 1
    assert tfx.rendering_lod(s, 0.5) == (r0, 1.0)                    # nearer than the start distance
    assert tfx.rendering_lod(s, 1.0) == (r0, 1.0)
    r, d = tfx.rendering_lod(s, 3.0)                                 # half way: ratio 0.5
    assert d == 0.75 and abs(r - 1.5 * r0) <= 1e-9
    r, d = tfx.rendering_lod(s, 50.0)                                # past the end: clamped at 1
    assert d == 0.5 and abs(r - 2 * r0) <= 1e-9
    s.lod_start_distance, s.lod_end_distance = 5.0, 1.0              # reversed range: min/max order it
    assert tfx.rendering_lod(s, 3.0)[1] == 0.75
    s.shadow_lod_start_distance, s.shadow_lod_end_distance, s.shadow_lod_percent = 2.0, 4.0, 0.2
    s.shadow_lod_width_multiplier = 3.0
    r, d = tfx.rendering_lod(s, 3.0, shadow_update=True)             # the shadow pass reads its own fields
    assert abs(d - 0.6) < 1e-7 and abs(r - 2 * r0) <= 1e-9
    s.lod_start_distance = s.lod_end_distance = 2.0                  # empty range: the 1e-5 floor
    assert tfx.rendering_lod(s, 2.0 + 1e-6)[1] < 1.0


def test_lod_matches_float32_restatement():
    rng = np.random.Generator(np.random.PCG64(3))
    s = tfx.RenderingLodSettings.defaults()
    s.enable_hair_lod = 1
    for _ in range(2000):
        s.lod_start_distance, s.lod_end_distance = rng.uniform(0, 10, 2)
        s.shadow_lod_start_distance, s.shadow_lod_end_distance = rng.uniform(0, 10, 2)
        s.lod_percent, s.shadow_lod_percent = rng.uniform(0, 1, 2)
        s.lod_width_multiplier, s.shadow_lod_width_multiplier = rng.uniform(1, 4, 2)
        s.fiber_radius = rng.uniform(1e-4, 1e-2)
        dist, shadow = float(rng.uniform(0, 12)), bool(rng.integers(0, 2))
        got = tfx.rendering_lod(s, dist, shadow)
        exp = lod_restated(s, dist, shadow)
        assert (f32(got[0]), f32(got[1])) == exp, (dist, shadow)


def test_draw_size_closed_forms():
    V, F, S = 16, 2, 300
    total = 6 * S * (V - 1)                                          # GetNumHairTriangleIndices
    assert tfx.num_indices_to_render(total, V, F, 1.0) == total
    assert tfx.num_indices_to_render(total, V, F, 0.5) == total // 2   # 50 whole groups of 3 strands
    # 9000 primitives * 0.33f = 2969.99.. -> 2969, up to a whole strand (2970 = 99 strands), down to whole
    # guide groups of 3 strands (99 strands = 33 groups) -> 2970 primitives
    assert tfx.num_indices_to_render(total, V, F, 0.33) == 2970 * 3
    assert tfx.num_indices_to_render(total, V, F, 0.0) == 0
    assert tfx.num_indices_to_render(total, V, F, 0.004) == 0         # under one group
    # without follow strands: 36 primitives round up to whole strands, 60 = two strands
    assert tfx.num_indices_to_render(total, V, 0, 0.004) == 2 * (V - 1) * 2 * 3
    with pytest.raises(tfx.TFXError):
        tfx.num_indices_to_render(total, 1, F, 0.5)


def test_draw_size_matches_uint32_restatement():
    rng = np.random.Generator(np.random.PCG64(5))
    for _ in range(3000):
        V = int(rng.choice([4, 8, 16, 32, 64]))
        F = int(rng.integers(0, 5))
        S = int(rng.integers(1, 200000)) * (F + 1)
        total = 6 * S * (V - 1)
        if total > 0xFFFFFFFF:
            continue
        d = float(f32(rng.uniform(0, 1.2)))
        assert tfx.num_indices_to_render(total, V, F, d) == draw_restated(total, V, F, d), (total, V, F, d)
