"""CPU tests of the C-ABI boundary: exports, POD packers, camera, errors.

Packers are checked against independent IEEE conversions (numpy float16 for
binary16, torch bfloat16 for bf16, both round-to-nearest-even) and the
byte layout of lib.rs:85-94 / camera.rs:8-11.
"""
import re
from pathlib import Path

import numpy as np
import pytest

from lambent_amd import api

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "lambent.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lmb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_function():
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(api._lib, n)]
    assert not missing, missing
    assert set(names) == set(api.HEADER_FUNCTIONS)


def f16(x):
    return int(np.asarray(x, np.float32).astype(np.float16).view(np.uint16))


def bf16(x):
    import torch
    return int(torch.tensor([x], dtype=torch.float32).to(torch.bfloat16).view(torch.int16)[0]) & 0xFFFF


@pytest.mark.parametrize("vals", [(0.0, 1.0), (0.5, 0.25), (1.0 / 3.0, 65504.0), (1e-5, -2.5),
                                  (70000.0, 6e-8), (2.0 ** -14, 1.0009765625 + 2 ** -12)])
def test_material_new_f16_packing(vals):
    rng = np.random.default_rng(0)
    m = api.Material.new(vals, vals[::-1], (0.75, 0.125), (3, 65535), 7, 2, None,
                         api.EmissionScale.from_f32(17.0, 0.3), (1.33, 1.3442),
                         api.MaterialType.Transparent)
    w = m.words()
    assert w[0] == f16(vals[0]) | (f16(vals[1]) << 16)
    assert w[1] == f16(vals[1]) | (f16(vals[0]) << 16)
    assert w[2] == f16(0.75) | (f16(0.125) << 16)
    assert w[3] == 3 | (65535 << 16)
    assert w[4] == 7 | (2 << 16)
    assert w[5] == 0xFFFF | (2 << 16)
    assert w[6] == bf16(17.0) | (bf16(0.3) << 16)
    assert w[7] == f16(1.33) | (f16(1.3442) << 16)
    assert len(m.to_bytes()) == 32


def test_f16_rounding_many_values():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.normal(size=2000) * 10 ** rng.uniform(-8, 5, 2000),
                         [np.float32(1 + 2 ** -11), np.float32(1 + 3 * 2 ** -11), 65520.0, 1e9]])
    for x in xs.astype(np.float32):
        m = api.Material.new((x, 0.0), (0, 0), (0, 0), (0, 0), 0, None, None,
                             api.EmissionScale(), None, api.MaterialType.Diffuse)
        assert m.words()[0] & 0xFFFF == f16(x), x


def test_bf16_rounding_many_values():
    rng = np.random.default_rng(2)
    xs = rng.normal(size=500) * 10 ** rng.uniform(-30, 30, 500)
    for x in xs.astype(np.float32):
        assert api.EmissionScale.from_emission_scale(float(x)).pack() & 0xFFFF == bf16(float(x))


def test_emission_scale_17_is_exact_in_bf16():
    p = api.EmissionScale.from_emission_scale(17.0).pack()
    assert np.array([p << 16], np.uint32).view(np.float32)[0] == 17.0
    assert p >> 16 == 0


def test_material_defaults_and_errors():
    m = api.Material.new((0, 0), (0, 0), (0, 0), (0, 0), 0, None, None, api.EmissionScale(),
                         None, api.MaterialType.Diffuse)
    assert m.words()[7] == f16(1.0) | (f16(1.0) << 16)  # VACUUM
    assert m.words()[4] >> 16 == 0xFFFF
    with pytest.raises(api.LambentError) as e:
        api.Material.new((0, 0), (0, 0), (0, 0), (0, 0), 0, 0xFFFF, None, api.EmissionScale(),
                         None, api.MaterialType.Diffuse)
    assert e.value.code == api.ERR_RESERVED_INDEX
    with pytest.raises(api.LambentError):
        api.Material.new((0, 0), (0, 0), (0, 0), (0, 0), 0, None, 0xFFFF, api.EmissionScale(),
                         None, api.MaterialType.Diffuse)


def test_refractive_index_between():
    r = api.refractive_indices.refractive_index_between(api.refractive_indices.AIR,
                                                        api.refractive_indices.DIAMOND)
    assert r[0] == pytest.approx(2.4063 / 1.00027531, rel=1e-6)
    assert r[1] == pytest.approx(2.4922 / 1.000287, rel=1e-6)


def test_dispatch_size():
    assert api.dispatch_size(256, 256) == api.DispatchSize(4, 256)
    assert api.dispatch_size(1920, 1080) == api.DispatchSize(30, 1080)
    assert api.dispatch_size(65, 1) == api.DispatchSize(2, 1)
    assert api.dispatch_size(0, 0) == api.DispatchSize(0, 0)


def test_camera_inverse_matches_float64():
    proj = api.perspective(45.0, 16 / 9, 0.1, 200.0)
    view = api.look_to_rh((0.0, 0.0, -25.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0))
    cam = api.Camera.from_proj_view(proj, view)
    # columns -> matrix: M[row, col] = cols[col][row]
    P = proj.astype(np.float64).T
    V = view.astype(np.float64).T
    np.testing.assert_allclose(cam.projection_inverse.T, np.linalg.inv(P), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(cam.view_inverse.T, np.linalg.inv(V), rtol=1e-6, atol=1e-6)
    assert len(cam.to_bytes()) == 128
    assert cam.to_bytes()[:64] == cam.projection_inverse.astype(np.float32).tobytes()


def test_camera_singular_errors():
    ok = np.eye(4, dtype=np.float32)
    bad = np.zeros((4, 4), np.float32)
    with pytest.raises(api.CameraError) as e:
        api.Camera.from_proj_view(bad, ok)
    assert e.value.code == api.ERR_SINGULAR_PROJECTION
    with pytest.raises(api.CameraError) as e:
        api.Camera.from_proj_view(ok, bad)
    assert e.value.code == api.ERR_SINGULAR_VIEW
    cam = api.Camera.from_proj_view(ok, ok)
    proj2 = api.perspective(60.0, 1.0, 0.1, 100.0)
    with pytest.raises(api.CameraError):
        cam.update_from_proj_view(proj2, bad)
    # camera.rs:27-42: the projection is updated before the view fails
    np.testing.assert_allclose(cam.projection_inverse.T,
                               np.linalg.inv(proj2.astype(np.float64).T), rtol=1e-6, atol=1e-9)
    np.testing.assert_array_equal(cam.view_inverse, np.eye(4, dtype=np.float32))


def test_tracer_options_validation():
    for bad in (dict(samples=0), dict(bounces=0), dict(bounces=65), dict(tmin=1.0, tmax=0.5)):
        with pytest.raises(api.LambentError) as e:
            api.RayTracer(api.Medium).create_pipeline(api.PathTracerOpts(**bad))
        assert e.value.code == api.ERR_INVALID_ARGUMENT
    api.RayTracer(api.High).create_pipeline(api.PathTracerOpts(samples=16, bounces=8, rr_start=3))


def test_textures_validation():
    with pytest.raises(api.TextureError):
        api.TextureSet(diffuse=[np.zeros((0, 4, 4), np.uint8)])
    ts = api.TextureSet(diffuse=[api.solid_rgba8(1, 2, 3)], background=api.uniform_background(1.0))
    assert ts.add(api.TextureKind.DIFFUSE, api.solid_rgba8(4, 5, 6)) == 1
    with pytest.raises(api.TextureError):
        api.TextureSet(background=np.zeros((5, 2, 2, 4), np.uint8))


def test_texture_loader_decodes_reference_pngs():
    # examples/cornell_box/src/*.png values (decoded RGBA8, textures.rs:31-38)
    pytest.importorskip("PIL")
    import io
    from PIL import Image
    buf = io.BytesIO()
    Image.new("RGB", (3, 2), (192, 0, 0)).save(buf, format="PNG")
    tl = api.TextureLoader()
    tl.load_from_bytes(buf.getvalue())
    assert tl.images[0].shape == (2, 3, 4)
    assert tuple(tl.images[0][0, 0]) == (192, 0, 0, 255)
    with pytest.raises(api.TextureError):
        tl.load_from_bytes(b"not an image")


def test_gpu_objects_fail_loudly_without_device(gpu_available):
    if gpu_available:
        pytest.skip("a GPU is present")
    with pytest.raises(api.LambentError) as e:
        api.DataBuffers(8, 8, 1)
    assert e.value.code == api.ERR_DEVICE


def test_instanced_scene_rejects_a_missing_geometry():
    """Instance validation happens before any device work."""
    v = np.zeros((3, 4), np.float32)
    with pytest.raises(api.LambentError, match="missing geometry"):
        api.Scene.from_instances([(v, None)], [(1, np.eye(3, 4), [0], 0xFF)],
                                 [api.Material.new([0, 0], [1, 0], [1, 1], [0, 0], 0, None, None,
                                                   api.EmissionScale(0.0, 0.0), None,
                                                   api.MaterialType.Diffuse)])


def test_aabb_geometry_is_validated_before_any_device_work():
    """lmb_geometry_desc boxes (6 floats, min then max) must be finite with
    min <= max, and a geometry holds triangles or boxes, not both; refused
    before the library touches a device (no GPU here)."""
    ident = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)
    mats = api.materials_array([api.Material.new(
        [0.5, 0.5], [0.5, 0.5], [0.5, 0.5], [0, 0], 0, None, None, api.EmissionScale(0.0, 0.0),
        None, api.MaterialType.Diffuse)])
    bad = [np.array([[0, 0, 0, 1, -1, 1]], np.float32),          # min > max
           np.array([[0, 0, np.nan, 1, 1, 1]], np.float32),      # not finite
           np.array([[0, 0, 0, np.inf, 1, 1]], np.float32)]
    for boxes in bad:
        with pytest.raises(api.LambentError, match="AABB 0 of geometry 0"):
            api.Scene.from_instances([api.AabbGeometry(boxes)],
                                     [(0, ident, np.zeros(1, np.uint32), 0xFF)], mats)


def test_readme_names_alias_the_code_names():
    """README.md:21-23, 36-50 names: Camera::new / descriptor, descriptor on a
    slice of materials, TextureLoader raw / create_texture."""
    proj = api.perspective(45.0, 1.5, 0.1, 200.0)
    view = api.look_to_rh((0.5, 0.5, -1.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0))
    a, b = api.Camera.new(proj, view), api.Camera.from_proj_view(proj, view)
    assert a.descriptor() == b.buffer_descriptor() == b.to_bytes() and len(a.descriptor()) == 128
    m = [api.Material.new([0.5, 0.5], [0.5, 0.5], [0.5, 0.5], [0, 0], k, None, None,
                          api.EmissionScale(0.0, 0.0), None, api.MaterialType.Diffuse) for k in range(3)]
    assert api.descriptor(m) == api.materials_buffer(m) and len(api.descriptor(m)) == 96
    tl = api.TextureLoader()
    tl.load_rgba8(api.solid_rgba8(1, 2, 3, 4, size=2))
    assert len(tl.raw()) == 1 and tl.raw()[0].shape == (2, 2, 4)
    assert api.TextureLoader.create_texture is api.TextureLoader.create_textures
