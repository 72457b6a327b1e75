"""The module aliasing that lets the reference's scene code import the
drop-in classes unchanged (`sys.modules["powerfoam.rasterize"] = ...`): the
reference's import lines then resolve to this package's classes."""

import importlib
import sys
import types


def test_reference_imports_resolve_to_the_drop_in(monkeypatch):
    import powerfoam_amd.bvh
    import powerfoam_amd.camera
    import powerfoam_amd.color_fn
    import powerfoam_amd.geometry
    import powerfoam_amd.rasterize
    import powerfoam_amd.raytrace
    pkg = types.ModuleType("powerfoam")
    pkg.__path__ = []
    monkeypatch.setitem(sys.modules, "powerfoam", pkg)
    for name in ("rasterize", "camera", "raytrace", "color_fn", "bvh", "geometry"):
        monkeypatch.setitem(sys.modules, f"powerfoam.{name}", importlib.import_module(f"powerfoam_amd.{name}"))
    from powerfoam.bvh import AABBTree
    from powerfoam.camera import TorchCamera
    from powerfoam.color_fn import SphericalVoronoi
    from powerfoam.geometry import InterpenetrationFunction, morton_sort
    from powerfoam.rasterize import Rasterizer
    from powerfoam.raytrace import RayTracer
    assert Rasterizer is powerfoam_amd.rasterize.Rasterizer
    assert RayTracer is powerfoam_amd.raytrace.RayTracer
    assert TorchCamera is powerfoam_amd.camera.TorchCamera
    assert SphericalVoronoi is powerfoam_amd.color_fn.SphericalVoronoi
    assert AABBTree is powerfoam_amd.bvh.AABBTree
    assert morton_sort is powerfoam_amd.geometry.morton_sort
    assert InterpenetrationFunction is powerfoam_amd.geometry.InterpenetrationFunction
