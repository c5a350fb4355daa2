"""MI355X-native Steel step path: scene packing, the reference's Sim/SceneBuilder API, HIP kernels via a C ABI."""
from .scene import SceneBuilder, convexHull, loadScene, LoadedScene, JT, GT
from .pack import pack_scene, UnsupportedError as UnsupportedSceneError

__all__ = ["SceneBuilder", "convexHull", "loadScene", "LoadedScene", "JT", "GT", "pack_scene", "UnsupportedSceneError"]
