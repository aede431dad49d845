"""Loads the product package `ray-tracer_amd/` (its directory name is not a
Python identifier) as the module `ray_tracer_amd` and re-exports rtg."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "ray-tracer_amd")


def load_package():
    if "ray_tracer_amd" in sys.modules:
        return sys.modules["ray_tracer_amd"]
    spec = importlib.util.spec_from_file_location("ray_tracer_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ray_tracer_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


pkg = load_package()
from ray_tracer_amd import rtg  # noqa: E402,F401
