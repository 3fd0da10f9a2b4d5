"""Import helper for the product package, whose directory name
(voxelengine3.0_amd) is not a Python identifier: it is registered as the
module `voxelengine_amd`."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "voxelengine3.0_amd")


def load_package():
    name = "voxelengine_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    return mod
