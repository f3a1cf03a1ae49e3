import os
import sys

import pytest

# torch (when present) is imported before the library is first loaded: both
# link a HIP runtime under the soname libamdhip64.so.7, the first one loaded
# serves the process, and torch's GPU tensors need its own (the torch binding
# of the adjoint is tested with them).
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_hashes.json")) as f:
        return json.load(f)


def load_package():
    """Import the rast-3dmhd_amd package (its directory name is not an identifier)."""
    import importlib.util
    if "rast3dmhd_amd" in sys.modules:
        return sys.modules["rast3dmhd_amd"]
    path = os.path.join(ROOT, "rast-3dmhd_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(
        "rast3dmhd_amd", path, submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rast3dmhd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
