"""Shared test set-up: the package (directory hosono-2dmhd_amd, imported as
hosono_2dmhd_amd), the oracle's Python front end, the `gpu` marker."""
import importlib.util
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "hosono-2dmhd_amd"


def load_package():
    if "hosono_2dmhd_amd" in sys.modules:
        return sys.modules["hosono_2dmhd_amd"]
    spec = importlib.util.spec_from_file_location(
        "hosono_2dmhd_amd", PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hosono_2dmhd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()
if str(ROOT / "oracle") not in sys.path:
    sys.path.insert(0, str(ROOT / "oracle"))
for extra in (ROOT / "tests", ROOT / "tests" / "golden"):
    if str(extra) not in sys.path:
        sys.path.insert(0, str(extra))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X")
