"""Shared fixtures: the `gpu` marker, loaders for the package, the product
library and the CPU oracle (the oracle is the checker, never the product)."""
import ctypes
import importlib.util
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG_DIR = ROOT / "sparse-eigs-mex_amd"
ORACLE_SO = ROOT / "oracle" / "build" / "liboracle.so"


def load_pkg():
    if "speigs_amd" in sys.modules:
        return sys.modules["speigs_amd"]
    spec = importlib.util.spec_from_file_location("speigs_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["speigs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: longer CPU cases")


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle(pkg):
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    lib = ctypes.CDLL(str(ORACLE_SO))
    pkg.bind_api(lib, "oracle_")
    return lib


@pytest.fixture(scope="session")
def hiplib(pkg):
    return pkg.load()


@pytest.hookimpl(hookwrapper=True)
def pytest_runtest_makereport(item, call):
    """A failed test's report carries speigs_hip_last_error(): where the
    library's most recent internal failure on this thread happened."""
    outcome = yield
    rep = outcome.get_result()
    if rep.when == "call" and rep.failed:
        mod = sys.modules.get("speigs_amd")
        if mod is not None and getattr(mod, "_lib", None) is not None:
            rep.sections.append(("speigs_hip_last_error", mod.last_error() or "(none recorded)"))
