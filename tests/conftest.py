import importlib
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def load_pkg(sub=None):
    name = "cfpq-on-gpgpu_amd" + ("." + sub if sub else "")
    return importlib.import_module(name)


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def gen():
    return load_pkg("gen")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.build()
    return o


@pytest.fixture(scope="session")
def engine(pkg):
    eng = pkg.Engine(0)
    yield eng
    eng.close()


@pytest.fixture(scope="session")
def host_cpu_bin(tmp_path_factory):
    """The C++ drop-in host (parsers, worklist driver, print_results) over a
    plain CPU bitset matrix: tests/host/cpu_backend_main.cpp."""
    d = tmp_path_factory.mktemp("hostbin")
    exe = str(d / "cfpq_host_cpu")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include", "cfpq"), "-o", exe,
                    os.path.join(ROOT, "tests", "host", "cpu_backend_main.cpp"),
                    os.path.join(ROOT, "cfpq-on-gpgpu_amd", "host", "Grammar.cpp"),
                    os.path.join(ROOT, "cfpq-on-gpgpu_amd", "host", "Graph.cpp")], check=True)
    return exe
