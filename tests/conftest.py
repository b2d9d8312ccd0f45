import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

ge.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return ge.load_package()


@pytest.fixture(scope="session")
def capi():
    from luna_jl_amd import _capi
    return _capi


@pytest.fixture(scope="session")
def orc(capi):
    oracle = ge.load_oracle()
    return oracle.Oracle(capi)


@pytest.fixture(scope="session")
def refvals():
    """Values the reference's own tests hold (tests/golden/reference_values.json)."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_values.json")) as f:
        return {k: v["value"] for k, v in json.load(f).items()}


@pytest.fixture(scope="session")
def cfg0(pkg):
    from luna_jl_amd import configs
    return configs.capillary_instance()


@pytest.fixture(scope="session")
def cfg_small(pkg):
    from luna_jl_amd import configs
    return configs.capillary_instance(energy=150e-6, pressure=2.0, trange=0.5e-12)
