import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

__graft_entry__.import_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


@pytest.fixture(scope="session")
def lib():
    """The HIP library (GPU tests only).  Fails loudly if it is not built."""
    import compound_ray_amd as cra

    L = cra.load_library()
    L.setVerbosity(False)
    return L


@pytest.fixture(scope="session")
def config1(tmp_path_factory):
    from compound_ray_amd import scenes

    return scenes.build_config(1, tmp_path_factory.mktemp("config1"))


@pytest.fixture(scope="session")
def oracle_c1(config1):
    import oracle

    return oracle.OracleScene.from_scene(config1["scene"])
