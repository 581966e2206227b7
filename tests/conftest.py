import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, 'oracle'))
sys.path.insert(0, os.path.join(ROOT, 'tests'))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device) to run')


def load_zav1():
    """Import the product package (directory name has a hyphen)."""
    if 'zav1' in sys.modules:
        return sys.modules['zav1']
    path = os.path.join(ROOT, 'zenav1-svt_amd', '__init__.py')
    spec = importlib.util.spec_from_file_location('zav1', path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules['zav1'] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope='session')
def zav1():
    return load_zav1()


@pytest.fixture(scope='session')
def gpu(zav1):
    """Initialise device 0 (GPU tests only)."""
    if zav1.device_count() <= 0:
        pytest.fail('no HIP device visible for a gpu-marked test')
    zav1.init(0)
    return zav1


@pytest.fixture(scope='session')
def oracle():
    import oracle as O
    O.lib()
    return O
