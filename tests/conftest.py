import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (gfx950) GPU and the built HIP library')


@pytest.fixture(scope='session', autouse=True)
def _oracle_built():
    """The oracle library is test infrastructure; build it when missing."""
    so = os.path.join(REPO, 'oracle', 'liboracle.so')
    if not os.path.exists(so):
        subprocess.check_call(['make', '-s', '-C', os.path.join(REPO, 'oracle')])
    yield
