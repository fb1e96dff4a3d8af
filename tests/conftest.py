"""Test configuration.

Backends: tests taking the ``backend`` fixture run on the CPU oracle
(``sequential``, registered from ``oracle/``) and on ``hip`` (marked ``gpu``),
as the reference's suite is parametrised by backend (test/unit/conftest.py).
``op2.init`` selects a backend once per process, so a process runs one of
them: ``-m "not gpu"`` the oracle, ``-m gpu`` the GPU backend.
"""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import sequential  # noqa: E402

sequential.register()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (hip backend)")


def pytest_generate_tests(metafunc):
    if 'backend' in metafunc.fixturenames:
        names = ['sequential', 'hip']
        skip = set(getattr(metafunc.module, 'skip_backends', ())) | \
            set(getattr(metafunc.cls, 'skip_backends', ()) if metafunc.cls else ())
        only = getattr(metafunc.module, 'backends', None)
        params = []
        for b in names:
            if b in skip or (only and b not in only):
                continue
            marks = [pytest.mark.gpu] if b == 'hip' else []
            params.append(pytest.param(b, marks=marks))
        metafunc.parametrize('backend', params, indirect=True, scope='session')


def pytest_collection_modifyitems(items):
    def key(item):
        # modules named test_zz_* run after everything else
        last = os.path.basename(str(item.fspath)).startswith('test_zz_')
        try:
            return (last, item.callspec.getparam('backend'))
        except (AttributeError, ValueError):
            return (last, '_nobackend')
    items.sort(key=key)


_current = {}


@pytest.fixture(scope='session')
def backend(request):
    from pyop2_amd import op2, backends
    name = request.param
    cur = backends.get_backend()
    if cur not in ('void', name):
        pytest.skip('backend %s cannot run after %s in one process' % (name, cur))
    if cur == 'void':
        op2.init(backend=name)
    return name


@pytest.fixture(scope='session')
def hip():
    """An initialised hip backend (GPU tests only)."""
    from pyop2_amd import op2, backends
    cur = backends.get_backend()
    if cur == 'void':
        op2.init(backend='hip')
    elif cur != 'hip':
        pytest.skip('hip backend unavailable in a process running %s' % cur)
    return op2
