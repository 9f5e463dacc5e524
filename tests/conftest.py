import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: test needs an MI355X GPU')


@pytest.fixture(scope='session')
def backend_factory():
    """Create backends on the GPU (one library load per session)."""
    from pyfr_amd.backend import AMDBackend
    from pyfr_amd.inifile import Inifile

    made = []

    def make(precision='double', graphs=True, fused=True, fuse_flux=True):
        """precision: 'double', 'single' or 'mixed'; fused: use the fused
        operator kernels where available; fuse_flux: also fold tdisf into
        them for straight-sided Euler elements."""
        cfg = Inifile.from_dict({'backend': {'precision': precision},
                                 'backend-hip': {'graphs': str(graphs),
                                                 'fused': str(fused),
                                                 'fuse-flux': str(fuse_flux)}})
        be = AMDBackend(cfg)
        made.append(be)
        return be

    yield make
