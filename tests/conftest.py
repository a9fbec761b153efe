import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
# the serving tests run the persistent work-queue kernel beside grid-wide launches on
# streams of their own: one hardware queue each (set before HIP starts)
if int(os.environ.get('GPU_MAX_HW_QUEUES', '4')) < 16:  # raised, never lowered
    os.environ['GPU_MAX_HW_QUEUES'] = '16'


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs through the HIP C-ABI)')
    config.addinivalue_line('markers', 'slow: long CPU test')


@pytest.fixture(scope='session')
def kats():
    import json

    with open(os.path.join(ROOT, 'tests', 'golden', 'reference_kats.json')) as f:
        return json.load(f)
