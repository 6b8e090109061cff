# SPDX-License-Identifier: Apache-2.0
import os
import sys

# torch bundles its own libamdhip64.so.7; loading it before libbenes_hip.so
# makes both share that one HIP runtime in the test process.
try:
    import torch  # noqa: F401
except ImportError:  # CPU-only environments without torch
    pass

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, os.path.join(ROOT, "benes-perm_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
