"""
Test configuration: import paths and the `gpu` marker.

Tests marked `gpu` need an MI355X and call the HIP path through
libzkgpu_hip.so; they are not skipped when the GPU is missing (the package
has no CPU fallback, so they fail loudly instead). Everything else runs on
the CPU: the oracle against known answers, the product's host-side logic,
and the C ABI's exported symbols.
"""

import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
for sub in ("zk-gpu-accelerator_amd", "oracle"):
    p = str(ROOT / sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running CPU oracle comparisons")
