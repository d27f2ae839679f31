import sys
from pathlib import Path

REPO = Path(__file__).resolve().parents[1]
for p in (REPO / "n-body-2d_amd", REPO / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
