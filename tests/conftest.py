import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

PATTERN_DIR = ROOT / "tests" / "golden" / "patterns"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtdfs on cuda:0)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def fixture_patterns():
    """The reference's 14 pattern files (data/pattern/1.g..14.g), copied as data."""
    return [(i, (PATTERN_DIR / f"{i}.g").read_text()) for i in range(1, 15)]


@pytest.fixture(scope="session")
def patterns():
    return fixture_patterns()
