"""Measured figures of the parity tests, written as JSON files into one report directory at the repository root
(created on first use), one file per test module."""
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPORT_DIR = os.path.join(ROOT, "gpurun_out")


def write(name: str, figures: dict) -> None:
    os.makedirs(REPORT_DIR, exist_ok=True)
    with open(os.path.join(REPORT_DIR, name), "w") as f:
        json.dump(figures, f, indent=1)
