"""pytest configuration: the `gpu` marker and shared helpers.

Tests marked `gpu` need an MI355X (they call the gfx950 kernels through the C ABI);
everything else runs on the CPU (the oracle, host planning, the ABI surface)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and libpis_hip.so")


def force_err(f, f_ref):
    """Normwise force error max_i |dF_i| / RMS_i |F_ref,i| (the 1e-10 FP64 criterion)."""
    return float(np.abs(np.asarray(f) - f_ref).max() / np.sqrt((f_ref ** 2).sum(axis=1).mean()))


def force_rms_err(f, f_ref):
    """RMS_i |dF_i| / RMS_i |F_ref,i| (the mixed-precision criterion)."""
    d = np.asarray(f) - f_ref
    return float(np.sqrt((d ** 2).sum(axis=1).mean() / (f_ref ** 2).sum(axis=1).mean()))
