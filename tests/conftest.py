import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


class ZeroReference(AssertionError):
    """The reference of a relative comparison is identically zero: the check
    would compare nothing (a run too short for the field to reach what it
    samples, say)."""


def rel_l2(a, b):
    """Relative L2 distance of `a` from the reference `b`; raises ZeroReference
    when `b` is all zero, so that no comparison passes on two zero arrays."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    if not den > 0:
        raise ZeroReference(f"reference of shape {b.shape} is identically zero (or not finite)")
    return float(np.linalg.norm((a - b).ravel()) / den)


# Model parameters after `sweep` calls of the reference's models_isFinish():
# multiLayerModel.c:235-269 (nextStructure1), zigzagModel.c:80-104,
# morphoScaleModel.c:256-285 (nextStructure).  The circle model has one point.
def swept_params(model, sweep):
    p = {}
    if model == 2:  # multilayer
        t0, t1, edge, branch, layer = 100, 90, 0.0, 0, 4
        for _ in range(sweep):
            t0 += 10
            if t0 > t1 + 50:
                t1 += 10
                t0 = t1 + 10
                if t1 > 150:
                    t1 = 90
                    edge += 0.5
                    if edge > 1.0:
                        edge = 0.0
                        branch += 50
                        if branch > 50:
                            branch = 0
                            layer += 4
        p = dict(ml_thick_nm0=t0, ml_thick_nm1=t1, ml_edge_rate=edge, ml_branch_nm=branch, ml_layer_num=layer)
    elif model == 5:  # zigzag
        deg, thick = 10, 80
        for _ in range(sweep):
            deg += 10
            if deg > 80:
                deg = 10
                thick += 10
        p = dict(zz_degree=deg, zz_thick_nm=thick)
    elif model == 3:  # morpho
        t, edge, branch, layer = 90, 0.0, 0, 6
        for _ in range(sweep):
            t += 10
            if t > 160:
                t = 90
                edge += 0.5
                if edge > 1.0:
                    edge = 0.0
                    branch += 25
                    if branch > 50:
                        layer += 2
                        branch = 0
        p = dict(mo_thick_nm0=t, mo_thick_nm1=t, mo_edge_rate=edge, mo_branch_nm=branch, mo_layer_num=layer)
    return p


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, f"ref_{name}.npz")))
