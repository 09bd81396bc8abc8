"""Helpers shared by the tests: reference case tables and tensor plumbing."""
import csv
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_cases(name):
    """Rows of tests/golden/cases/<name> (the reference's CSV case table) as dicts."""
    with open(os.path.join(GOLDEN, "cases", name), newline="") as f:
        return list(csv.DictReader(f))


def status_code(name):
    import aclblas
    return aclblas.STATUS_CODE[name.replace("ACLBLAS_STATUS_", "")]


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()
