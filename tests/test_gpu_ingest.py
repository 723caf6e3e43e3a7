"""A dataset read from a CSV or JSON file on the device equals the same rows given in
memory (rsr_dataset_from_file vs rsr_dataset_create): identical MSEs for a program batch,
with and without z-score normalisation."""
import json
import os

import numpy as np
import pytest

from rsr_amd import api
from rsr_amd import workloads as W
from rsr_amd.isa import VT

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", ["csv", "json"])
@pytest.mark.parametrize("normalize", [False, True])
def test_dataset_from_file_equals_in_memory(gpu_ctx, tmp_path, fmt, normalize):
    X, y, t = W.data_c1(1001, seed=77)
    p = os.path.join(tmp_path, "d." + fmt)
    with open(p, "w") as f:
        if fmt == "csv":
            f.write(",".join(f"c{i}" for i in range(X.shape[1])) + ",y\n")
            for r, v in zip(X, y):
                f.write(",".join(repr(float(a)) for a in r) + "," + repr(float(v)) + "\n")
        else:
            json.dump([{"x": [float(a) for a in r], "y": float(v)} for r, v in zip(X, y)], f)
    progs = W.gen_c1(64, seed=78)
    P = api.Programs(gpu_ctx, progs)
    a = api.Dataset.new(gpu_ctx, X, y, t, normalize=normalize)
    b = api.Dataset.from_file(gpu_ctx, p, t, fmt=fmt, normalize=normalize)
    assert a.info() == b.info()
    ma = api.compute_mse(gpu_ctx, P, a, allow_rejected=True)
    mb = api.compute_mse(gpu_ctx, P, b, allow_rejected=True)
    assert np.array_equal(ma.view(np.uint32), mb.view(np.uint32))
