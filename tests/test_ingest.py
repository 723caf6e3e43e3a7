"""Datasets on disk (Dataset::from_csv / from_json, dataset.rs:265-357), host readers.

Pins: known answers written by hand (target column in the middle, quoted fields, CRLF,
blank lines, a header-only file), Python's csv and json modules as an independent parser
of random files (numbers written with repr of float32 values, which parse back to the
same float32 bits), and the reference's error cases (too few columns, a record longer or
shorter than the header, a non-number, whitespace inside a field, hex, a wrong x length).
"""
import csv
import json
import os

import numpy as np
import pytest

from rsr_amd import api
from rsr_amd.isa import VT

F1 = [VT.Float]
F2 = [VT.Float, VT.Float]


def _write(tmp_path, name, text):
    p = os.path.join(tmp_path, name)
    with open(p, "w", newline="") as f:
        f.write(text)
    return p


def test_csv_known_answers(tmp_path):
    p = _write(tmp_path, "a.csv", "a,b,y\n1.5,-2,3\r\n\n\"4\",5e-1,\"-6.25\"\n")
    rows, y = api.read_dataset_file(p, F2)
    assert rows.tolist() == [[1.5, -2.0], [4.0, 0.5]]
    assert y.tolist() == [3.0, -6.25]
    # target in the middle: features are the other columns, left to right
    rows, y = api.read_dataset_file(p, F2, target_col=1)
    assert rows.tolist() == [[1.5, 3.0], [4.0, -6.25]]
    assert y.tolist() == [-2.0, 0.5]
    # a Vec3 feature takes three columns
    p = _write(tmp_path, "v.csv", "x0,x1,x2,t\n1,2,3,4\n")
    rows, y = api.read_dataset_file(p, [VT.Vec3])
    assert rows.tolist() == [[1, 2, 3]] and y.tolist() == [4]
    assert api.read_dataset_file(_write(tmp_path, "h.csv", "a,y\n"), F1)[1].shape == (0,)


def test_csv_special_numbers(tmp_path):
    p = _write(tmp_path, "s.csv", "a,y\ninf,-Infinity\nNaN,+1.\n.25,1E2\n")
    rows, y = api.read_dataset_file(p, F1)
    assert np.isinf(rows[0, 0]) and y[0] == -np.inf
    assert np.isnan(rows[1, 0]) and y[1] == 1.0
    assert rows[2, 0] == 0.25 and y[2] == 100.0


@pytest.mark.parametrize("text", [
    "a,b,y\n1,2\n",          # shorter than the header
    "a,y\n1,2,3\n",          # longer than the header
    "a,y\n1,x\n",            # not a number
    "a,y\n1, 2\n",           # whitespace inside a field
    "a,y\n0x10,2\n",         # hex
    "a,y\n,2\n",             # empty field
])
def test_csv_errors(tmp_path, text):
    with pytest.raises(api.RsrError):
        api.read_dataset_file(_write(tmp_path, "e.csv", text), F1)


def test_csv_too_few_columns_for_the_schema(tmp_path):
    with pytest.raises(api.RsrError):
        api.read_dataset_file(_write(tmp_path, "e.csv", "a,y\n1,2\n"), F2)


def test_csv_random_files_match_python_csv(tmp_path):
    rng = np.random.default_rng(3)
    for trial in range(5):
        n, F = int(rng.integers(1, 200)), int(rng.integers(1, 6))
        vals = (rng.normal(size=(n, F + 1)) * 10.0 ** rng.integers(-3, 4, size=(n, F + 1))).astype(np.float32)
        tc = int(rng.integers(0, F + 1))
        p = os.path.join(tmp_path, f"r{trial}.csv")
        with open(p, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow([f"c{i}" for i in range(F + 1)])
            for r in vals:
                w.writerow([repr(float(v)) for v in r])
        rows, y = api.read_dataset_file(p, [VT.Float] * F, target_col=tc)
        with open(p, newline="") as f:
            recs = list(csv.reader(f))[1:]
        ref_y = np.array([np.float32(float(r[tc])) for r in recs], np.float32)
        ref_x = np.array([[np.float32(float(v)) for i, v in enumerate(r) if i != tc] for r in recs], np.float32)
        assert np.array_equal(y.view(np.uint32), ref_y.view(np.uint32))
        assert np.array_equal(rows.view(np.uint32), ref_x.reshape(n, F).view(np.uint32))
        assert np.array_equal(rows.view(np.uint32), np.delete(vals, tc, axis=1).view(np.uint32))


def test_json_known_answers_and_random(tmp_path):
    p = _write(tmp_path, "a.json", '[{"x": [1.5, -2], "y": 3, "id": "a"}, {"y": -0.5e1, "x": [4, 0.25]}]')
    rows, y = api.read_dataset_file(p, F2, fmt="json")
    assert rows.tolist() == [[1.5, -2.0], [4.0, 0.25]] and y.tolist() == [3.0, -5.0]
    assert api.read_dataset_file(_write(tmp_path, "e.json", " [ ] "), F2, fmt="json")[1].shape == (0,)
    rng = np.random.default_rng(4)
    vals = rng.normal(size=(50, 4)).astype(np.float32)
    recs = [{"x": [float(v) for v in r[:3]], "y": float(r[3]), "meta": {"k": [1, None, True]}} for r in vals]
    p = _write(tmp_path, "r.json", json.dumps(recs))
    rows, y = api.read_dataset_file(p, [VT.Vec3], fmt="json")
    assert np.array_equal(rows, vals[:, :3]) and np.array_equal(y, vals[:, 3])


@pytest.mark.parametrize("text", ['[{"x": [1], "y": 2}]', '[{"x": [1, 2]}]', '{"x": [1, 2], "y": 1}',
                                  '[{"x": [1, 2], "y": 1}] x', '[{"x": [1, "a"], "y": 1}]',
                                  '[{"x": [1, 2], "y": +1}]'])
def test_json_errors(tmp_path, text):
    with pytest.raises(api.RsrError):
        api.read_dataset_file(_write(tmp_path, "e.json", text), F2, fmt="json")
