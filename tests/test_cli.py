"""Dataset import (gdsj_csv_load, import_dataset.cpp:49-95) and the
command-line join (gdsjoin_main <file> <eps> <dim> 3|1, main.cu:66-331),
with the table and outlier-score files (main.cu:544-626)."""
import os
import re
import subprocess

import numpy as np
import pytest

from gpu_self_join_amd import _lib
from oracle import oracle as O
from tests.helpers import README_COUNT, load_test_dat_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu_self_join_amd", "gdsjoin_main")


@pytest.fixture(scope="module")
def test_dat_path(tmp_path_factory):
    p = tmp_path_factory.mktemp("data") / "test_dat.txt"
    p.write_bytes(load_test_dat_text())
    return str(p)


@pytest.mark.parametrize("dtype", ["float", "double"])
def test_csv_load_matches_oracle_parse_on_test_dat(test_dat_path, dtype):
    npdt = np.float32 if dtype == "float" else np.float64
    v = _lib.csv_load(test_dat_path, dtype)
    ref = O.parse_text(load_test_dat_text(), npdt)
    assert v.dtype == npdt and v.shape == (200_000,)
    assert np.array_equal(v.view(np.uint32 if dtype == "float" else np.uint64),
                          ref.view(np.uint32 if dtype == "float" else np.uint64))


def _load_text(tmp_path, text, dtype="float"):
    p = tmp_path / "d.txt"
    p.write_bytes(text.encode())
    return _lib.csv_load(str(p), dtype)


def test_csv_pieces_and_newlines(tmp_path):
    # pieces are cut at ',' only: newlines (and spaces) inside a piece separate values
    v = _load_text(tmp_path, "1.5,2.25\n3.0,4\n5e-1 , -6\n")
    assert v.tolist() == [1.5, 2.25, 3.0, 4.0, 0.5, -6.0]


def test_csv_failed_extraction_ends_the_piece(tmp_path):
    # "7 x 8": 7 is read, x fails, 8 is lost with the rest of that piece;
    # an empty piece reads nothing; "+9" is a valid extraction
    v = _load_text(tmp_path, "1,7 x 8,,+9,2")
    assert v.tolist() == [1.0, 7.0, 9.0, 2.0]


def test_csv_decimal_to_float_directly(tmp_path):
    # just above the midpoint between 1 and the next float: read straight into
    # float it rounds up; through double it would land on the midpoint and
    # round to even (1.0)
    s = "1.00000005960464477539062500000001"
    v = _load_text(tmp_path, s)
    assert v[0] == np.float32(1.0 + 2.0 ** -23)
    assert np.float32(float(s)) == np.float32(1.0)
    assert v[0] == O.parse_text(s.encode(), np.float32)[0]


def test_csv_missing_file_is_an_error(tmp_path):
    with pytest.raises(_lib.GdsjError):
        _lib.csv_load(str(tmp_path / "nope.txt"))


def test_cli_usage_errors():
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode != 0 and "Incorrect number of input parameters" in r.stderr
    r = subprocess.run([CLI, "x.txt", "1.0", "2", "9"], capture_output=True, text=True)
    assert r.returncode != 0 and "not part of this build" in r.stderr


@pytest.mark.gpu
def test_cli_mode3_readme_count(gpu, test_dat_path, tmp_path):
    args = [CLI, test_dat_path, "1.0", "2", "3", "--k", "2", "--ilp", "0", "--shortcircuit", "0", "--reorder", "0",
            "--queryreorder", "0"]
    r = subprocess.run(args, capture_output=True, text=True, cwd=str(tmp_path), timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Total neighbors: (\d+)", r.stdout)
    assert m and int(m.group(1)) == README_COUNT
    line = (tmp_path / "gpu_stats.txt").read_text().strip().splitlines()[-1]
    f = [x.strip() for x in line.split(",")]
    assert f[1] == test_dat_path and float(f[2]) == 1.0 and int(f[3]) == README_COUNT
    # "..., GPUNUMDIM/.../DTYPE(float/double): 2, 2, 0, 0, 0, 0, 0, 0, float" (main.cu:260)
    assert f[4] == "GPUNUMDIM/NUMINDEXEDDIM/ILP/STAMP/SORT/REORDER/SHORTCIRCUIT/QUERYREORDER/DTYPE(float/double): 2"
    assert f[5:] == ["2", "0", "0", "0", "0", "0", "0", "float"]


@pytest.mark.gpu
def test_cli_print_matches_oracle(gpu, tmp_path):
    rng = np.random.default_rng(11)
    x = rng.uniform(0, 10, (300, 3)).astype(np.float32)
    p = tmp_path / "pts.txt"
    p.write_text("\n".join(",".join(repr(float(v)) for v in row) for row in x) + "\n")
    r = subprocess.run([CLI, str(p), "1.25", "3", "3", "--print", "--reorder", "0"], capture_output=True, text=True,
                       cwd=str(tmp_path), timeout=120)
    assert r.returncode == 0, r.stderr
    rows = re.findall(r"point id: (\d+), neighbors: ([\d,]*)", r.stdout)
    assert len(rows) == 300
    xs = _lib.csv_load(str(p)).reshape(-1, 3)
    ref = O.join(xs, float(np.float32(1.25)), k=3, ilp=8, shortcircuit=1, reorder=0)
    off = 0
    for i, (pid, ids) in enumerate(rows):
        assert int(pid) == i
        got = [int(t) for t in ids.split(",") if t]
        c = int(ref["counts"][i])
        assert got == sorted(ref["ids"][off:off + c].tolist())
        off += c


def _write_points(path, x):
    path.write_text("\n".join(",".join(repr(float(v)) for v in row) for row in x) + "\n")


@pytest.mark.gpu
def test_cli_table_and_outlier_files(gpu, tmp_path):
    rng = np.random.default_rng(12)
    x = rng.uniform(0, 10, (400, 3)).astype(np.float32)
    p = tmp_path / "pts.txt"
    _write_points(p, x)
    r = subprocess.run([CLI, str(p), "1.5", "3", "3", "--table-file", "--outlier-scores", "--reorder", "0"],
                       capture_output=True, text=True, cwd=str(tmp_path), timeout=120)
    assert r.returncode == 0, r.stderr
    xs = _lib.csv_load(str(p)).reshape(-1, 3)
    ref = O.join(xs, float(np.float32(1.5)), k=3, ilp=8, shortcircuit=1, reorder=0)
    counts = ref["counts"].astype(np.int64)
    # DSSJ_out.txt: header, then per point id its ascending ids, each followed by ", "
    lines = (tmp_path / "DSSJ_out.txt").read_text().split("\n")
    assert lines[0] == "#data point (line is the point id), neighbor point ids"
    off = 0
    for i in range(400):
        c = int(counts[i])
        want = "".join("%d, " % v for v in sorted(ref["ids"][off:off + c].tolist()))
        assert lines[1 + i] == want
        off += c
    # DSSJ_outlier_scores.txt: count, rank by ascending count (ties by id)
    sl = (tmp_path / "DSSJ_outlier_scores.txt").read_text().strip().split("\n")
    assert sl[0].startswith("#data point (line is the point id), col0: Number of neighbors")
    order = np.argsort(counts, kind="stable")
    rank = np.empty(400, np.int64)
    rank[order] = np.arange(400)
    got = np.array([[int(t) for t in ln.split(",")] for ln in sl[1:]])
    assert np.array_equal(got[:, 0], counts) and np.array_equal(got[:, 1], rank)


@pytest.mark.gpu
def test_cli_mode1_bruteforce_total(gpu, tmp_path):
    rng = np.random.default_rng(13)
    x = rng.uniform(0, 5, (500, 2)).astype(np.float32)
    p = tmp_path / "pts.txt"
    _write_points(p, x)
    r = subprocess.run([CLI, str(p), "0.75", "2", "1"], capture_output=True, text=True, cwd=str(tmp_path),
                       timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Total neighbours in table: (\d+)", r.stdout)
    xs = _lib.csv_load(str(p)).reshape(-1, 2)
    ref = O.bruteforce(xs, float(np.float32(0.75)), ilp=0, shortcircuit=0, reorder=0)
    assert m and int(m.group(1)) == int(ref["counts"].sum())


@pytest.mark.gpu
def test_cli_mode0_cpu_bruteforce_table(gpu, tmp_path):
    """Mode 0 (main.cu:121-141): the table of the reference's CPU brute
    force, whose sums are sequential and unfused (main.cu:963-967), listed
    point by point; checked against a float32 numpy restatement of those sums
    and against the restated CPU brute force's counts."""
    rng = np.random.default_rng(17)
    x = (rng.random((1500, 3)) * 6.0).astype(np.float32)
    path = tmp_path / "pts.txt"
    path.write_text("\n".join(",".join("%.9g" % v for v in row) for row in x) + "\n")
    xs = _lib.csv_load(str(path)).reshape(-1, 3)
    eps = np.float32(0.7)
    r = subprocess.run([CLI, str(path), "0.7", "3", "0"], capture_output=True, text=True, cwd=str(tmp_path),
                       timeout=120)
    assert r.returncode == 0, r.stderr
    # the CPU brute force's arithmetic: t = p_i - p_j, acc = acc + round(t*t)
    want = []
    for i in range(len(xs)):
        acc = np.zeros(len(xs), np.float32)
        for k in range(3):
            t = xs[i, k] - xs[:, k]
            acc = acc + t * t
        want.append(np.nonzero(np.sqrt(acc) <= eps)[0])
    total = sum(len(w) for w in want)
    assert "Total neighbors in table: %d" % total in r.stdout
    rows = [ln for ln in r.stdout.splitlines() if ln.startswith("point id: ")]
    assert len(rows) == len(xs)
    for i, ln in enumerate(rows):
        head, tail = ln.split("neighbors: ")
        assert head == "point id: %d, " % i
        got = np.array([int(v) for v in tail.split(",") if v], dtype=np.int64)
        assert np.array_equal(got, want[i]), i
    c, _ = O.cpu_bruteforce(xs, float(eps), np.arange(len(xs), dtype=np.uint32), 4)
    assert np.array_equal(c, np.array([len(w) for w in want]))
