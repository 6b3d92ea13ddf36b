"""Dataset import (ImportDataset grammar, main.cu:469-498) and label files
(cluster-check.cpp:201-224 format)."""
import numpy as np
import pytest

import dclust


def test_import_ex_txt_format(tmp_path):
    p = tmp_path / "ex.txt"
    p.write_text("1,1\n2,2\n3,3\n10,10\n11,11\n12,12\n20,20\n21,21\n22,22")
    x = dclust.import_dataset(str(p), 2)
    assert x.shape == (9, 2) and x[3].tolist() == [10.0, 10.0] and x[-1].tolist() == [22, 22]


def test_import_flattens_rows_like_the_reference(tmp_path):
    # values are flattened row-major regardless of line breaks
    p = tmp_path / "d.txt"
    p.write_text("0.5,1.5,2.5\n3.5,4.5,5.5\n")
    assert dclust.import_dataset(str(p), 2).tolist() == [[0.5, 1.5], [2.5, 3.5], [4.5, 5.5]]
    assert dclust.import_dataset(str(p), 3).shape == (2, 3)


def test_import_max_points_and_exponents(tmp_path):
    p = tmp_path / "d.txt"
    p.write_text("1e-3, -2.5E+2\n7,8\r\n9,10\n")
    x = dclust.import_dataset(str(p), 2, max_points=2)
    assert x.tolist() == [[1e-3, -250.0], [7.0, 8.0]]


@pytest.mark.parametrize("text", ["1,2\nabc,3\n", "1,2,3\n"])
def test_import_rejects_bad_or_partial(tmp_path, text):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(dclust.DclustError) as e:
        dclust.import_dataset(str(p), 2)
    assert e.value.code == dclust.E_IO


def test_import_missing_file():
    with pytest.raises(dclust.DclustError) as e:
        dclust.import_dataset("/nonexistent/file.txt", 2)
    assert e.value.code == dclust.E_IO


def test_label_file_round_trip(tmp_path):
    lab = np.array([0, 3, 1, 1, 0, 2], dtype=np.int32)
    p = tmp_path / "labels.txt"
    dclust.write_labels(str(p), lab)
    assert p.read_text() == "0\n3\n1\n1\n0\n2\n"
    assert np.loadtxt(p, dtype=np.int32).tolist() == lab.tolist()


def test_import_long_lines(tmp_path):
    vals = np.arange(6000, dtype=np.float64)
    p = tmp_path / "long.txt"
    p.write_text(",".join(f"{v:.1f}" for v in vals) + "\n")
    x = dclust.import_dataset(str(p), 3)
    assert x.shape == (2000, 3) and x.ravel().tolist() == vals.tolist()


@pytest.mark.parametrize("dim", [2, 3])
def test_binary_round_trip(tmp_path, dim):
    rng = np.random.default_rng(dim)
    x = rng.normal(0, 1e3, (5001, dim))
    x[7] = [np.pi] * dim
    p = tmp_path / "d.f64"
    dclust.write_dataset_f64(str(p), x)
    assert p.stat().st_size == 24 + x.size * 8
    assert p.read_bytes()[:8] == b"DCLUSTF8"
    y = dclust.import_dataset(str(p), dim)
    assert y.shape == x.shape and np.array_equal(y.view(np.uint64), x.view(np.uint64))
    assert np.array_equal(dclust.import_dataset(str(p), dim, max_points=10), x[:10])


def test_binary_errors(tmp_path):
    x = np.arange(30, dtype=np.float64).reshape(10, 3)
    p = tmp_path / "d.f64"
    dclust.write_dataset_f64(str(p), x)
    with pytest.raises(dclust.DclustError) as e:
        dclust.import_dataset(str(p), 2)  # the file says 3-D
    assert e.value.code == dclust.E_PARAM
    data = p.read_bytes()
    (tmp_path / "short.f64").write_bytes(data[:-8])
    (tmp_path / "long.f64").write_bytes(data + b"\0")
    for name in ("short.f64", "long.f64"):
        with pytest.raises(dclust.DclustError) as e:
            dclust.import_dataset(str(tmp_path / name), 3)
        assert e.value.code == dclust.E_IO


@pytest.mark.parametrize("chunk", ["7", "64", "0"])
def test_text_parse_in_chunks_is_exact(tmp_path, monkeypatch, chunk):
    # many threads over small chunks (cut at line ends) give exactly the
    # values strtod gives the sequential reader: repr() round-trips
    if chunk != "0":
        monkeypatch.setenv("DCLUST_IO_CHUNK", chunk)
    rng = np.random.default_rng(5)
    x = rng.normal(0, 10, (4000, 3)) * np.exp(rng.uniform(-20, 20, (4000, 3)))
    lines = []
    for i, row in enumerate(x):
        # mixed layouts: spaces, CRLF, two points on one line, empty fields
        if i % 7 == 3:
            lines.append(" " + " , ".join(repr(float(v)) for v in row) + " \r")
        elif i % 11 == 5:
            lines.append(",".join(repr(float(v)) for v in row) + ",,")
        else:
            lines.append(",".join(repr(float(v)) for v in row))
    p = tmp_path / "big.txt"
    p.write_text("\n".join(lines) + "\n")
    y = dclust.import_dataset(str(p), 3)
    assert np.array_equal(y.view(np.uint64), x.view(np.uint64))


@pytest.mark.parametrize("chunk", ["5", "0"])
def test_text_limit_stops_before_a_later_bad_line(tmp_path, monkeypatch, chunk):
    # reading stops after the line that reaches max_points, as the sequential
    # reader does: a bad line after it is not an error, one before it is
    if chunk != "0":
        monkeypatch.setenv("DCLUST_IO_CHUNK", chunk)
    p = tmp_path / "d.txt"
    p.write_text("1,2\n3,4\n5,6\nabc\n7,8\n")
    assert dclust.import_dataset(str(p), 2, max_points=3).tolist() == [[1, 2], [3, 4], [5, 6]]
    for mp in (0, 4):
        with pytest.raises(dclust.DclustError) as e:
            dclust.import_dataset(str(p), 2, max_points=mp)
        assert e.value.code == dclust.E_IO
    # the limit reached inside a line whose later field is bad: an error
    q = tmp_path / "e.txt"
    q.write_text("1,2,3,x\n")
    with pytest.raises(dclust.DclustError):
        dclust.import_dataset(str(q), 2, max_points=1)


def test_cli_converts_text_to_binary(tmp_path):
    # `dclust <file> --write-f64 out` converts without touching a GPU
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(dclust.__file__), "bin", "dclust")
    src = tmp_path / "ex.txt"
    src.write_text("1,1\n2,2\n3,3\n10,10\n11,11\n12,12\n20,20\n21,21\n22,22")
    dst = tmp_path / "ex.f64"
    r = subprocess.run([exe, str(src), "--dim", "2", "--write-f64", str(dst)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(dclust.import_dataset(str(dst), 2), dclust.import_dataset(str(src), 2))
