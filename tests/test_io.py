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
