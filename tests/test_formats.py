"""Data formats around the path (host-only, no GPU): the CSV reader
(readers/csv.rs), the .xy dataset layout (dataset/mod.rs:56-118) and the
estimators' save / load (the serde round trips of knn_classifier.rs:352-363
and knn_regressor.rs), with the device copy of the rows made lazily."""
import os

import numpy as np
import pytest

from smartcore_amd import smartcore as sc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_csv_reference_kats():
    # csv.rs read_simple_csv: one header row, three data rows
    m = sc.matrix_from_csv_source("'sepal.length','sepal.width','petal.length','petal.width'\n"
                                  "5.1,3.5,1.4,0.2\n4.9,3.0,1.4,0.2\n4.7,3.2,1.3,0.2")
    np.testing.assert_array_equal(m, [[5.1, 3.5, 1.4, 0.2], [4.9, 3.0, 1.4, 0.2],
                                      [4.7, 3.2, 1.3, 0.2]])
    # csv.rs read_csv_semicolon_as_seperator: two header rows, ';'
    m = sc.matrix_from_csv_source("'sepal.length';'sepal.width';'petal.length';'petal.width'\n"
                                  "'Length of sepals.';'Width of Sepals';'Length of petals';"
                                  "'Width of petals'\n5.1;3.5;1.4;0.2\n4.9;3.0;1.4;0.2\n"
                                  "4.7;3.2;1.3;0.2", n_rows_header=2, field_seperator=";")
    assert m.shape == (3, 4) and m[2, 2] == 1.3
    # extract_row_vectors_from_csv_text read_default_csv
    m = sc.matrix_from_csv_source("column 1, column 2, column3\n1.0,2.0,3.0\n4.0,5.0,6.0")
    np.testing.assert_array_equal(m, [[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])


def test_csv_reference_error_texts():
    with pytest.raises(sc.ReadingError) as e:
        sc.matrix_from_csv_source("'sepal.length','sepal.width','petal.length','petal.width'\n"
                                  "5.1,3.5,1.4,0.2\n4.9,invalid,1.4,0.2\n4.7,3.2,1.3,0.2")
    assert e.value.kind == "InvalidField"
    assert e.value.msg == "Value 'invalid' could not be read. Column: 1, Row: 1."
    with pytest.raises(sc.ReadingError) as e:
        sc.matrix_from_csv_source("'field_1','field_2'\n5.1,3.5\n4.9,3.0,1.4")
    assert e.value.kind == "InvalidField"
    assert e.value.msg == "3 fields found but expected 2, Row: 1."
    with pytest.raises(sc.ReadingError) as e:
        sc.matrix_from_csv_source("only a header\n")
    assert e.value.kind == "NoRowsProvided" and e.value.msg is None


def test_csv_values_parse_like_rust_floats():
    # whitespace around fields is trimmed; exponents, signs, inf and NaN
    # parse; the f32 reader rounds the decimal text once, straight to f32
    m = sc.matrix_from_csv_source("h\n 1e3 \n-2.5E-1\n+inf\nNaN\n.5\n7.\n")
    assert m[:4, 0].tolist()[:3] == [1000.0, -0.25, float("inf")] and np.isnan(m[3, 0])
    assert m[4, 0] == 0.5 and m[5, 0] == 7.0
    f = sc.matrix_from_csv_source("h\n0.1\n16777217\n", f32=True)
    assert f.dtype == np.float32
    assert f[0, 0] == np.float32(0.1) and f[1, 0] == np.float32(16777216.0)
    for bad in ("0x10", "1e", "e5", "1.2.3", "", "--1"):
        with pytest.raises(sc.ReadingError):
            sc.matrix_from_csv_source("h\n" + bad + "\n")


def test_xy_layout_and_round_trip():
    # the header is num_features then num_samples, u64 little-endian, then
    # the f32 values row by row and the f32 targets
    d = np.load(os.path.join(GOLDEN, "digits.npz"))
    x = np.ascontiguousarray(d["x"], dtype=np.float32)
    y = np.ascontiguousarray(d["y"], dtype=np.float32)
    b = sc.serialize_data(x, y)
    want = (np.array([x.shape[1], x.shape[0]], "<u8").tobytes() + x.astype("<f4").tobytes() +
            y.astype("<f4").tobytes())
    assert b == want
    x2, y2 = sc.deserialize_data(want)
    np.testing.assert_array_equal(x2, x)
    np.testing.assert_array_equal(y2, y)
    with pytest.raises(sc.Failed):
        sc.deserialize_data(want[:100])


def _knn_data():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(200, 5))
    y = rng.integers(0, 4, size=200).astype(np.int32)
    return x, y


@pytest.mark.parametrize("metric", ["euclidian", "manhattan", "minkowski"])
def test_classifier_round_trip_without_device(metric):
    x, y = _knn_data()
    dist = {"euclidian": sc.Distances.euclidian(), "manhattan": sc.Distances.manhattan(),
            "minkowski": sc.Distances.minkowski(3)}[metric]
    p = sc.KNNClassifierParameters().with_k(4).with_distance(dist) \
        .with_weight(sc.KNNWeightFunction.Distance)
    knn = sc.KNNClassifier.fit(x, y, p)
    assert not knn.on_device()  # the rows go to the device on the first predict
    back = sc.KNNClassifier.deserialize(knn.serialize())
    assert back == knn
    assert not back.on_device()
    other = sc.KNNClassifier.fit(x, y, p.with_k(5))
    assert not (other == knn)
    # f32 rows keep their type
    knn32 = sc.KNNClassifier.fit(x.astype(np.float32), y, p)
    assert sc.KNNClassifier.deserialize(knn32.serialize()) == knn32
    with pytest.raises(sc.Failed):
        sc.KNNClassifier.deserialize(knn.serialize()[:-3])


def test_regressor_round_trip_without_device():
    x, y = _knn_data()
    reg = sc.KNNRegressor.fit(x, y.astype(np.float64) * 0.5,
                              sc.KNNClassifierParameters().with_k(3))
    back = sc.KNNRegressor.deserialize(reg.serialize())
    assert back == reg and not back.on_device()
    with pytest.raises(sc.Failed):
        sc.KNNClassifier.deserialize(reg.serialize())
