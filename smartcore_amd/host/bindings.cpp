// Python bindings of the C++ estimator mirror (include/smartcore_amd.hpp).
//
// Exposes smartcore's names -- KNNClassifier, KNNRegressor, KMeans, their
// parameter structs with the with_* builders, Distances, KNNAlgorithmName,
// KNNWeightFunction, LinearKNNSearch, KMeansSearchParameters -- for f64 and
// f32 rows, so the parity tests can read like the reference's own.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <variant>

#include "../../include/smartcore_amd.hpp"
#include "../../include/smartcore_amd_io.hpp"

namespace py = pybind11;
using namespace smartcore;

namespace {

struct PyDistance {
  int metric = SC_EUCLIDIAN;
  uint16_t p = 2;
};

struct PyKNNParams {
  PyDistance distance{};
  KNNAlgorithmName algorithm = KNNAlgorithmName::CoverTree;
  KNNWeightFunction weight = KNNWeightFunction::Uniform;
  size_t k = 3;
  SearchMode mode = SearchMode::PlainFp64;
};

template <typename T>
DenseMatrix<T> to_matrix(py::array_t<T, py::array::c_style | py::array::forcecast> a) {
  if (a.ndim() != 2) throw Failed::input("x must be a 2-D array");
  return DenseMatrix<T>::from_row_major(a.shape(0), a.shape(1), a.data());
}

template <typename T, typename D>
KNNClassifierParameters<T, D> params_for(const PyKNNParams& p, D d) {
  KNNClassifierParameters<T, D> out{std::move(d)};
  out.algorithm = p.algorithm;
  out.weight = p.weight;
  out.k = p.k;
  out.mode = p.mode;
  return out;
}

bool is_f32(const py::array& a) { return py::dtype::of<float>().is(a.dtype()); }

template <template <typename, typename, typename> class Est, typename TY>
class PyKNN {
 public:
  using V = std::variant<Est<double, TY, Euclidian<double>>, Est<double, TY, Manhattan<double>>,
                         Est<double, TY, Minkowski<double>>, Est<float, TY, Euclidian<float>>,
                         Est<float, TY, Manhattan<float>>, Est<float, TY, Minkowski<float>>>;

  static PyKNN fit(py::array x, py::array_t<TY, py::array::c_style | py::array::forcecast> y,
                   const PyKNNParams& p) {
    std::vector<TY> yv(y.data(), y.data() + y.size());
    if (is_f32(x)) return PyKNN(fit_t<float>(x, yv, p), true);
    return PyKNN(fit_t<double>(x, yv, p), false);
  }

  py::bytes serialize() const {
    const auto b = std::visit([](const auto& m) { return m.serialize(); }, m_);
    return py::bytes(reinterpret_cast<const char*>(b.data()), b.size());
  }
  // The element type and distance are read from the bytes: every
  // alternative is tried and the one whose layout matches is kept.
  static PyKNN deserialize(py::bytes data) {
    const std::string s = data;
    const std::vector<uint8_t> b(s.begin(), s.end());
    return try_load<0>(b);
  }
  bool equals(const PyKNN& o) const { return m_ == o.m_; }
  bool on_device() const {
    return std::visit([](const auto& m) { return m.on_device(); }, m_);
  }

  py::array_t<TY> predict(py::array x) const {
    std::vector<TY> out = std::visit(
        [&](const auto& m) -> std::vector<TY> {
          using TX = typename std::decay_t<decltype(m)>::value_type;
          auto a = py::array_t<TX, py::array::c_style | py::array::forcecast>::ensure(x);
          return m.predict(to_matrix<TX>(a));
        },
        m_);
    py::array_t<TY> r(out.size());
    std::copy(out.begin(), out.end(), r.mutable_data());
    return r;
  }

 private:
  PyKNN(V m, bool f32) : m_(std::move(m)), f32_(f32) {}

  template <size_t I>
  static PyKNN try_load(const std::vector<uint8_t>& b) {
    using M = std::variant_alternative_t<I, V>;
    try {
      return PyKNN(V(std::in_place_index<I>, M::deserialize(b)),
                   std::is_same<typename M::value_type, float>::value);
    } catch (const Failed&) {
      if constexpr (I + 1 < std::variant_size_v<V>)
        return try_load<I + 1>(b);
      else
        throw;
    }
  }

  template <typename T>
  static V fit_t(py::array x, const std::vector<TY>& y, const PyKNNParams& p) {
    auto mx = to_matrix<T>(py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(x));
    switch (p.distance.metric) {
      case SC_EUCLIDIAN:
        return Est<T, TY, Euclidian<T>>::fit(mx, y, params_for<T>(p, Euclidian<T>{}));
      case SC_MANHATTAN:
        return Est<T, TY, Manhattan<T>>::fit(mx, y, params_for<T>(p, Manhattan<T>{}));
      default:
        return Est<T, TY, Minkowski<T>>::fit(mx, y, params_for<T>(p, Minkowski<T>(p.distance.p)));
    }
  }

  V m_;
  bool f32_;
};

struct PyDBSCANParams {
  PyDistance distance{};
  size_t min_samples = 5;
  double eps = 0.5;
  KNNAlgorithmName algorithm = KNNAlgorithmName::CoverTree;
};

class PyDBSCAN {
 public:
  using V = std::variant<DBSCAN<double, int32_t, Euclidian<double>>,
                         DBSCAN<double, int32_t, Manhattan<double>>,
                         DBSCAN<double, int32_t, Minkowski<double>>,
                         DBSCAN<float, int32_t, Euclidian<float>>,
                         DBSCAN<float, int32_t, Manhattan<float>>,
                         DBSCAN<float, int32_t, Minkowski<float>>>;

  static PyDBSCAN fit(py::array x, const PyDBSCANParams& p) {
    if (is_f32(x)) return PyDBSCAN(fit_t<float>(x, p), true);
    return PyDBSCAN(fit_t<double>(x, p), false);
  }
  py::array_t<int32_t> predict(py::array x) const {
    std::vector<int32_t> out = std::visit(
        [&](const auto& m) -> std::vector<int32_t> {
          if (f32_) return predict_t<float>(m, x);
          return predict_t<double>(m, x);
        },
        m_);
    py::array_t<int32_t> r(out.size());
    std::copy(out.begin(), out.end(), r.mutable_data());
    return r;
  }
  py::array_t<int32_t> cluster_labels() const {
    return std::visit(
        [](const auto& m) {
          const auto& v = m.cluster_labels();
          py::array_t<int32_t> r(v.size());
          std::copy(v.begin(), v.end(), r.mutable_data());
          return r;
        },
        m_);
  }
  size_t num_classes() const {
    return std::visit([](const auto& m) { return m.num_classes(); }, m_);
  }

 private:
  PyDBSCAN(V m, bool f32) : m_(std::move(m)), f32_(f32) {}

  template <typename T, typename M>
  static std::vector<int32_t> predict_t(const M& m, py::array x) {
    if constexpr (std::is_same<typename M::value_type, T>::value)
      return m.predict(to_matrix<T>(py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(x)));
    else
      return {};
  }

  template <typename T>
  static V fit_t(py::array x, const PyDBSCANParams& p) {
    auto mx = to_matrix<T>(py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(x));
    auto base = DBSCANParameters<T>{}.with_min_samples(p.min_samples).with_eps(p.eps)
                    .with_algorithm(p.algorithm);
    switch (p.distance.metric) {
      case SC_EUCLIDIAN:
        return DBSCAN<T, int32_t, Euclidian<T>>::fit(mx, base);
      case SC_MANHATTAN:
        return DBSCAN<T, int32_t, Manhattan<T>>::fit(mx, base.with_distance(Manhattan<T>{}));
      default:
        return DBSCAN<T, int32_t, Minkowski<T>>::fit(mx, base.with_distance(Minkowski<T>(p.distance.p)));
    }
  }

  V m_;
  bool f32_;
};

class PyKMeans {
 public:
  using V = std::variant<KMeans<double, uint32_t>, KMeans<float, uint32_t>>;

  static PyKMeans fit(py::array x, const KMeansParameters& p) {
    if (is_f32(x))
      return PyKMeans(KMeans<float, uint32_t>::fit(to_matrix<float>(ensure<float>(x)), p));
    return PyKMeans(KMeans<double, uint32_t>::fit(to_matrix<double>(ensure<double>(x)), p));
  }
  static PyKMeans fit_with_centroids(py::array x, const KMeansParameters& p,
                                     py::array_t<double, py::array::c_style | py::array::forcecast> c) {
    std::vector<double> cv(c.data(), c.data() + c.size());
    if (is_f32(x))
      return PyKMeans(KMeans<float, uint32_t>::fit_with_centroids(to_matrix<float>(ensure<float>(x)), p, cv));
    return PyKMeans(KMeans<double, uint32_t>::fit_with_centroids(to_matrix<double>(ensure<double>(x)), p, cv));
  }
  py::array_t<uint32_t> predict(py::array x) const {
    std::vector<uint32_t> out = std::visit(
        [&](const auto& m) {
          using M = std::decay_t<decltype(m)>;
          if constexpr (std::is_same<M, KMeans<float, uint32_t>>::value)
            return m.predict(to_matrix<float>(ensure<float>(x)));
          else
            return m.predict(to_matrix<double>(ensure<double>(x)));
        },
        m_);
    return vec(out);
  }
  py::array_t<uint32_t> labels() const {
    return std::visit([](const auto& m) { return vec(m.labels()); }, m_);
  }
  py::array_t<uint64_t> size() const {
    return std::visit([](const auto& m) { return vec(m.size()); }, m_);
  }
  py::array_t<double> centroids() const {
    return std::visit(
        [](const auto& m) {
          auto v = vec(m.centroids());
          v.resize({static_cast<py::ssize_t>(m.k()), static_cast<py::ssize_t>(m.centroids().size() / m.k())});
          return v;
        },
        m_);
  }
  double distortion() const {
    return std::visit([](const auto& m) { return m.distortion(); }, m_);
  }
  py::bytes serialize() const {
    const auto b = std::visit([](const auto& m) { return m.serialize(); }, m_);
    return py::bytes(reinterpret_cast<const char*>(b.data()), b.size());
  }
  static PyKMeans deserialize(py::bytes data) {
    const std::string s = data;
    const std::vector<uint8_t> b(s.begin(), s.end());
    try {
      return PyKMeans(KMeans<double, uint32_t>::deserialize(b));
    } catch (const Failed&) {
      return PyKMeans(KMeans<float, uint32_t>::deserialize(b));
    }
  }
  bool equals(const PyKMeans& o) const { return m_ == o.m_; }
  uint32_t iterations() const {
    return std::visit([](const auto& m) { return m.iterations(); }, m_);
  }

 private:
  explicit PyKMeans(V m) : m_(std::move(m)) {}
  template <typename T>
  static py::array_t<T, py::array::c_style | py::array::forcecast> ensure(py::array x) {
    return py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(x);
  }
  template <typename T>
  static py::array_t<T> vec(const std::vector<T>& v) {
    py::array_t<T> r(v.size());
    std::copy(v.begin(), v.end(), r.mutable_data());
    return r;
  }
  V m_;
};

// LinearKNNSearch over f64 or f32 rows with a chosen metric.
class PyLinearKNNSearch {
 public:
  PyLinearKNNSearch(py::array data, const PyDistance& dist) {
    if (is_f32(data)) {
      init<float>(data, dist);
    } else {
      init<double>(data, dist);
    }
  }
  py::list find(py::array from, size_t k) const { return f_find(from, k); }
  py::list find_radius(py::array from, double radius) const { return f_radius(from, radius); }

 private:
  template <typename T>
  void init(py::array data, const PyDistance& dist) {
    auto a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(data);
    const size_t n = a.shape(0), d = a.shape(1);
    std::vector<std::vector<T>> rows(n, std::vector<T>(d));
    for (size_t i = 0; i < n; ++i) std::copy(a.data() + i * d, a.data() + (i + 1) * d, rows[i].begin());
    auto bind = [&](auto metric) {
      using D = decltype(metric);
      auto s = std::make_shared<LinearKNNSearch<T, D>>(rows, metric);
      f_find = [s](py::array from, size_t k) {
        auto q = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(from);
        std::vector<T> qv(q.data(), q.data() + q.size());
        py::list out;
        for (auto& t : s->find(qv, k)) out.append(py::make_tuple(std::get<0>(t), std::get<1>(t)));
        return out;
      };
      f_radius = [s](py::array from, double radius) {
        auto q = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(from);
        std::vector<T> qv(q.data(), q.data() + q.size());
        py::list out;
        for (auto& t : s->find_radius(qv, radius)) out.append(py::make_tuple(std::get<0>(t), std::get<1>(t)));
        return out;
      };
    };
    if (dist.metric == SC_EUCLIDIAN)
      bind(Euclidian<T>{});
    else if (dist.metric == SC_MANHATTAN)
      bind(Manhattan<T>{});
    else
      bind(Minkowski<T>(dist.p));
  }
  std::function<py::list(py::array, size_t)> f_find;
  std::function<py::list(py::array, double)> f_radius;
};

}  // namespace

PYBIND11_MODULE(smartcore, m) {
  m.doc() = "smartcore's kNN / k-means estimator API on the MI355X C ABI";
  static py::exception<Failed> failed_exc(m, "Failed");
  static py::exception<Panic> panic_exc(m, "Panic");
  py::register_exception_translator([](std::exception_ptr p) {
    try {
      if (p) std::rethrow_exception(p);
    } catch (const Failed& e) {
      py::object err = failed_exc;
      py::object inst = err(e.what());
      inst.attr("code") = static_cast<int>(e.error());
      inst.attr("msg") = e.message();
      PyErr_SetObject(failed_exc.ptr(), inst.ptr());
    } catch (const Panic& e) {
      PyErr_SetString(panic_exc.ptr(), e.what());
    }
  });

  py::class_<PyDistance>(m, "Distance").def_readonly("metric", &PyDistance::metric).def_readonly("p", &PyDistance::p);
  m.def("Euclidian", []() { return PyDistance{SC_EUCLIDIAN, 2}; });
  m.def("Manhattan", []() { return PyDistance{SC_MANHATTAN, 1}; });
  m.def("Minkowski", [](uint16_t p) { return PyDistance{SC_MINKOWSKI, p}; });
  py::class_<Distances>(m, "Distances")
      .def_static("euclidian", []() { return PyDistance{SC_EUCLIDIAN, 2}; })
      .def_static("manhattan", []() { return PyDistance{SC_MANHATTAN, 1}; })
      .def_static("minkowski", [](uint16_t p) { return PyDistance{SC_MINKOWSKI, p}; });

  py::enum_<KNNAlgorithmName>(m, "KNNAlgorithmName")
      .value("LinearSearch", KNNAlgorithmName::LinearSearch)
      .value("CoverTree", KNNAlgorithmName::CoverTree);
  py::enum_<KNNWeightFunction>(m, "KNNWeightFunction")
      .value("Uniform", KNNWeightFunction::Uniform)
      .value("Distance", KNNWeightFunction::Distance);
  py::enum_<SearchMode>(m, "SearchMode")
      .value("PlainFp64", SearchMode::PlainFp64)
      .value("Mixed", SearchMode::Mixed);

  auto params = [&](const char* name) {
    py::class_<PyKNNParams>(m, name)
        .def(py::init<>())
        .def_readwrite("distance", &PyKNNParams::distance)
        .def_readwrite("algorithm", &PyKNNParams::algorithm)
        .def_readwrite("weight", &PyKNNParams::weight)
        .def_readwrite("k", &PyKNNParams::k)
        .def_readwrite("mode", &PyKNNParams::mode)
        .def("with_k", [](PyKNNParams p, size_t k) { p.k = k; return p; })
        .def("with_distance", [](PyKNNParams p, PyDistance d) { p.distance = d; return p; })
        .def("with_algorithm", [](PyKNNParams p, KNNAlgorithmName a) { p.algorithm = a; return p; })
        .def("with_weight", [](PyKNNParams p, KNNWeightFunction w) { p.weight = w; return p; })
        .def("with_mode", [](PyKNNParams p, SearchMode md) { p.mode = md; return p; });
  };
  params("KNNClassifierParameters");
  m.attr("KNNRegressorParameters") = m.attr("KNNClassifierParameters");

  using Cls = PyKNN<KNNClassifier, int32_t>;
  py::class_<Cls>(m, "KNNClassifier")
      .def_static("fit", &Cls::fit, py::arg("x"), py::arg("y"),
                  py::arg("parameters") = PyKNNParams{})
      .def("predict", &Cls::predict)
      .def("serialize", &Cls::serialize)
      .def_static("deserialize", &Cls::deserialize)
      .def("on_device", &Cls::on_device)
      .def("__eq__", &Cls::equals);
  using Reg = PyKNN<KNNRegressor, double>;
  py::class_<Reg>(m, "KNNRegressor")
      .def_static("fit", &Reg::fit, py::arg("x"), py::arg("y"),
                  py::arg("parameters") = PyKNNParams{})
      .def("predict", &Reg::predict)
      .def("serialize", &Reg::serialize)
      .def_static("deserialize", &Reg::deserialize)
      .def("on_device", &Reg::on_device)
      .def("__eq__", &Reg::equals);

  py::class_<KMeansParameters>(m, "KMeansParameters")
      .def(py::init<>())
      .def_readwrite("k", &KMeansParameters::k)
      .def_readwrite("max_iter", &KMeansParameters::max_iter)
      .def_readwrite("seed", &KMeansParameters::seed)
      .def_readwrite("mode", &KMeansParameters::mode)
      .def("with_k", &KMeansParameters::with_k)
      .def("with_max_iter", &KMeansParameters::with_max_iter)
      .def("with_mode", &KMeansParameters::with_mode);
  py::class_<KMeansSearchParameters>(m, "KMeansSearchParameters")
      .def(py::init<>())
      .def_readwrite("k", &KMeansSearchParameters::k)
      .def_readwrite("max_iter", &KMeansSearchParameters::max_iter)
      .def_readwrite("seed", &KMeansSearchParameters::seed)
      .def("__iter__", [](const KMeansSearchParameters& s) {
        return py::iter(py::cast(s.grid()));
      });
  py::class_<PyKMeans>(m, "KMeans")
      .def_static("fit", &PyKMeans::fit, py::arg("x"), py::arg("parameters") = KMeansParameters{})
      .def_static("fit_with_centroids", &PyKMeans::fit_with_centroids)
      .def("predict", &PyKMeans::predict)
      .def_property_readonly("labels", &PyKMeans::labels)
      .def_property_readonly("size", &PyKMeans::size)
      .def_property_readonly("centroids", &PyKMeans::centroids)
      .def_property_readonly("distortion", &PyKMeans::distortion)
      .def_property_readonly("iterations", &PyKMeans::iterations)
      .def("serialize", &PyKMeans::serialize)
      .def_static("deserialize", &PyKMeans::deserialize)
      .def("__eq__", &PyKMeans::equals);

  // readers (src/readers/csv.rs) and the .xy datasets (src/dataset/mod.rs)
  static py::exception<readers::ReadingError> reading_exc(m, "ReadingError");
  py::register_exception_translator([](std::exception_ptr p) {
    try {
      if (p) std::rethrow_exception(p);
    } catch (const readers::ReadingError& e) {
      static const char* kinds[] = {"CouldNotReadFileSystem", "NoRowsProvided", "InvalidField",
                                    "InvalidRow", "ShapesDoNotMatch"};
      py::object err = reading_exc;
      py::object inst = err(e.what());
      inst.attr("kind") = kinds[static_cast<int>(e.kind())];
      inst.attr("msg") = e.message() ? py::cast(*e.message()) : py::none();
      PyErr_SetObject(reading_exc.ptr(), inst.ptr());
    }
  });
  m.def(
      "matrix_from_csv_source",
      [](const std::string& text, size_t n_rows_header, const std::string& sep, bool f32) -> py::array {
        readers::CSVDefinition def{n_rows_header, sep};
        auto out = [](auto mx) {
          using T = typename std::decay_t<decltype(mx.row_major())>::value_type;
          const auto sh = mx.shape();
          py::array_t<T> a({static_cast<py::ssize_t>(sh.first), static_cast<py::ssize_t>(sh.second)});
          const auto v = mx.row_major();
          std::copy(v.begin(), v.end(), a.mutable_data());
          return py::array(a);
        };
        if (f32) return out(readers::matrix_from_csv_source<float>(text, def));
        return out(readers::matrix_from_csv_source<double>(text, def));
      },
      py::arg("text"), py::arg("n_rows_header") = 1, py::arg("field_seperator") = ",",
      py::arg("f32") = false);
  m.def("deserialize_data", [](py::bytes data) {
    const std::string s = data;
    const auto ds = dataset::deserialize_data(std::vector<uint8_t>(s.begin(), s.end()));
    py::array_t<float> x({static_cast<py::ssize_t>(ds.num_samples), static_cast<py::ssize_t>(ds.num_features)});
    std::copy(ds.data.begin(), ds.data.end(), x.mutable_data());
    py::array_t<float> y(ds.target.size());
    std::copy(ds.target.begin(), ds.target.end(), y.mutable_data());
    return py::make_tuple(x, y);
  });
  m.def("serialize_data", [](py::array_t<float, py::array::c_style | py::array::forcecast> x,
                             py::array_t<float, py::array::c_style | py::array::forcecast> y) {
    if (x.ndim() != 2 || y.ndim() != 1 || y.shape(0) != x.shape(0))
      throw Failed::input("x must be n x f and y of length n");
    dataset::Dataset ds;
    ds.num_samples = x.shape(0);
    ds.num_features = x.shape(1);
    ds.data.assign(x.data(), x.data() + x.size());
    ds.target.assign(y.data(), y.data() + y.size());
    const auto b = dataset::serialize_data(ds);
    return py::bytes(reinterpret_cast<const char*>(b.data()), b.size());
  });

  py::class_<PyDBSCANParams>(m, "DBSCANParameters")
      .def(py::init<>())
      .def_readwrite("distance", &PyDBSCANParams::distance)
      .def_readwrite("min_samples", &PyDBSCANParams::min_samples)
      .def_readwrite("eps", &PyDBSCANParams::eps)
      .def_readwrite("algorithm", &PyDBSCANParams::algorithm)
      .def("with_distance", [](PyDBSCANParams p, PyDistance d) { p.distance = d; return p; })
      .def("with_min_samples", [](PyDBSCANParams p, size_t v) { p.min_samples = v; return p; })
      .def("with_eps", [](PyDBSCANParams p, double v) { p.eps = v; return p; })
      .def("with_algorithm", [](PyDBSCANParams p, KNNAlgorithmName a) { p.algorithm = a; return p; });
  using DSP = DBSCANSearchParameters<double, PyDistance>;
  py::class_<DSP>(m, "DBSCANSearchParameters")
      .def(py::init<>())
      .def_readwrite("distance", &DSP::distance)
      .def_readwrite("min_samples", &DSP::min_samples)
      .def_readwrite("eps", &DSP::eps)
      .def_readwrite("algorithm", &DSP::algorithm)
      .def("__iter__", [](const DSP& s) {
        std::vector<PyDBSCANParams> out;
        for (const auto& g : s.grid()) out.push_back(PyDBSCANParams{g.distance, g.min_samples, g.eps, g.algorithm});
        return py::iter(py::cast(out));
      });
  py::class_<PyDBSCAN>(m, "DBSCAN")
      .def_static("fit", &PyDBSCAN::fit, py::arg("x"), py::arg("parameters") = PyDBSCANParams{})
      .def("predict", &PyDBSCAN::predict)
      .def_property_readonly("cluster_labels", &PyDBSCAN::cluster_labels)
      .def_property_readonly("num_classes", &PyDBSCAN::num_classes);

  py::class_<PyLinearKNNSearch>(m, "LinearKNNSearch")
      .def(py::init<py::array, const PyDistance&>(), py::arg("data"),
           py::arg("distance") = PyDistance{})
      .def("find", &PyLinearKNNSearch::find)
      .def("find_radius", &PyLinearKNNSearch::find_radius);
}
