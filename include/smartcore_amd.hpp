// smartcore_amd.hpp -- C++ mirror of smartcore's estimator API for the
// distance-bound path, built on the C ABI in smartcore_amd.h.
//
// Names, parameters, defaults and error texts follow smartcore v0.4.0 (paths
// relative to its crate root); `Result<T, Failed>` becomes a thrown
// smartcore::Failed.  The device does the searching and the Lloyd steps;
// everything the reference does around them (label maps, votes, weights,
// centroid bookkeeping) is plain host code here, in the reference's order.
#ifndef SMARTCORE_AMD_HPP
#define SMARTCORE_AMD_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "smartcore_amd.h"

namespace smartcore {

// ---------------------------------------------------------------- errors
// src/error/mod.rs:11-37
enum class FailedError : int {
  FitFailed = 1,
  PredictFailed,
  TransformFailed,
  FindFailed,
  DecompositionFailed,
  SolutionFailed,
  ParametersError,
  InvalidStateError,
};

inline const char* to_string(FailedError e) {
  // src/error/mod.rs:108-121
  switch (e) {
    case FailedError::FitFailed: return "Fit failed";
    case FailedError::PredictFailed: return "Predict failed";
    case FailedError::TransformFailed: return "Transform failed";
    case FailedError::FindFailed: return "Find failed";
    case FailedError::DecompositionFailed: return "Decomposition failed";
    case FailedError::SolutionFailed: return "Can't find solution";
    case FailedError::ParametersError: return "Error in input, check parameters";
    case FailedError::InvalidStateError: return "Invalid state, this should never happen";
  }
  return "Unknown";
}

// Failed{err, msg}; what() is "{err}: {msg}" (error/mod.rs:124-127).
class Failed : public std::runtime_error {
 public:
  Failed(FailedError err, std::string msg)
      : std::runtime_error(std::string(to_string(err)) + ": " + msg), err_(err), msg_(std::move(msg)) {}
  static Failed fit(const std::string& m) { return Failed(FailedError::FitFailed, m); }
  static Failed predict(const std::string& m) { return Failed(FailedError::PredictFailed, m); }
  static Failed input(const std::string& m) { return Failed(FailedError::ParametersError, m); }
  static Failed because(FailedError e, const std::string& m) { return Failed(e, m); }
  FailedError error() const { return err_; }
  const std::string& message() const { return msg_; }
  bool operator==(const Failed& o) const { return err_ == o.err_ && msg_ == o.msg_; }

 private:
  FailedError err_;
  std::string msg_;
};

// A condition the reference panics on ("Input vector sizes are different.",
// "p must be at least 1"), or a device failure.
class Panic : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

namespace detail {
inline std::string last_error() {
  const size_t n = sc_last_error(nullptr, 0);
  std::string s(n + 1, '\0');
  sc_last_error(&s[0], s.size());
  s.resize(n);
  return s;
}
inline void check(int rc) {
  if (rc == SC_OK) return;
  const std::string m = last_error();
  if (rc >= 1 && rc <= 8) throw Failed(static_cast<FailedError>(rc), m);
  throw Panic(m);
}
template <typename T>
constexpr sc_dtype dtype_of() {
  static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                "device path supports f64 and f32 rows");
  return std::is_same<T, double>::value ? SC_F64 : SC_F32;
}

// Little-endian byte stream of the estimators' save / load (the analogue of
// the reference's serde derives, knn_classifier.rs:72-73, kmeans.rs:71-72).
// Layout: "SCAMD" | version u8 | kind u8 | the estimator's fields in
// declaration order; vectors are a u64 length and their elements.
class ByteWriter {
 public:
  void u8(uint8_t v) { b_.push_back(v); }
  void u16(uint16_t v) { raw(&v, 2); }
  void u32(uint32_t v) { raw(&v, 4); }
  void u64(uint64_t v) { raw(&v, 8); }
  void f64(double v) { raw(&v, 8); }
  void i64(int64_t v) { raw(&v, 8); }
  template <typename T>
  void num(T v) {
    raw(&v, sizeof(T));
  }
  template <typename T>
  void vec(const std::vector<T>& v) {
    u64(v.size());
    for (const T& x : v) num<T>(x);
  }
  void header(uint8_t kind) {
    for (char c : std::string("SCAMD")) u8(static_cast<uint8_t>(c));
    u8(1);
    u8(kind);
  }
  std::vector<uint8_t> take() { return std::move(b_); }

 private:
  // the hosts this library runs on are little-endian
  void raw(const void* p, size_t n) {
    const uint8_t* c = static_cast<const uint8_t*>(p);
    b_.insert(b_.end(), c, c + n);
  }
  std::vector<uint8_t> b_;
};

class ByteReader {
 public:
  explicit ByteReader(const std::vector<uint8_t>& b) : b_(b) {}
  uint8_t u8() { uint8_t v; raw(&v, 1); return v; }
  uint16_t u16() { uint16_t v; raw(&v, 2); return v; }
  uint32_t u32() { uint32_t v; raw(&v, 4); return v; }
  uint64_t u64() { uint64_t v; raw(&v, 8); return v; }
  double f64() { double v; raw(&v, 8); return v; }
  template <typename T>
  T num() {
    T v;
    raw(&v, sizeof(T));
    return v;
  }
  template <typename T>
  std::vector<T> vec() {
    const uint64_t n = u64();
    if (n > (b_.size() - at_) / std::max<size_t>(sizeof(T), 1)) throw_short();
    std::vector<T> v(static_cast<size_t>(n));
    for (auto& x : v) x = num<T>();
    return v;
  }
  void header(uint8_t kind);
  bool done() const { return at_ == b_.size(); }

 private:
  [[noreturn]] static void throw_short();
  void raw(void* p, size_t n) {
    if (n > b_.size() - at_) throw_short();
    std::copy(b_.begin() + at_, b_.begin() + at_ + n, static_cast<uint8_t*>(p));
    at_ += n;
  }
  const std::vector<uint8_t>& b_;
  size_t at_ = 0;
};
}  // namespace detail

namespace detail {
inline void ByteReader::throw_short() { throw Failed::input("serialized model is truncated"); }
inline void ByteReader::header(uint8_t kind) {
  std::string magic;
  for (int i = 0; i < 5; ++i) magic.push_back(static_cast<char>(u8()));
  if (magic != "SCAMD") throw Failed::input("not a serialized smartcore_amd model");
  if (u8() != 1) throw Failed::input("unsupported serialized model version");
  if (u8() != kind) throw Failed::input("serialized model is of another estimator");
}
}  // namespace detail

// ---------------------------------------------------------------- linalg
// DenseMatrix (src/linalg/basic/matrix.rs:27-32): column-major by default
// (from_2d_vec sets column_major = true, matrix.rs:215-237).
template <typename T>
class DenseMatrix {
 public:
  DenseMatrix() = default;
  DenseMatrix(size_t nrows, size_t ncols, std::vector<T> values, bool column_major)
      : nrows_(nrows), ncols_(ncols), values_(std::move(values)), column_major_(column_major) {
    if (values_.size() != nrows_ * ncols_) throw Failed::input("The specified shape does not match the values");
  }
  static DenseMatrix from_2d_vec(const std::vector<std::vector<T>>& rows) {
    if (rows.empty()) throw Failed::input("The 2d vec provided is empty.");
    const size_t nr = rows.size(), nc = rows[0].size();
    std::vector<T> v(nr * nc);
    for (size_t c = 0; c < nc; ++c)
      for (size_t r = 0; r < nr; ++r) {
        if (rows[r].size() != nc) throw Failed::input("Row sizes differ");
        v[c * nr + r] = rows[r][c];
      }
    return DenseMatrix(nr, nc, std::move(v), true);
  }
  // Row-major host block (the packing the device index uses).
  static DenseMatrix from_row_major(size_t nrows, size_t ncols, const T* data) {
    return DenseMatrix(nrows, ncols, std::vector<T>(data, data + nrows * ncols), false);
  }
  std::pair<size_t, size_t> shape() const { return {nrows_, ncols_}; }
  // matrix.rs:367-380
  const T& get(size_t r, size_t c) const {
    return column_major_ ? values_[c * nrows_ + r] : values_[r * ncols_ + c];
  }
  std::vector<T> get_row(size_t r) const {
    std::vector<T> row(ncols_);
    for (size_t c = 0; c < ncols_; ++c) row[c] = get(r, c);
    return row;
  }
  std::vector<T> row_major() const {
    if (!column_major_) return values_;
    std::vector<T> out(nrows_ * ncols_);
    for (size_t r = 0; r < nrows_; ++r)
      for (size_t c = 0; c < ncols_; ++c) out[r * ncols_ + c] = values_[c * nrows_ + r];
    return out;
  }

 private:
  size_t nrows_ = 0, ncols_ = 0;
  std::vector<T> values_;
  bool column_major_ = true;
};

// ---------------------------------------------------------------- distances
// trait Distance<T> (metrics/distance/mod.rs:38-41) with the seam the device
// path needs: a metric that maps onto a device kernel says which.
struct DeviceMetric {
  sc_metric metric;
  uint16_t p;
};

// euclidian.rs:51-72
template <typename T>
struct Euclidian {
  static double squared_distance(const std::vector<T>& x, const std::vector<T>& y) {
    if (x.size() != y.size()) throw Panic("Input vector sizes are different.");
    double sum = 0.0;
    for (size_t i = 0; i < x.size(); ++i) {
      const T r = static_cast<T>(x[i] - y[i]);
      sum += static_cast<double>(static_cast<T>(r * r));
    }
    return sum;
  }
  double distance(const std::vector<T>& x, const std::vector<T>& y) const {
    return std::sqrt(squared_distance(x, y));
  }
  std::optional<DeviceMetric> device_metric() const { return DeviceMetric{SC_EUCLIDIAN, 2}; }
};

// manhattan.rs:50-62
template <typename T>
struct Manhattan {
  double distance(const std::vector<T>& x, const std::vector<T>& y) const {
    if (x.size() != y.size()) throw Panic("Input vector sizes are different");
    double dist = 0.0;
    for (size_t i = 0; i < x.size(); ++i) dist += std::fabs(static_cast<double>(static_cast<T>(x[i] - y[i])));
    return dist;
  }
  std::optional<DeviceMetric> device_metric() const { return DeviceMetric{SC_MANHATTAN, 1}; }
};

// minkowski.rs:50-67
template <typename T>
struct Minkowski {
  uint16_t p;
  explicit Minkowski(uint16_t p_) : p(p_) {}
  double distance(const std::vector<T>& x, const std::vector<T>& y) const {
    if (x.size() != y.size()) throw Panic("Input vector sizes are different");
    if (p < 1) throw Panic("p must be at least 1");
    const double pt = p;
    double dist = 0.0;
    for (size_t i = 0; i < x.size(); ++i)
      dist += std::pow(std::fabs(static_cast<double>(static_cast<T>(x[i] - y[i]))), pt);
    return std::pow(dist, 1.0 / pt);
  }
  std::optional<DeviceMetric> device_metric() const { return DeviceMetric{SC_MINKOWSKI, p}; }
};

// Rebuilds a distance from what device_metric() reports (model load).
template <typename D>
struct DistanceFrom;
template <typename T>
struct DistanceFrom<Euclidian<T>> {
  static Euclidian<T> make(uint16_t) { return {}; }
};
template <typename T>
struct DistanceFrom<Manhattan<T>> {
  static Manhattan<T> make(uint16_t) { return {}; }
};
template <typename T>
struct DistanceFrom<Minkowski<T>> {
  static Minkowski<T> make(uint16_t p) { return Minkowski<T>(p); }
};

// metrics/distance/mod.rs:46-61
struct Distances {
  template <typename T>
  static Euclidian<T> euclidian() { return {}; }
  template <typename T>
  static Minkowski<T> minkowski(uint16_t p) { return Minkowski<T>(p); }
  template <typename T>
  static Manhattan<T> manhattan() { return {}; }
};

// ---------------------------------------------------------------- search
// KNNAlgorithmName (algorithm/neighbour/mod.rs:53-60).  CoverTree is the
// reference default; this build has no cover tree and serves it with the
// same device brute-force search (identical neighbour sets off exact ties;
// the output follows LinearSearch's heap-array order).
enum class KNNAlgorithmName { LinearSearch, CoverTree };

// Device precision form of the search (see sc_mode).
enum class SearchMode { PlainFp64 = SC_PLAIN_FP64, Mixed = SC_MIXED };

// LinearKNNSearch (linear_search.rs:36-108) over a device-resident copy of
// the rows.  The host rows stay alive next to it because find() returns
// references into them (linear_search.rs:82).
template <typename T, typename D>
class LinearKNNSearch {
 public:
  // linear_search.rs:45-47 (row-major rows, one vector per row).  The device
  // copy of the rows is made on the first search, so a model that was fitted
  // or loaded but never searched holds no device memory.
  static LinearKNNSearch create(std::vector<std::vector<T>> data, D distance) {
    return LinearKNNSearch(std::move(data), std::move(distance));
  }
  LinearKNNSearch(std::vector<std::vector<T>> data, D distance)
      : data_(std::move(data)), distance_(std::move(distance)), dev_(std::make_shared<Device>()) {
    dim_ = data_.empty() ? 0 : data_[0].size();
    for (const auto& r : data_)
      if (r.size() != dim_) throw Panic("Input vector sizes are different.");
  }

  using Neighbour = std::tuple<size_t, double, const std::vector<T>*>;

  // linear_search.rs:52-84 for one query
  std::vector<Neighbour> find(const std::vector<T>& from, size_t k) const {
    return find_batch({from}, k).at(0);
  }

  // The loop of KNNClassifier::predict batched into one device call.
  std::vector<std::vector<Neighbour>> find_batch(const std::vector<std::vector<T>>& from, size_t k,
                                                 SearchMode mode = SearchMode::PlainFp64) const {
    const auto dm = distance_.device_metric();
    if (!dm) throw Failed::because(FailedError::FindFailed, "metric has no device kernel");
    std::vector<T> q(from.size() * dim_);
    for (size_t i = 0; i < from.size(); ++i) {
      if (from[i].size() != dim_) throw Panic("Input vector sizes are different.");
      std::copy(from[i].begin(), from[i].end(), q.begin() + i * dim_);
    }
    std::vector<int64_t> idx(from.size() * std::max<size_t>(k, 1));
    std::vector<double> dist(idx.size());
    std::vector<uint32_t> cnt(std::max<size_t>(from.size(), 1));
    detail::check(sc_knn_search(index(), q.data(), from.size(), k, dm->metric, dm->p,
                                static_cast<sc_mode>(mode), idx.data(), dist.data(), cnt.data()));
    std::vector<std::vector<Neighbour>> out(from.size());
    for (size_t i = 0; i < from.size(); ++i) {
      out[i].reserve(cnt[i]);
      for (size_t j = 0; j < cnt[i]; ++j) {
        const size_t r = static_cast<size_t>(idx[i * k + j]);
        out[i].emplace_back(r, dist[i * k + j], &data_[r]);
      }
    }
    return out;
  }

  // linear_search.rs:89-108
  std::vector<Neighbour> find_radius(const std::vector<T>& from, double radius) const {
    const auto dm = distance_.device_metric();
    if (!dm) throw Failed::because(FailedError::FindFailed, "metric has no device kernel");
    if (from.size() != dim_) throw Panic("Input vector sizes are different.");
    sc_radius_result* r = nullptr;
    detail::check(sc_radius_search(index(), from.data(), 1, radius, dm->metric, dm->p, &r));
    const size_t total = sc_radius_result_total(r);
    std::vector<uint64_t> off(2);
    std::vector<int64_t> idx(std::max<size_t>(total, 1));
    std::vector<double> dist(idx.size());
    sc_radius_result_copy(r, off.data(), idx.data(), dist.data());
    sc_radius_result_free(r);
    std::vector<Neighbour> out;
    for (size_t j = 0; j < total; ++j)
      out.emplace_back(static_cast<size_t>(idx[j]), dist[j], &data_[static_cast<size_t>(idx[j])]);
    return out;
  }

  size_t len() const { return data_.size(); }
  // The device-resident rows (uploaded on first use).
  const sc_index* index() const {
    std::lock_guard<std::mutex> lk(dev_->mu);
    if (!dev_->ix) {
      std::vector<T> packed(data_.size() * dim_);
      for (size_t i = 0; i < data_.size(); ++i)
        std::copy(data_[i].begin(), data_[i].end(), packed.begin() + i * dim_);
      sc_index* ix = nullptr;
      detail::check(sc_index_create(packed.data(), detail::dtype_of<T>(), data_.size(), dim_, &ix));
      dev_->ix = ix;
    }
    return dev_->ix;
  }
  bool on_device() const {
    std::lock_guard<std::mutex> lk(dev_->mu);
    return dev_->ix != nullptr;
  }

  // The rows and the distance (serde's view of LinearKNNSearch,
  // linear_search.rs:34-39); the device copy is not part of it.
  void save(detail::ByteWriter& w) const {
    const auto dm = distance_.device_metric();
    w.u8(static_cast<uint8_t>(detail::dtype_of<T>()));
    w.u8(static_cast<uint8_t>(dm ? dm->metric : SC_EUCLIDIAN));
    w.u16(dm ? dm->p : 2);
    w.u64(data_.size());
    w.u64(dim_);
    for (const auto& r : data_)
      for (const T& v : r) w.num<T>(v);
  }
  static LinearKNNSearch load(detail::ByteReader& r) {
    if (r.u8() != static_cast<uint8_t>(detail::dtype_of<T>()))
      throw Failed::input("serialized rows are of another type");
    const auto metric = static_cast<sc_metric>(r.u8());
    const uint16_t p = r.u16();
    D dist = DistanceFrom<D>::make(p);
    const auto dm = dist.device_metric();
    if (!dm || dm->metric != metric) throw Failed::input("serialized model uses another distance");
    const uint64_t n = r.u64(), d = r.u64();
    std::vector<std::vector<T>> rows(static_cast<size_t>(n), std::vector<T>(static_cast<size_t>(d)));
    for (auto& row : rows)
      for (T& v : row) v = r.num<T>();
    return LinearKNNSearch(std::move(rows), std::move(dist));
  }
  bool operator==(const LinearKNNSearch& o) const {
    const auto a = distance_.device_metric(), b = o.distance_.device_metric();
    return data_ == o.data_ && a.has_value() == b.has_value() &&
           (!a || (a->metric == b->metric && a->p == b->p));
  }

 private:
  struct Device {
    std::mutex mu;
    sc_index* ix = nullptr;
    ~Device() {
      if (ix) sc_index_destroy(ix);
    }
  };
  std::vector<std::vector<T>> data_;
  D distance_;
  std::shared_ptr<Device> dev_;
  size_t dim_ = 0;
};

// ---------------------------------------------------------------- neighbors
// KNNWeightFunction (neighbors/mod.rs:53-78)
enum class KNNWeightFunction { Uniform, Distance };

inline std::vector<double> calc_weights(KNNWeightFunction w, const std::vector<double>& distances) {
  if (w == KNNWeightFunction::Distance) {
    const bool any_zero =
        std::any_of(distances.begin(), distances.end(), [](double e) { return e == 0.0; });
    std::vector<double> out(distances.size());
    for (size_t i = 0; i < distances.size(); ++i)
      out[i] = any_zero ? (distances[i] == 0.0 ? 1.0 : 0.0) : 1.0 / distances[i];
    return out;
  }
  return std::vector<double>(distances.size(), 1.0);
}

// KNNClassifierParameters (knn_classifier.rs:51-69, Default :144-154)
template <typename T, typename D = Euclidian<T>>
struct KNNClassifierParameters {
  D distance{};
  KNNAlgorithmName algorithm = KNNAlgorithmName::CoverTree;
  KNNWeightFunction weight = KNNWeightFunction::Uniform;
  size_t k = 3;
  SearchMode mode = SearchMode::PlainFp64;

  KNNClassifierParameters with_k(size_t k_) const { auto p = *this; p.k = k_; return p; }
  template <typename DD>
  KNNClassifierParameters<T, DD> with_distance(DD d) const {
    KNNClassifierParameters<T, DD> p{std::move(d)};
    p.algorithm = algorithm;
    p.weight = weight;
    p.k = k;
    p.mode = mode;
    return p;
  }
  KNNClassifierParameters with_algorithm(KNNAlgorithmName a) const { auto p = *this; p.algorithm = a; return p; }
  KNNClassifierParameters with_weight(KNNWeightFunction w) const { auto p = *this; p.weight = w; return p; }
  KNNClassifierParameters with_mode(SearchMode m) const { auto p = *this; p.mode = m; return p; }
};

template <typename T, typename D>
std::vector<std::vector<T>> rows_of(const DenseMatrix<T>& x) {
  const auto sh = x.shape();
  std::vector<std::vector<T>> rows(sh.first);
  for (size_t i = 0; i < sh.first; ++i) rows[i] = x.get_row(i);
  return rows;
}

// KNNClassifier (knn_classifier.rs:74-299)
template <typename TX, typename TY, typename D = Euclidian<TX>>
class KNNClassifier {
 public:
  using value_type = TX;
  static KNNClassifier fit(const DenseMatrix<TX>& x, const std::vector<TY>& y,
                           const KNNClassifierParameters<TX, D>& parameters) {
    const size_t y_n = y.size();
    const size_t x_n = x.shape().first;
    auto data = rows_of<TX, D>(x);
    // arrays.rs:223-231 unique (sort + dedup); position map knn_classifier.rs:229-235
    std::vector<TY> classes(y.begin(), y.end());
    std::sort(classes.begin(), classes.end());
    classes.erase(std::unique(classes.begin(), classes.end()), classes.end());
    std::vector<size_t> yi(y_n);
    for (size_t i = 0; i < y_n; ++i)
      yi[i] = static_cast<size_t>(std::find(classes.begin(), classes.end(), y[i]) - classes.begin());
    if (x_n != y_n)
      throw Failed::fit("Size of x should equal size of y; |x|=[" + std::to_string(x_n) +
                        "], |y|=[" + std::to_string(y_n) + "]");
    if (parameters.k <= 1) throw Failed::fit("k should be > 1, k=[" + std::to_string(parameters.k) + "]");
    KNNClassifier m;
    m.classes_ = std::move(classes);
    m.y_ = std::move(yi);
    m.k_ = parameters.k;
    m.weight_ = parameters.weight;
    m.mode_ = parameters.mode;
    m.search_ = std::make_shared<LinearKNNSearch<TX, D>>(std::move(data), parameters.distance);
    return m;
  }

  // knn_classifier.rs:265-277, all rows in one device call
  std::vector<TY> predict(const DenseMatrix<TX>& x) const {
    const auto rows = rows_of<TX, D>(x);
    const auto found = search_->find_batch(rows, k_, mode_);
    std::vector<TY> result(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) result[i] = classes_[predict_for_row(found[i])];
    return result;
  }

  const std::vector<TY>& classes() const { return classes_; }
  // Whether the rows have been uploaded (by a predict since fit or load).
  bool on_device() const { return search_->on_device(); }

  // Save / load (serde round trip, knn_classifier.rs:352-363): the classes,
  // the label map, k, the weight, the mode, the rows and the distance.  A
  // loaded model uploads its rows on its first predict.
  std::vector<uint8_t> serialize() const {
    detail::ByteWriter w;
    w.header(1);
    w.vec(classes_);
    w.vec(std::vector<uint64_t>(y_.begin(), y_.end()));
    w.u64(k_);
    w.u8(static_cast<uint8_t>(weight_));
    w.u8(static_cast<uint8_t>(mode_));
    search_->save(w);
    return w.take();
  }
  static KNNClassifier deserialize(const std::vector<uint8_t>& bytes) {
    detail::ByteReader r(bytes);
    r.header(1);
    KNNClassifier m;
    m.classes_ = r.vec<TY>();
    const auto y64 = r.vec<uint64_t>();
    m.y_.assign(y64.begin(), y64.end());
    m.k_ = static_cast<size_t>(r.u64());
    m.weight_ = static_cast<KNNWeightFunction>(r.u8());
    m.mode_ = static_cast<SearchMode>(r.u8());
    m.search_ = std::make_shared<LinearKNNSearch<TX, D>>(LinearKNNSearch<TX, D>::load(r));
    if (!r.done()) throw Failed::input("trailing bytes after the serialized model");
    return m;
  }
  bool operator==(const KNNClassifier& o) const {
    return classes_ == o.classes_ && y_ == o.y_ && k_ == o.k_ && weight_ == o.weight_ &&
           mode_ == o.mode_ && *search_ == *o.search_;
  }

 private:
  // knn_classifier.rs:279-299
  size_t predict_for_row(const std::vector<typename LinearKNNSearch<TX, D>::Neighbour>& sr) const {
    std::vector<double> d(sr.size());
    for (size_t i = 0; i < sr.size(); ++i) d[i] = std::get<1>(sr[i]);
    const auto weights = calc_weights(weight_, d);
    double w_sum = 0.0;
    for (double w : weights) w_sum += w;
    std::vector<double> c(classes_.size(), 0.0);
    double max_c = 0.0;
    size_t max_i = 0;
    for (size_t i = 0; i < sr.size(); ++i) {
      const size_t cls = y_[std::get<0>(sr[i])];
      c[cls] += weights[i] / w_sum;
      if (c[cls] > max_c) {
        max_c = c[cls];
        max_i = cls;
      }
    }
    return max_i;
  }

  std::vector<TY> classes_;
  std::vector<size_t> y_;
  size_t k_ = 0;
  KNNWeightFunction weight_ = KNNWeightFunction::Uniform;
  SearchMode mode_ = SearchMode::PlainFp64;
  std::shared_ptr<LinearKNNSearch<TX, D>> search_;
};

// KNNRegressorParameters (knn_regressor.rs, same fields and defaults)
template <typename T, typename D = Euclidian<T>>
using KNNRegressorParameters = KNNClassifierParameters<T, D>;

// KNNRegressor (knn_regressor.rs:212-281)
template <typename TX, typename TY, typename D = Euclidian<TX>>
class KNNRegressor {
 public:
  using value_type = TX;
  static KNNRegressor fit(const DenseMatrix<TX>& x, const std::vector<TY>& y,
                          const KNNRegressorParameters<TX, D>& parameters) {
    const size_t y_n = y.size();
    const size_t x_n = x.shape().first;
    auto data = rows_of<TX, D>(x);
    if (x_n != y_n)
      throw Failed::fit("Size of x should equal size of y; |x|=[" + std::to_string(x_n) +
                        "], |y|=[" + std::to_string(y_n) + "]");
    if (parameters.k < 1) throw Failed::fit("k should be > 0, k=[" + std::to_string(parameters.k) + "]");
    KNNRegressor m;
    m.y_ = y;
    m.k_ = parameters.k;
    m.weight_ = parameters.weight;
    m.mode_ = parameters.mode;
    m.search_ = std::make_shared<LinearKNNSearch<TX, D>>(std::move(data), parameters.distance);
    return m;
  }

  std::vector<TY> predict(const DenseMatrix<TX>& x) const {
    const auto rows = rows_of<TX, D>(x);
    const auto found = search_->find_batch(rows, k_, mode_);
    std::vector<TY> result(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) {
      // knn_regressor.rs:268-281
      const auto& sr = found[i];
      std::vector<double> d(sr.size());
      for (size_t j = 0; j < sr.size(); ++j) d[j] = std::get<1>(sr[j]);
      const auto weights = calc_weights(weight_, d);
      double w_sum = 0.0;
      for (double w : weights) w_sum += w;
      TY acc = TY(0);
      for (size_t j = 0; j < sr.size(); ++j)
        acc += y_[std::get<0>(sr[j])] * static_cast<TY>(weights[j] / w_sum);
      result[i] = acc;
    }
    return result;
  }

  bool on_device() const { return search_->on_device(); }
  // Save / load, as KNNClassifier's.
  std::vector<uint8_t> serialize() const {
    detail::ByteWriter w;
    w.header(2);
    w.vec(y_);
    w.u64(k_);
    w.u8(static_cast<uint8_t>(weight_));
    w.u8(static_cast<uint8_t>(mode_));
    search_->save(w);
    return w.take();
  }
  static KNNRegressor deserialize(const std::vector<uint8_t>& bytes) {
    detail::ByteReader r(bytes);
    r.header(2);
    KNNRegressor m;
    m.y_ = r.vec<TY>();
    m.k_ = static_cast<size_t>(r.u64());
    m.weight_ = static_cast<KNNWeightFunction>(r.u8());
    m.mode_ = static_cast<SearchMode>(r.u8());
    m.search_ = std::make_shared<LinearKNNSearch<TX, D>>(LinearKNNSearch<TX, D>::load(r));
    if (!r.done()) throw Failed::input("trailing bytes after the serialized model");
    return m;
  }
  bool operator==(const KNNRegressor& o) const {
    return y_ == o.y_ && k_ == o.k_ && weight_ == o.weight_ && mode_ == o.mode_ &&
           *search_ == *o.search_;
  }

 private:
  std::vector<TY> y_;
  size_t k_ = 0;
  KNNWeightFunction weight_ = KNNWeightFunction::Uniform;
  SearchMode mode_ = SearchMode::PlainFp64;
  std::shared_ptr<LinearKNNSearch<TX, D>> search_;
};

// ---------------------------------------------------------------- cluster
// KMeansParameters (kmeans.rs:112-146)
struct KMeansParameters {
  size_t k = 2;
  size_t max_iter = 100;
  std::optional<uint64_t> seed;
  SearchMode mode = SearchMode::PlainFp64;
  KMeansParameters with_k(size_t k_) const { auto p = *this; p.k = k_; return p; }
  KMeansParameters with_max_iter(size_t m) const { auto p = *this; p.max_iter = m; return p; }
  KMeansParameters with_mode(SearchMode m) const { auto p = *this; p.mode = m; return p; }
};

// KMeansSearchParameters and its grid iterator (kmeans.rs:151-232)
struct KMeansSearchParameters {
  std::vector<size_t> k{2};
  std::vector<size_t> max_iter{100};
  std::vector<std::optional<uint64_t>> seed{std::nullopt};

  std::vector<KMeansParameters> grid() const {
    std::vector<KMeansParameters> out;
    size_t ck = 0, cm = 0, cs = 0;
    for (;;) {
      if (ck == k.size() && cm == max_iter.size() && cs == seed.size()) break;
      KMeansParameters next;
      next.k = k[ck];
      next.max_iter = max_iter[cm];
      next.seed = seed[cs];
      if (ck + 1 < k.size()) {
        ck += 1;
      } else if (cm + 1 < max_iter.size()) {
        ck = 0;
        cm += 1;
      } else if (cs + 1 < seed.size()) {
        ck = 0;
        cm = 0;
        cs += 1;
      } else {
        ck += 1;
        cm += 1;
        cs += 1;
      }
      out.push_back(next);
    }
    return out;
  }
};

// KMeans (kmeans.rs:73-352)
template <typename TX, typename TY = uint32_t>
class KMeans {
 public:
  // kmeans.rs:254-323.  Initial labels come from k-means++ (kmeans.rs:354-413)
  // run on the device with a restatement of the reference's SmallRng;
  // initial centroids are their means in row order (kmeans.rs:275-292).
  static KMeans fit(const DenseMatrix<TX>& data, const KMeansParameters& parameters) {
    const auto sh = data.shape();
    const auto rows = data.row_major();
    if (parameters.k < 2) throw Failed::fit("invalid number of clusters: " + std::to_string(parameters.k));
    if (parameters.max_iter == 0)
      throw Failed::fit("invalid maximum number of iterations: " + std::to_string(parameters.max_iter));
    sc_index* ix = nullptr;
    detail::check(sc_index_create(rows.data(), detail::dtype_of<TX>(), sh.first, sh.second, &ix));
    std::unique_ptr<sc_index, int (*)(sc_index*)> guard(ix, sc_index_destroy);
    std::vector<uint32_t> y(sh.first);
    detail::check(sc_kmeans_plus_plus(ix, parameters.k, parameters.seed.has_value() ? 1 : 0,
                                      parameters.seed.value_or(0), y.data()));
    std::vector<double> centroids(parameters.k * sh.second, 0.0);
    std::vector<uint64_t> size(parameters.k, 0);
    for (size_t i = 0; i < sh.first; ++i) size[y[i]] += 1;
    for (size_t i = 0; i < sh.first; ++i)
      for (size_t j = 0; j < sh.second; ++j)
        centroids[y[i] * sh.second + j] += static_cast<double>(rows[i * sh.second + j]);
    for (size_t i = 0; i < parameters.k; ++i)
      for (size_t j = 0; j < sh.second; ++j) centroids[i * sh.second + j] /= static_cast<double>(size[i]);
    return fit_from(ix, sh.second, parameters, std::move(centroids));
  }

  // The Lloyd loop from given initial centroids (k x d, row-major).
  static KMeans fit_with_centroids(const DenseMatrix<TX>& data, const KMeansParameters& parameters,
                                   std::vector<double> centroids) {
    const auto sh = data.shape();
    const auto rows = data.row_major();
    if (parameters.k < 2) throw Failed::fit("invalid number of clusters: " + std::to_string(parameters.k));
    if (parameters.max_iter == 0)
      throw Failed::fit("invalid maximum number of iterations: " + std::to_string(parameters.max_iter));
    sc_index* ix = nullptr;
    detail::check(sc_index_create(rows.data(), detail::dtype_of<TX>(), sh.first, sh.second, &ix));
    std::unique_ptr<sc_index, int (*)(sc_index*)> guard(ix, sc_index_destroy);
    return fit_from(ix, sh.second, parameters, std::move(centroids));
  }

  // kmeans.rs:327-352
  std::vector<TY> predict(const DenseMatrix<TX>& x) const {
    const auto sh = x.shape();
    const auto rows = x.row_major();
    std::vector<uint32_t> lab(std::max<size_t>(sh.first, 1));
    detail::check(sc_kmeans_predict(rows.data(), detail::dtype_of<TX>(), sh.first, sh.second,
                                    centroids_.data(), k_, lab.data()));
    std::vector<TY> out(sh.first);
    for (size_t i = 0; i < sh.first; ++i) out[i] = static_cast<TY>(lab[i]);
    return out;
  }

  size_t k() const { return k_; }
  const std::vector<uint32_t>& labels() const { return y_; }
  const std::vector<uint64_t>& size() const { return size_; }
  const std::vector<double>& centroids() const { return centroids_; }
  double distortion() const { return distortion_; }
  uint32_t iterations() const { return iters_; }

  // Save / load (serde round trip, kmeans.rs:512-545): k, the labels, the
  // sizes, the distortion, the iteration count and the centroids.  KMeans
  // keeps nothing on the device between calls (predict uploads its rows).
  std::vector<uint8_t> serialize() const {
    detail::ByteWriter w;
    w.header(3);
    w.u8(static_cast<uint8_t>(detail::dtype_of<TX>()));
    w.u64(k_);
    w.vec(y_);
    w.vec(size_);
    w.f64(distortion_);
    w.u32(iters_);
    w.vec(centroids_);
    return w.take();
  }
  static KMeans deserialize(const std::vector<uint8_t>& bytes) {
    detail::ByteReader r(bytes);
    r.header(3);
    if (r.u8() != static_cast<uint8_t>(detail::dtype_of<TX>()))
      throw Failed::input("serialized rows are of another type");
    KMeans m;
    m.k_ = static_cast<size_t>(r.u64());
    m.y_ = r.vec<uint32_t>();
    m.size_ = r.vec<uint64_t>();
    m.distortion_ = r.f64();
    m.iters_ = r.u32();
    m.centroids_ = r.vec<double>();
    if (!r.done()) throw Failed::input("trailing bytes after the serialized model");
    if (m.size_.size() != m.k_ || (m.k_ && m.centroids_.size() % m.k_ != 0))
      throw Failed::input("serialized k-means model is inconsistent");
    return m;
  }
  bool operator==(const KMeans& o) const {
    return k_ == o.k_ && y_ == o.y_ && size_ == o.size_ && centroids_ == o.centroids_ &&
           (distortion_ == o.distortion_ || (std::isnan(distortion_) && std::isnan(o.distortion_))) &&
           iters_ == o.iters_;
  }

 private:
  static KMeans fit_from(sc_index* ix, size_t d, const KMeansParameters& p, std::vector<double> c) {
    if (c.size() != p.k * d) throw Failed::input("initial centroids must be k x d");
    KMeans m;
    m.k_ = p.k;
    m.y_.resize(std::max<size_t>(sc_index_rows(ix), 1));
    m.size_.resize(p.k);
    detail::check(sc_kmeans_run(ix, p.k, p.max_iter, static_cast<sc_mode>(p.mode), c.data(),
                                m.y_.data(), m.size_.data(), &m.distortion_, &m.iters_));
    m.y_.resize(sc_index_rows(ix));
    m.centroids_ = std::move(c);
    return m;
  }

  size_t k_ = 0;
  std::vector<uint32_t> y_;
  std::vector<uint64_t> size_;
  double distortion_ = 0.0;
  uint32_t iters_ = 0;
  std::vector<double> centroids_;
};

// DBSCANParameters (dbscan.rs:74-131, Default :245-255)
template <typename T, typename D = Euclidian<T>>
struct DBSCANParameters {
  D distance{};
  size_t min_samples = 5;
  double eps = 0.5;
  KNNAlgorithmName algorithm = KNNAlgorithmName::CoverTree;
  template <typename DD>
  DBSCANParameters<T, DD> with_distance(DD d) const {
    return DBSCANParameters<T, DD>{std::move(d), min_samples, eps, algorithm};
  }
  DBSCANParameters with_min_samples(size_t m) const { auto p = *this; p.min_samples = m; return p; }
  DBSCANParameters with_eps(double e) const { auto p = *this; p.eps = e; return p; }
  DBSCANParameters with_algorithm(KNNAlgorithmName a) const { auto p = *this; p.algorithm = a; return p; }
};

// DBSCANSearchParameters and its grid iterator (dbscan.rs:133-229): the
// distance varies fastest, then min_samples, eps and algorithm.
template <typename T, typename D = Euclidian<T>>
struct DBSCANSearchParameters {
  std::vector<D> distance{D{}};
  std::vector<size_t> min_samples{5};
  std::vector<double> eps{0.5};
  std::vector<KNNAlgorithmName> algorithm{KNNAlgorithmName::CoverTree};

  std::vector<DBSCANParameters<T, D>> grid() const {
    std::vector<DBSCANParameters<T, D>> out;
    size_t cd = 0, cm = 0, ce = 0, ca = 0;
    for (;;) {
      if (cd == distance.size() && cm == min_samples.size() && ce == eps.size() &&
          ca == algorithm.size())
        break;
      DBSCANParameters<T, D> next;
      next.distance = distance[cd];
      next.min_samples = min_samples[cm];
      next.eps = eps[ce];
      next.algorithm = algorithm[ca];
      if (cd + 1 < distance.size()) {
        cd += 1;
      } else if (cm + 1 < min_samples.size()) {
        cd = 0;
        cm += 1;
      } else if (ce + 1 < eps.size()) {
        cd = 0;
        cm = 0;
        ce += 1;
      } else if (ca + 1 < algorithm.size()) {
        cd = 0;
        cm = 0;
        ce = 0;
        ca += 1;
      } else {
        cd += 1;
        cm += 1;
        ce += 1;
        ca += 1;
      }
      out.push_back(next);
    }
    return out;
  }
};

// DBSCAN (dbscan.rs:59-397).  Neighbourhoods come from the device radius
// search (the LinearSearch result; CoverTree is served by the same search);
// the expansion loop runs in the reference's order inside sc_dbscan_fit.
template <typename TX, typename TY = int32_t, typename D = Euclidian<TX>>
class DBSCAN {
 public:
  using value_type = TX;
  static DBSCAN fit(const DenseMatrix<TX>& x, const DBSCANParameters<TX, D>& parameters) {
    const auto dm = parameters.distance.device_metric();
    if (!dm) throw Failed::because(FailedError::FindFailed, "metric has no device kernel");
    const auto sh = x.shape();
    const auto rows = x.row_major();
    sc_index* ix = nullptr;
    detail::check(sc_index_create(rows.data(), detail::dtype_of<TX>(), sh.first, sh.second, &ix));
    DBSCAN m;
    m.index_.reset(ix, [](sc_index* p) { sc_index_destroy(p); });
    m.labels_.resize(std::max<size_t>(sh.first, 1));
    detail::check(sc_dbscan_fit(ix, parameters.eps, parameters.min_samples, dm->metric, dm->p,
                                m.labels_.data(), &m.num_classes_));
    m.labels_.resize(sh.first);
    m.eps_ = parameters.eps;
    m.metric_ = dm->metric;
    m.p_ = dm->p;
    m.dim_ = sh.second;
    return m;
  }

  // dbscan.rs:358-397
  std::vector<TY> predict(const DenseMatrix<TX>& x) const {
    const auto sh = x.shape();
    if (sh.second != dim_) throw Panic("Input vector sizes are different.");
    const auto rows = x.row_major();
    std::vector<int32_t> out(std::max<size_t>(sh.first, 1));
    detail::check(sc_dbscan_predict(index_.get(), labels_.data(), num_classes_, eps_, metric_, p_,
                                    rows.data(), sh.first, out.data()));
    std::vector<TY> y(sh.first);
    for (size_t i = 0; i < sh.first; ++i) y[i] = static_cast<TY>(out[i]);
    return y;
  }

  const std::vector<int32_t>& cluster_labels() const { return labels_; }
  size_t num_classes() const { return num_classes_; }
  double eps() const { return eps_; }

 private:
  std::shared_ptr<sc_index> index_;
  std::vector<int32_t> labels_;
  size_t num_classes_ = 0;
  double eps_ = 0.0;
  sc_metric metric_ = SC_EUCLIDIAN;
  uint16_t p_ = 2;
  size_t dim_ = 0;
};

}  // namespace smartcore

#endif
