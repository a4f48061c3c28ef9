// smartcore_amd_io.hpp -- the data formats around the distance-bound path:
// the CSV reader that produces a DenseMatrix (smartcore v0.4.0
// src/readers/csv.rs, errors src/readers/error.rs) and the .xy dataset files
// (src/dataset/mod.rs:56-118).  Host-only code; it builds on the DenseMatrix
// of smartcore_amd.hpp.
#ifndef SMARTCORE_AMD_IO_HPP
#define SMARTCORE_AMD_IO_HPP

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "smartcore_amd.hpp"

namespace smartcore {
namespace readers {

// ReadingError (readers/error.rs:5-32)
enum class ReadingErrorKind {
  CouldNotReadFileSystem,
  NoRowsProvided,
  InvalidField,
  InvalidRow,
  ShapesDoNotMatch,
};

class ReadingError : public std::runtime_error {
 public:
  ReadingError(ReadingErrorKind kind, std::optional<std::string> msg)
      : std::runtime_error(msg ? *msg : std::string("NoRowsProvided")), kind_(kind), msg_(std::move(msg)) {}
  ReadingErrorKind kind() const { return kind_; }
  // ReadingError::message (error.rs:40-50): every kind but NoRowsProvided
  // carries one
  const std::optional<std::string>& message() const { return msg_; }
  bool operator==(const ReadingError& o) const { return kind_ == o.kind_ && msg_ == o.msg_; }

 private:
  ReadingErrorKind kind_;
  std::optional<std::string> msg_;
};

// CSVDefinition (csv.rs:26-41): one header row, fields separated by ",".
struct CSVDefinition {
  size_t n_rows_header = 1;
  std::string field_seperator = ",";
};

namespace detail {

// str::lines: split on '\n', drop one trailing '\r' per line, and no empty
// line after a final '\n'.
inline std::vector<std::string> lines(const std::string& text) {
  std::vector<std::string> out;
  size_t start = 0;
  while (start < text.size()) {
    size_t end = text.find('\n', start);
    const bool last = end == std::string::npos;
    if (last) end = text.size();
    std::string line = text.substr(start, end - start);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    out.push_back(std::move(line));
    start = end + 1;
    if (last) break;
  }
  return out;
}

// str::split(sep) (a non-empty separator): every separator occurrence cuts,
// so n occurrences give n + 1 fields, empty ones included.
inline std::vector<std::string> split(const std::string& s, const std::string& sep) {
  std::vector<std::string> out;
  size_t start = 0;
  for (;;) {
    const size_t at = sep.empty() ? std::string::npos : s.find(sep, start);
    if (at == std::string::npos) {
      out.push_back(s.substr(start));
      return out;
    }
    out.push_back(s.substr(start, at - start));
    start = at + sep.size();
  }
}

// str::trim (ASCII and Unicode whitespace; this reader's inputs are ASCII)
inline std::string trim(const std::string& s) {
  const char* ws = " \t\n\r\f\v";
  const size_t a = s.find_first_not_of(ws);
  if (a == std::string::npos) return std::string();
  const size_t b = s.find_last_not_of(ws);
  return s.substr(a, b - a + 1);
}

// str::parse::<f64 / f32>: an optional sign, then decimal digits with an
// optional point and exponent, or "inf" / "infinity" / "nan" (any case).
// The whole string must be consumed; hexadecimal forms are not accepted.
template <typename T>
inline bool parse_float(const std::string& s, T* out) {
  if (s.empty()) return false;
  size_t i = 0;
  if (s[0] == '+' || s[0] == '-') i = 1;
  std::string body = s.substr(i);
  std::string lower;
  for (char c : body) lower.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(c))));
  const bool neg = s[0] == '-';
  if (lower == "inf" || lower == "infinity") {
    *out = neg ? -std::numeric_limits<T>::infinity() : std::numeric_limits<T>::infinity();
    return true;
  }
  if (lower == "nan") {
    *out = std::numeric_limits<T>::quiet_NaN();
    return true;
  }
  // digits [. digits] [e [sign] digits], at least one digit in the mantissa
  size_t j = 0, mant = 0;
  while (j < body.size() && std::isdigit(static_cast<unsigned char>(body[j]))) ++j, ++mant;
  if (j < body.size() && body[j] == '.') {
    ++j;
    while (j < body.size() && std::isdigit(static_cast<unsigned char>(body[j]))) ++j, ++mant;
  }
  if (mant == 0) return false;
  if (j < body.size() && (body[j] == 'e' || body[j] == 'E')) {
    ++j;
    if (j < body.size() && (body[j] == '+' || body[j] == '-')) ++j;
    size_t ed = 0;
    while (j < body.size() && std::isdigit(static_cast<unsigned char>(body[j]))) ++j, ++ed;
    if (ed == 0) return false;
  }
  if (j != body.size()) return false;
  errno = 0;
  char* endp = nullptr;
  // strtod / strtof round correctly; an out-of-range value gives +-inf or
  // 0, as Rust's parse does
  if (sizeof(T) == sizeof(float))
    *out = static_cast<T>(std::strtof(s.c_str(), &endp));
  else
    *out = static_cast<T>(std::strtod(s.c_str(), &endp));
  return endp == s.c_str() + s.size();
}

}  // namespace detail

// matrix_from_csv_source (csv.rs:76-101) over the text of a CSV source: the
// row format is detected from the first row after the header; every row must
// have that many fields; the result is a row-major DenseMatrix
// (from_iterator with axis 0, matrix.rs:477-479).  Errors carry the
// reference's messages ("Value 'x' could not be read. Column: 1, Row: 1.",
// "3 fields found but expected 2, Row: 1.").
template <typename T>
DenseMatrix<T> matrix_from_csv_source(const std::string& csv_text, const CSVDefinition& def = {}) {
  const auto all = detail::lines(csv_text);
  // detect_row_format (csv.rs:59-73)
  if (all.size() <= def.n_rows_header) throw ReadingError(ReadingErrorKind::NoRowsProvided, std::nullopt);
  const size_t n_fields = detail::split(all[def.n_rows_header], def.field_seperator).size();
  std::vector<T> values;
  size_t nrows = 0;
  for (size_t li = def.n_rows_header; li < all.size(); ++li) {
    const size_t row_index = li - def.n_rows_header;
    const std::string row_suffix = ", Row: " + std::to_string(row_index) + ".";
    const auto fields = detail::split(all[li], def.field_seperator);
    // validate_csv_row (csv.rs:172-186), then enrich_reading_error (:191-203)
    if (fields.size() != n_fields)
      throw ReadingError(ReadingErrorKind::InvalidField,
                         std::to_string(fields.size()) + " fields found but expected " +
                             std::to_string(n_fields) + row_suffix);
    for (size_t f = 0; f < fields.size(); ++f) {
      const std::string v = detail::trim(fields[f]);
      T x{};
      // extract_value_from_csv_field (csv.rs:206-220), enriched twice
      if (!detail::parse_float<T>(v, &x))
        throw ReadingError(ReadingErrorKind::InvalidField,
                           "Value '" + v + "' could not be read. Column: " + std::to_string(f) +
                               row_suffix);
      values.push_back(x);
    }
    ++nrows;
  }
  return DenseMatrix<T>(nrows, n_fields, std::move(values), false);
}

// matrix_from_csv_source over a file; a file-system failure is
// CouldNotReadFileSystem with std::io::Error's text ("No such file or
// directory (os error 2)").
template <typename T>
DenseMatrix<T> matrix_from_csv_file(const std::string& path, const CSVDefinition& def = {}) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    const int e = errno;
    throw ReadingError(ReadingErrorKind::CouldNotReadFileSystem,
                       std::string(std::strerror(e)) + " (os error " + std::to_string(e) + ")");
  }
  std::stringstream ss;
  ss << f.rdbuf();
  return matrix_from_csv_source<T>(ss.str(), def);
}

}  // namespace readers

namespace dataset {

// Dataset<f32, f32> (dataset/mod.rs:20-36), the fields the .xy files hold.
struct Dataset {
  std::vector<float> data;    // num_samples x num_features, row-major
  std::vector<float> target;  // num_samples
  size_t num_samples = 0;
  size_t num_features = 0;

  // Dataset::as_matrix (mod.rs:39-52) as a row-major DenseMatrix.
  DenseMatrix<float> as_dense_matrix() const {
    return DenseMatrix<float>(num_samples, num_features, data, false);
  }
};

namespace detail {
inline uint64_t le64(const uint8_t* p) {
  uint64_t v = 0;
  for (int b = 7; b >= 0; --b) v = (v << 8) | p[b];
  return v;
}
inline float le_f32(const uint8_t* p) {
  uint32_t u = 0;
  for (int b = 3; b >= 0; --b) u = (u << 8) | p[b];
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
}  // namespace detail

// deserialize_data (mod.rs:86-118): num_features (u64 LE), num_samples
// (u64 LE), then num_samples * num_features f32 LE values and num_samples
// f32 LE targets.  A buffer shorter than its header says fails here (the
// reference would panic on the slice).
inline Dataset deserialize_data(const std::vector<uint8_t>& bytes) {
  if (bytes.size() < 16) throw Failed::input(".xy data shorter than its 16-byte header");
  Dataset ds;
  ds.num_features = static_cast<size_t>(detail::le64(bytes.data()));
  ds.num_samples = static_cast<size_t>(detail::le64(bytes.data() + 8));
  const size_t nx = ds.num_samples * ds.num_features;
  if (bytes.size() < 16 + 4 * (nx + ds.num_samples))
    throw Failed::input(".xy data shorter than its header's sizes");
  ds.data.resize(nx);
  ds.target.resize(ds.num_samples);
  size_t c = 16;
  for (size_t i = 0; i < nx; ++i, c += 4) ds.data[i] = detail::le_f32(bytes.data() + c);
  for (size_t i = 0; i < ds.num_samples; ++i, c += 4) ds.target[i] = detail::le_f32(bytes.data() + c);
  return ds;
}

// serialize_data (mod.rs:56-84): the same layout.
inline std::vector<uint8_t> serialize_data(const Dataset& ds) {
  std::vector<uint8_t> out;
  out.reserve(16 + 4 * (ds.data.size() + ds.target.size()));
  auto put64 = [&](uint64_t v) {
    for (int b = 0; b < 8; ++b) out.push_back(static_cast<uint8_t>(v >> (8 * b)));
  };
  auto put32 = [&](float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    for (int b = 0; b < 4; ++b) out.push_back(static_cast<uint8_t>(u >> (8 * b)));
  };
  put64(ds.num_features);
  put64(ds.num_samples);
  for (float v : ds.data) put32(v);
  for (float v : ds.target) put32(v);
  return out;
}

inline Dataset load_xy(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Failed::input("couldn't open " + path);
  std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  return deserialize_data(bytes);
}

inline void save_xy(const Dataset& ds, const std::string& path) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw Failed::input("couldn't create " + path);
  const auto bytes = serialize_data(ds);
  f.write(reinterpret_cast<const char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
}

}  // namespace dataset
}  // namespace smartcore

#endif
