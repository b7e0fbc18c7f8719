// tos_bio.hpp -- raw binary records in the reference's `bio` format (utils/binaryIO.h:70-111),
// header-only, so that a driver written against the reference keeps its record calls:
//   bio::write_vector(file, v)     the bare array, no header
//   bio::write_vectors(file, vs)   N equal-length arrays, component-interleaved
//   bio::read_vector(file, v)      floor(file size / sizeof(T)) items; false if unreadable
// Native byte order (little-endian on the hosts this library runs on).
#ifndef TOS_BIO_HPP
#define TOS_BIO_HPP

#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

namespace bio {

template <class T, class = typename std::enable_if<std::is_scalar<T>::value>::type>
bool write_vector(const std::string& file, const std::vector<T>& v) {
  std::FILE* fh = std::fopen(file.c_str(), "wb");
  if (!fh) return false;
  const size_t n = v.empty() ? 0 : std::fwrite(v.data(), sizeof(T), v.size(), fh);
  return std::fclose(fh) == 0 && n == v.size();
}

template <class T, int N, class = typename std::enable_if<std::is_scalar<T>::value>::type>
bool write_vectors(const std::string& file, const std::vector<T> (&vs)[N]) {
  const size_t n = vs[0].size();
  std::vector<T> inter(n * N);
  for (int j = 0; j < N; ++j) {
    if (vs[j].size() != n) return false;
    for (size_t i = 0; i < n; ++i) inter[i * N + j] = vs[j][i];
  }
  return write_vector(file, inter);
}

template <class T, class = typename std::enable_if<std::is_scalar<T>::value>::type>
bool read_vector(const std::string& file, std::vector<T>& v) {
  std::FILE* fh = std::fopen(file.c_str(), "rb");
  if (!fh) return false;
  std::fseek(fh, 0, SEEK_END);
  const long len = std::ftell(fh);
  std::fseek(fh, 0, SEEK_SET);
  v.assign(len > 0 ? (size_t)len / sizeof(T) : 0, T(0));
  const size_t got = v.empty() ? 0 : std::fread(v.data(), sizeof(T), v.size(), fh);
  std::fclose(fh);
  return got == v.size();
}

}  // namespace bio

#endif
