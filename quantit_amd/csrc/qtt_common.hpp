// Common types of the MI355X block-sparse tensor library: status codes,
// error classes, HIP checks and the packed conserved-charge type.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace qtt {

// Status codes returned across the C ABI (mirrors the exception types the
// reference throws; see include/quantit_amd.h).
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,  // std::invalid_argument
  kOutOfRange = 2,      // std::out_of_range
  kDomainError = 3,      // std::domain_error
  kLogicError = 4,       // std::logic_error (NaN in the 2x2 eigensolve)
  kTorchCheck = 5,       // c10::Error raised by TORCH_CHECK in the reference
  kOutOfMemory = 6,     // device arena exhausted
  kHipError = 7,         // HIP runtime failure
  kRuntimeError = 8,     // anything else
};

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};
struct InvalidArgument : Error {
  explicit InvalidArgument(const std::string& m) : Error(kInvalidArgument, m) {}
};
struct OutOfRange : Error {
  explicit OutOfRange(const std::string& m) : Error(kOutOfRange, m) {}
};
struct DomainError : Error {
  explicit DomainError(const std::string& m) : Error(kDomainError, m) {}
};
struct LogicError : Error {
  explicit LogicError(const std::string& m) : Error(kLogicError, m) {}
};
struct CheckError : Error {
  explicit CheckError(const std::string& m) : Error(kTorchCheck, m) {}
};
struct OutOfMemory : Error {
  explicit OutOfMemory(const std::string& m) : Error(kOutOfMemory, m) {}
};

#define QTT_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t qtt_e_ = (call);                                                          \
    if (qtt_e_ != hipSuccess)                                                            \
      throw ::qtt::Error(::qtt::kHipError, std::string(#call) + ": " +             \
                                                   hipGetErrorString(qtt_e_));           \
  } while (0)

// ---------------------------------------------------------------------------
// Conserved charges.  A charge is a tuple of up to kMaxGroups group elements
// stored as int16: Z components add with int16 wrap-around
// (include/Conserved/quantity.h:175-239), C<N> components add mod N
// (quantity.h:62-165).  The order is the lexicographic tuple order
// (Composite/quantity_impl.h:326); C<N> values live in [0,N) so a signed
// compare gives the same order as the reference's uint16 compare.
// ---------------------------------------------------------------------------
constexpr int kMaxGroups = 4;
constexpr int kMaxRank = 8;

struct Groups {
  int n = 1;                                   // number of group components
  std::array<uint16_t, kMaxGroups> mod{};      // 0 = Z, N = C<N>
  bool operator==(const Groups& o) const {
    if (n != o.n) return false;
    for (int i = 0; i < n; ++i)
      if (mod[i] != o.mod[i]) return false;
    return true;
  }
  bool operator!=(const Groups& o) const { return !(*this == o); }
};

struct Charge {
  std::array<int16_t, kMaxGroups> v{};
  bool operator==(const Charge& o) const { return v == o.v; }
  bool operator!=(const Charge& o) const { return v != o.v; }
  bool operator<(const Charge& o) const { return v < o.v; }
};

inline Charge q_add(const Groups& g, Charge a, const Charge& b) {
  for (int i = 0; i < g.n; ++i) {
    if (g.mod[i]) {
      int s = int(uint16_t(a.v[i])) + int(uint16_t(b.v[i]));
      if (s >= g.mod[i]) s -= g.mod[i];
      a.v[i] = int16_t(s);
    } else {
      a.v[i] = int16_t(uint16_t(a.v[i]) + uint16_t(b.v[i]));
    }
  }
  return a;
}
inline Charge q_inv(const Groups& g, Charge a) {
  for (int i = 0; i < g.n; ++i) {
    if (g.mod[i])
      a.v[i] = a.v[i] ? int16_t(g.mod[i] - uint16_t(a.v[i])) : int16_t(0);
    else
      a.v[i] = int16_t(-uint16_t(a.v[i]));
  }
  return a;
}
inline Charge q_neutral() { return Charge{}; }

std::string charge_str(const Groups& g, const Charge& c);

// Block index: section index per dimension.
struct BIdx {
  std::array<int32_t, kMaxRank> i{};
};
inline bool bidx_less(const BIdx& a, const BIdx& b, int r) {
  for (int k = 0; k < r; ++k) {
    if (a.i[k] != b.i[k]) return a.i[k] < b.i[k];
  }
  return false;
}
inline bool bidx_eq(const BIdx& a, const BIdx& b, int r) {
  for (int k = 0; k < r; ++k)
    if (a.i[k] != b.i[k]) return false;
  return true;
}

}  // namespace qtt
