// Tensors drop-in for MI355X: common types, messages and device helpers.
//
// Mirrors the vocabulary the reference's sparse path uses (cptr/mptr/cref/mref,
// VarSize, Parallel, Side, eprint/wprint -- provided there by the Tools
// submodule included from Base.hpp:10) so code written against
// HenrikSchumacher/Tensors compiles unchanged for the paths this drop-in covers.
// Device work goes through the C ABI in tensors_amd.h (libtensors_amd.so).
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../tensors_amd.h"

namespace Tensors {

using Size_T = std::size_t;
constexpr Size_T VarSize = 0;

template <typename T> using cptr = const T*;
template <typename T> using mptr = T*;
template <typename T> using cref = const T&;
template <typename T> using mref = T&;

using Int32 = std::int32_t;
using Int64 = std::int64_t;
using Real32 = float;
using Real64 = double;

enum class Parallel_T : bool { Sequential = false, Parallel = true };
constexpr Parallel_T Sequential = Parallel_T::Sequential;
constexpr Parallel_T Parallel = Parallel_T::Parallel;

enum class Side : int { Left = 0, Right = 1 };

// Operator tags of MatrixCSR::Op and Sparse::Dot (Tools::Op in the reference).
enum class Op : int { Id = 0, Conj, Trans, ConjTrans, Re, ReTrans, Im, ImTrans };
constexpr bool TransposedQ(Op op) {
    return op == Op::Trans || op == Op::ConjTrans || op == Op::ReTrans || op == Op::ImTrans;
}
constexpr bool NotTransposedQ(Op op) { return !TransposedQ(op); }

template <typename I> concept IntQ = std::is_integral_v<I>;

inline void eprint(const std::string& s) { std::fprintf(stderr, "ERROR: %s\n", s.c_str()); }
inline void wprint(const std::string& s) { std::fprintf(stderr, "WARNING: %s\n", s.c_str()); }
inline void logprint(const std::string&) {}

template <typename T> inline std::string ToString(const T& v) { return std::to_string(v); }

template <typename T> constexpr const char* TypeName() {
    if constexpr (std::is_same_v<T, double>) return "Real64";
    else if constexpr (std::is_same_v<T, float>) return "Real32";
    else if constexpr (std::is_same_v<T, std::int32_t>) return "Int32";
    else if constexpr (std::is_same_v<T, std::int64_t>) return "Int64";
    else if constexpr (std::is_same_v<T, std::uint64_t>) return "UInt64";
    else return "?";
}

namespace Device {

inline void Check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// tg_* statuses that the reference treats as "print and do nothing" are turned
// into eprint; every other failure is a GPU/runtime error and throws.
inline int Status(int st, const std::string& where) {
    if (st == TG_OK) return st;
    if (st == TG_ERR_ALIAS || st == TG_ERR_NRHS || st == TG_ERR_SHAPE) {
        eprint(where + ": " + tg_last_error());
        return st;
    }
    throw std::runtime_error(where + ": " + tg_last_error());
}

inline bool IsDevicePointer(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.isManaged;
}

template <typename T> constexpr tg_dtype DType() {
    if constexpr (std::is_same_v<T, double>) return TG_F64;
    else if constexpr (std::is_same_v<T, float>) return TG_F32;
    else if constexpr (std::is_same_v<T, std::int32_t>) return TG_I32;
    else return TG_I64;
}

// Owning device allocation.
template <typename T> class Buffer {
  public:
    Buffer() = default;
    explicit Buffer(Size_T n) { Resize(n); }
    ~Buffer() { Free(); }
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    Buffer(Buffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    Buffer& operator=(Buffer&& o) noexcept {
        if (this != &o) { Free(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void Resize(Size_T n) {
        if (n <= n_) return;
        Free();
        Check(hipMalloc((void**)&p_, std::max<Size_T>(n, 1) * sizeof(T)), "hipMalloc");
        n_ = n;
    }
    void Free() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* data() const { return p_; }
    Size_T size() const { return n_; }
    void Upload(const T* h, Size_T n) {
        Resize(n);
        if (n) Check(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void Download(T* h, Size_T n) const {
        if (n) Check(hipMemcpy(h, p_, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }

  private:
    T* p_ = nullptr;
    Size_T n_ = 0;
};

// A mutex owned by one object: a copied or moved object gets a fresh, unlocked
// one.  It guards a matrix's lazily built device mirror and its staging buffers,
// so that const calls on one matrix from several host threads (the first of
// which uploads the mirror) run one after another instead of racing.
struct OwnMutex {
    OwnMutex() = default;
    OwnMutex(const OwnMutex&) {}
    OwnMutex& operator=(const OwnMutex&) { return *this; }
    std::unique_lock<std::recursive_mutex> Lock() const { return std::unique_lock<std::recursive_mutex>(m); }

  private:
    mutable std::recursive_mutex m;
};

// Copies `count` contiguous elements between host and device (synchronous).
template <typename T>
inline void Copy(T* dst, const T* src, Size_T count, hipMemcpyKind kind) {
    if (count == 0) return;
    Check(hipMemcpy(dst, src, count * sizeof(T), kind), "hipMemcpy");
}

// Copies a rows x cols block with leading dimensions between host and device.
template <typename T>
inline void Copy2D(T* dst, Size_T ldd, const T* src, Size_T lds, Size_T rows, Size_T cols, hipMemcpyKind kind) {
    if (rows == 0 || cols == 0) return;
    Check(hipMemcpy2D(dst, ldd * sizeof(T), src, lds * sizeof(T), cols * sizeof(T), rows, kind), "hipMemcpy2D");
}

}  // namespace Device
}  // namespace Tensors
