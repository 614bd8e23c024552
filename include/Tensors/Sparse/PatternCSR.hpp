// Sparse::PatternCSR drop-in: the sparsity pattern (outer, inner) of an m x n
// matrix (PatternCSR.hpp:33-47), as far as the block product needs it --
// constructors from arrays, the accessors, and a lazily uploaded device mirror
// (like the reference's lazily built job pointers, PatternCSR.hpp:389-398).
#pragma once

#include <vector>

#include "../Tensor.hpp"

namespace Tensors {
namespace Sparse {

template <typename Int_, typename LInt_, Parallel_T parQ_ = Parallel> class PatternCSR {
  public:
    using Int = Int_;
    using LInt = LInt_;
    static constexpr Parallel_T parQ = parQ_;

    PatternCSR() = default;

    template <IntQ J0, IntQ J1, IntQ I0, IntQ I1, IntQ I3>
    PatternCSR(const J0* outer_, const J1* inner_, I0 m_, I1 n_, I3 thread_count_)
        : outer(outer_, (Int)m_ + 1), inner(inner_, (LInt)outer_[m_]), m((Int)m_), n((Int)n_),
          thread_count((Int)std::max<I3>(1, thread_count_)) {}

    // The pattern of any CSR matrix with Outer() / Inner() (e.g. Sparse::MatrixCSR).
    template <typename CSR_T>
    explicit PatternCSR(const CSR_T& A)
        : PatternCSR(A.Outer().data(), A.Inner().data(), A.RowCount(), A.ColCount(), A.ThreadCount()) {}

    PatternCSR(const PatternCSR& o) : outer(o.outer), inner(o.inner), m(o.m), n(o.n), thread_count(o.thread_count) {}
    PatternCSR& operator=(const PatternCSR& o) {
        if (this != &o) {
            outer = o.outer;
            inner = o.inner;
            m = o.m;
            n = o.n;
            thread_count = o.thread_count;
            dev_valid = false;
        }
        return *this;
    }

    Int RowCount() const { return m; }
    Int ColCount() const { return n; }
    LInt NonzeroCount() const { return outer.Size() > 0 ? outer.Last() : LInt(0); }
    Int ThreadCount() const { return thread_count; }
    const Tensor1<LInt, Int>& Outer() const { return outer; }
    const Tensor1<Int, LInt>& Inner() const { return inner; }

    // PatternCSR::WellFormedQ (PatternCSR.hpp:1146-1167)
    bool WellFormedQ() const { return outer.Size() > Int(1) && outer.Last() > LInt(0); }

    static constexpr bool kRp64 = sizeof(LInt) > 4;
    const void* DeviceRowPointers() const {
        RequireDevice();
        return kRp64 ? (const void*)d_rp64.data() : (const void*)d_rp32.data();
    }
    tg_dtype DeviceRowPointerType() const { return kRp64 ? TG_I64 : TG_I32; }
    const int32_t* DeviceColumnIndices() const {
        RequireDevice();
        return d_ci.data();
    }

  private:
    void RequireDevice() const {
        const auto lk = mirror_mutex.Lock();
        if (dev_valid) return;
        if (kRp64) {
            std::vector<int64_t> r(outer.Size());
            for (Int i = 0; i < outer.Size(); ++i) r[(Size_T)i] = (int64_t)outer[i];
            d_rp64.Upload(r.data(), r.size());
        } else {
            std::vector<int32_t> r(outer.Size());
            for (Int i = 0; i < outer.Size(); ++i) r[(Size_T)i] = (int32_t)outer[i];
            d_rp32.Upload(r.data(), r.size());
        }
        std::vector<int32_t> c((Size_T)NonzeroCount());
        for (Size_T e = 0; e < c.size(); ++e) c[e] = (int32_t)inner[(LInt)e];
        d_ci.Upload(c.data(), c.size());
        dev_valid = true;
    }

    Tensor1<LInt, Int> outer;
    Tensor1<Int, LInt> inner;
    Int m = 0;
    Int n = 0;
    Int thread_count = 1;
    Device::OwnMutex mirror_mutex;
    mutable bool dev_valid = false;
    mutable Device::Buffer<int64_t> d_rp64;
    mutable Device::Buffer<int32_t> d_rp32;
    mutable Device::Buffer<int32_t> d_ci;
};

}  // namespace Sparse
}  // namespace Tensors
