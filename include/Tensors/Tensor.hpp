// Host containers Tensor1 / Tensor2 of the drop-in (row-major, as
// src/Tensor1.hpp:9-10 and src/Tensor2.hpp:9-30 of the reference).  They hold
// the host side of Dot / solver arguments; device staging is done by the
// callers.  Random() draws uniform[-1,1] from the library's counter-based
// generator (the reference seeds std::mt19937_64 from std::random_device,
// src/Tensor_Common.hpp:141-211, which cannot be reproduced).
#pragma once

#include "Base.hpp"

namespace Tensors {

namespace detail {
inline std::uint64_t& RandomCounter() {
    static std::uint64_t c = 0x7e11507e11507ull;
    return c;
}
}  // namespace detail

template <typename Scal_, typename Int_ = Int64> class Tensor1 {
  public:
    using Scal = Scal_;
    using Int = Int_;

    Tensor1() = default;
    template <IntQ I> explicit Tensor1(I d0) : a_((Size_T)std::max<I>(d0, 0)) {}
    template <IntQ I> Tensor1(I d0, const Scal& init) : a_((Size_T)std::max<I>(d0, 0), init) {}
    template <typename S, IntQ I> Tensor1(const S* p, I d0) : a_((Size_T)std::max<I>(d0, 0)) { Read(p); }

    Scal* data() { return a_.data(); }
    const Scal* data() const { return a_.data(); }
    Int Size() const { return (Int)a_.size(); }
    Int Dim(int i) const { return i == 0 ? (Int)a_.size() : 0; }
    Int Dimension(int i) const { return Dim(i); }
    Scal& operator[](Int i) { return a_[(Size_T)i]; }
    const Scal& operator[](Int i) const { return a_[(Size_T)i]; }
    Scal& operator()(Int i) { return a_[(Size_T)i]; }
    const Scal& operator()(Int i) const { return a_[(Size_T)i]; }
    Scal& First() { return a_.front(); }
    Scal& Last() { return a_.back(); }
    const Scal& First() const { return a_.front(); }
    const Scal& Last() const { return a_.back(); }

    void Fill(const Scal& v) { std::fill(a_.begin(), a_.end(), v); }
    void SetZero() { Fill(Scal(0)); }
    template <typename S> void Read(const S* p) { for (Size_T i = 0; i < a_.size(); ++i) a_[i] = (Scal)p[i]; }
    template <typename S> void Write(S* p) const { for (Size_T i = 0; i < a_.size(); ++i) p[i] = (S)a_[i]; }
    void iota() { for (Size_T i = 0; i < a_.size(); ++i) a_[i] = (Scal)i; }
    void Random(Int = 1) {
        static_assert(std::is_floating_point_v<Scal>, "Random needs a real type");
        auto& c = detail::RandomCounter();
        Device::Status(tg_gen_uniform(0x5eed, c, (int64_t)a_.size(), Device::DType<Scal>(), a_.data()),
                       "Tensor1::Random");
        c += a_.size();
    }
    Int CountNaNs() const {
        Int c = 0;
        for (const auto& v : a_) c += std::isnan((double)v) ? 1 : 0;
        return c;
    }
    Scal Max() const { return a_.empty() ? Scal(0) : *std::max_element(a_.begin(), a_.end()); }
    Scal MaxNorm() const {
        Scal m = 0;
        for (const auto& v : a_) m = std::max<Scal>(m, (Scal)std::abs(v));
        return m;
    }
    bool operator==(const Tensor1& o) const { return a_ == o.a_; }

  private:
    std::vector<Scal> a_;
};

template <typename Scal_, typename Int_ = Int64> class Tensor2 {
  public:
    using Scal = Scal_;
    using Int = Int_;

    Tensor2() = default;
    template <IntQ I0, IntQ I1> Tensor2(I0 d0, I1 d1) : d0_((Int)d0), d1_((Int)d1), a_((Size_T)d0 * (Size_T)d1) {}
    template <IntQ I0, IntQ I1> Tensor2(I0 d0, I1 d1, const Scal& init)
        : d0_((Int)d0), d1_((Int)d1), a_((Size_T)d0 * (Size_T)d1, init) {}
    template <typename S, IntQ I0, IntQ I1> Tensor2(const S* p, I0 d0, I1 d1) : Tensor2(d0, d1) { Read(p); }

    Scal* data() { return a_.data(); }
    const Scal* data() const { return a_.data(); }
    Scal* data(Int i) { return a_.data() + (Size_T)i * (Size_T)d1_; }
    const Scal* data(Int i) const { return a_.data() + (Size_T)i * (Size_T)d1_; }
    Int Dim(int i) const { return i == 0 ? d0_ : (i == 1 ? d1_ : 0); }
    Int Dimension(int i) const { return Dim(i); }
    Int Size() const { return (Int)a_.size(); }
    Scal& operator()(Int i, Int j) { return a_[(Size_T)i * (Size_T)d1_ + (Size_T)j]; }
    const Scal& operator()(Int i, Int j) const { return a_[(Size_T)i * (Size_T)d1_ + (Size_T)j]; }
    Scal* operator[](Int i) { return data(i); }
    const Scal* operator[](Int i) const { return data(i); }

    void Fill(const Scal& v) { std::fill(a_.begin(), a_.end(), v); }
    void SetZero() { Fill(Scal(0)); }
    template <typename S> void Read(const S* p) { for (Size_T i = 0; i < a_.size(); ++i) a_[i] = (Scal)p[i]; }
    template <typename S> void Write(S* p) const { for (Size_T i = 0; i < a_.size(); ++i) p[i] = (S)a_[i]; }
    // Strided read / write (src/Tensor2.hpp:120-270).
    template <typename S> void Read(const S* B, Int ldB, Int = 1) {
        for (Int i = 0; i < d0_; ++i)
            for (Int j = 0; j < d1_; ++j) (*this)(i, j) = (Scal)B[(Size_T)i * (Size_T)ldB + (Size_T)j];
    }
    template <typename S> void Write(S* B, Int ldB, Int = 1) const {
        for (Int i = 0; i < d0_; ++i)
            for (Int j = 0; j < d1_; ++j) B[(Size_T)i * (Size_T)ldB + (Size_T)j] = (S)(*this)(i, j);
    }
    void Random(Int = 1) {
        static_assert(std::is_floating_point_v<Scal>, "Random needs a real type");
        auto& c = detail::RandomCounter();
        Device::Status(tg_gen_uniform(0x5eed, c, (int64_t)a_.size(), Device::DType<Scal>(), a_.data()),
                       "Tensor2::Random");
        c += a_.size();
    }
    Scal MaxNorm() const {
        Scal m = 0;
        for (const auto& v : a_) m = std::max<Scal>(m, (Scal)std::abs(v));
        return m;
    }
    bool operator==(const Tensor2& o) const { return d0_ == o.d0_ && d1_ == o.d1_ && a_ == o.a_; }

  private:
    Int d0_ = 0, d1_ = 0;
    std::vector<Scal> a_;
};

// Three-index row-major container: entry (i, j, k) at (i*d1 + j)*d2 + k
// (src/Tensor3.hpp:116-165).  In the drop-in it carries host copies of the
// GMRES basis and Hessenberg matrix.
template <typename Scal_, typename Int_ = Int64> class Tensor3 {
  public:
    using Scal = Scal_;
    using Int = Int_;

    Tensor3() = default;
    template <IntQ I0, IntQ I1, IntQ I2>
    Tensor3(I0 d0, I1 d1, I2 d2) : d_{(Int)d0, (Int)d1, (Int)d2}, a_((Size_T)d0 * (Size_T)d1 * (Size_T)d2) {}

    Scal* data() { return a_.data(); }
    const Scal* data() const { return a_.data(); }
    Scal* data(Int i) { return a_.data() + (Size_T)i * (Size_T)d_[1] * (Size_T)d_[2]; }
    const Scal* data(Int i) const { return a_.data() + (Size_T)i * (Size_T)d_[1] * (Size_T)d_[2]; }
    Scal* data(Int i, Int j) { return data(i) + (Size_T)j * (Size_T)d_[2]; }
    const Scal* data(Int i, Int j) const { return data(i) + (Size_T)j * (Size_T)d_[2]; }
    Int Dim(int i) const { return (i >= 0 && i < 3) ? d_[i] : 0; }
    Int Dimension(int i) const { return Dim(i); }
    Int Size() const { return (Int)a_.size(); }
    Scal& operator()(Int i, Int j, Int k) { return data(i, j)[k]; }
    const Scal& operator()(Int i, Int j, Int k) const { return data(i, j)[k]; }

    void Fill(const Scal& v) { std::fill(a_.begin(), a_.end(), v); }
    void SetZero() { Fill(Scal(0)); }
    template <typename S> void Read(const S* p) { for (Size_T i = 0; i < a_.size(); ++i) a_[i] = (Scal)p[i]; }
    template <typename S> void Write(S* p) const { for (Size_T i = 0; i < a_.size(); ++i) p[i] = (S)a_[i]; }

  private:
    Int d_[3] = {0, 0, 0};
    std::vector<Scal> a_;
};

}  // namespace Tensors
