// Sparse::MatrixCSR drop-in: same class name, template parameters, constructors
// and Dot overloads as the reference (src/Sparse/MatrixCSR.hpp:9-10, 67-328,
// 965-1086), with the products running as HIP kernels on the GPU.
//
// Host arrays (outer, inner, values) stay authoritative, as in the reference.
// A device mirror is uploaded on first use and reused -- the same lazy-cache
// pattern as PatternCSR::RequireJobPtr (src/Sparse/PatternCSR.hpp:389-398);
// the non-const accessors Outer()/Inner()/Values() drop the mirror.  Const calls
// on one matrix from several host threads (Dot, the first of which uploads the
// mirror) are serialised by a per-matrix mutex (mirror_mutex).  Writing through
// a reference that Values(), Inner() or Outer() returned earlier while another
// thread calls Dot is still a data race, as it is in the reference.
//
// Dot accepts host or device pointers for X and Y (told apart with
// hipPointerGetAttributes).  Device pointers are used in place; host blocks
// are staged through device buffers the matrix keeps.  Like the reference's
// Dot, a call returns with Y complete: it waits for the null stream it ran on
// (not for the whole device, so work on the caller's other non-blocking streams
// keeps running).  Error behaviour is the
// reference's: shape mismatch, X == Y and NRHS != nrhs print a message and do
// nothing.  Supported: Scal = double or float with X, Y of the same type;
// complex scalars and base-1 indices are not on this path.
#pragma once

#include <filesystem>
#include <vector>

#include "../Tensor.hpp"

namespace Tensors {
namespace Sparse {

// Triple buffer used by assembly routines (src/TripleAggregator.hpp:24-62).
template <typename I, typename J, typename S, typename LInt> class TripleAggregator {
  public:
    TripleAggregator() = default;
    explicit TripleAggregator(LInt reserve) {
        i_.reserve((Size_T)reserve);
        j_.reserve((Size_T)reserve);
        a_.reserve((Size_T)reserve);
    }
    void Push(I i, J j, S a) {
        i_.push_back(i);
        j_.push_back(j);
        a_.push_back(a);
    }
    LInt Size() const { return (LInt)i_.size(); }
    const I* data_0() const { return i_.data(); }
    const J* data_1() const { return j_.data(); }
    const S* data_2() const { return a_.data(); }

  private:
    std::vector<I> i_;
    std::vector<J> j_;
    std::vector<S> a_;
};

template <typename Scal_, typename Int_, typename LInt_, Parallel_T parQ_ = Parallel> class MatrixCSR {
  public:
    using Scal = Scal_;
    using Int = Int_;
    using LInt = LInt_;
    static constexpr Parallel_T parQ = parQ_;
    static_assert(std::is_same_v<Scal, double> || std::is_same_v<Scal, float>,
                  "the GPU drop-in supports real double/float matrices");

    // ---- constructors (MatrixCSR.hpp:67-126) --------------------------------
    MatrixCSR() = default;

    template <IntQ I0, IntQ I1, IntQ I3>
    MatrixCSR(I0 m_, I1 n_, I3 thread_count_) : outer((Int)m_ + 1, LInt(0)), m((Int)m_), n((Int)n_),
                                                thread_count((Int)std::max<I3>(1, thread_count_)) {}

    template <IntQ I0, IntQ I1, IntQ I2, IntQ I3>
    MatrixCSR(I0 m_, I1 n_, I2 nnz_, I3 thread_count_)
        : outer((Int)m_ + 1, LInt(0)), inner((LInt)nnz_), values((LInt)nnz_), m((Int)m_), n((Int)n_),
          thread_count((Int)std::max<I3>(1, thread_count_)) {}

    template <typename S, IntQ J0, IntQ J1, IntQ I0, IntQ I1, IntQ I3>
    MatrixCSR(const J0* outer_, const J1* inner_, const S* values_, I0 m_, I1 n_, I3 thread_count_)
        : outer(outer_, (Int)m_ + 1), inner(inner_, (LInt)outer_[m_]), values(values_, (LInt)outer_[m_]),
          m((Int)m_), n((Int)n_), thread_count((Int)std::max<I3>(1, thread_count_)) {}

    template <IntQ I0, IntQ I1, IntQ I3>
    MatrixCSR(Tensor1<LInt, Int>&& outer_, Tensor1<Int, LInt>&& inner_, Tensor1<Scal, LInt>&& values_, I0 m_, I1 n_,
              I3 thread_count_)
        : outer(std::move(outer_)), inner(std::move(inner_)), values(std::move(values_)), m((Int)m_), n((Int)n_),
          thread_count((Int)std::max<I3>(1, thread_count_)) {}

    // From triples (MatrixCSR.hpp:203-235); assembled on the GPU.
    template <IntQ ExtInt, typename ExtScal>
    MatrixCSR(LInt nnz_, const ExtInt* i, const ExtInt* j, const ExtScal* a, Int m_, Int n_, Int thread_count_,
              bool compressQ = true, int symmetrizeQ = 0, bool assemblerQ = false)
        : m(m_), n(n_), thread_count(std::max<Int>(1, thread_count_)) {
        if (assemblerQ) wprint(ClassName() + ": assemblerQ is not supported by the GPU drop-in; ignored.");
        FromTriples<ExtInt, ExtScal>((int64_t)nnz_, {i}, {j}, {a}, {(int64_t)nnz_}, compressQ, symmetrizeQ);
    }

    template <IntQ ExtInt, typename ExtScal>
    MatrixCSR(const TripleAggregator<ExtInt, ExtInt, ExtScal, LInt>& triples, Int m_, Int n_, Int final_thread_count,
              bool compressQ = true, int symmetrizeQ = 0, bool assemblerQ = false)
        : m(m_), n(n_), thread_count(std::max<Int>(1, final_thread_count)) {
        if (assemblerQ) wprint(ClassName() + ": assemblerQ is not supported by the GPU drop-in; ignored.");
        FromTriples<ExtInt, ExtScal>((int64_t)triples.Size(), {triples.data_0()}, {triples.data_1()}, {triples.data_2()},
                    {(int64_t)triples.Size()}, compressQ, symmetrizeQ);
    }

    template <IntQ ExtInt, typename ExtScal>
    MatrixCSR(const std::vector<TripleAggregator<ExtInt, ExtInt, ExtScal, LInt>>& triples, Int m_, Int n_,
              Int final_thread_count, bool compressQ = true, int symmetrizeQ = 0)
        : m(m_), n(n_), thread_count(std::max<Int>(1, final_thread_count)) {
        std::vector<const ExtInt*> I, J;
        std::vector<const ExtScal*> A;
        std::vector<int64_t> c;
        int64_t total = 0;
        for (const auto& t : triples) {
            I.push_back(t.data_0());
            J.push_back(t.data_1());
            A.push_back(t.data_2());
            c.push_back((int64_t)t.Size());
            total += (int64_t)t.Size();
        }
        FromTriples<ExtInt, ExtScal>(total, I, J, A, c, compressQ, symmetrizeQ);
    }

    // ---- MatrixMarket files (MatrixCSR/MatrixMarket.hpp:3-306) ------------------
    // The entries are parsed on the host (tg_mm_info / tg_mm_read) and assembled on
    // the GPU like the triples constructor: duplicates summed, off-diagonal entries
    // mirrored for "symmetric" files.  A file the reference refuses (missing, not
    // MatrixMarket, dense, complex, skew-symmetric, hermitian) or a malformed entry
    // prints the reason and leaves the matrix unchanged.
    void LoadFromMatrixMarket(const std::filesystem::path& file, Int thread_count_) {
        const std::string where = MethodName("LoadFromMatrixMarket");
        int32_t mm = 0, nn = 0, field = 0, symm = 0;
        int64_t count = 0;
        if (tg_mm_info(file.c_str(), &mm, &nn, &count, &field, &symm) != TG_OK) {
            eprint(where + ": " + tg_last_error() + ". Doing nothing.");
            return;
        }
        std::vector<int32_t> I((Size_T)std::max<int64_t>(count, 1)), J(I.size());
        std::vector<double> V(I.size());
        if (tg_mm_read(file.c_str(), count, I.data(), J.data(), V.data(), 0) != TG_OK) {
            eprint(where + ": " + tg_last_error() + ". Doing nothing.");
            return;
        }
        MatrixCSR A((LInt)count, I.data(), J.data(), V.data(), (Int)mm, (Int)nn, thread_count_, true, symm);
        *this = std::move(A);
    }

    // "coordinate real general", one line per nonzero in row order (every row; the
    // reference's writer keeps only two of its four thread buffers, :283-295).
    void WriteToMatrixMarket(const std::filesystem::path& file) const {
        const std::string where = MethodName("WriteToMatrixMarket");
        if (!WellFormedQ()) {
            eprint(where + ": Matrix is not well-formed. Doing nothing.");
            return;
        }
        const int64_t nnz = (int64_t)NonzeroCount();
        std::vector<int64_t> r(outer.Size());
        for (Int i = 0; i < outer.Size(); ++i) r[(Size_T)i] = (int64_t)outer[i];
        std::vector<int32_t> c((Size_T)nnz);
        for (int64_t e = 0; e < nnz; ++e) c[(Size_T)e] = (int32_t)inner[(LInt)e];
        if (tg_mm_write(file.c_str(), (int32_t)m, (int32_t)n, r.data(), TG_I64, c.data(), values.data(),
                        Device::DType<Scal>(), 0) != TG_OK)
            eprint(where + ": " + tg_last_error());
    }

    MatrixCSR(const MatrixCSR& o)
        : outer(o.outer), inner(o.inner), values(o.values), m(o.m), n(o.n), thread_count(o.thread_count) {}
    MatrixCSR& operator=(const MatrixCSR& o) {
        if (this != &o) {
            outer = o.outer;
            inner = o.inner;
            values = o.values;
            m = o.m;
            n = o.n;
            thread_count = o.thread_count;
            dev_valid = false;
        }
        return *this;
    }
    MatrixCSR(MatrixCSR&&) = default;
    MatrixCSR& operator=(MatrixCSR&&) = default;

    // ---- accessors ----------------------------------------------------------
    Int RowCount() const { return m; }
    Int ColCount() const { return n; }
    LInt NonzeroCount() const { return outer.Size() > 0 ? outer[m] : LInt(0); }
    Int ThreadCount() const { return thread_count; }
    void SetThreadCount(Int t) { thread_count = std::max<Int>(1, t); }  // kept for API parity
    Tensor1<LInt, Int>& Outer() { dev_valid = false; return outer; }
    const Tensor1<LInt, Int>& Outer() const { return outer; }
    Tensor1<Int, LInt>& Inner() { dev_valid = false; return inner; }
    const Tensor1<Int, LInt>& Inner() const { return inner; }
    Tensor1<Scal, LInt>& Values() { dev_valid = false; return values; }
    const Tensor1<Scal, LInt>& Values() const { return values; }
    Tensor1<Scal, LInt>& Value() { return Values(); }
    const Tensor1<Scal, LInt>& Value() const { return values; }

    // PatternCSR::WellFormedQ (PatternCSR.hpp:1146-1167)
    bool WellFormedQ() const {
        const bool ok = outer.Size() > Int(1) && outer.Last() > LInt(0);
        if (!ok) wprint(ClassName() + "::WellFormedQ: Matrix is not well formed.");
        return ok;
    }

    // ---- device mirror ---------------------------------------------------------
    static constexpr bool kRp64 = sizeof(LInt) > 4;
    void RequireDevice() const {
        const auto lk = mirror_mutex.Lock();
        if (dev_valid) return;
        plan.Reset();  // it points at the device arrays about to be replaced
        const int64_t nnz = (int64_t)NonzeroCount();
        if (kRp64) {
            std::vector<int64_t> r(outer.Size());
            for (Int i = 0; i < outer.Size(); ++i) r[(Size_T)i] = (int64_t)outer[i];
            d_rp64.Upload(r.data(), r.size());
        } else {
            std::vector<int32_t> r(outer.Size());
            for (Int i = 0; i < outer.Size(); ++i) r[(Size_T)i] = (int32_t)outer[i];
            d_rp32.Upload(r.data(), r.size());
        }
        std::vector<int32_t> c((Size_T)nnz);
        for (int64_t e = 0; e < nnz; ++e) c[(Size_T)e] = (int32_t)inner[(LInt)e];
        d_ci.Upload(c.data(), c.size());
        d_val.Upload(values.data(), (Size_T)std::min<int64_t>(nnz, (int64_t)values.Size()));
        dev_valid = true;
    }
    // Column-slice plan of the device pattern (tg_csr_plan_create), built on first
    // use like the reference's lazily built job pointers (PatternCSR.hpp:389-398).
    tg_csr_plan_t DevicePlan() const {
        const auto lk = mirror_mutex.Lock();
        RequireDevice();
        if (!plan.p)
            Device::Status(tg_csr_plan_create((int32_t)m, (int32_t)n, (int64_t)NonzeroCount(), DeviceRowPointers(),
                                              DeviceRowPointerType(), d_ci.data(), &plan.p, nullptr),
                           MethodName("RequirePlan"));
        return plan.p;
    }
    const void* DeviceRowPointers() const { RequireDevice(); return kRp64 ? (const void*)d_rp64.data() : (const void*)d_rp32.data(); }
    tg_dtype DeviceRowPointerType() const { return kRp64 ? TG_I64 : TG_I32; }
    const int32_t* DeviceColumnIndices() const { RequireDevice(); return d_ci.data(); }
    const Scal* DeviceValues() const { RequireDevice(); return d_val.data(); }

    // ---- products (MatrixCSR.hpp:965-1086) -----------------------------------
    template <Int NRHS = (Int)VarSize, typename a_T, typename X_T, typename b_T, typename Y_T>
    void Dot(const a_T alpha, cptr<X_T> X, const Int ldX, const b_T beta, mptr<Y_T> Y, const Int ldY,
             const Int nrhs = Int(1)) const {
        Dot_<NRHS>(nullptr, alpha, X, ldX, beta, Y, ldY, nrhs);
    }
    template <Int NRHS = (Int)VarSize, typename a_T, typename X_T, typename b_T, typename Y_T>
    void Dot(const a_T alpha, cptr<X_T> X, const b_T beta, mptr<Y_T> Y, const Int nrhs = Int(1)) const {
        Dot_<NRHS>(nullptr, alpha, X, nrhs, beta, Y, nrhs, nrhs);
    }
    template <typename a_T, typename X_T, typename b_T, typename Y_T, typename I1, typename I2>
    void Dot(const a_T alpha, const Tensor1<X_T, I1>& X, const b_T beta, Tensor1<Y_T, I2>& Y) const {
        if ((Int)X.Dim(0) == n && (Int)Y.Dim(0) == m)
            Dot_<1>(nullptr, alpha, X.data(), 1, beta, Y.data(), 1, 1);
        else
            eprint(MethodName("Dot") + ": shapes of matrix, input, and output do not match.");
    }
    template <Int NRHS = (Int)VarSize, typename a_T, typename X_T, typename b_T, typename Y_T, typename I1,
              typename I2>
    void Dot(const a_T alpha, const Tensor2<X_T, I1>& X, const b_T beta, Tensor2<Y_T, I2>& Y) const {
        if ((Int)X.Dim(0) == n && (Int)Y.Dim(0) == m && X.Dim(1) == Y.Dim(1)) {
            const Int k = (Int)X.Dim(1);
            Dot_<NRHS>(nullptr, alpha, X.data(), k, beta, Y.data(), k, k);
        } else {
            eprint(MethodName("Dot") + ": shapes of matrix, input, and output do not match.");
        }
    }
    // external values (MatrixCSR.hpp:1026-1086)
    template <Int NRHS = (Int)VarSize, typename T_ext, typename a_T, typename b_T>
    void Dot(cptr<T_ext> ext_values, const a_T alpha, cptr<T_ext> X, const Int ldX, const b_T beta, mptr<T_ext> Y,
             const Int ldY, const Int nrhs = Int(1)) const {
        static_assert(std::is_same_v<T_ext, Scal>, "external values must have the matrix's scalar type");
        Dot_<NRHS>(ext_values, alpha, X, ldX, beta, Y, ldY, nrhs);
    }
    template <Int NRHS = (Int)VarSize, typename T_ext, typename a_T, typename b_T>
    void Dot(cptr<T_ext> ext_values, const a_T alpha, cptr<T_ext> X, const b_T beta, mptr<T_ext> Y,
             const Int nrhs = Int(1)) const {
        static_assert(std::is_same_v<T_ext, Scal>, "external values must have the matrix's scalar type");
        Dot_<NRHS>(ext_values, alpha, X, nrhs, beta, Y, nrhs, nrhs);
    }

    // Transpose (MatrixCSR.hpp:445-495), on the GPU.
    MatrixCSR Transpose() const {
        if (!WellFormedQ()) return MatrixCSR(n, m, 0, 1);
        RequireDevice();
        const int64_t nnz = (int64_t)NonzeroCount();
        Device::Buffer<char> ws(tg_csr_transpose_workspace_bytes((int32_t)m, (int32_t)n, nnz));
        Device::Buffer<int64_t> rp64;
        Device::Buffer<int32_t> rp32;
        Device::Buffer<int32_t> ci((Size_T)nnz);
        Device::Buffer<Scal> va((Size_T)nnz);
        void* orp = kRp64 ? (void*)(rp64.Resize((Size_T)n + 1), rp64.data()) : (void*)(rp32.Resize((Size_T)n + 1), rp32.data());
        Device::Status(tg_csr_transpose((int32_t)m, (int32_t)n, nnz, DeviceRowPointers(), DeviceRowPointerType(),
                                        d_ci.data(), d_val.data(), Device::DType<Scal>(), orp, ci.data(), va.data(),
                                        ws.data(), ws.size(), nullptr),
                       MethodName("Transpose"));
        return FromDevice(n, m, orp, ci.data(), va.data(), nnz);
    }

    // MatrixCSR::Op<op> (MatrixCSR.hpp:389-511) for real scalars: Id / Conj / Re copy,
    // Trans / ConjTrans / ReTrans transpose (on the GPU), Im / ImTrans are empty.
    template <Tensors::Op op> MatrixCSR Op() const {
        if constexpr (op == Tensors::Op::Im) return MatrixCSR(m, n, 0, 1);
        else if constexpr (op == Tensors::Op::ImTrans) return MatrixCSR(n, m, 0, 1);
        else if constexpr (Tensors::TransposedQ(op)) return Transpose();
        else {
            if (!WellFormedQ()) return MatrixCSR(m, n, 0, 1);
            return MatrixCSR(*this);
        }
    }

    // Sparse x sparse (MatrixCSR.hpp:835-921), on the GPU.
    MatrixCSR Dot(const MatrixCSR& B) const {
        if (!WellFormedQ() || !B.WellFormedQ()) {
            eprint(MethodName("Dot") + ": Matrix is not well-formed.");
            return MatrixCSR();
        }
        RequireDevice();
        B.RequireDevice();
        Device::Buffer<char> ws0((Size_T)(3 * (m + 1) * 8 + (1 << 20)));
        int64_t E = 0;
        Device::Status(tg_csr_spgemm_expanded_count((int32_t)m, DeviceRowPointers(), d_ci.data(),
                                                    B.DeviceRowPointers(), DeviceRowPointerType(), &E, ws0.data(),
                                                    ws0.size(), nullptr),
                       MethodName("Dot"));
        Device::Buffer<char> ws(tg_csr_spgemm_workspace_bytes((int32_t)m, (int32_t)B.n, E, Device::DType<Scal>()));
        Device::Buffer<int64_t> rp64;
        Device::Buffer<int32_t> rp32;
        Device::Buffer<int32_t> ci((Size_T)std::max<int64_t>(E, 1));
        Device::Buffer<Scal> va((Size_T)std::max<int64_t>(E, 1));
        void* orp = kRp64 ? (void*)(rp64.Resize((Size_T)m + 1), rp64.data()) : (void*)(rp32.Resize((Size_T)m + 1), rp32.data());
        int64_t nnz = 0;
        Device::Status(tg_csr_spgemm((int32_t)m, (int32_t)n, (int32_t)B.n, DeviceRowPointers(), d_ci.data(),
                                     d_val.data(), B.DeviceRowPointers(), B.d_ci.data(), B.d_val.data(),
                                     DeviceRowPointerType(), Device::DType<Scal>(), E, orp, ci.data(), va.data(),
                                     ws.data(), ws.size(), &nnz, nullptr),
                       MethodName("Dot"));
        return FromDevice(m, B.n, orp, ci.data(), va.data(), nnz);
    }

    static MatrixCSR IdentityMatrix(const Int n_, const Int thread_count_ = 1) {
        MatrixCSR A(n_, n_, n_, thread_count_);
        A.outer.iota();
        A.inner.iota();
        A.values.Fill(Scal(1));
        return A;
    }

    static std::string ClassName() {
        return std::string("Sparse::MatrixCSR<") + TypeName<Scal>() + "," + TypeName<Int>() + "," + TypeName<LInt>() +
               ">";
    }
    static std::string MethodName(const std::string& t) { return ClassName() + "::" + t; }

  protected:
    template <Int NRHS, typename a_T, typename X_T, typename b_T, typename Y_T>
    void Dot_(const Scal* ext_values, const a_T alpha, const X_T* X, const Int ldX, const b_T beta, Y_T* Y,
              const Int ldY, const Int nrhs) const {
        static_assert(std::is_same_v<X_T, Scal> && std::is_same_v<Y_T, Scal>,
                      "GPU Dot: X and Y must have the matrix's scalar type");
        if (!WellFormedQ()) wprint(MethodName("Dot_") + ": Not WellFormedQ(). Doing nothing.");
        if ((const void*)X == (const void*)Y && (double)alpha != 0.0) {
            eprint(MethodName("Dot") + ": Input and output pointer coincide. This is not safe. Aborting.");
            return;
        }
        // the staging buffers and the plan's partial blocks belong to the matrix
        const auto lk = mirror_mutex.Lock();
        RequireDevice();
        const int64_t nnz = (int64_t)NonzeroCount();
        const Scal* vals = d_val.data();
        if (ext_values) {
            if (Device::IsDevicePointer(ext_values)) {
                vals = (const Scal*)ext_values;
            } else {
                d_ext.Upload((const Scal*)ext_values, (Size_T)nnz);
                vals = d_ext.data();
            }
        }
        const bool xd = X == nullptr || Device::IsDevicePointer(X), yd = Device::IsDevicePointer(Y);
        const std::string where = MethodName("Dot");
        // the matrix's own values get the plan's re-laid copy on the first product whose
        // shape uses it (nrhs <= 4, X of 2-64 MB); changing the values drops the plan
        const double xbytes = (double)n * (double)nrhs * (double)sizeof(Scal);
        if (!ext_values && !plan.split && nrhs <= 4 && xbytes > 2.0 * (1 << 20) && xbytes <= 64.0 * (1 << 20)) {
            Device::Status(tg_csr_plan_set_values(DevicePlan(), d_val.data(), Device::DType<Scal>(), nullptr), where);
            plan.split = true;
        }
        if (xd && yd) {
            Device::Status(tg_csr_spmm_plan(DevicePlan(), vals, Device::DType<Scal>(), (double)alpha, X, (int64_t)ldX,
                                            (double)beta, Y, (int64_t)ldY, (int32_t)nrhs, (int32_t)NRHS, nullptr),
                           where);
            Device::Check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
            return;
        }
        if (NRHS > 0 && nrhs != NRHS) {
            eprint(where + ": nrhs != NRHS. Doing nothing.");
            return;
        }
        // host blocks: stage through device buffers (ld = nrhs)
        const X_T* Xd = X;
        if (!xd) {
            d_x.Resize((Size_T)n * (Size_T)nrhs);
            Device::Copy2D(d_x.data(), (Size_T)nrhs, X, (Size_T)ldX, (Size_T)n, (Size_T)nrhs, hipMemcpyHostToDevice);
            Xd = d_x.data();
        }
        Y_T* Yd = Y;
        if (!yd) {
            d_y.Resize((Size_T)m * (Size_T)nrhs);
            Yd = d_y.data();
            if ((double)beta != 0.0)
                Device::Copy2D(Yd, (Size_T)nrhs, Y, (Size_T)ldY, (Size_T)m, (Size_T)nrhs, hipMemcpyHostToDevice);
        }
        const int st = Device::Status(
            tg_csr_spmm_plan(DevicePlan(), vals, Device::DType<Scal>(), (double)alpha, Xd,
                             xd ? (int64_t)ldX : (int64_t)nrhs, (double)beta, Yd, yd ? (int64_t)ldY : (int64_t)nrhs,
                             (int32_t)nrhs, (int32_t)NRHS, nullptr),
            where);
        if (st == TG_OK && !yd)
            Device::Copy2D(Y, (Size_T)ldY, Yd, (Size_T)nrhs, (Size_T)m, (Size_T)nrhs, hipMemcpyDeviceToHost);
        else
            Device::Check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    }

    template <typename ExtInt, typename ExtScal>
    void FromTriples(int64_t total, const std::vector<const ExtInt*>& I, const std::vector<const ExtInt*>& J,
                     const std::vector<const ExtScal*>& A, const std::vector<int64_t>& counts, bool compressQ,
                     int symmetrizeQ) {
        std::vector<int32_t> hi((Size_T)total), hj((Size_T)total);
        std::vector<Scal> ha((Size_T)total);
        int64_t o = 0;
        for (Size_T l = 0; l < counts.size(); ++l)
            for (int64_t k = 0; k < counts[l]; ++k, ++o) {
                hi[(Size_T)o] = (int32_t)I[l][k];
                hj[(Size_T)o] = (int32_t)J[l][k];
                ha[(Size_T)o] = (Scal)A[l][k];
            }
        Device::Buffer<int32_t> di, dj;
        Device::Buffer<Scal> da;
        di.Upload(hi.data(), hi.size());
        dj.Upload(hj.data(), hj.size());
        da.Upload(ha.data(), ha.size());
        const int64_t cap = std::max<int64_t>(1, symmetrizeQ ? 2 * total : total);
        Device::Buffer<char> ws(tg_csr_from_triples_workspace_bytes(total, (int32_t)m, symmetrizeQ ? 1 : 0));
        Device::Buffer<int64_t> rp64;
        Device::Buffer<int32_t> rp32;
        Device::Buffer<int32_t> ci((Size_T)cap);
        Device::Buffer<Scal> va((Size_T)cap);
        void* orp = kRp64 ? (void*)(rp64.Resize((Size_T)m + 1), rp64.data()) : (void*)(rp32.Resize((Size_T)m + 1), rp32.data());
        int64_t nnz = 0;
        Device::Status(tg_csr_from_triples(total, di.data(), dj.data(), da.data(), Device::DType<Scal>(), (int32_t)m,
                                           (int32_t)n, compressQ ? 1 : 0, symmetrizeQ ? 1 : 0, orp,
                                           kRp64 ? TG_I64 : TG_I32, ci.data(), va.data(), cap, ws.data(), ws.size(),
                                           &nnz, nullptr),
                       MethodName("FromTriples"));
        *this = FromDevice(m, n, orp, ci.data(), va.data(), nnz, thread_count);
    }

    static MatrixCSR FromDevice(Int m_, Int n_, const void* rp, const int32_t* ci, const Scal* va, int64_t nnz,
                                Int threads = 1) {
        MatrixCSR B(m_, n_, (LInt)nnz, threads);
        if (kRp64) {
            std::vector<int64_t> r((Size_T)m_ + 1);
            Device::Check(hipMemcpy(r.data(), rp, r.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
            for (Size_T i = 0; i < r.size(); ++i) B.outer[(Int)i] = (LInt)r[i];
        } else {
            std::vector<int32_t> r((Size_T)m_ + 1);
            Device::Check(hipMemcpy(r.data(), rp, r.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
            for (Size_T i = 0; i < r.size(); ++i) B.outer[(Int)i] = (LInt)r[i];
        }
        std::vector<int32_t> c((Size_T)nnz);
        Device::Check(hipMemcpy(c.data(), ci, c.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        for (int64_t e = 0; e < nnz; ++e) B.inner[(LInt)e] = (Int)c[(Size_T)e];
        if (nnz) Device::Check(hipMemcpy(B.values.data(), va, (Size_T)nnz * sizeof(Scal), hipMemcpyDeviceToHost),
                               "hipMemcpy");
        return B;
    }

    Tensor1<LInt, Int> outer;
    Tensor1<Int, LInt> inner;
    Tensor1<Scal, LInt> values;
    Int m = 0;
    Int n = 0;
    Int thread_count = 1;

    // Owns a tg_csr_plan_t; a copied matrix starts without one (the plan points at
    // the device arrays of the matrix that built it), a moved one takes it along.
    struct PlanHolder {
        mutable tg_csr_plan_t p = nullptr;
        mutable bool split = false;  // the own values are registered (tg_csr_plan_set_values)
        PlanHolder() = default;
        PlanHolder(const PlanHolder&) {}
        PlanHolder(PlanHolder&& o) noexcept : p(o.p), split(o.split) {
            o.p = nullptr;
            o.split = false;
        }
        PlanHolder& operator=(const PlanHolder&) {
            Reset();
            return *this;
        }
        PlanHolder& operator=(PlanHolder&& o) noexcept {
            if (this != &o) {
                Reset();
                p = o.p;
                split = o.split;
                o.p = nullptr;
                o.split = false;
            }
            return *this;
        }
        ~PlanHolder() { Reset(); }
        void Reset() const {
            if (p) tg_csr_plan_destroy(p);
            p = nullptr;
            split = false;
        }
    };

    Device::OwnMutex mirror_mutex;
    mutable bool dev_valid = false;
    mutable PlanHolder plan;
    mutable Device::Buffer<int64_t> d_rp64;
    mutable Device::Buffer<int32_t> d_rp32;
    mutable Device::Buffer<int32_t> d_ci;
    mutable Device::Buffer<Scal> d_val;
    mutable Device::Buffer<Scal> d_ext;
    mutable Device::Buffer<Scal> d_x;
    mutable Device::Buffer<Scal> d_y;
};

}  // namespace Sparse
}  // namespace Tensors
