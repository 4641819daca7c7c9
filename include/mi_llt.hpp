/*
 * mi_llt.hpp -- C++ host class mirroring Eigen::LLT<Matrix<Scalar,Dynamic,Dynamic>, UpLo>
 * (Eigen/src/Cholesky/LLT.h:66-233) over the C ABI of mi_llt.h, so that C++ code written
 * against Eigen's LLT switches to the MI355X by changing the type name:
 *
 *     Eigen::LLT<Eigen::MatrixXd> llt(A);     ->     mi::LLT<double> llt(A);
 *     Eigen::MatrixXd X = llt.solve(B);               Eigen::MatrixXd X = llt.solve(B);
 *
 * Host matrices are any column-major dense type with rows(), cols(), data() and
 * outerStride() (Eigen::MatrixXd / MatrixXf and their Maps, or mi::HostMatrix below); results
 * are written into such a type (resize(rows, cols) is called when it has one).  The factor is
 * kept in device memory (the object's own copy, as LLT keeps m_matrix, LLT.h:438-440), so
 * repeated solves move only the right-hand sides.  Device-pointer variants take data that is
 * already in HBM.
 *
 * Semantics kept from the reference: compute() copies the matrix, records the L1 norm of the
 * symmetric view (m_l1_norm, LLT.h:442-453) and factors the UpLo triangle only; info() is
 * Success or NumericalIssue (the leading minor that is not positive definite, LLT.h:327-330);
 * solve() is forward then backward substitution (LLT.h:483-498); matrixLLT() is the whole
 * storage (the other triangle still holds the input, LLT.h:225-229); rcond() (LLT.h:166-171)
 * and rankUpdate(v, sigma) (LLT.h:467-480) as in the reference.  Misuse that Eigen catches
 * with eigen_assert (use before compute, size mismatch) throws std::logic_error here; a
 * failure of the device or the runtime throws mi::Error.
 *
 * Scalar = double: FP64 throughout.  Scalar = float: LLT<MatrixXf> (mi_spotrf / mi_spotrs);
 * rcond() and rankUpdate() are FP64 only.
 */
#ifndef MI_LLT_HPP
#define MI_LLT_HPP

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "mi_llt.h"

namespace mi {

// Eigen's ComputationInfo values (Eigen/src/Core/util/Constants.h:430-440) and UpLo flags.
enum ComputationInfo { Success = 0, NumericalIssue = 1, NoConvergence = 2, InvalidInput = 3 };
enum UpLoType { Lower = 0x1, Upper = 0x2 };

class Error : public std::runtime_error {
public:
    Error(const std::string &what, int code) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// A minimal column-major host matrix (for callers without Eigen).
template <typename Scalar>
class HostMatrix {
public:
    HostMatrix() = default;
    HostMatrix(int64_t r, int64_t c) : r_(r), c_(c), v_((size_t)(r * c)) {}
    int64_t rows() const { return r_; }
    int64_t cols() const { return c_; }
    int64_t outerStride() const { return r_ > 0 ? r_ : 1; }
    Scalar *data() { return v_.data(); }
    const Scalar *data() const { return v_.data(); }
    Scalar &operator()(int64_t i, int64_t j) { return v_[(size_t)(i + j * r_)]; }
    const Scalar &operator()(int64_t i, int64_t j) const { return v_[(size_t)(i + j * r_)]; }
    void resize(int64_t r, int64_t c)
    {
        r_ = r;
        c_ = c;
        v_.assign((size_t)(r * c), Scalar(0));
    }

private:
    int64_t r_ = 0, c_ = 0;
    std::vector<Scalar> v_;
};

namespace detail {

inline void check(int rc, mi_llt_handle h, const char *what)
{
    if (rc != 0) throw Error(std::string(what) + ": " + (h ? mi_llt_last_error(h) : "") + " (" + std::to_string(rc) + ")", rc);
}
inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e), MI_LLT_E_HIP);
}

// resize(rows, cols) when the output type has one
template <class M>
auto resize_if(M &m, int64_t r, int64_t c, int) -> decltype(m.resize(r, c), void())
{
    if ((int64_t)m.rows() != r || (int64_t)m.cols() != c) m.resize(r, c);
}
template <class M>
void resize_if(M &m, int64_t r, int64_t c, long)
{
    if ((int64_t)m.rows() != r || (int64_t)m.cols() != c) throw std::logic_error("mi::LLT: output has the wrong size");
}

inline int potrf(mi_llt_handle h, char u, int64_t n, double *a, int64_t ld, int *info) { return mi_dpotrf(h, u, n, a, ld, info); }
inline int potrf(mi_llt_handle h, char u, int64_t n, float *a, int64_t ld, int *info) { return mi_spotrf(h, u, n, a, ld, info); }
inline int potrs(mi_llt_handle h, char u, int64_t n, int64_t r, const double *a, int64_t lda, double *b, int64_t ldb)
{
    return mi_dpotrs(h, u, n, r, a, lda, b, ldb);
}
inline int potrs(mi_llt_handle h, char u, int64_t n, int64_t r, const float *a, int64_t lda, float *b, int64_t ldb)
{
    return mi_spotrs(h, u, n, r, a, lda, b, ldb);
}

}  // namespace detail

template <typename Scalar_ = double, int UpLo_ = Lower>
class LLT {
    static_assert(std::is_same<Scalar_, double>::value || std::is_same<Scalar_, float>::value,
                  "mi::LLT: double or float");
    static_assert(UpLo_ == Lower || UpLo_ == Upper, "mi::LLT: Lower or Upper");

public:
    typedef Scalar_ Scalar;
    typedef Scalar_ RealScalar;
    typedef int64_t Index;
    static constexpr int UpLo = UpLo_;

    // LLT.h:86-95: default construction and preallocation
    LLT() = default;
    explicit LLT(Index size) { reserve(size); }
    // LLT.h:104-110: construction computes
    template <class M>
    explicit LLT(const M &a)
    {
        compute(a);
    }
    ~LLT() { release(); }
    LLT(const LLT &) = delete;
    LLT &operator=(const LLT &) = delete;
    LLT(LLT &&o) noexcept { *this = std::move(o); }
    LLT &operator=(LLT &&o) noexcept
    {
        if (this != &o) {
            release();
            h_ = o.h_;
            d_ = o.d_;
            cap_ = o.cap_;
            n_ = o.n_;
            l1_ = o.l1_;
            init_ = o.init_;
            info_ = o.info_;
            stream_ = o.stream_;
            o.h_ = nullptr;
            o.d_ = nullptr;
            o.cap_ = 0;
            o.init_ = false;
        }
        return *this;
    }

    // All device work of this object goes to `stream` (the default: the null stream).
    void setStream(hipStream_t stream)
    {
        stream_ = stream;
        if (h_) detail::check(mi_llt_set_stream(h_, stream), h_, "mi_llt_set_stream");
    }

    // LLT.h:430-460 from a host matrix (copied to the object's device storage).
    template <class M>
    LLT &compute(const M &a)
    {
        const Index n = (Index)a.rows();
        if ((Index)a.cols() != n) throw std::logic_error("mi::LLT::compute: square matrix required");
        reserve(n);
        if (n > 0)
            detail::hip_check(hipMemcpy2DAsync(d_, (size_t)ld() * sizeof(Scalar), a.data(),
                                           (size_t)a.outerStride() * sizeof(Scalar), (size_t)n * sizeof(Scalar),
                                           (size_t)n, hipMemcpyHostToDevice, stream_),
                          "mi::LLT::compute (copy)");
        return factor();
    }

    // The same from a device matrix (n x n, leading dimension lda, column-major).
    LLT &computeDevice(const Scalar *dA, Index n, Index lda)
    {
        reserve(n);
        if (n > 0)
            detail::hip_check(hipMemcpy2DAsync(d_, (size_t)ld() * sizeof(Scalar), dA, (size_t)lda * sizeof(Scalar),
                                           (size_t)n * sizeof(Scalar), (size_t)n, hipMemcpyDeviceToDevice, stream_),
                          "mi::LLT::computeDevice (copy)");
        return factor();
    }

    // LLT.h:191-195
    ComputationInfo info() const
    {
        require_init();
        return info_;
    }
    Index rows() const { return n_; }
    Index cols() const { return n_; }
    const LLT &adjoint() const { return *this; }  // LLT.h:204

    // LLT.h:483-498: X = A^{-1} B for a host matrix B (a copy of B, solved).
    template <class M>
    M solve(const M &b) const
    {
        M x = b;
        solveInPlace(x);
        return x;
    }

    // LLT.h:510-522: B <- A^{-1} B in place (host matrix).
    template <class M>
    void solveInPlace(M &b) const
    {
        require_init();
        const Index r = (Index)b.cols();
        if ((Index)b.rows() != n_) throw std::logic_error("mi::LLT::solve: wrong number of rows");
        if (n_ == 0 || r == 0) return;
        Scalar *db = nullptr;
        detail::hip_check(hipMalloc(&db, (size_t)n_ * r * sizeof(Scalar)), "mi::LLT::solve (workspace)");
        struct Free {
            Scalar *p;
            ~Free() { (void)hipFree(p); }
        } guard{db};
        detail::hip_check(hipMemcpy2DAsync(db, (size_t)n_ * sizeof(Scalar), b.data(),
                                           (size_t)b.outerStride() * sizeof(Scalar), (size_t)n_ * sizeof(Scalar),
                                           (size_t)r, hipMemcpyHostToDevice, stream_),
                          "mi::LLT::solve (copy in)");
        solveInPlaceDevice(db, r, n_);
        detail::hip_check(hipMemcpy2DAsync(b.data(), (size_t)b.outerStride() * sizeof(Scalar), db,
                                           (size_t)n_ * sizeof(Scalar), (size_t)n_ * sizeof(Scalar), (size_t)r,
                                           hipMemcpyDeviceToHost, stream_),
                          "mi::LLT::solve (copy out)");
        detail::hip_check(hipStreamSynchronize(stream_), "mi::LLT::solve");
    }

    // B <- A^{-1} B for a device matrix (n x nrhs, ld ldb); waits for the stream and checks
    // the device status.
    void solveInPlaceDevice(Scalar *dB, Index nrhs, Index ldb) const
    {
        require_init();
        detail::check(detail::potrs(h_, uplo(), n_, nrhs, d_, ld(), dB, ldb), h_, "mi::LLT::solve");
        int st = 0;
        detail::check(mi_llt_status(h_, &st), h_, "mi_llt_status");
        if (st) throw Error("mi::LLT::solve: device status " + std::to_string(st), MI_LLT_E_DEVICE);
    }

    // LLT.h:178-183: the whole storage (the other triangle still holds the input).
    template <class M>
    void matrixLLT(M &out) const
    {
        require_init();
        detail::resize_if(out, n_, n_, 0);
        if (n_ == 0) return;
        detail::hip_check(hipMemcpy2DAsync(out.data(), (size_t)out.outerStride() * sizeof(Scalar), d_,
                                           (size_t)ld() * sizeof(Scalar), (size_t)n_ * sizeof(Scalar), (size_t)n_,
                                           hipMemcpyDeviceToHost, stream_),
                          "mi::LLT::matrixLLT");
        detail::hip_check(hipStreamSynchronize(stream_), "mi::LLT::matrixLLT");
    }

    // LLT.h:128-139, 401-419: the triangular factors, materialised (zeros in the other triangle,
    // as assigning Eigen's TriangularView to a dense matrix does).
    template <class M>
    void matrixL(M &out) const
    {
        matrixLLT(out);
        for (Index j = 0; j < n_; ++j)
            for (Index i = 0; i < n_; ++i) {
                Scalar &e = out.data()[i + j * (Index)out.outerStride()];
                if (UpLo == (int)Lower) {
                    if (i < j) e = Scalar(0);
                } else if (i > j) {
                    e = Scalar(0);
                }
            }
        if (UpLo == (int)Upper) transpose_square(out);
    }
    template <class M>
    void matrixU(M &out) const
    {
        matrixL(out);
        transpose_square(out);
    }

    // LLT.h:166-171
    RealScalar rcond() const
    {
        static_assert(std::is_same<Scalar, double>::value, "mi::LLT::rcond: double only");
        require_init();
        if (info_ != Success) throw std::logic_error("LLT failed because matrix appears to be negative");
        double r = 0.0;
        detail::check(mi_dpocon(h_, uplo(), n_, d_, ld(), l1_, &r), h_, "mi_dpocon");
        return r;
    }

    // LLT.h:467-480 with a host vector (size() == rows(), contiguous data()).
    template <class V>
    LLT &rankUpdate(const V &v, const RealScalar &sigma = 1)
    {
        static_assert(std::is_same<Scalar, double>::value, "mi::LLT::rankUpdate: double only");
        require_init();
        if ((Index)v.size() != n_) throw std::logic_error("mi::LLT::rankUpdate: vector of size rows() required");
        if (n_ == 0) return *this;
        double *dv = nullptr;
        int *dinfo = nullptr;
        detail::hip_check(hipMalloc(&dv, (size_t)n_ * sizeof(double) + 256), "mi::LLT::rankUpdate (workspace)");
        struct Free {
            double *p;
            ~Free() { (void)hipFree(p); }
        } guard{dv};
        dinfo = (int *)(dv + n_);
        detail::hip_check(hipMemcpyAsync(dv, v.data(), (size_t)n_ * sizeof(double), hipMemcpyHostToDevice, stream_),
                          "mi::LLT::rankUpdate (copy)");
        detail::check(mi_dllt_rank_update(h_, uplo(), n_, d_, ld(), dv, 1, sigma, dinfo), h_, "mi_dllt_rank_update");
        int info = 0;
        detail::hip_check(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, stream_),
                          "mi::LLT::rankUpdate (info)");
        int st = 0;
        detail::check(mi_llt_status(h_, &st), h_, "mi_llt_status");
        if (st) throw Error("mi::LLT::rankUpdate: device status " + std::to_string(st), MI_LLT_E_DEVICE);
        info_ = info == 0 ? Success : NumericalIssue;
        return *this;
    }

    // m_l1_norm (LLT.h:442-453)
    RealScalar l1Norm() const
    {
        require_init();
        return (RealScalar)l1_;
    }
    // The device storage (n x n, leading dimension rows()) and the library handle.
    const Scalar *deviceData() const { return d_; }
    mi_llt_handle handle() const { return h_; }

private:
    char uplo() const { return UpLo == (int)Lower ? 'L' : 'U'; }
    Index ld() const { return n_ > 0 ? n_ : 1; }
    void require_init() const
    {
        if (!init_) throw std::logic_error("LLT is not initialized.");
    }

    void reserve(Index n)
    {
        if (!h_) {
            detail::check(mi_llt_create(&h_), nullptr, "mi_llt_create");
            detail::check(mi_llt_set_stream(h_, stream_), h_, "mi_llt_set_stream");
        }
        const size_t need = (size_t)(n > 0 ? n : 1) * (size_t)(n > 0 ? n : 1) * sizeof(Scalar);
        if (need > cap_) {
            detail::hip_check(hipStreamSynchronize(stream_), "mi::LLT (sync)");
            if (d_) (void)hipFree(d_);
            d_ = nullptr;
            cap_ = 0;
            detail::hip_check(hipMalloc(&d_, need), "mi::LLT (storage)");
            cap_ = need;
        }
        n_ = n;
    }

    // m_l1_norm of the input in the device storage: the device kernel for FP64 storage
    double l1_norm_of(const double *dA)
    {
        double *dn = nullptr;
        detail::hip_check(hipMalloc(&dn, sizeof(double)), "mi::LLT (norm)");
        struct Free {
            double *p;
            ~Free() { (void)hipFree(p); }
        } guard{dn};
        detail::check(mi_dsy_l1norm(h_, uplo(), n_, dA, ld(), dn), h_, "mi_dsy_l1norm");
        double v = 0.0;
        detail::hip_check(hipMemcpyAsync(&v, dn, sizeof(double), hipMemcpyDeviceToHost, stream_), "l1 norm");
        detail::hip_check(hipStreamSynchronize(stream_), "l1 norm");
        return v;
    }
    double l1_norm_of(const float *) { return host_l1_norm(); }

    LLT &factor()
    {
        l1_ = l1_norm_of(d_);
        int info = 0;
        detail::check(detail::potrf(h_, uplo(), n_, d_, ld(), &info), h_, "mi::LLT::compute");
        info_ = info == 0 ? Success : NumericalIssue;
        init_ = true;
        return *this;
    }

    // FP32 storage: the norm of the symmetric view from a host copy of the stored triangle.
    double host_l1_norm()
    {
        std::vector<Scalar> h((size_t)n_ * n_);
        if (n_ == 0) return 0.0;
        detail::hip_check(hipMemcpyAsync(h.data(), d_, h.size() * sizeof(Scalar), hipMemcpyDeviceToHost, stream_),
                          "mi::LLT (norm copy)");
        detail::hip_check(hipStreamSynchronize(stream_), "mi::LLT (norm)");
        double best = 0.0;
        for (Index c = 0; c < n_; ++c) {
            double s = 0.0;
            for (Index i = 0; i < n_; ++i) {
                const bool stored = UpLo == (int)Lower ? i >= c : i <= c;
                s += std::fabs((double)(stored ? h[(size_t)(i + c * n_)] : h[(size_t)(c + i * n_)]));
            }
            best = s > best ? s : best;
        }
        return best;
    }

    template <class M>
    void transpose_square(M &m) const
    {
        const Index ldm = (Index)m.outerStride();
        for (Index j = 0; j < n_; ++j)
            for (Index i = j + 1; i < n_; ++i) std::swap(m.data()[i + j * ldm], m.data()[j + i * ldm]);
    }

    void release()
    {
        if (h_) {
            (void)hipStreamSynchronize(stream_);
            mi_llt_destroy(h_);
        }
        if (d_) (void)hipFree(d_);
        h_ = nullptr;
        d_ = nullptr;
        cap_ = 0;
    }

    mi_llt_handle h_ = nullptr;
    Scalar *d_ = nullptr;
    size_t cap_ = 0;
    Index n_ = 0;
    double l1_ = 0.0;
    bool init_ = false;
    ComputationInfo info_ = Success;
    hipStream_t stream_ = nullptr;
};

}  // namespace mi

#endif  // MI_LLT_HPP
