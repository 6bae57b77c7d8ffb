/*
 * pyext.cpp — the `_curfp_C` Python extension: a thin pybind11 layer over
 * the C ABI in include/curfp.h, with the surface of the reference's
 * curfp/_ext.cpp:672-878: Handle with set_stream_ptr, one function per C
 * entry under the reference's name and keyword names (handle, transr, uplo,
 * ..., A_ptr, ...), tensors passed as integer device pointers, enums as ints
 * with the same module constants, a non-zero status raised as
 * RuntimeError("curfp error: status=N").
 */
#include <pybind11/pybind11.h>
#include <cstdint>
#include <stdexcept>
#include <string>
#include "../../include/curfp.h"

namespace py = pybind11;

static void check_status(curfpStatus_t s)
{
    if (s != CURFP_STATUS_SUCCESS)
        throw std::runtime_error("curfp error: status=" + std::to_string((int)s));
}

class Handle {
    curfpHandle_t h_ = nullptr;

public:
    Handle() { check_status(curfpCreate(&h_)); }
    ~Handle() { if (h_) curfpDestroy(h_); }
    Handle(const Handle &) = delete;
    Handle &operator=(const Handle &) = delete;

    void set_stream_ptr(int64_t ptr) { check_status(curfpSetStream(h_, reinterpret_cast<hipStream_t>(ptr))); }
    void synchronize() { check_status(curfpSynchronize(h_)); }
    void set_sweep_wait_limit(int polls) { check_status(curfpSetSweepWaitLimit(h_, polls)); }
    int64_t get_stream_ptr() const
    {
        hipStream_t s = nullptr;
        check_status(curfpGetStream(h_, &s));
        return reinterpret_cast<int64_t>(s);
    }
    void set_math_mode(int mode) { check_status(curfpSetMathMode(h_, static_cast<curfpMathMode_t>(mode))); }
    int get_math_mode() const
    {
        curfpMathMode_t m = CURFP_MATH_FP32;
        check_status(curfpGetMathMode(h_, &m));
        return (int)m;
    }
    size_t workspace_bytes() const
    {
        size_t b = 0;
        check_status(curfpGetWorkspaceSize(h_, &b));
        return b;
    }
    void set_profiling(bool on) { check_status(curfpSetProfiling(h_, on ? 1 : 0)); }
    py::dict kernel_stats(int cls, bool reset)
    {
        int64_t launches = 0;
        double flops = 0, bytes = 0, ms = 0;
        check_status(curfpGetKernelStats(h_, static_cast<curfpKernelClass_t>(cls), &launches, &flops, &bytes, &ms,
                                         reset ? 1 : 0));
        py::dict d;
        d["launches"] = launches;
        d["flops"] = flops;
        d["bytes"] = bytes;
        d["ms"] = ms;
        return d;
    }
    curfpHandle_t get() const { return h_; }
};

template <typename T> static T *P(int64_t p) { return reinterpret_cast<T *>(p); }
#define OPS(x) static_cast<curfpOperation_t>(x)
#define FILL(x) static_cast<curfpFillMode_t>(x)

PYBIND11_MODULE(_curfp_C, m)
{
    m.doc() = "curfp C ABI for MI355X (gfx950), HIP kernels";

    py::class_<Handle>(m, "Handle")
        .def(py::init<>())
        .def("set_stream_ptr", &Handle::set_stream_ptr, py::arg("ptr"),
             "Set the HIP stream from its integer value (stream.cuda_stream).")
        .def("synchronize", &Handle::synchronize,
             "Wait for the handle's stream; raises if an earlier asynchronous call failed on the device.")
        .def("set_sweep_wait_limit", &Handle::set_sweep_wait_limit, py::arg("polls"))
        .def("get_stream_ptr", &Handle::get_stream_ptr)
        .def("set_math_mode", &Handle::set_math_mode)
        .def("get_math_mode", &Handle::get_math_mode)
        .def("set_profiling", &Handle::set_profiling)
        .def("workspace_bytes", &Handle::workspace_bytes)
        .def("kernel_stats", &Handle::kernel_stats, py::arg("cls"), py::arg("reset") = false)
        .def("__enter__", [](Handle &h) -> Handle & { return h; }, py::return_value_policy::reference)
        .def("__exit__", [](Handle &, py::object, py::object, py::object) {});

    m.attr("OP_N") = (int)CURFP_OP_N;
    m.attr("OP_T") = (int)CURFP_OP_T;
    m.attr("FILL_LOWER") = (int)CURFP_FILL_MODE_LOWER;
    m.attr("FILL_UPPER") = (int)CURFP_FILL_MODE_UPPER;
    m.attr("NORM_MAX") = (int)CURFP_NORM_MAX;
    m.attr("NORM_ONE") = (int)CURFP_NORM_ONE;
    m.attr("NORM_FRO") = (int)CURFP_NORM_FRO;
    m.attr("SIDE_LEFT") = (int)CURFP_SIDE_LEFT;
    m.attr("SIDE_RIGHT") = (int)CURFP_SIDE_RIGHT;
    m.attr("MATH_FP32") = (int)CURFP_MATH_FP32;
    m.attr("MATH_BF16") = (int)CURFP_MATH_BF16;
    m.attr("MATH_BF16X3") = (int)CURFP_MATH_BF16X3;
    py::dict kc;
    kc["gemm_f32"] = (int)CURFP_KERNEL_GEMM_F32;
    kc["gemm_bf16"] = (int)CURFP_KERNEL_GEMM_BF16;
    kc["gemm_f64"] = (int)CURFP_KERNEL_GEMM_F64;
    kc["potrf_tile"] = (int)CURFP_KERNEL_POTRF_TILE;
    kc["trtri"] = (int)CURFP_KERNEL_TRTRI;
    kc["pack"] = (int)CURFP_KERNEL_PACK;
    kc["diag"] = (int)CURFP_KERNEL_DIAG;
    kc["other"] = (int)CURFP_KERNEL_OTHER;
    kc["trsv"] = (int)CURFP_KERNEL_TRSV;
    kc["trsm"] = (int)CURFP_KERNEL_TRSM;
    m.attr("KERNEL_CLASSES") = kc;

    /* ---- FP32 ---- */
    m.def("ssfrk", [](Handle &h, int transr, int uplo, int trans, int n, int k, float alpha, int64_t A, int lda,
                      float beta, int64_t C) {
        check_status(curfpSsfrk(h.get(), OPS(transr), FILL(uplo), OPS(trans), n, k, &alpha, P<const float>(A), lda,
                                &beta, P<float>(C)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("trans"), py::arg("n"), py::arg("k"),
       py::arg("alpha"), py::arg("A_ptr"), py::arg("lda"), py::arg("beta"), py::arg("C_ptr"));
    m.def("ssfrk_bf16", [](Handle &h, int transr, int uplo, int trans, int n, int k, float alpha, int64_t A,
                           int lda, float beta, int64_t C) {
        check_status(curfpSsfrkBf16(h.get(), OPS(transr), FILL(uplo), OPS(trans), n, k, &alpha,
                                    P<const uint16_t>(A), lda, &beta, P<float>(C)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("trans"), py::arg("n"), py::arg("k"),
       py::arg("alpha"), py::arg("A_ptr"), py::arg("lda"), py::arg("beta"), py::arg("C_ptr"));
    m.def("spftrf", [](Handle &h, int transr, int uplo, int n, int64_t A) {
        int info = 0;
        check_status(curfpSpftrf(h.get(), OPS(transr), FILL(uplo), n, P<float>(A), &info));
        return info;
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("A_ptr"));
    m.def("spftrs", [](Handle &h, int transr, int uplo, int n, int nrhs, int64_t A, int64_t B, int ldb) {
        check_status(curfpSpftrs(h.get(), OPS(transr), FILL(uplo), n, nrhs, P<const float>(A), P<float>(B), ldb));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("nrhs"), py::arg("A_ptr"),
       py::arg("B_ptr"), py::arg("ldb"));
    m.def("spftrs_rm", [](Handle &h, int transr, int uplo, int n, int nrhs, int64_t A, int64_t B, int ldb) {
        check_status(
            curfpSpftrsRowMajor(h.get(), OPS(transr), FILL(uplo), n, nrhs, P<const float>(A), P<float>(B), ldb));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("nrhs"), py::arg("A_ptr"),
       py::arg("B_ptr"), py::arg("ldb"));
    m.def("strttf", [](Handle &h, int transr, int uplo, int n, int64_t A, int lda, int64_t arf) {
        check_status(curfpSstrttf(h.get(), OPS(transr), FILL(uplo), n, P<const float>(A), lda, P<float>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("A_ptr"), py::arg("lda"),
       py::arg("arf_ptr"));
    m.def("stfttr", [](Handle &h, int transr, int uplo, int n, int64_t arf, int64_t A, int lda) {
        check_status(curfpSstfttr(h.get(), OPS(transr), FILL(uplo), n, P<const float>(arf), P<float>(A), lda));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("arf_ptr"), py::arg("A_ptr"),
       py::arg("lda"));
    m.def("sadd_to_diagonal", [](Handle &h, int transr, int uplo, int n, float value, int64_t arf) {
        check_status(curfpSaddToDiagonal(h.get(), OPS(transr), FILL(uplo), n, &value, P<float>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("value"), py::arg("arf_ptr"));
    m.def("sgen_residual", [](Handle &h, int trans, int n, int k, int nrhs, double alpha, int64_t A, int lda,
                              double lambda, int64_t X, int ldx, int64_t B, int ldb, int64_t R, int ldr) {
        check_status(curfpSgenResidual(h.get(), OPS(trans), n, k, nrhs, &alpha, P<const float>(A), lda, &lambda,
                                       P<const double>(X), ldx, P<const double>(B), ldb, P<double>(R), ldr));
    }, py::arg("handle"), py::arg("trans"), py::arg("n"), py::arg("k"), py::arg("nrhs"), py::arg("alpha"),
       py::arg("A_ptr"), py::arg("lda"), py::arg("lam"), py::arg("X_ptr"), py::arg("ldx"), py::arg("B_ptr"),
       py::arg("ldb"), py::arg("R_ptr"), py::arg("ldr"));

    m.def("spftrs_refine", [](Handle &h, int transr, int uplo, int trans, int n, int k, int nrhs, double alpha,
                              int64_t A, int lda, double lambda, int64_t F, int64_t B, int ldb, int64_t X, int ldx,
                              double tol, int max_iter) {
        int iters = 0;
        double resid = 0.0;
        check_status(curfpSpftrsRefine(h.get(), OPS(transr), FILL(uplo), OPS(trans), n, k, nrhs, &alpha,
                                       P<const float>(A), lda, &lambda, P<const float>(F), P<const double>(B), ldb,
                                       P<double>(X), ldx, tol, max_iter, &iters, &resid));
        return py::make_tuple(iters, resid);
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("trans"), py::arg("n"), py::arg("k"),
       py::arg("nrhs"), py::arg("alpha"), py::arg("A_ptr"), py::arg("lda"), py::arg("lam"), py::arg("F_ptr"),
       py::arg("B_ptr"), py::arg("ldb"), py::arg("X_ptr"), py::arg("ldx"), py::arg("tol"), py::arg("max_iter"));

    m.def("sgemm_strided", [](Handle &h, int M, int N, int K, float alpha, int64_t A, int64_t a_rs, int64_t a_cs,
                              int64_t B, int64_t b_rs, int64_t b_cs, float beta, int64_t C, int64_t c_rs,
                              int64_t c_cs) {
        check_status(curfpSgemmStrided(h.get(), M, N, K, &alpha, P<const float>(A), a_rs, a_cs, P<const float>(B),
                                       b_rs, b_cs, &beta, P<float>(C), c_rs, c_cs));
    }, py::arg("handle"), py::arg("m"), py::arg("n"), py::arg("k"), py::arg("alpha"), py::arg("A_ptr"),
       py::arg("a_rs"), py::arg("a_cs"), py::arg("B_ptr"), py::arg("b_rs"), py::arg("b_cs"), py::arg("beta"),
       py::arg("C_ptr"), py::arg("c_rs"), py::arg("c_cs"));
    m.def("dgemm_strided", [](Handle &h, int M, int N, int K, double alpha, int64_t A, int64_t a_rs, int64_t a_cs,
                              int64_t B, int64_t b_rs, int64_t b_cs, double beta, int64_t C, int64_t c_rs,
                              int64_t c_cs) {
        check_status(curfpDgemmStrided(h.get(), M, N, K, &alpha, P<const double>(A), a_rs, a_cs, P<const double>(B),
                                       b_rs, b_cs, &beta, P<double>(C), c_rs, c_cs));
    }, py::arg("handle"), py::arg("m"), py::arg("n"), py::arg("k"), py::arg("alpha"), py::arg("A_ptr"),
       py::arg("a_rs"), py::arg("a_cs"), py::arg("B_ptr"), py::arg("b_rs"), py::arg("b_cs"), py::arg("beta"),
       py::arg("C_ptr"), py::arg("c_rs"), py::arg("c_cs"));

    /* ---- FP64 ---- */
    m.def("dsfrk", [](Handle &h, int transr, int uplo, int trans, int n, int k, double alpha, int64_t A, int lda,
                      double beta, int64_t C) {
        check_status(curfpDsfrk(h.get(), OPS(transr), FILL(uplo), OPS(trans), n, k, &alpha, P<const double>(A), lda,
                                &beta, P<double>(C)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("trans"), py::arg("n"), py::arg("k"),
       py::arg("alpha"), py::arg("A_ptr"), py::arg("lda"), py::arg("beta"), py::arg("C_ptr"));
    m.def("dpftrf", [](Handle &h, int transr, int uplo, int n, int64_t A) {
        int info = 0;
        check_status(curfpDpftrf(h.get(), OPS(transr), FILL(uplo), n, P<double>(A), &info));
        return info;
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("A_ptr"));
    m.def("dpftrs", [](Handle &h, int transr, int uplo, int n, int nrhs, int64_t A, int64_t B, int ldb) {
        check_status(curfpDpftrs(h.get(), OPS(transr), FILL(uplo), n, nrhs, P<const double>(A), P<double>(B), ldb));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("nrhs"), py::arg("A_ptr"),
       py::arg("B_ptr"), py::arg("ldb"));
    m.def("dpftrs_rm", [](Handle &h, int transr, int uplo, int n, int nrhs, int64_t A, int64_t B, int ldb) {
        check_status(
            curfpDpftrsRowMajor(h.get(), OPS(transr), FILL(uplo), n, nrhs, P<const double>(A), P<double>(B), ldb));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("nrhs"), py::arg("A_ptr"),
       py::arg("B_ptr"), py::arg("ldb"));
    m.def("dstrttf", [](Handle &h, int transr, int uplo, int n, int64_t A, int lda, int64_t arf) {
        check_status(curfpDstrttf(h.get(), OPS(transr), FILL(uplo), n, P<const double>(A), lda, P<double>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("A_ptr"), py::arg("lda"),
       py::arg("arf_ptr"));
    m.def("dstfttr", [](Handle &h, int transr, int uplo, int n, int64_t arf, int64_t A, int lda) {
        check_status(curfpDstfttr(h.get(), OPS(transr), FILL(uplo), n, P<const double>(arf), P<double>(A), lda));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("arf_ptr"), py::arg("A_ptr"),
       py::arg("lda"));
    m.def("dadd_to_diagonal", [](Handle &h, int transr, int uplo, int n, double value, int64_t arf) {
        check_status(curfpDaddToDiagonal(h.get(), OPS(transr), FILL(uplo), n, &value, P<double>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("value"), py::arg("arf_ptr"));
    m.def("ssfmv", [](Handle &h, int transr, int uplo, int n, float alpha, int64_t arf, int64_t x, int incx,
                      float beta, int64_t y, int incy) {
        check_status(curfpSsfmv(h.get(), OPS(transr), FILL(uplo), n, &alpha, P<const float>(arf), P<const float>(x),
                                incx, &beta, P<float>(y), incy));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("alpha"), py::arg("arf_ptr"),
       py::arg("x_ptr"), py::arg("incx"), py::arg("beta"), py::arg("y_ptr"), py::arg("incy"));
    m.def("dsfmv", [](Handle &h, int transr, int uplo, int n, double alpha, int64_t arf, int64_t x, int incx,
                      double beta, int64_t y, int incy) {
        check_status(curfpDsfmv(h.get(), OPS(transr), FILL(uplo), n, &alpha, P<const double>(arf),
                                P<const double>(x), incx, &beta, P<double>(y), incy));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("alpha"), py::arg("arf_ptr"),
       py::arg("x_ptr"), py::arg("incx"), py::arg("beta"), py::arg("y_ptr"), py::arg("incy"));
    m.def("spfcon", [](Handle &h, int transr, int uplo, int n, int64_t arf, float anorm) {
        float r = 0.f;
        check_status(curfpSpfcon(h.get(), OPS(transr), FILL(uplo), n, P<const float>(arf), anorm, &r));
        return r;
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("arf_ptr"), py::arg("anorm"));
    m.def("dpfcon", [](Handle &h, int transr, int uplo, int n, int64_t arf, double anorm) {
        double r = 0.0;
        check_status(curfpDpfcon(h.get(), OPS(transr), FILL(uplo), n, P<const double>(arf), anorm, &r));
        return r;
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("arf_ptr"), py::arg("anorm"));
    m.def("ssfr", [](Handle &h, int transr, int uplo, int n, float alpha, int64_t x, int incx, int64_t arf) {
        check_status(curfpSsfr(h.get(), OPS(transr), FILL(uplo), n, &alpha, P<const float>(x), incx, P<float>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("alpha"), py::arg("x_ptr"),
       py::arg("incx"), py::arg("arf_ptr"));
    m.def("dsfr", [](Handle &h, int transr, int uplo, int n, double alpha, int64_t x, int incx, int64_t arf) {
        check_status(curfpDsfr(h.get(), OPS(transr), FILL(uplo), n, &alpha, P<const double>(x), incx, P<double>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("alpha"), py::arg("x_ptr"),
       py::arg("incx"), py::arg("arf_ptr"));
    m.def("ssfr2", [](Handle &h, int transr, int uplo, int n, float alpha, int64_t x, int incx, int64_t y, int incy,
                      int64_t arf) {
        check_status(curfpSsfr2(h.get(), OPS(transr), FILL(uplo), n, &alpha, P<const float>(x), incx,
                                P<const float>(y), incy, P<float>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("alpha"), py::arg("x_ptr"),
       py::arg("incx"), py::arg("y_ptr"), py::arg("incy"), py::arg("arf_ptr"));
    m.def("dsfr2", [](Handle &h, int transr, int uplo, int n, double alpha, int64_t x, int incx, int64_t y, int incy,
                      int64_t arf) {
        check_status(curfpDsfr2(h.get(), OPS(transr), FILL(uplo), n, &alpha, P<const double>(x), incx,
                                P<const double>(y), incy, P<double>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("alpha"), py::arg("x_ptr"),
       py::arg("incx"), py::arg("y_ptr"), py::arg("incy"), py::arg("arf_ptr"));
    m.def("ssfr2k", [](Handle &h, int transr, int uplo, int trans, int n, int k, float alpha, int64_t A, int lda,
                       int64_t B, int ldb, float beta, int64_t C) {
        check_status(curfpSsfr2k(h.get(), OPS(transr), FILL(uplo), OPS(trans), n, k, &alpha, P<const float>(A), lda,
                                 P<const float>(B), ldb, &beta, P<float>(C)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("trans"), py::arg("n"), py::arg("k"),
       py::arg("alpha"), py::arg("A_ptr"), py::arg("lda"), py::arg("B_ptr"), py::arg("ldb"), py::arg("beta"),
       py::arg("C_ptr"));
    m.def("dsfr2k", [](Handle &h, int transr, int uplo, int trans, int n, int k, double alpha, int64_t A, int lda,
                       int64_t B, int ldb, double beta, int64_t C) {
        check_status(curfpDsfr2k(h.get(), OPS(transr), FILL(uplo), OPS(trans), n, k, &alpha, P<const double>(A), lda,
                                 P<const double>(B), ldb, &beta, P<double>(C)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("trans"), py::arg("n"), py::arg("k"),
       py::arg("alpha"), py::arg("A_ptr"), py::arg("lda"), py::arg("B_ptr"), py::arg("ldb"), py::arg("beta"),
       py::arg("C_ptr"));
    m.def("spftri", [](Handle &h, int transr, int uplo, int n, int64_t arf) {
        check_status(curfpSpftri(h.get(), OPS(transr), FILL(uplo), n, P<float>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("arf_ptr"));
    m.def("dpftri", [](Handle &h, int transr, int uplo, int n, int64_t arf) {
        check_status(curfpDpftri(h.get(), OPS(transr), FILL(uplo), n, P<double>(arf)));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("arf_ptr"));
    m.def("ssfmm", [](Handle &h, int transr, int uplo, int side, int m, int n, float alpha, int64_t arf, int64_t B,
                      int ldb, float beta, int64_t C, int ldc) {
        check_status(curfpSsfmm(h.get(), OPS(transr), FILL(uplo), static_cast<curfpSideMode_t>(side), m, n, &alpha,
                                P<const float>(arf), P<const float>(B), ldb, &beta, P<float>(C), ldc));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("side"), py::arg("m"), py::arg("n"),
       py::arg("alpha"), py::arg("arf_ptr"), py::arg("B_ptr"), py::arg("ldb"), py::arg("beta"), py::arg("C_ptr"),
       py::arg("ldc"));
    m.def("dsfmm", [](Handle &h, int transr, int uplo, int side, int m, int n, double alpha, int64_t arf, int64_t B,
                      int ldb, double beta, int64_t C, int ldc) {
        check_status(curfpDsfmm(h.get(), OPS(transr), FILL(uplo), static_cast<curfpSideMode_t>(side), m, n, &alpha,
                                P<const double>(arf), P<const double>(B), ldb, &beta, P<double>(C), ldc));
    }, py::arg("handle"), py::arg("transr"), py::arg("uplo"), py::arg("side"), py::arg("m"), py::arg("n"),
       py::arg("alpha"), py::arg("arf_ptr"), py::arg("B_ptr"), py::arg("ldb"), py::arg("beta"), py::arg("C_ptr"),
       py::arg("ldc"));
    m.def("slansf", [](Handle &h, int norm, int transr, int uplo, int n, int64_t arf) {
        float r = 0.f;
        check_status(curfpSlansf(h.get(), static_cast<curfpNormType_t>(norm), OPS(transr), FILL(uplo), n,
                                 P<const float>(arf), &r));
        return r;
    }, py::arg("handle"), py::arg("norm"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("arf_ptr"));
    m.def("dlansf", [](Handle &h, int norm, int transr, int uplo, int n, int64_t arf) {
        double r = 0.0;
        check_status(curfpDlansf(h.get(), static_cast<curfpNormType_t>(norm), OPS(transr), FILL(uplo), n,
                                 P<const double>(arf), &r));
        return r;
    }, py::arg("handle"), py::arg("norm"), py::arg("transr"), py::arg("uplo"), py::arg("n"), py::arg("arf_ptr"));
}
