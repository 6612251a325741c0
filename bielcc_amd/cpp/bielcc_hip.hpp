// Drop-in C++ host layer: BIELcc's public names and signatures (namespace
// bielcc) over the C ABI of include/bielcc_hip.h.  Every computation runs on
// the GPU through the ABI; this header only marshals arguments.
//
// Reference interfaces kept (file:line of the reference):
//   NumParam                  electrodynamics/num_param.h:28-55
//   KernelParam<KType>        integrators/kernel_param.h:25-57
//   IntegralParam             integrators/integral_param.h:16-45
//   F_* integrands            electrodynamics/kernel_lib.h:62-492 (objects here, callable)
//   IntegralUniversalPnt      integrators/integral_universal.h:38-220
//   IntegralUniversalSegPnt   integrators/integral_universal_seg.h:38-98
//   K0, K0_Common, K0_GradSimplePot_H_HS   electrodynamics/K0.h:36-249
//   integral1Divr, integralGrad1DivrSeg, IntegralSurf_SimplePot_H_HS,
//   IntegralSeg_GradSimplePot_H_HS         integrators/integrals_analytic.{h,cpp}
//   K_RotRot_*, R_Rot_*       electrodynamics/K.h:47-301, R.h:46-175
//   CellQuad, CellTriangle    mesh/cell.{h,cpp}
//   SurfMesh<CellType>        mesh/surf_mesh.h (.dat / .vtk readers, accessors)
//   GetFieldE_Ideal           electrodynamics/get_field.h:35-78
//   SurfaceCurrents_byBasis   electrodynamics/currents.h:22-45
//   PlaneWaveFieldE           electrodynamics/incident_field.h:20-29
//   PlaneWaveFieldH, MaterialParam (its permeability)  incident_field.h:40-59, material.h:73-190
//   IntegralUniversal (F(y) forms)         integrators/integral_universal.h:252-431
//   IntegralUniversalSegPnt (F(y) form)    integrators/integral_universal_seg.h:120-179
// Semantics kept: outputs are overwritten; KernelParam& is mutated (smoothR)
// exactly as the reference's integrators do; NaN/Inf propagate.  Errors of the
// GPU layer throw std::runtime_error (the reference throws only from IO).
#pragma once

#include "bielcc_hip.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <variant>
#include <vector>

namespace bielcc {

namespace Tolerance_Constants {
const double MACHINE_ZERO = std::numeric_limits<double>::epsilon();
const double POINT_TOLERANCE = 1e-5;
const double POINT_TOLERANCE_SQ = POINT_TOLERANCE * POINT_TOLERANCE;
}  // namespace Tolerance_Constants

namespace Calculation_Constants {
const double MACHINE_ZERO = std::numeric_limits<double>::epsilon();
const double INTEGRAL_ACCURACY = 1e-6;
const double ANALYTIC_CALC_DIST = 3.0;
const double SMOOTHING_DIST_SURF_INT = 2.;
const double SMOOTHING_DIST_SEG_INT = 2.;
const int START_CELL_SPLIT = 8;
const int START_SEG_SPLIT = 10;
const int PMAX_CELL_SPLIT = 1;
const int PMAX_SEG_SPLIT = 1;
}  // namespace Calculation_Constants

struct NumParam {
    double eps = 0., rs = 0., rs_seg = 0.;
    double analytic_calc = 0.;
    int n_start = 0, n_start_seg = 0;
    int p_max = 0, p_max_seg = 0;
    int k = 0;
    double T = 0., dt = 0.;
    double kappa = 0., M = 0.;
    NumParam() = default;
    NumParam(double _eps, double _rs, double _rs_seg, double _analytic_calc, int _n_start, int _n_start_seg,
             int _p_max, int _p_max_seg)
        : eps(_eps), rs(_rs), rs_seg(_rs_seg), analytic_calc(_analytic_calc), n_start(_n_start),
          n_start_seg(_n_start_seg), p_max(_p_max), p_max_seg(_p_max_seg) {}
};

template <typename KType>
struct KernelParam {
    double smoothR = 0.;
    double smoothRMin = 0.;
    KType k{};
    std::complex<double> vec_cmplx[3]{};
    double vec_dbl[3]{};
    double n[3]{};
    double ort[3]{};
    double e0[3]{};
    std::complex<double> current[3]{};
    double freq = 0.;
    KernelParam() = default;
    KernelParam(KType k_, double smoothR_) : smoothR(smoothR_), smoothRMin(0.), k(k_) {}
    KernelParam(KType k_, double smoothR_, double smoothRMin_) : smoothR(smoothR_), smoothRMin(smoothRMin_), k(k_) {}
};

class IntegralParam {
    int _idim, _NStart, _PMax;
    double _epsAccuracy;

public:
    IntegralParam() : _idim(1), _NStart(8), _PMax(1), _epsAccuracy(0.0) {}
    IntegralParam(int idim, int NStart, int PMax, double epsAccuracy)
        : _idim(idim), _NStart(NStart), _PMax(PMax), _epsAccuracy(epsAccuracy) {}
    int GetIDim() const { return _idim; }
    int GetNStart() const { return _NStart; }
    int GetPMax() const { return _PMax; }
    double GetEpsAccur() const { return _epsAccuracy; }
    void SetNStart(int n) { _NStart = n; }
    void SetPMax(int p) { _PMax = p; }
    void SetEpsAccur(double a) { _epsAccuracy = a; }
};

// ------------------------------------------------------------------ GPU plumbing
namespace hip {

inline void check(bielcc_ctx ctx, int rc, const char* what) {
    if (rc == BIELCC_OK) return;
    char buf[512] = {0};
    if (ctx) bielcc_last_error(ctx, buf, sizeof buf);
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + buf);
}

// One context per process by default (device from BIELCC_DEVICE, else 0).
class Context {
public:
    explicit Context(int device = 0) { check(nullptr, bielcc_ctx_create(device, &ctx_), "bielcc_ctx_create"); }
    ~Context() { if (ctx_) bielcc_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    bielcc_ctx get() const { return ctx_; }
    void sync() { check(ctx_, bielcc_sync(ctx_), "bielcc_sync"); }

private:
    bielcc_ctx ctx_ = nullptr;
};

inline Context& default_context() {
    static Context ctx([] {
        const char* d = std::getenv("BIELCC_DEVICE");
        return d ? std::atoi(d) : 0;
    }());
    return ctx;
}

// A device buffer of doubles owned by the default context.
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n_doubles) : n_(n_doubles) {
        auto& c = default_context();
        check(c.get(), bielcc_malloc(c.get(), std::max<size_t>(n_, 1) * sizeof(double), &p_), "bielcc_malloc");
    }
    DevBuf(const double* host, size_t n) : DevBuf(n) { upload(host, n); }
    ~DevBuf() { if (p_) bielcc_free(default_context().get(), p_); }
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; }
    DevBuf& operator=(DevBuf&& o) noexcept { std::swap(p_, o.p_); std::swap(n_, o.n_); return *this; }
    DevBuf(const DevBuf&) = delete;
    double* ptr() const { return static_cast<double*>(p_); }
    void upload(const double* host, size_t n) {
        auto& c = default_context();
        check(c.get(), bielcc_memcpy(c.get(), p_, host, n * sizeof(double), 0), "bielcc_memcpy");
    }
    void download(double* host, size_t n) const {
        auto& c = default_context();
        check(c.get(), bielcc_memcpy(c.get(), host, p_, n * sizeof(double), 1), "bielcc_memcpy");
    }

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Staging of one single-function call: the inputs packed into one host
// array, one upload into a per-thread grow-only device buffer, the outputs
// after the inputs, one download.  A caller that loops over points does no
// device allocation per call (the reference functions are called per pair).
class Call {
public:
    size_t in(const double* h, size_t n) {
        const size_t o = host_.size();
        host_.insert(host_.end(), h, h + n);
        n_in_ = host_.size();
        return o;
    }
    size_t out(size_t n) {
        const size_t o = host_.size();
        host_.resize(o + n, 0.);
        return o;
    }
    // uploads the inputs; returns the device base (valid until the next call on this thread)
    double* upload() {
        auto& c = default_context();
        buf_ = &buffer(host_.size());
        check(c.get(), bielcc_memcpy(c.get(), buf_->p, host_.data(), n_in_ * sizeof(double), 0), "bielcc_memcpy");
        return static_cast<double*>(buf_->p);
    }
    // downloads everything after the inputs; returns the host copy
    const double* download() {
        auto& c = default_context();
        const size_t n = host_.size() - n_in_;
        if (n)
            check(c.get(), bielcc_memcpy(c.get(), host_.data() + n_in_, static_cast<double*>(buf_->p) + n_in_,
                                         n * sizeof(double), 1), "bielcc_memcpy");
        return host_.data();
    }

private:
    struct Buf {
        void* p = nullptr;
        size_t cap = 0;
        ~Buf() { if (p) bielcc_free(default_context().get(), p); }
    };
    static Buf& buffer(size_t n_doubles) {
        thread_local Buf b;
        const size_t need = std::max<size_t>(n_doubles, 1) * sizeof(double);
        if (b.cap < need) {
            auto& c = default_context();
            if (b.p) bielcc_free(c.get(), b.p);
            b.p = nullptr;
            b.cap = 0;
            const size_t cap = std::max<size_t>(need, 4096);
            check(c.get(), bielcc_malloc(c.get(), cap, &b.p), "bielcc_malloc");
            b.cap = cap;
        }
        return b;
    }
    std::vector<double> host_;
    size_t n_in_ = 0;
    Buf* buf_ = nullptr;
};

inline double kre(double k) { return k; }
inline double kim(double) { return 0.; }
inline double kre(const std::complex<double>& k) { return k.real(); }
inline double kim(const std::complex<double>& k) { return k.imag(); }

template <typename T>
inline void put_c(const T& v, double* o) { std::complex<double> z(v); o[0] = z.real(); o[1] = z.imag(); }
template <typename P>
inline P get_c(const double* o) {
    if constexpr (std::is_same<P, double>::value) return o[0];
    else return P(o[0], o[1]);
}

template <typename KType>
inline bielcc_kernel_param to_kp(const KernelParam<KType>& p) {
    bielcc_kernel_param k{};
    k.smoothR = p.smoothR; k.smoothRMin = p.smoothRMin; k.k_re = kre(p.k); k.k_im = kim(p.k);
    for (int i = 0; i < 3; i++) {
        k.vec_dbl[i] = p.vec_dbl[i];
        k.n[i] = p.n[i];
        k.current[2 * i] = p.current[i].real();
        k.current[2 * i + 1] = p.current[i].imag();
    }
    return k;
}

inline bielcc_num_param to_np(const NumParam& p) {
    bielcc_num_param n;
    n.eps = p.eps; n.rs = p.rs; n.rs_seg = p.rs_seg; n.analytic_calc = p.analytic_calc;
    n.n_start = p.n_start; n.n_start_seg = p.n_start_seg; n.p_max = p.p_max; n.p_max_seg = p.p_max_seg;
    return n;
}

template <size_t NV>
inline void flat(const double (&cell)[NV][3], double* o) {
    for (size_t v = 0; v < NV; v++) for (int d = 0; d < 3; d++) o[3 * v + d] = cell[v][d];
}

}  // namespace hip

// ------------------------------------------------------------------ integrands
// Each F_* of kernel_lib.h is an object here; passing it where the reference
// passes the function template selects the GPU integrand, and calling it
// evaluates the integrand at one (x, y) pair on the GPU.
template <int Id, bool Real>
struct Integrand {
    static constexpr int id = Id;
    static constexpr bool real = Real;
    template <typename KType, typename P>
    void operator()(const double* x, const double* y, const KernelParam<KType>& kp, P* res) const {
        auto& c = hip::default_context();
        hip::Call call;
        const size_t ix = call.in(x, 3), iy = call.in(y, 3), ir = call.out(6);
        double* d = call.upload();
        bielcc_kernel_param k = hip::to_kp(kp);
        hip::check(c.get(), bielcc_kernel_batch(c.get(), Id, 1, d + ix, d + iy, &k, d + ir), "bielcc_kernel_batch");
        const double* o = call.download() + ir;
        for (int g = 0; g < 3; g++) res[g] = hip::get_c<P>(o + 2 * g);
    }
};

inline constexpr Integrand<BIELCC_F_SIMPLEPOT_L, true> F_SimplePot_L{};
inline constexpr Integrand<BIELCC_F_GRADSIMPLEPOT_L, true> F_GradSimplePot_L{};
inline constexpr Integrand<BIELCC_F_DOUBLEPOT_L, true> F_DoublePot_L{};
inline constexpr Integrand<BIELCC_F_VECTORPOT_L, true> F_VectorPot_L{};
inline constexpr Integrand<BIELCC_F_SIMPLEPOT_H, false> F_SimplePot_H{};
inline constexpr Integrand<BIELCC_NUF_SIMPLEPOT_H, false> nuF_SimplePot_H{};
inline constexpr Integrand<BIELCC_F_SIMPLEPOT_HM1, false> F_SimplePot_Hm1{};
inline constexpr Integrand<BIELCC_F_GRADSIMPLEPOT_H, false> F_GradSimplePot_H{};
inline constexpr Integrand<BIELCC_F_GRADSIMPLEPOT_HM1, false> F_GradSimplePot_Hm1{};
inline constexpr Integrand<BIELCC_F_DOUBLEPOT_H, false> F_DoublePot_H{};
inline constexpr Integrand<BIELCC_NF_DOUBLEPOT_H, false> nF_DoublePot_H{};
inline constexpr Integrand<BIELCC_F_ROTROTSIMPLEPOT_H, false> F_RotRotSimplePot_H{};

// ------------------------------------------------------------------ integrators
namespace detail {
template <typename P, typename KType, int Id, bool R>
void integral(int nv, const double* cellv, const double* x, Integrand<Id, R>, KernelParam<KType>& kp,
              const IntegralParam& ip, P* res) {
    auto& c = hip::default_context();
    hip::Call call;
    const size_t ic = call.in(cellv, (size_t)3 * (nv == 2 ? 2 : nv)), ix = call.in(x, 3);
    const size_t ir = call.out(6), is = call.out(1);
    double* d = call.upload();
    bielcc_kernel_param k = hip::to_kp(kp);
    bielcc_integral_param p{ip.GetIDim(), ip.GetNStart(), ip.GetPMax(), ip.GetEpsAccur()};
    hip::check(c.get(), bielcc_integral_batch(c.get(), Id, nv, 1, d + ic, d + ix, &k, &p, d + ir, d + is),
               "bielcc_integral_batch");
    const double* h = call.download();
    for (int g = 0; g < ip.GetIDim(); g++) res[g] = hip::get_c<P>(h + ir + 2 * g);
    kp.smoothR = h[is];
}
}  // namespace detail

template <typename P, typename KType, int Id, bool R>
void IntegralUniversalPnt(const double* x, const double (&cell)[4][3], Integrand<Id, R> f, KernelParam<KType>& kp,
                          const IntegralParam& ip, P* res) {
    double c[12];
    hip::flat(cell, c);
    detail::integral(4, c, x, f, kp, ip, res);
}

template <typename P, typename KType, int Id, bool R>
void IntegralUniversalPnt(const double* x, const double (&cell)[3][3], Integrand<Id, R> f, KernelParam<KType>& kp,
                          const IntegralParam& ip, P* res) {
    double c[9];
    hip::flat(cell, c);
    detail::integral(3, c, x, f, kp, ip, res);
}

template <typename P, typename KType, int Id, bool R>
void IntegralUniversalSegPnt(const double* a, const double* b, const double* x, Integrand<Id, R> f,
                             KernelParam<KType>& kp, const IntegralParam& ip, P* res) {
    double s[6] = {a[0], a[1], a[2], b[0], b[1], b[2]};
    detail::integral(2, s, x, f, kp, ip, res);
}

// ------------------------------------------------------------------ contour operators
namespace detail {
template <typename P, typename KType, size_t CP>
void k0(int variant, int id, const P* j, const double* x, const double (&cell)[CP][3], KernelParam<KType>& kp,
        const IntegralParam& ip, double analyticR, P* res) {
    auto& c = hip::default_context();
    double cv[12], jj[6] = {0, 0, 0, 0, 0, 0};
    hip::flat(cell, cv);
    if (j) for (int g = 0; g < 3; g++) hip::put_c(j[g], jj + 2 * g);
    hip::Call call;
    const size_t ic = call.in(cv, 3 * CP), ix = call.in(x, 3), ij = call.in(jj, 6);
    const size_t ir = call.out(6), is = call.out(1);
    double* d = call.upload();
    bielcc_kernel_param k = hip::to_kp(kp);
    bielcc_integral_param p{ip.GetIDim(), ip.GetNStart(), ip.GetPMax(), ip.GetEpsAccur()};
    hip::check(c.get(), bielcc_k0_batch(c.get(), variant, id, (int)CP, 1, d + ic, d + ix, d + ij, &k, &p,
                                        analyticR, d + ir, d + is), "bielcc_k0_batch");
    const double* h = call.download();
    for (int g = 0; g < 3; g++) res[g] = hip::get_c<P>(h + ir + 2 * g);
    kp.smoothR = h[is];
}
}  // namespace detail

template <typename P, typename KType, size_t CP, int Id, bool R>
void K0(const P* j, const double* x, const double (&cell)[CP][3], Integrand<Id, R>, KernelParam<KType>& kp,
        const IntegralParam& ip, P* res) {
    detail::k0(0, Id, j, x, cell, kp, ip, 0., res);
}

template <typename P, typename KType, size_t CP, int Id, bool R>
void K0_Common(const double* x, const double (&cell)[CP][3], Integrand<Id, R>, KernelParam<KType>& kp,
               const IntegralParam& ip, P* res) {
    detail::k0<P>(1, Id, nullptr, x, cell, kp, ip, 0., res);
}

template <typename P, typename KType, size_t CP>
void K0_GradSimplePot_H_HS(const P* j, const double* x, const double (&cell)[CP][3], const IntegralParam& ip,
                           double analyticR_r, KType k_wave, std::complex<double>* res) {
    KernelParam<KType> kp(k_wave, Tolerance_Constants::MACHINE_ZERO);
    P tmp[3];
    detail::k0(2, 0, j, x, cell, kp, ip, analyticR_r, tmp);
    for (int g = 0; g < 3; g++) res[g] = tmp[g];
}

// ------------------------------------------------------------------ analytic
namespace detail {
inline void analytic(int which, int nv, const double* cellv, const double* x, const IntegralParam* ip,
                     double analyticR, double kr, double ki, double* o) {
    auto& c = hip::default_context();
    hip::Call call;
    const size_t ic = call.in(cellv, (size_t)3 * nv), ix = call.in(x, 3), ir = call.out(6);
    double* d = call.upload();
    bielcc_integral_param p{1, 8, 1, 0.};
    if (ip) p = bielcc_integral_param{ip->GetIDim(), ip->GetNStart(), ip->GetPMax(), ip->GetEpsAccur()};
    hip::check(c.get(), bielcc_analytic_batch(c.get(), which, nv == 2 ? 4 : nv, 1, d + ic, d + ix, &p, analyticR,
                                              kr, ki, d + ir), "bielcc_analytic_batch");
    const double* h = call.download() + ir;
    for (int i = 0; i < 6; i++) o[i] = h[i];
}
}  // namespace detail

inline double integral1Divr(const double (&rut0)[4][3], const double* x) {
    double c[12], o[6];
    hip::flat(rut0, c);
    detail::analytic(0, 4, c, x, nullptr, 0., 0., 0., o);
    return o[0];
}
inline double integral1Divr(const double (&rut0)[3][3], const double* x) {
    double c[9], o[6];
    hip::flat(rut0, c);
    detail::analytic(0, 3, c, x, nullptr, 0., 0., 0., o);
    return o[0];
}
inline void integralGrad1DivrSeg(const double* A, const double* B, const double* x, double* res) {
    double s[6] = {A[0], A[1], A[2], B[0], B[1], B[2]}, o[6];
    detail::analytic(1, 2, s, x, nullptr, 0., 0., 0., o);
    for (int i = 0; i < 3; i++) res[i] = o[2 * i];
}
template <size_t CP, typename KType>
std::complex<double> IntegralSurf_SimplePot_H_HS(const double (&cell)[CP][3], const double* x, double analyticR_r,
                                                 const IntegralParam& ip, KType k_wave) {
    double c[12], o[6];
    hip::flat(cell, c);
    detail::analytic(2, (int)CP, c, x, &ip, analyticR_r, hip::kre(k_wave), hip::kim(k_wave), o);
    return {o[0], o[1]};
}
template <typename KType>
void IntegralSeg_GradSimplePot_H_HS(const double* A, const double* B, const double* x, double analyticR_r,
                                    const IntegralParam& ip, KType k_wave, std::complex<double>* res) {
    double s[6] = {A[0], A[1], A[2], B[0], B[1], B[2]}, o[6];
    detail::analytic(3, 2, s, x, &ip, analyticR_r, hip::kre(k_wave), hip::kim(k_wave), o);
    for (int i = 0; i < 3; i++) res[i] = {o[2 * i], o[2 * i + 1]};
}

// ------------------------------------------------------------------ operators
namespace detail {
template <typename JType, typename KType, size_t CP>
void op(int id, const JType* j, const double* x, const double (&cell)[CP][3], const NumParam& np, KType k,
        std::complex<double>* res) {
    auto& c = hip::default_context();
    double cv[12], jj[6];
    hip::flat(cell, cv);
    for (int g = 0; g < 3; g++) hip::put_c(j[g], jj + 2 * g);
    hip::Call call;
    const size_t ic = call.in(cv, 3 * CP), ix = call.in(x, 3), ij = call.in(jj, 6), ir = call.out(6);
    double* d = call.upload();
    bielcc_num_param p = hip::to_np(np);
    hip::check(c.get(), bielcc_op_batch(c.get(), id, (int)CP, 1, d + ic, d + ix, d + ij, &p, hip::kre(k),
                                        hip::kim(k), d + ir), "bielcc_op_batch");
    const double* o = call.download() + ir;
    for (int g = 0; g < 3; g++) res[g] = {o[2 * g], o[2 * g + 1]};
}
}  // namespace detail

#define BIELCC_OPERATOR(NAME, ID)                                                                        \
    template <typename JType, typename KType, size_t CellPoints>                                         \
    void NAME(const JType* j, const double* x, const double (&cell)[CellPoints][3], const NumParam& np,  \
              const KType k, std::complex<double>* res) {                                                \
        detail::op(ID, j, x, cell, np, k, res);                                                          \
    }
BIELCC_OPERATOR(K_RotRot_Near_Smooth, BIELCC_OP_K_NEAR_SMOOTH)
BIELCC_OPERATOR(K_RotRot_HS, BIELCC_OP_K_HS)
BIELCC_OPERATOR(K_RotRot_HS_Colloc, BIELCC_OP_K_HS_COLLOC)
BIELCC_OPERATOR(K_RotRot_HS2_Colloc, BIELCC_OP_K_HS2_COLLOC)
BIELCC_OPERATOR(K_RotRot_Far, BIELCC_OP_K_FAR)
BIELCC_OPERATOR(R_Rot_Smooth, BIELCC_OP_R_SMOOTH)
BIELCC_OPERATOR(R_Rot_Colloc, BIELCC_OP_R_COLLOC)
BIELCC_OPERATOR(R_Rot_Near, BIELCC_OP_R_NEAR)
#undef BIELCC_OPERATOR

// ------------------------------------------------------------------ cells and meshes
namespace detail {
// CellFill data of n cells from the device tables (bielcc_mesh_tables).
inline std::vector<double> cell_tables(int nv, int n, const double* verts) {
    auto& c = hip::default_context();
    int id = 0;
    hip::check(c.get(), bielcc_mesh_create(c.get(), n, nv, verts, &id), "bielcc_mesh_create");
    std::vector<double> t((size_t)n * 20);
    int rc = bielcc_mesh_tables(c.get(), id, t.data());
    bielcc_mesh_destroy(c.get(), id);
    hip::check(c.get(), rc, "bielcc_mesh_tables");
    return t;
}
}  // namespace detail

template <size_t NV>
class CellBase {
protected:
    double _cell[NV][3];
    double _tau[2][3];
    double _rkt[3];
    double _norm[3];
    double _s;

    void set_from_table(const double (&cell_in)[NV][3], const double* t) {
        for (size_t i = 0; i < NV; i++) for (int d = 0; d < 3; d++) _cell[i][d] = cell_in[i][d];
        for (int d = 0; d < 3; d++) {
            _rkt[d] = t[d]; _norm[d] = t[3 + d]; _tau[0][d] = t[6 + d]; _tau[1][d] = t[9 + d];
        }
        _s = t[12];
    }

public:
    // CellQuad::CellFill / CellTriangle::CellFill (mesh/cell.cpp:8-100)
    void CellFill(const double (&cell_in)[NV][3]) {
        double v[12];
        hip::flat(cell_in, v);
        auto t = detail::cell_tables((int)NV, 1, v);
        set_from_table(cell_in, t.data());
    }
    void FillFromTable(const double (&cell_in)[NV][3], const double* t) { set_from_table(cell_in, t); }
    void GetCellCoords(double (&cell)[NV][3]) const {
        for (size_t i = 0; i < NV; i++) for (int d = 0; d < 3; d++) cell[i][d] = _cell[i][d];
    }
    const double (&GetCellCoords() const)[NV][3] { return _cell; }
    const double (&GetTau() const)[2][3] { return _tau; }
    const double* GetTau1() const { return _tau[0]; }
    const double* GetTau2() const { return _tau[1]; }
    const double* GetRkt() const { return _rkt; }
    const double* GetNorm() const { return _norm; }
    const double* GetVertex(int i) const { return _cell[i]; }
    double GetArea() const { return _s; }
};

class CellQuad : public CellBase<4> {};
class CellTriangle : public CellBase<3> {};

template <typename CellType>
class SurfMesh {
    static constexpr int _NVertex = std::is_same<CellType, CellQuad>::value ? 4 : 3;
    std::vector<CellType> _cell_list;
    std::vector<double> _verts;
    double _grid_step = 0.;
    int _mesh_id = 0;

    static bool ends_with(const std::string& s, const std::string& suf) {
        return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
    }

    void read_dat(const std::string& fn) {   // mesh/surf_mesh.cpp:18-175
        std::ifstream fin(fn);
        if (!fin.is_open()) throw std::runtime_error("Read geodat.dat error: " + fn);
        int n_obj, kmod, sign, nfm, unused;
        fin >> n_obj;
        for (int o = 0; o < n_obj; o++) {
            fin >> kmod;
            for (int m = 0; m < kmod; m++) {
                fin >> sign >> nfm >> unused >> unused;
                for (int c = 0; c < nfm; c++)
                    for (int v = 0; v < _NVertex; v++)
                        for (int d = 0; d < 3; d++) { double x; fin >> x; _verts.push_back(x); }
            }
        }
        if (!fin) throw std::runtime_error("truncated .dat file: " + fn);
    }

    void read_vtk(const std::string& fn) {   // mesh/surf_mesh.h:367-555 (cells with _NVertex vertices)
        std::ifstream file(fn);
        if (!file.is_open()) throw std::runtime_error("cannot open " + fn);
        std::string line;
        std::vector<std::array<double, 3>> pts;
        std::vector<std::vector<int>> polys;
        while (std::getline(file, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream iss(line);
            std::string kw;
            iss >> kw;
            if (kw == "POINTS") {
                int n; iss >> n; pts.resize(n);
                for (int i = 0; i < n; i++) {
                    if (!std::getline(file, line)) throw std::runtime_error("unexpected end of points");
                    std::istringstream p(line);
                    p >> pts[i][0] >> pts[i][1] >> pts[i][2];
                }
            } else if (kw == "CELLS" || kw == "POLYGONS") {
                int n; iss >> n; polys.resize(n);
                for (int i = 0; i < n; i++) {
                    if (!std::getline(file, line)) throw std::runtime_error("unexpected end of cells");
                    std::istringstream p(line);
                    int m; p >> m; polys[i].resize(m);
                    for (int j = 0; j < m; j++) p >> polys[i][j];
                }
            } else if (kw == "CELL_TYPES") {
                int n; iss >> n;
                for (int i = 0; i < n; i++) std::getline(file, line);
            } else if (kw == "CELL_DATA" || kw == "POINT_DATA") {
                break;
            }
        }
        if (pts.empty() || polys.empty()) throw std::runtime_error("VTK file contains no points or cells");
        for (auto& p : polys) {
            if ((int)p.size() != _NVertex) continue;
            for (int v : p) {
                if (v < 0 || v >= (int)pts.size()) throw std::runtime_error("point index out of range");
                for (int d = 0; d < 3; d++) _verts.push_back(pts[v][d]);
            }
        }
    }

    void upload() {
        int n = (int)(_verts.size() / (3 * _NVertex));
        auto& c = hip::default_context();
        hip::check(c.get(), bielcc_mesh_create(c.get(), n, _NVertex, _verts.data(), &_mesh_id), "bielcc_mesh_create");
        std::vector<double> t((size_t)n * 20);
        hip::check(c.get(), bielcc_mesh_tables(c.get(), _mesh_id, t.data()), "bielcc_mesh_tables");
        hip::check(c.get(), bielcc_mesh_grid_step(c.get(), _mesh_id, &_grid_step), "bielcc_mesh_grid_step");
        _cell_list.resize(n);
        for (int i = 0; i < n; i++) {
            double cell[_NVertex][3];
            for (int v = 0; v < _NVertex; v++) for (int d = 0; d < 3; d++) cell[v][d] = _verts[(size_t)(i * _NVertex + v) * 3 + d];
            _cell_list[i].FillFromTable(cell, t.data() + 20 * (size_t)i);
        }
    }

public:
    SurfMesh() = default;
    explicit SurfMesh(const std::string& filename) { ReadFromFile(filename); }
    // cells given directly as [n][nv][3] coordinates
    SurfMesh(const double* verts, int n) : _verts(verts, verts + (size_t)n * _NVertex * 3) { upload(); }
    ~SurfMesh() { if (_mesh_id) bielcc_mesh_destroy(hip::default_context().get(), _mesh_id); }
    SurfMesh(const SurfMesh&) = delete;
    SurfMesh& operator=(const SurfMesh&) = delete;

    void ReadFromFile(const std::string& fn) {
        _verts.clear();
        if (ends_with(fn, ".dat")) read_dat(fn);
        else if (ends_with(fn, ".vtk")) read_vtk(fn);
        else throw std::runtime_error("File type error: ReadFromFile");
        upload();
    }
    int GetNFrm() const { return (int)_cell_list.size(); }
    double GetGridStep() const { return _grid_step; }
    const CellType& GetCell(int i) const { return _cell_list[i]; }
    int DeviceMeshId() const { return _mesh_id; }
    std::vector<std::array<double, 3>> FormCollocPoints() const {
        std::vector<std::array<double, 3>> r(_cell_list.size());
        for (size_t i = 0; i < r.size(); i++) for (int d = 0; d < 3; d++) r[i][d] = _cell_list[i].GetRkt()[d];
        return r;
    }
    std::vector<std::array<double, 3>> FormNormalVectors() const {
        std::vector<std::array<double, 3>> r(_cell_list.size());
        for (size_t i = 0; i < r.size(); i++) for (int d = 0; d < 3; d++) r[i][d] = _cell_list[i].GetNorm()[d];
        return r;
    }
    std::vector<std::array<std::array<double, 3>, 2>> FormSurfaceBasis() const {
        std::vector<std::array<std::array<double, 3>, 2>> r(_cell_list.size());
        for (size_t i = 0; i < r.size(); i++)
            for (int d = 0; d < 3; d++) { r[i][0][d] = _cell_list[i].GetTau1()[d]; r[i][1][d] = _cell_list[i].GetTau2()[d]; }
        return r;
    }
};

// ------------------------------------------------------------------ drivers
// PlaneWaveFieldE (incident_field.h:20-29)
template <typename KType>
void PlaneWaveFieldE(const double* x, const KernelParam<KType>& param, std::complex<double>* e_inc) {
    auto& c = hip::default_context();
    hip::Call call;
    const size_t ix = call.in(x, 3), ir = call.out(6);
    double* d = call.upload();
    hip::check(c.get(), bielcc_plane_wave_e(c.get(), 1, d + ix, param.vec_dbl, param.e0, d + ir),
               "bielcc_plane_wave_e");
    const double* o = call.download() + ir;
    for (int g = 0; g < 3; g++) e_inc[g] = {o[2 * g], o[2 * g + 1]};
}

// MaterialParam (electrodynamics/material.h:73-190): the permeability
// PlaneWaveFieldH reads.  Calc(freq) sets mu = mu0 mu_r (material.cpp:105-116).
enum class ChangeMaterialTypeDir { Isotropic, Anisotropic };
enum class ChangeMaterialTypeFreq { NonDispersive, Dispersive, DispersiveByRefractIndex };
class MaterialParam {
public:
    MaterialParam() = default;
    MaterialParam(std::string name, ChangeMaterialTypeDir, ChangeMaterialTypeFreq, double eps_r, double mu_r,
                  double sigma, double tand, double tanm)
        : _name(std::move(name)), _eps_r(eps_r), _mu_r(mu_r), _sigma(sigma), _tand(tand), _tanm(tanm) {}
    void Calc(double /*freq*/) { _mu = 4 * M_PI * 1e-7 * _mu_r; _tanm = 0.; }
    std::variant<double, std::complex<double>> GetMu() const { return _mu; }
    void SetMu(std::variant<double, std::complex<double>> mu) { _mu = mu; }

private:
    std::string _name;
    double _eps_r = 0., _mu_r = 0., _sigma = 0., _tand = 0., _tanm = 0.;
    std::variant<double, std::complex<double>> _mu = 0.;
};

// PlaneWaveFieldH (incident_field.h:40-59)
template <typename KType>
void PlaneWaveFieldH(const double* x, const KernelParam<KType>& param, const MaterialParam& material_param,
                     std::complex<double>* h_inc) {
    auto& c = hip::default_context();
    const auto mu = material_param.GetMu();
    const std::complex<double> m = std::holds_alternative<double>(mu) ? std::complex<double>(std::get<double>(mu), 0.)
                                                                      : std::get<std::complex<double>>(mu);
    hip::Call call;
    const size_t ix = call.in(x, 3), ir = call.out(6);
    double* d = call.upload();
    hip::check(c.get(), bielcc_plane_wave_h(c.get(), 1, d + ix, param.vec_dbl, param.e0, param.freq, m.real(),
                                            m.imag(), d + ir), "bielcc_plane_wave_h");
    const double* o = call.download() + ir;
    for (int g = 0; g < 3; g++) h_inc[g] = {o[2 * g], o[2 * g + 1]};
}

// The constant F(y) integrand (the area test of examples/integrators/main0_integrals.cpp:77-81),
// the F(y) integrand besides PlaneWaveFieldE that the integrators below run on the GPU.
template <typename KType>
void FY_Unit(const double*, const KernelParam<KType>&, double* ff) { ff[0] = 1.; }

// IntegralUniversal, F(y) forms (integral_universal.h:252-431), and the F(y) form of
// IntegralUniversalSegPnt (integral_universal_seg.h:120-179).  The integrand must be
// PlaneWaveFieldE<KType> or FY_Unit<KType>; any other host function throws.
namespace detail {
template <typename P, typename KType>
void integral_y(int nv, const double* cellv, void (*f_0)(const double*, const KernelParam<KType>&, P*),
                KernelParam<KType>& kp, const IntegralParam& ip, P* res) {
    int f = -1;
    if constexpr (std::is_same<P, std::complex<double>>::value) {
        if (f_0 == &PlaneWaveFieldE<KType>) f = BIELCC_FY_PLANE_WAVE_E;
    } else {
        if (f_0 == &FY_Unit<KType>) f = BIELCC_FY_UNIT;
    }
    if (f < 0) throw std::runtime_error("IntegralUniversal: only PlaneWaveFieldE and FY_Unit run on the GPU");
    auto& c = hip::default_context();
    hip::Call call;
    const size_t ic = call.in(cellv, (size_t)3 * (nv == 2 ? 2 : nv)), ir = call.out(6), is = call.out(1);
    double* d = call.upload();
    bielcc_kernel_param k = hip::to_kp(kp);
    bielcc_integral_param p{ip.GetIDim(), ip.GetNStart(), ip.GetPMax(), ip.GetEpsAccur()};
    hip::check(c.get(), bielcc_integral_y_batch(c.get(), f, nv, 1, d + ic, &k, kp.e0, &p, d + ir, d + is),
               "bielcc_integral_y_batch");
    const double* h = call.download();
    for (int g = 0; g < ip.GetIDim(); g++) res[g] = hip::get_c<P>(h + ir + 2 * g);
    kp.smoothR = h[is];
}
}  // namespace detail

template <typename P, typename KType>
void IntegralUniversal(const double (&cell)[4][3], void (*f_0)(const double*, const KernelParam<KType>&, P*),
                       KernelParam<KType>& ker_param, const IntegralParam& int_param, P* res) {
    double c[12];
    hip::flat(cell, c);
    detail::integral_y(4, c, f_0, ker_param, int_param, res);
}

template <typename P, typename KType>
void IntegralUniversal(const double (&cell)[3][3], void (*f_0)(const double*, const KernelParam<KType>&, P*),
                       KernelParam<KType>& ker_param, const IntegralParam& int_param, P* res) {
    double c[9];
    hip::flat(cell, c);
    detail::integral_y(3, c, f_0, ker_param, int_param, res);
}

template <typename P, typename KType>
void IntegralUniversalSegPnt(const double* a, const double* b,
                             void (*f_0)(const double*, const KernelParam<KType>&, P*),
                             KernelParam<KType>& ker_param, const IntegralParam& int_param, P* res) {
    double s[6] = {a[0], a[1], a[2], b[0], b[1], b[2]};
    detail::integral_y(2, s, f_0, ker_param, int_param, res);
}

// GetFieldE_Ideal (get_field.h:35-78).  The scattered sum runs on the GPU; the
// incident term is the caller's EInc_Func, applied per point afterwards as the
// reference does (on the GPU when it is PlaneWaveFieldE).
template <typename CellType, typename KType, typename P>
void GetFieldE_Ideal(const SurfMesh<CellType>& surf_mesh, const std::vector<std::array<std::complex<double>, 3>>& j_E,
                     std::complex<double> mult, const KernelParam<KType>& EInc_param,
                     void (*EInc_Func)(const double*, const KernelParam<KType>&, P*), const NumParam& num_param,
                     const std::vector<std::array<double, 3>>& points,
                     std::vector<std::array<std::complex<double>, 3>>& field_E) {
    auto& c = hip::default_context();
    const int n = surf_mesh.GetNFrm(), M = (int)points.size();
    if ((int)j_E.size() != n) throw std::runtime_error("GetFieldE_Ideal: j_E size differs from the cell count");
    std::vector<double> jj((size_t)6 * n), pp((size_t)3 * M), ee((size_t)6 * M);
    for (int i = 0; i < n; i++) for (int g = 0; g < 3; g++) hip::put_c(j_E[i][g], jj.data() + 6 * i + 2 * g);
    for (int i = 0; i < M; i++) for (int d = 0; d < 3; d++) pp[3 * i + d] = points[i][d];
    hip::DevBuf dj(jj.data(), jj.size()), dp(pp.data(), pp.size()), de(ee.size());
    bielcc_num_param np = hip::to_np(num_param);
    const bool plane = EInc_Func == static_cast<void (*)(const double*, const KernelParam<KType>&, P*)>(
                                        &PlaneWaveFieldE<KType>);
    if (plane) {
        hip::check(c.get(), bielcc_get_field_e(c.get(), surf_mesh.DeviceMeshId(), dj.ptr(), mult.real(), mult.imag(),
                                               EInc_param.vec_dbl, EInc_param.e0, &np, hip::kre(EInc_param.k),
                                               hip::kim(EInc_param.k), M, dp.ptr(), de.ptr()), "bielcc_get_field_e");
    } else {
        hip::check(c.get(), bielcc_field(c.get(), surf_mesh.DeviceMeshId(), &np, hip::kre(EInc_param.k),
                                         hip::kim(EInc_param.k), dj.ptr(), nullptr, mult.real(), mult.imag(), 1., 0.,
                                         nullptr, nullptr, 0, M, dp.ptr(), de.ptr(), nullptr), "bielcc_field");
    }
    de.download(ee.data(), ee.size());
    field_E.resize(M);
    for (int i = 0; i < M; i++) {
        for (int g = 0; g < 3; g++) field_E[i][g] = {ee[6 * i + 2 * g], ee[6 * i + 2 * g + 1]};
        if (!plane) {
            P e_inc[3]{};
            EInc_Func(points[i].data(), EInc_param, e_inc);
            for (int g = 0; g < 3; g++) field_E[i][g] += e_inc[g];
        }
    }
}

// SurfaceCurrents_byBasis (currents.h:22-33) on a device mesh: b has 2N entries.
template <typename CellType>
void SurfaceCurrents_byBasis(const SurfMesh<CellType>& mesh, const std::complex<double>* b,
                             std::vector<std::array<std::complex<double>, 3>>& j_vec) {
    auto& c = hip::default_context();
    const int n = mesh.GetNFrm();
    hip::DevBuf db(reinterpret_cast<const double*>(b), (size_t)4 * n), dj((size_t)6 * n);
    hip::check(c.get(), bielcc_currents_by_basis(c.get(), mesh.DeviceMeshId(), 1, db.ptr(), 2 * n, dj.ptr()),
               "bielcc_currents_by_basis");
    std::vector<double> o((size_t)6 * n);
    dj.download(o.data(), o.size());
    j_vec.resize(n);
    for (int i = 0; i < n; i++) for (int g = 0; g < 3; g++) j_vec[i][g] = {o[6 * i + 2 * g], o[6 * i + 2 * g + 1]};
}

// ------------------------------------------------------------------ writers
// Host-side output in the reference's file formats (io/io_vtk.h, io/io_aereco.h):
// same layouts and the stream's default number formatting, so files written
// here and by the reference are byte-identical for the same values.
struct IO_VTK {
    // IO_VTK::WritePntsWithField_C (io/io_vtk.cpp:482-551): POLYDATA points with
    // real part, imaginary part, per-component magnitude and phase.
    void WritePntsWithField_C(const std::string& filename, const std::vector<std::array<double, 3>>& points,
                              const std::vector<std::array<std::complex<double>, 3>>& field_vectors) {
        if (points.size() != field_vectors.size())
            throw std::runtime_error("Points and field vectors must have same size");
        std::ofstream f(filename);
        if (!f.is_open()) throw std::runtime_error("Cannot open file: " + filename);
        f << "# vtk DataFile Version 3.0\nPoints with complex vector field\nASCII\nDATASET POLYDATA\n";
        f << "POINTS " << points.size() << " double\n";
        for (const auto& p : points) f << p[0] << " " << p[1] << " " << p[2] << "\n";
        f << "POINT_DATA " << field_vectors.size() << "\n";
        f << "VECTORS real_part double\n";
        for (const auto& v : field_vectors) f << v[0].real() << " " << v[1].real() << " " << v[2].real() << "\n";
        f << "VECTORS image_part double\n";
        for (const auto& v : field_vectors) f << v[0].imag() << " " << v[1].imag() << " " << v[2].imag() << "\n";
        f << "VECTORS magnitude double\n";
        for (const auto& v : field_vectors) f << std::abs(v[0]) << " " << std::abs(v[1]) << " " << std::abs(v[2]) << "\n";
        const char* comp[3] = {"x", "y", "z"};
        for (int g = 0; g < 3; g++) {
            f << "SCALARS phase_" << comp[g] << " double 1\nLOOKUP_TABLE default\n";
            for (const auto& v : field_vectors) f << std::arg(v[g]) << "\n";
        }
    }
};

struct IO_AERECO {
    // IO_AERECO::WriteVecs_C (io/io_aereco.cpp:98-128): header "2 n/2", one
    // vector per line, real and imaginary parts in two files.
    void WriteVecs_C(const std::string& filenameReal, const std::string& filenameImage,
                     const std::vector<std::array<std::complex<double>, 3>>& vals) {
        std::ofstream re(filenameReal), im(filenameImage);
        if (!re.is_open() || !im.is_open())
            throw std::runtime_error("IO_AERECO::WriteVecs_C error. Can't open files:" + filenameReal + ", " +
                                     filenameImage);
        const int n = (int)vals.size();
        re << 2 << " " << n / 2 << std::endl;
        im << 2 << " " << n / 2 << std::endl;
        for (const auto& v : vals) {
            re << v[0].real() << " " << v[1].real() << " " << v[2].real() << std::endl;
            im << v[0].imag() << " " << v[1].imag() << " " << v[2].imag() << std::endl;
        }
    }
    // IO_AERECO::WriteField_C (io/io_aereco.cpp:135-173): header "n1 n1" with
    // n1 = size/2 and n1^2 rows (real, imaginary, |v|).  The reference reads
    // past the end of the vector when n1^2 > size; that case is an error here.
    void WriteField_C(const std::string& filenameReal, const std::string& filenameImage,
                      const std::string& filenameAbs, const std::vector<std::array<std::complex<double>, 3>>& field_val) {
        const int n1 = (int)field_val.size() / 2;
        if ((size_t)n1 * (size_t)n1 > field_val.size())
            throw std::runtime_error("IO_AERECO::WriteField_C: (size/2)^2 rows exceed the values given");
        std::ofstream re(filenameReal), im(filenameImage), ab(filenameAbs);
        if (!ab.is_open() || !im.is_open() || !re.is_open())
            throw std::runtime_error("IO_AERECO::WriteField_C error. Can't open files:" + filenameReal + ", " +
                                     filenameImage + ", " + filenameAbs);
        re << n1 << " " << n1 << std::endl;
        im << n1 << " " << n1 << std::endl;
        ab << n1 << " " << n1 << std::endl;
        for (int i = 0; i < n1 * n1; i++) {
            const auto& v = field_val[i];
            re << v[0].real() << " " << v[1].real() << " " << v[2].real() << std::endl;
            im << v[0].imag() << " " << v[1].imag() << " " << v[2].imag() << std::endl;
            ab << std::sqrt(std::abs(v[0]) * std::abs(v[0]) + std::abs(v[1]) * std::abs(v[1]) +
                            std::abs(v[2]) * std::abs(v[2])) << std::endl;
        }
    }
};

// ------------------------------------------------------------------ whole-problem drivers (new)
// Solve the collocation system for a plane wave (kvec, e0) and return the
// basis coefficients b (2N) - assembly, LU and solve stay on the device.
template <typename CellType, typename KType>
std::vector<std::complex<double>> SolveCurrents_PlaneWave(const SurfMesh<CellType>& mesh, const NumParam& num_param,
                                                          KType k, const double* kvec, const double* e0) {
    auto& c = hip::default_context();
    const int n2 = 2 * mesh.GetNFrm();
    hip::DevBuf Z((size_t)2 * n2 * n2), B((size_t)2 * n2);
    void* piv = nullptr;
    hip::check(c.get(), bielcc_malloc(c.get(), (size_t)n2 * sizeof(int), &piv), "bielcc_malloc");
    std::unique_ptr<void, void (*)(void*)> piv_guard(piv, [](void* p) { bielcc_free(hip::default_context().get(), p); });
    bielcc_num_param np = hip::to_np(num_param);
    hip::check(c.get(), bielcc_assemble(c.get(), mesh.DeviceMeshId(), 0, &np, hip::kre(k), hip::kim(k), 0, n2, Z.ptr(), n2),
               "bielcc_assemble");
    hip::check(c.get(), bielcc_rhs_planewave(c.get(), mesh.DeviceMeshId(), 1, kvec, e0, B.ptr(), n2), "bielcc_rhs_planewave");
    int info = 0;
    hip::check(c.get(), bielcc_lu_factor(c.get(), n2, Z.ptr(), n2, static_cast<int*>(piv), &info), "bielcc_lu_factor");
    hip::check(c.get(), bielcc_lu_solve(c.get(), n2, Z.ptr(), n2, static_cast<int*>(piv), 1, B.ptr(), n2), "bielcc_lu_solve");
    std::vector<std::complex<double>> b(n2);
    B.download(reinterpret_cast<double*>(b.data()), (size_t)2 * n2);
    return b;
}

}  // namespace bielcc
