// edb.hpp -- header-only C++ host side over the C ABI (edb.h), with the
// reference's class names and member signatures so that code written against
// CalQ/QCalculator.h, CalQ/fft3d.h, CalQ/EntropyFix.h and CalQ/BKW.h can
// switch by changing its includes:
//
//   QCalculator<c_type, r_type>   QCalculator.h:17-62  (genB, calQ, closeFilter)
//   fft3d                         fft3d.h:18-48        (fft, ifft)
//   EntropyFix<i_type, r_type>    EntropyFix.h:13-46   (EntropyCal, setAverage, FixSol)
//   BKW3D<c_type, r_type>         BKW.h:22-62          (FEuler, RK2, SSP3, RK4, CFE, DUFORT)
//
// Differences a caller sees: every computation runs on the GPU; c_type must
// be std::complex<double> and r_type double (the only instantiation the
// reference's mains use, test.cpp:11-12, testBKW3D.cpp:12-13); failures
// throw std::runtime_error carrying edb_last_error() where the reference
// asserts or returns garbage (even N, Nf != Nt).  QCalculator takes the
// velocity half-width L that the reference hard-codes (QCalculator.h:176).
#pragma once

#include <cmath>
#include <complex>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "edb.h"

namespace edb {

inline void check(int status, const char* where) {
  if (status != EDB_OK) throw std::runtime_error(std::string(where) + ": " + edb_last_error());
}

template <class c_type, class r_type>
class QCalculator {
  static_assert(std::is_same<c_type, std::complex<double>>::value, "c_type must be std::complex<double>");
  static_assert(std::is_same<r_type, double>::value, "r_type must be double");

 public:
  QCalculator(int N, int alpha, r_type L = 6.62132) : N_(N) {
    check(edb_qcalc_create(N, alpha, L, &h_), "QCalculator");
  }
  virtual ~QCalculator() { edb_qcalc_destroy(h_); }
  QCalculator(const QCalculator&) = delete;
  QCalculator& operator=(const QCalculator&) = delete;

  void closeFilter() { check(edb_qcalc_close_filter(h_), "closeFilter"); }
  void genB() { check(edb_qcalc_genB(h_), "genB"); }
  // Q for one spectrum, N^3 entries in FFTW order (QCalculator.h:231-257).
  virtual void calQ(c_type* Qout, c_type* fin) {
    check(edb_qcalc_calQ(h_, reinterpret_cast<double*>(Qout), reinterpret_cast<const double*>(fin)),
          "calQ");
  }
  // Batched FP32 evaluation of `batch` spectra stored one after another.
  void genB_batched() { check(edb_qcalc_genB_batched(h_), "genB_batched"); }
  void calQ_batched(c_type* Qout, const c_type* fin, int batch, bool hermitian) {
    check(edb_qcalc_calQ_batched(h_, batch, hermitian ? 1 : 0, reinterpret_cast<double*>(Qout),
                                 reinterpret_cast<const double*>(fin)),
          "calQ_batched");
  }
  // Moment correction of every Q (zero discrete mass, momentum and energy
  // moments); no reference counterpart.
  void setConservative(bool on = true) { check(edb_qcalc_set_conservative(h_, on ? 1 : 0), "setConservative"); }
  edb_qcalc* handle() { return h_; }
  int size() const { return N_; }

 private:
  edb_qcalc* h_ = nullptr;
  int N_;
};

class fft3d {
 public:
  typedef double complex_t[2];  // fftw_complex
  void fft(int Nx, int Ny, int Nz, complex_t* re, complex_t* fre) {
    check(edb_fft3d_fft(Nx, Ny, Nz, &re[0][0], &fre[0][0]), "fft3d::fft");
  }
  // fft3d.h:32: output first, like the reference.
  void ifft(int Nx, int Ny, int Nz, complex_t* re, complex_t* fre) {
    check(edb_fft3d_ifft(Nx, Ny, Nz, &re[0][0], &fre[0][0]), "fft3d::ifft");
  }
};

template <class i_type, class r_type>
class EntropyFix {
  static_assert(std::is_same<r_type, double>::value, "r_type must be double");

 public:
  explicit EntropyFix(i_type Nt, r_type eps = 1.0e-10) { check(edb_efix_init(&st_, (long)Nt, eps), "EntropyFix"); }

  r_type EntropyCal(const std::vector<r_type>& fin) { return EntropyCal((i_type)fin.size(), fin.data()); }
  r_type EntropyCal(i_type Nf, const r_type* fin) {
    double H = 0.;
    check(edb_efix_entropy_cal(&st_, (long)Nf, fin, &H), "EntropyCal");
    return H;
  }
  void setAverage(const std::vector<r_type>& fin) { setAverage((i_type)fin.size(), fin.data()); }
  void setAverage(i_type Nf, const r_type* fin) {
    check(edb_efix_set_average(&st_, (long)Nf, fin), "setAverage");
  }
  void FixSol(std::vector<r_type>& fin, r_type AimEntropy) { FixSol((i_type)fin.size(), fin.data(), AimEntropy); }
  void FixSol(i_type Nf, r_type* fin, r_type AimEntropy) {
    check(edb_efix_fixsol(&st_, (long)Nf, fin, AimEntropy), "FixSol");
  }
  r_type average() const { return st_.average; }

 private:
  edb_efix_state st_{};
};

// BKW3D drivers on one instance.  Initial data: a user function of
// (t, x, y, z) on x_i = -R + i*2R/N (BKW.h:121-138); by default the BKW
// profile of testfft.cpp:11-16, s = 1 - 0.4 exp(-t/6).
template <class c_type, class r_type>
class BKW3D {
 public:
  BKW3D(int N0, r_type R, int alpha = 0, bool filter = false) : N0_(N0), R_(R), alpha_(alpha), filter_(filter) {}
  void setFixEntropy(bool on = true) { fix_ = on; }
  void setFixApplied(bool on = true) { applied_ = on; }
  void setInitial(std::function<r_type(r_type, r_type, r_type, r_type)> f) { init_ = f; }
  void setPrecision(int precision) { precision_ = precision; }
  void setConservative(bool on = true) { conserve_ = on; }

  void FEuler(r_type T0, r_type TEnd, r_type dt) { run(EDB_SCHEME_FEULER, T0, TEnd, dt, 0.); }
  void RK2(r_type T0, r_type TEnd, r_type dt) { run(EDB_SCHEME_RK2, T0, TEnd, dt, 0.); }
  void SSP3(r_type T0, r_type TEnd, r_type dt) { run(EDB_SCHEME_SSP3, T0, TEnd, dt, 0.); }
  void RK4(r_type T0, r_type TEnd, r_type dt) { run(EDB_SCHEME_RK4, T0, TEnd, dt, 0.); }
  void CFE(r_type T0, r_type TEnd, r_type dt, r_type beta) { run(EDB_SCHEME_CFE, T0, TEnd, dt, beta); }
  void DUFORT(r_type T0, r_type TEnd, r_type dt, r_type beta) { run(EDB_SCHEME_DUFORT, T0, TEnd, dt, beta); }

  // Real-space F after the last step (N^3 values).
  const std::vector<r_type>& F() const { return F_; }
  const std::vector<r_type>& entropy() const { return H_; }

 private:
  static r_type bkw(r_type t, r_type x, r_type y, r_type z) {
    r_type v2 = x * x + y * y + z * z;
    r_type s = 1. - 2. * std::exp(-t / 6.) / 5.;
    return 1. / std::pow(2. * M_PI * s, 1.5) * std::exp(-v2 / (2. * s)) *
           ((5. * s - 3.) / (2. * s) + (1. - s) / (2. * s * s) * v2);
  }

  void run(int scheme, r_type T0, r_type TEnd, r_type dt, r_type beta) {
    edb_solver_params p{};
    p.N = N0_;
    p.alpha = alpha_;
    p.L = R_;
    p.filter = filter_ ? 1 : 0;
    p.batch = 1;
    p.precision = precision_;
    p.scheme = scheme;
    p.dt = dt;
    p.beta = beta;
    p.fix_mode = fix_ ? (applied_ ? EDB_FIX_APPLIED : EDB_FIX_DIAGNOSTIC) : EDB_FIX_NONE;
    p.hermitian = 1;
    p.conserve = conserve_ ? 1 : 0;
    double est = (TEnd - T0) / dt;
    p.history = (int)(est > 0 ? (est < 1e6 ? est : 1e6) : 0) + 16;
    edb_solver* s = nullptr;
    check(edb_solver_create(&p, nullptr, &s), "BKW3D");
    long N3 = (long)N0_ * N0_ * N0_;
    std::vector<r_type> F0(N3);
    r_type h = 2 * R_ / r_type(N0_);
    auto f = init_ ? init_ : std::function<r_type(r_type, r_type, r_type, r_type)>(bkw);
    long l = 0;
    for (int i = 0; i < N0_; ++i)
      for (int j = 0; j < N0_; ++j)
        for (int k = 0; k < N0_; ++k) F0[l++] = f(T0, -R_ + i * h, -R_ + j * h, -R_ + k * h);
    try {
      check(edb_solver_set_initial_real(s, T0, F0.data()), "setInitialValue");
      int n = 0;
      check(edb_solver_run(s, TEnd, &n), "run");
      std::vector<double> H(n), Hf(n);
      std::vector<int> fx(n);
      check(edb_solver_entropy_history(s, n < p.history ? n : p.history, H.data(), Hf.data(), fx.data()),
            "entropy");
      H_ = Hf;
      F_.resize(N3);
      check(edb_solver_get_real(s, F_.data()), "F");
    } catch (...) {
      edb_solver_destroy(s);
      throw;
    }
    edb_solver_destroy(s);
  }

  int N0_;
  r_type R_;
  int alpha_;
  bool filter_;
  bool fix_ = false, applied_ = false, conserve_ = false;
  int precision_ = EDB_PREC_FP64_PLAIN;
  std::function<r_type(r_type, r_type, r_type, r_type)> init_;
  std::vector<r_type> F_, H_;
};

}  // namespace edb
