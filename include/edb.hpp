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

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <functional>
#include <iomanip>
#include <ostream>
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
  // Compressed beta (F over |l+m|^2, |l-m|^2 plus per-mode sigma and Bmm) in
  // place of genB's M^6 array, and calQ with beta formed from it: every beta
  // equals genB's, the sum runs in another fixed order (FP64 rounding).  No
  // reference counterpart; runs for N up to 64.
  void genBCompressed() { check(edb_qcalc_genB_compressed(h_), "genBCompressed"); }
  void calQCompressed(c_type* Qout, c_type* fin) {
    check(edb_qcalc_calQ_compressed(h_, reinterpret_cast<double*>(Qout), reinterpret_cast<const double*>(fin)),
          "calQCompressed");
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

// BKW3D drivers on one instance.  The collision operator is the drivers'
// QCalculator(N0, 0) (BKW.h:201, 462): Jackson filter on (IS_FILTER = 1,
// QCalculator.h:34) unless the constructor's `filter` is false.  Initial
// data: a user function of (t, x, y, z) on x_i = -R + i*2R/N
// (BKW.h:121-138); by default the BKW
// profile of testfft.cpp:11-16, s = 1 - 0.4 exp(-t/6).  Entropy fix as in
// the reference drivers: RK4 and SSP3 always run it (BKW.h:432, 499),
// FEuler and CFE when setFixEntropy is on (:228, :294), RK2 and DUFORT never
// (their fix is compiled out, :359, :604); it is diagnostic (F fixed, the
// spectral state untouched) unless setFixApplied.  With setOutput the
// drivers print the reference's progress lines and per-step tables
// (BKW.h:212-633) and FEuler writes F.dat (:217, :243-246); timing lines
// carry this run's own clock.
template <class c_type, class r_type>
class BKW3D {
 public:
  BKW3D(int N0, r_type R, int alpha = 0, bool filter = true) : N0_(N0), R_(R), alpha_(alpha), filter_(filter) {}
  void setFixEntropy(bool on = true) { fix_ = on; }
  void setFixApplied(bool on = true) { applied_ = on; }
  void setInitial(std::function<r_type(r_type, r_type, r_type, r_type)> f) { init_ = f; }
  void setPrecision(int precision) { precision_ = precision; }
  void setConservative(bool on = true) { conserve_ = on; }
  void setOutput(std::ostream* os, bool writeF = true) {
    out_ = os;
    writeF_ = writeF;
  }

  void FEuler(r_type T0, r_type TEnd, r_type dt) { run(EDB_SCHEME_FEULER, T0, TEnd, dt, 0.); }
  void RK2(r_type T0, r_type TEnd, r_type dt) { run(EDB_SCHEME_RK2, T0, TEnd, dt, 0.); }
  void SSP3(r_type T0, r_type TEnd, r_type dt) { run(EDB_SCHEME_SSP3, T0, TEnd, dt, 0.); }
  void RK4(r_type T0, r_type TEnd, r_type dt) { run(EDB_SCHEME_RK4, T0, TEnd, dt, 0.); }
  void CFE(r_type T0, r_type TEnd, r_type dt, r_type beta) { run(EDB_SCHEME_CFE, T0, TEnd, dt, beta); }
  void DUFORT(r_type T0, r_type TEnd, r_type dt, r_type beta) { run(EDB_SCHEME_DUFORT, T0, TEnd, dt, beta); }

  // The driver's real-space F after the last step (N^3 values; fixed values
  // where the last step's fix ran) and the entropy after each step's fix.
  const std::vector<r_type>& F() const { return F_; }
  const std::vector<r_type>& entropy() const { return H_; }

  // BKW3D::outV1Direction (BKW.h:163-180): x, F(x, h/2, h/2) at the grid
  // point (i, N/2, N/2), and the initial function at (0, x, h/2, h/2).
  void outV1Direction(const std::string& filename) const {
    std::ofstream w(filename);
    if (!w) throw std::runtime_error("outV1Direction: cannot open " + filename);
    const r_type h = 2. * R_ / r_type(N0_);
    auto f = init_ ? init_ : std::function<r_type(r_type, r_type, r_type, r_type)>(bkw);
    w << "# x\tf(t)\tf_0\n";
    for (int i = 0; i < N0_; ++i) {
      const r_type x = -R_ + i * h;
      const long l = ((long)i * N0_ + N0_ / 2) * N0_ + N0_ / 2;
      const r_type fx = F_.empty() ? r_type(0) : F_[l];
      w << std::fixed << std::setprecision(8) << x << "\t" << fx << "\t" << f(0, x, h / 2, h / 2) << "\n";
    }
  }

 private:
  static r_type bkw(r_type t, r_type x, r_type y, r_type z) {
    r_type v2 = x * x + y * y + z * z;
    r_type s = 1. - 2. * std::exp(-t / 6.) / 5.;
    return 1. / std::pow(2. * M_PI * s, 1.5) * std::exp(-v2 / (2. * s)) *
           ((5. * s - 3.) / (2. * s) + (1. - s) / (2. * s * s) * v2);
  }

  // TimeWatch's report (TimeWatch.h:24-33) for a span of this run.
  struct Watch {
    std::ostream* os;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::clock_t c0 = std::clock();
    ~Watch() {
      if (!os) return;
      double w = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      double c = double(std::clock() - c0) / CLOCKS_PER_SEC;
      *os << std::fixed << std::setprecision(2) << "#\tWTime: " << w << "s\t"
          << "CTime: " << c << "s\n"
          << "#\tSpeedup: " << (w > 0 ? c / w : 0.) << "\n";
    }
  };

  // One table row of the reference's drivers for step k (0-based) of the
  // current phase; prev is the entropy after the previous step's fix.
  void row(int scheme, int k, int step_no, r_type t, r_type H, r_type Hf, bool fixed, r_type mean,
           r_type mean_fix, r_type prev) {
    std::ostream& o = *out_;
    o << std::fixed << std::setprecision(8);
    switch (scheme) {
      case EDB_SCHEME_FEULER:  // BKW.h:226-241
        if (step_no > 1)
          o << step_no << "\t" << t << "\t" << H << "\t" << prev - H << "\t" << Hf << "\t" << prev - Hf
            << std::endl;
        else
          o << step_no << "\t" << t << "\t" << H << "\t-\t-\t-\t\n";
        break;
      case EDB_SCHEME_CFE:  // BKW.h:293-309
        if (k) {
          o << t << "\t" << mean << "\t" << H << "\t" << prev - H << "\t";
          if (fixed) o << Hf << "\t" << prev - Hf << "\t" << mean_fix;
          o << std::endl;
        } else {
          o << t << "\t" << H << "\t-\n";
        }
        break;
      case EDB_SCHEME_RK2:  // BKW.h:356-381, 579-588, 603-628
      case EDB_SCHEME_DUFORT:
        if (k)
          o << t << "\t" << mean << "\t" << H << "\t" << prev - H << "\t" << std::endl;
        else
          o << t << "\t" << mean << "\t" << H << "\t-\n";
        break;
      default:  // SSP3, RK4: BKW.h:430-444, 497-514
        if (k) {
          o << t << "\t" << H << "\t" << prev - H << "\t";
          if (fixed) o << Hf << "\t" << prev - Hf;
          o << std::endl;
        } else {
          o << t << "\t" << H << "\t-\n";
        }
    }
  }

  void run(int scheme, r_type T0, r_type TEnd, r_type dt, r_type beta) {
    const bool fixing = (scheme == EDB_SCHEME_RK4 || scheme == EDB_SCHEME_SSP3) ? true : fix_;
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
    p.fix_mode = fixing ? (applied_ ? EDB_FIX_APPLIED : EDB_FIX_DIAGNOSTIC) : EDB_FIX_NONE;
    p.hermitian = 1;
    p.conserve = conserve_ ? 1 : 0;
    double est = (TEnd - T0) / dt;
    p.history = (int)(est > 0 ? (est < 1e6 ? est : 1e6) : 0) + 16;
    if (out_) *out_ << "# Initialization...\n" << "# QCalculator calculate B...\n";
    edb_solver* s = nullptr;
    {
      Watch w{out_};
      check(edb_solver_create(&p, nullptr, &s), "BKW3D");
    }
    const long N3 = (long)N0_ * N0_ * N0_;
    std::vector<r_type> F0(N3);
    r_type h = 2 * R_ / r_type(N0_);
    auto f = init_ ? init_ : std::function<r_type(r_type, r_type, r_type, r_type)>(bkw);
    long l = 0;
    for (int i = 0; i < N0_; ++i)
      for (int j = 0; j < N0_; ++j)
        for (int k = 0; k < N0_; ++k) F0[l++] = f(T0, -R_ + i * h, -R_ + j * h, -R_ + k * h);
    try {
      check(edb_solver_set_initial_real(s, T0, F0.data()), "setInitialValue");
      F_.resize(N3);
      if (!out_) {
        int n = 0;
        check(edb_solver_run(s, TEnd, &n), "run");
        std::vector<double> H(n), Hf(n);
        std::vector<int> fx(n);
        const int kept = n < p.history ? n : p.history;
        check(edb_solver_entropy_history(s, kept, H.data(), Hf.data(), fx.data()), "entropy");
        H_.assign(Hf.begin(), Hf.begin() + kept);
      } else {
        run_reporting(s, scheme, p, T0, TEnd, dt, F0);
      }
      check(edb_solver_get_driver_real(s, F_.data()), "F");
    } catch (...) {
      edb_solver_destroy(s);
      throw;
    }
    edb_solver_destroy(s);
  }

  // Stepping with the reference's output: one step at a time when FEuler
  // writes F.dat, else the whole run and the table from the device history.
  void run_reporting(edb_solver* s, int scheme, const edb_solver_params& p, r_type T0, r_type TEnd, r_type dt,
                     const std::vector<r_type>& F0) {
    std::ostream& o = *out_;
    const long N3 = (long)N0_ * N0_ * N0_;
    static const char* kName[] = {"Forward Euler", "RK2", "SSP3", "RK4", "C-Forward Euler", "DUFORT"};
    static const char* kHead[] = {
        "# Step\tt\tentropy_{n}\tentropy_{n-1}-entropy_{n}\tfixedEntropy\tfixedDiff\n",
        "# t\tentropy_{n}\tentropy_{n-1}-entropy_{n}\n",
        "# t\tentropy_{n}\tentropy_{n-1}-entropy_{n}\tNew entropy\tnew entropy_{n-1}-entropy_{n}\n",
        "# t\tentropy_{n}\tentropy_{n-1}-entropy_{n}\tNew entropy\tnew entropy_{n-1}-entropy_{n}\n",
        "# t\taverage\tentropy_{n}\tentropy_{n-1}-entropy_{n}\n",
        "# t\tentropy_{n}\tentropy_{n-1}-entropy_{n}\n"};
    const bool dufort = scheme == EDB_SCHEME_DUFORT;
    o << "# Start the evolution using " << kName[dufort ? EDB_SCHEME_RK2 : scheme] << "...\n"
      << kHead[dufort ? EDB_SCHEME_RK2 : scheme];
    Watch w{out_};
    r_type prev = 1.;  // entropy0 = 1 (BKW.h:478 and the like)
    int step_no = 0;
    if (scheme == EDB_SCHEME_FEULER) {  // entropy0 = calEntropy() of the initial F, BKW.h:214-216
      r_type e = 0.;
      for (long i = 0; i < N3; ++i) e += std::abs(F0[i]) < 1.0e-6 ? 0. : F0[i] * std::log(F0[i]) - F0[i];
      prev = e * std::pow(2 * R_, 3.) / r_type(N3);
      o << std::fixed << std::setprecision(8) << "00\t" << T0 << "\t" << prev << "\t-\t-\t-\t\n";
      step_no = 1;
    }
    std::ofstream writeF;
    const bool perstep = scheme == EDB_SCHEME_FEULER && writeF_;
    if (perstep) writeF.open("F.dat");
    // steps and times of the drivers' loops (DUFORT: 8 RK2 steps of dt/8 first)
    std::vector<r_type> ts;
    std::vector<int> phase;
    {
      r_type t = T0;
      if (dufort)
        for (int k = 0; k < 8; ++k) {
          t += dt / 8;
          ts.push_back(t);
          phase.push_back(0);
        }
      do {
        t += dt;
        ts.push_back(t);
        phase.push_back(1);
      } while (t < TEnd);
    }
    const int n = (int)ts.size();
    if (n > p.history) throw std::runtime_error("BKW3D: run longer than the history kept");
    std::vector<double> H(n), Hf(n), mean(n), mean_fix(n);
    std::vector<int> fx(n);
    std::vector<double> Fs(N3);
    if (perstep) {
      for (int k = 0; k < n; ++k) {
        check(edb_solver_step(s, 1), "step");
        check(edb_solver_get_driver_real(s, Fs.data()), "F");
        writeF << ts[k] << "\t";
        for (long i = 0; i < N3; ++i) writeF << Fs[i] << "\t";
        writeF << "\n";
      }
    } else {
      check(edb_solver_step(s, n), "step");
    }
    check(edb_solver_entropy_history(s, n, H.data(), Hf.data(), fx.data()), "entropy");
    check(edb_solver_mean_history(s, n, mean.data(), mean_fix.data()), "average");
    H_.assign(Hf.begin(), Hf.end());
    int k0 = 0;
    for (int k = 0; k < n; ++k) {
      if (dufort && k == 8) {
        o << "# Start the evolution using DUFORT...\n" << kHead[EDB_SCHEME_DUFORT];
        k0 = 8;
      }
      const int sch = dufort ? (phase[k] ? EDB_SCHEME_DUFORT : EDB_SCHEME_RK2) : scheme;
      row(sch, k - k0, step_no, ts[k], H[k], Hf[k], fx[k] != 0, mean[k], mean_fix[k], prev);
      prev = Hf[k];
      ++step_no;
    }
  }

  int N0_;
  r_type R_;
  int alpha_;
  bool filter_;
  bool fix_ = false, applied_ = false, conserve_ = false;
  int precision_ = EDB_PREC_FP64_PLAIN;
  std::function<r_type(r_type, r_type, r_type, r_type)> init_;
  std::vector<r_type> F_, H_;
  std::ostream* out_ = nullptr;
  bool writeF_ = true;
};

}  // namespace edb
