// mirror_main.cpp -- the reference's test mains (CalQ/test.cpp,
// testEntropyFix.cpp, testfft.cpp, testBKW3D.cpp) written against the C++
// host side (include/edb.hpp) of the GPU library.  Writes its results to the
// directory given as argv[1] for tests/test_gpu_cpp.py to check against the
// oracle.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "edb.hpp"

typedef double r_type;
typedef std::complex<double> c_type;

static void dump(const std::string& path, const void* p, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(3);
  }
  std::fclose(f);
}

// The six drivers with the reference's per-step output (BKW3D::setOutput),
// each table written to <dir>/table_<scheme>.txt: N = 11, R = 6, filter off, BKW data at
// s = 0.8 (T0 = 6 ln 2 for s(t) = 1 - 0.4 exp(-t/6)), dt = 0.05 to T0 + 0.4,
// beta = 2 for CFE and DUFORT, fix on.
static int driver_tables(const std::string& dir) {
  const double T0 = 6. * std::log(2.), dt = 0.05, TEnd = T0 + 0.4;
  const char* names[] = {"FEuler", "RK2", "SSP3", "RK4", "CFE", "DUFORT"};
  for (int sc = 0; sc < 6; ++sc) {
    std::ofstream out(dir + "/table_" + names[sc] + ".txt");
    edb::BKW3D<c_type, r_type> bkw(11, 6.0, 0, /*filter=*/false);
    bkw.setFixEntropy(true);
    bkw.setOutput(&out, false);
    switch (sc) {
      case 0: bkw.FEuler(T0, TEnd, dt); break;
      case 1: bkw.RK2(T0, TEnd, dt); break;
      case 2: bkw.SSP3(T0, TEnd, dt); break;
      case 3: bkw.RK4(T0, TEnd, dt); break;
      case 4: bkw.CFE(T0, TEnd, dt, 2.0); break;
      default: bkw.DUFORT(T0, TEnd, dt, 2.0); break;
    }
    dump(dir + "/table_F_" + names[sc] + ".bin", bkw.F().data(), sizeof(double) * bkw.F().size());
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::string dir = argv[1];
  if (argc > 2 && std::string(argv[2]) == "tables") return driver_tables(dir);
  // CalQ/test.cpp: genB + calQ on inputs rand()%10 (here a fixed LCG).
  int N0 = 9, N = N0 * N0 * N0;
  std::vector<c_type> fin(N), Qout(N);
  unsigned x = 12345u;
  for (auto& v : fin) {
    x = x * 1103515245u + 12345u;
    double a = (double)((x >> 16) % 10);
    x = x * 1103515245u + 12345u;
    double b = (double)((x >> 16) % 10);
    v = c_type{a, b};
  }
  std::vector<c_type> Qc(N);
  {
    edb::QCalculator<c_type, r_type> qc(N0, 0);
    qc.genB();
    qc.calQ(Qout.data(), fin.data());
    qc.genBCompressed();
    qc.calQCompressed(Qc.data(), fin.data());
  }
  dump(dir + "/calq_fin.bin", fin.data(), sizeof(c_type) * N);
  dump(dir + "/calq_q.bin", Qout.data(), sizeof(c_type) * N);
  dump(dir + "/calq_q_compressed.bin", Qc.data(), sizeof(c_type) * N);

  // CalQ/testEntropyFix.cpp with N = 1000.
  {
    int n = 1000;
    std::vector<r_type> f1(n), f2(n);
    for (int i = 0; i < n; i++) f1[i] = (r_type)i;
    for (int i = 0; i < n / 2; i++) f2[i] = std::floor(((r_type)n - 1.0) / 2.0);
    for (int i = n / 2; i < n; i++) f2[i] = std::ceil(((r_type)n - 1.0) / 2.0);
    edb::EntropyFix<int, r_type> ef(n);
    r_type h[3];
    h[0] = ef.EntropyCal(f1);
    h[1] = ef.EntropyCal(f2);
    ef.FixSol(f1, h[1]);
    h[2] = ef.EntropyCal(f1);
    std::printf("H(f^{n+1}) = %.2f\nH(f(t_{n+1})) = %.2f\nH(hatf^{n+1}) = %.2f\n", h[0], h[1], h[2]);
    dump(dir + "/efix_h.bin", h, sizeof(h));
  }

  // CalQ/testfft.cpp: a round trip of the BKW profile through fft3d.
  {
    int n0 = 8, n = n0 * n0 * n0;
    std::vector<double> F(2 * n), Fhat(2 * n), Ftest(2 * n);
    double h = 12. / double(n0 - 1);
    int l = 0;
    for (int i = 0; i < n0; ++i)
      for (int j = 0; j < n0; ++j)
        for (int k = 0; k < n0; ++k) {
          double xx = -6. + i * h, y = -6. + j * h, z = -6. + k * h;
          double v2 = xx * xx + y * y + z * z, s = 1. - 2. / 5.;
          F[2 * l] = 1. / std::pow(2. * M_PI * s, 1.5) * std::exp(-v2 / (2. * s)) *
                     ((5. * s - 3.) / (2. * s) + (1. - s) / (2. * s * s) * v2);
          F[2 * l + 1] = 0.;
          ++l;
        }
    edb::fft3d f3d;
    f3d.fft(n0, n0, n0, (edb::fft3d::complex_t*)F.data(), (edb::fft3d::complex_t*)Fhat.data());
    f3d.ifft(n0, n0, n0, (edb::fft3d::complex_t*)Ftest.data(), (edb::fft3d::complex_t*)Fhat.data());
    dump(dir + "/fft_F.bin", F.data(), sizeof(double) * 2 * n);
    dump(dir + "/fft_Fhat.bin", Fhat.data(), sizeof(double) * 2 * n);
    dump(dir + "/fft_Ftest.bin", Ftest.data(), sizeof(double) * 2 * n);
  }

  // CalQ/testBKW3D.cpp-like: a short RK4 run with the diagnostic fix, the
  // default-constructed driver (filter on, as the reference's drivers).
  {
    edb::BKW3D<c_type, r_type> bkw(9, 8.0);
    bkw.setFixEntropy(true);
    bkw.RK4(0., 0.1, 0.02);
    dump(dir + "/bkw_F.bin", bkw.F().data(), sizeof(double) * bkw.F().size());
    dump(dir + "/bkw_H.bin", bkw.entropy().data(), sizeof(double) * bkw.entropy().size());
  }
  // error behaviour: an argument the reference would assert on throws here
  try {
    edb::EntropyFix<int, r_type> ef(10);
    std::vector<r_type> f(5, 1.0);
    ef.EntropyCal(f);
    std::printf("no exception\n");
    return 4;
  } catch (const std::runtime_error& e) {
    std::printf("rejected: %s\n", e.what());
  }
  return 0;
}
