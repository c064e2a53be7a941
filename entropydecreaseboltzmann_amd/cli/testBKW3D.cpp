// testBKW3D -- the reference's command-line driver (CalQ/testBKW3D.cpp) on
// the GPU library:  testBKW3D N dt TEnd
//
// Same arguments, header, per-step table on stdout, F.dat (t and the N^3
// real values of F after every step) and N<N>.dt<dt>.dat (outV1Direction:
// x, F(x, h/2, h/2), f_0) as the reference's main: forward Euler from t = 0
// on [-R, R]^3 with R = 6.62132, the Jackson filter on (the QCalculator
// default), the entropy fix on, and the reference's initial data -- the
// smooth periodic profile that BKW3D::funExact returns in its active branch
// (BKW.h:96-112), 3.2 + sum_{j<11} (j/55) [sin(j pi s1) + sin(j pi s2) +
// sin(j pi s3)] with s_d = (x_d + R)/(2R) - 1/2.  Differences: a missing or
// bad argument prints the usage and exits with status 2 (the reference reads
// argv unchecked), and even N runs the defined even-N operator.
#include <complex>
#include <cstdlib>
#include <iostream>
#include <string>

#include "edb.hpp"

typedef double r_type;
typedef std::complex<double> c_type;

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: testBKW3D N dt TEnd\n";
    return 2;
  }
  const int N = std::atoi(argv[1]);
  const r_type dt = std::atof(argv[2]);
  const r_type TEnd = std::atof(argv[3]);
  if (N < 3 || N > 33 || !(dt > 0.) || !(TEnd > 0.)) {
    std::cerr << "usage: testBKW3D N dt TEnd  (3 <= N <= 33, dt > 0, TEnd > 0)\n";
    return 2;
  }
  const r_type R = 6.62132;
  const bool FixEntropy = true;

  std::cout << "##################################################\n"
            << "# Entropy fixed method for BKW model\n"
            << "# spatial domain  :\t[-" << R << ", " << R << "]\n"
            << "# number of grid N:\t" << N << "^3\n"
            << "# time interval   :\t[0, " << TEnd << "]\n"
            << "# time step     dt:\t" << dt << "\n";
  if (FixEntropy)
    std::cout << "# Entropy fixing stratege is adopted.\n";
  else
    std::cout << "# Entropy fixing stratege is not adopted.\n";
  std::cout << "##################################################\n";

  auto initial = [R](r_type, r_type x, r_type y, r_type z) {
    const r_type s[3] = {(x + R) / (2. * R) - 0.5, (y + R) / (2. * R) - 0.5, (z + R) / (2. * R) - 0.5};
    r_type v = 3.2;
    for (int j = 0; j < 11; ++j)
      for (int d = 0; d < 3; ++d) v += r_type(j) / 55. * std::sin(j * s[d] * M_PI);
    return v;
  };
  try {
    edb::BKW3D<c_type, r_type> bkw3d(N, R, 0, /*filter=*/true);
    bkw3d.setFixEntropy(FixEntropy);
    bkw3d.setInitial(initial);
    bkw3d.setOutput(&std::cout);
    bkw3d.FEuler(0, TEnd, dt);
    const std::string filename = "N" + std::to_string(N) + ".dt" + std::to_string(dt) + ".dat";
    bkw3d.outV1Direction(filename);
  } catch (const std::exception& e) {
    std::cerr << "testBKW3D: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
