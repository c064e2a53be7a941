// conserve.hip -- conservative moment correction of collision spectra.
//
// No counterpart in the reference: its Fourier-Galerkin Q conserves mass
// (Q_0 == 0) but not momentum or energy.  With the forward transforms
// P_i = DFT(phi_i) of the five moment functions phi in {1, v1, v2, v3, |v|^2}
// on the grid x_j = -L + j h, the discrete moments of a spectrum are
//   m_i = h^3/N^3 sum_k Qhat_k conj(P_i,k),
// and the corrected spectrum is the least-squares nearest one on the Galerkin
// modes G with all five moments zero:
//   Qhat' = Qhat - sum_j mu_j P_j   (k in G),   A mu = m,
//   A_ij = sum_{k in G} P_j,k conj(P_i,k).
// Every phi is a sum of products of one-dimensional factors, so each P_i is
// supported on the three axis lines of the spectrum (two frequencies zero):
//   P_1 = N^3 [k = 0],
//   P_{v_d} = N^2 chi1(k_d) [other two k zero],
//   P_{|v|^2} = N^2 sum_d chi2(k_d) [other two k zero],
//   chi_p(k) = sum_j (-L + j h)^p exp(-2 pi i j k / N).
// The host builds the 3M-2 axis entries and A^-1 once per (N, L); the kernel
// (lane = instance) forms m from those entries, mu = A^-1 m, and updates them.
// The h^3/N^3 factor is common to m and A and cancels.
#include <cmath>
#include <utility>
#include <vector>

#include "edb_internal.hpp"
#include "ops.hpp"

namespace edb {
namespace {

typedef const double __attribute__((address_space(4)))* cdouble_p;

constexpr int kEnt = 11;  // centred index, then P_0..P_4 (re, im)

__global__ __launch_bounds__(64) void moment_correct_kernel(int E, int batch, int ld, double2* Q,
                                                            const double* tab_) {
  const cdouble_p tab = (cdouble_p)tab_;
  const int col = blockIdx.x * 64 + threadIdx.x;
  if (col >= batch) return;
  double mr[5] = {0., 0., 0., 0., 0.}, mi[5] = {0., 0., 0., 0., 0.};
  for (int e = 0; e < E; ++e) {
    const long c = (long)tab[e * kEnt];
    const double2 q = Q[c * ld + col];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const double pr = tab[e * kEnt + 1 + 2 * i], pi = tab[e * kEnt + 2 + 2 * i];
      mr[i] += q.x * pr + q.y * pi;  // q * conj(p)
      mi[i] += q.y * pr - q.x * pi;
    }
  }
  const long ai = (long)E * kEnt;
  double ur[5], ui[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    double sr = 0., si = 0.;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const double a = tab[ai + i * 5 + j];
      sr += a * mr[j];
      si += a * mi[j];
    }
    ur[i] = sr;
    ui[i] = si;
  }
  for (int e = 0; e < E; ++e) {
    const long c = (long)tab[e * kEnt];
    double2 q = Q[c * ld + col];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const double pr = tab[e * kEnt + 1 + 2 * j], pi = tab[e * kEnt + 2 + 2 * j];
      q.x -= ur[j] * pr - ui[j] * pi;
      q.y -= ur[j] * pi + ui[j] * pr;
    }
    Q[c * ld + col] = q;
  }
}

}  // namespace

int build_moment_table(int N, double L, std::vector<double>* tab, int* E_out) {
  Geometry g(N);
  const int n = g.n, M = g.M;
  const double h = 2. * L / (double)N;
  auto chi = [&](int p, int k, double* re, double* im) {
    long double sr = 0., si = 0.;
    for (int j = 0; j < N; ++j) {
      const long double x = -(long double)L + (long double)j * h;
      const long double w = p == 1 ? x : x * x;
      const long double a = -2.0L * (long double)M_PI * (long double)(((long)j * k % N + N) % N) / N;
      sr += w * cosl(a);
      si += w * sinl(a);
    }
    *re = (double)sr;
    *im = (double)si;
  };
  const double N2 = (double)N * N, N3 = N2 * N;
  tab->clear();
  auto centred = [&](int k1, int k2, int k3) {
    return (double)((((long)(k1 + n) * M) + (k2 + n)) * M + (k3 + n));
  };
  // origin
  {
    double c1r, c1i, c2r, c2i;
    chi(1, 0, &c1r, &c1i);
    chi(2, 0, &c2r, &c2i);
    const double row[kEnt] = {centred(0, 0, 0), N3, 0., N2 * c1r, N2 * c1i, N2 * c1r, N2 * c1i,
                              N2 * c1r, N2 * c1i, 3. * N2 * c2r, 3. * N2 * c2i};
    tab->insert(tab->end(), row, row + kEnt);
  }
  for (int d = 0; d < 3; ++d)
    for (int k = -n; k <= n; ++k) {
      if (k == 0) continue;
      double c1r, c1i, c2r, c2i;
      chi(1, k, &c1r, &c1i);
      chi(2, k, &c2r, &c2i);
      double row[kEnt] = {0.};
      int kk[3] = {0, 0, 0};
      kk[d] = k;
      row[0] = centred(kk[0], kk[1], kk[2]);
      row[1 + 2 * (1 + d)] = N2 * c1r;
      row[2 + 2 * (1 + d)] = N2 * c1i;
      row[9] = N2 * c2r;
      row[10] = N2 * c2i;
      tab->insert(tab->end(), row, row + kEnt);
    }
  const int E = (int)(tab->size() / kEnt);
  // A_ij = sum_e Re(P_j conj(P_i)), inverted on the diagonally scaled system
  long double A[5][5], Inv[5][5], d[5];
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      long double s = 0.;
      for (int e = 0; e < E; ++e) {
        const double* r = tab->data() + (long)e * kEnt;
        s += (long double)r[1 + 2 * j] * r[1 + 2 * i] + (long double)r[2 + 2 * j] * r[2 + 2 * i];
      }
      A[i][j] = s;
    }
  for (int i = 0; i < 5; ++i) {
    if (!(A[i][i] > 0.)) {
      set_error("moment correction: singular moment system");
      return kInvalid;
    }
    d[i] = 1.0L / sqrtl(A[i][i]);
  }
  long double W[5][10];
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      W[i][j] = d[i] * A[i][j] * d[j];
      W[i][5 + j] = (i == j) ? 1.0L : 0.0L;
    }
  for (int c = 0; c < 5; ++c) {
    int piv = c;
    for (int r = c + 1; r < 5; ++r)
      if (fabsl(W[r][c]) > fabsl(W[piv][c])) piv = r;
    for (int j = 0; j < 10; ++j) std::swap(W[c][j], W[piv][j]);
    if (fabsl(W[c][c]) < 1e-14L) {
      set_error("moment correction: singular moment system");
      return kInvalid;
    }
    const long double inv = 1.0L / W[c][c];
    for (int j = 0; j < 10; ++j) W[c][j] *= inv;
    for (int r = 0; r < 5; ++r) {
      if (r == c) continue;
      const long double f = W[r][c];
      for (int j = 0; j < 10; ++j) W[r][j] -= f * W[c][j];
    }
  }
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) Inv[i][j] = d[i] * W[i][5 + j] * d[j];
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) tab->push_back((double)Inv[i][j]);
  *E_out = E;
  return kOk;
}

int launch_moment_correct(const double* d_tab, int E, int batch, int ld, double* d_q, hipStream_t s) {
  EDB_REQUIRE(d_tab && E >= 1, "moment correction: table not built");
  EDB_REQUIRE(batch >= 1 && batch <= ld, "moment correction: bad batch");
  hipLaunchKernelGGL(moment_correct_kernel, dim3((batch + 63) / 64), dim3(64), 0, s, E, batch, ld,
                     (double2*)d_q, d_tab);
  EDB_HIP_TRY(hipGetLastError());
  return kOk;
}

}  // namespace edb
