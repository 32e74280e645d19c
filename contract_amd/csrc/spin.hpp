// spin.hpp -- host-side construction of the DeGrand-Rossi gamma matrices and
// their sparse (one nonzero per row) encodings for kernel arguments.
//
// G(n) = g1^{n&1} g2^{(n>>1)&1} g3^{(n>>2)&1} g4^{(n>>3)&1}, the indexing of the
// reference's sparse tables (include/gamma.cuh:58-127) and of PyQUDA's
// gamma.Gamma(n) (baryon_2pt.py:72-81).  gamma_5 = G(15) = diag(1, 1, -1, -1).
#pragma once

#include <complex>
#include <cstring>

#include "device.hpp"

namespace contract_amd
{

using zhost = std::complex<double>;

struct Mat4 {
  zhost v[4][4];
};

inline Mat4 mat_identity()
{
  Mat4 m{};
  for (int r = 0; r < 4; ++r) m.v[r][r] = 1.0;
  return m;
}

inline Mat4 mat_mul(const Mat4 &a, const Mat4 &b)
{
  Mat4 c{};
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k)
      for (int s = 0; s < 4; ++s) c.v[r][s] += a.v[r][k] * b.v[k][s];
  return c;
}

inline Mat4 mat_transpose(const Mat4 &a)
{
  Mat4 t{};
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) t.v[r][c] = a.v[c][r];
  return t;
}

// Euclidean DeGrand-Rossi basis, rows = sink spin.
inline Mat4 dirac_basis(int mu)
{
  const zhost I(0, 1);
  Mat4 g{};
  switch (mu) {
  case 0: g.v[0][3] = I; g.v[1][2] = I; g.v[2][1] = -I; g.v[3][0] = -I; break;
  case 1: g.v[0][3] = -1.0; g.v[1][2] = 1.0; g.v[2][1] = 1.0; g.v[3][0] = -1.0; break;
  case 2: g.v[0][2] = I; g.v[1][3] = -I; g.v[2][0] = -I; g.v[3][1] = I; break;
  default: g.v[0][2] = 1.0; g.v[1][3] = 1.0; g.v[2][0] = 1.0; g.v[3][1] = 1.0; break;
  }
  return g;
}

inline Mat4 gamma_matrix(int n)
{
  Mat4 m = mat_identity();
  for (int mu = 0; mu < 4; ++mu)
    if ((n >> mu) & 1) m = mat_mul(m, dirac_basis(mu));
  return m;
}

inline Mat4 gamma5_times(int n) { return mat_mul(gamma_matrix(15), gamma_matrix(n)); }

// Encode a matrix with exactly one nonzero per row.  Returns false if the
// matrix is not of that form.
inline bool to_sparse(const Mat4 &m, SparseSpin &s)
{
  for (int r = 0; r < 4; ++r) {
    int col = -1;
    for (int c = 0; c < 4; ++c)
      if (m.v[r][c] != zhost(0, 0)) {
        if (col >= 0) return false;
        col = c;
      }
    if (col < 0) return false;
    s.col[r] = col;
    s.re[r] = m.v[r][col].real();
    s.im[r] = m.v[r][col].imag();
  }
  return true;
}

inline SparseSpin sparse_of(const Mat4 &m)
{
  SparseSpin s;
  std::memset(&s, 0, sizeof(s));
  to_sparse(m, s);
  return s;
}

} // namespace contract_amd
