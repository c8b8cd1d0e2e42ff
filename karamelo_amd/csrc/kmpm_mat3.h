// kmpm_mat3.h -- 3x3 matrix functions of the TLMPM step (row-major), for
// the device kernels and the host upload alike.  Products sum k = 0, 1, 2
// in turn, the order Eigen's fixed-size expressions evaluate.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace kmpm {

__host__ __device__ inline void m3_mul(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

__host__ __device__ inline void m3_transpose(const double *A, double *T) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[3 * j + i] = A[3 * i + j];
}

// Cofactor matrix: C[i][j] is the cofactor of A[i][j]; A^-1 = C^T / det A.
__host__ __device__ inline void m3_cofactors(const double *A, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      C[3 * i + j] = A[3 * i1 + j1] * A[3 * i2 + j2] - A[3 * i1 + j2] * A[3 * i2 + j1];
    }
}

// Matrix3d::inverse for 3x3 (solid.cpp:1186): the cofactors of column 0 give
// the determinant, every entry is a cofactor times 1 / det.
__host__ __device__ inline void m3_inverse(const double *A, double *Ai) {
  double C[9];
  m3_cofactors(A, C);
  const double det = C[0] * A[0] + C[3] * A[3] + C[6] * A[6];
  const double invdet = 1.0 / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Ai[3 * i + j] = C[3 * j + i] * invdet;
}

__host__ __device__ inline double m3_det(const double *m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Rotation R of the polar decomposition F = R U (PolDec, mpm_math.h:104-133,
// takes R = U V^T of an SVD; R is unique for det F > 0).  Newton's
// iteration X <- (X + X^-T) / 2 from X = F until the largest change is at
// most 1e-9, then one more iteration (quadratic convergence reaches rounding
// level).  Returns false when det X <= 0.
__host__ __device__ inline bool polar_rotation(const double *F, double *R) {
  double X[9], C[9];
  for (int e = 0; e < 9; e++) X[e] = F[e];
  bool last = false;
  for (int it = 0; it < 100; it++) {
    m3_cofactors(X, C);
    const double det = X[0] * C[0] + X[1] * C[1] + X[2] * C[2];
    if (!(det > 0)) return false;
    const double inv = 1.0 / det;
    double diff = 0;
    for (int e = 0; e < 9; e++) {
      const double xn = 0.5 * (X[e] + C[e] * inv);   // X^-T = C / det
      const double d = fabs(xn - X[e]);
      diff = d > diff ? d : diff;
      X[e] = xn;
    }
    if (last) break;
    if (diff <= 1e-9) last = true;
  }
  for (int e = 0; e < 9; e++) R[e] = X[e];
  return m3_det(R) > 0.0;
}

// Real parts of the eigenvalues of F (EigenSolver<Matrix3d>,
// solid.cpp:1401-1408): with s = tr F / 3 the characteristic polynomial of
// F - s I is the depressed cubic t^3 + p t + q (p the sum of the principal
// 2x2 minors, q = -det); three real roots in trigonometric form, otherwise
// one real root (Cardano) and a complex pair of real part -t/2.
__host__ __device__ inline void eig_real_parts(const double *F, double *re) {
  const double s = (F[0] + F[4] + F[8]) / 3.0;
  double G[9];
  for (int e = 0; e < 9; e++) G[e] = F[e];
  G[0] -= s;
  G[4] -= s;
  G[8] -= s;
  const double p =
      (G[0] * G[4] - G[1] * G[3]) + (G[0] * G[8] - G[2] * G[6]) + (G[4] * G[8] - G[5] * G[7]);
  const double q = -m3_det(G);
  if (p == 0 && q == 0) {
    re[0] = re[1] = re[2] = s;
    return;
  }
  const double disc = q * q / 4.0 + p * p * p / 27.0;
  if (disc > 0) {
    const double sq = sqrt(disc);
    const double t = cbrt(-q / 2.0 + sq) + cbrt(-q / 2.0 - sq);
    re[0] = s + t;
    re[1] = re[2] = s - t / 2.0;
  } else {
    const double m = 2.0 * sqrt(-p / 3.0);
    double arg = (3.0 * q / (2.0 * p)) * sqrt(-3.0 / p);
    arg = arg < -1.0 ? -1.0 : (arg > 1.0 ? 1.0 : arg);
    const double th = acos(arg) / 3.0;
    for (int k = 0; k < 3; k++) re[k] = s + m * cos(th - 2.0943951023931957 * k);
  }
}

// vol0PK1 = vol0 J (R sigma R^T) F^-T (solid.cpp:1277-1280,1368-1372).
__host__ __device__ inline void pk1(double vol0, double J, const double *R, const double *sig,
                                    const double *Finv, double *out) {
  double Rt[9], RS[9], RSRt[9], FiT[9];
  const double a = vol0 * J;
  m3_transpose(R, Rt);
  m3_mul(R, sig, RS);
  m3_mul(RS, Rt, RSRt);
  for (int e = 0; e < 9; e++) RSRt[e] = a * RSRt[e];
  m3_transpose(Finv, FiT);
  m3_mul(RSRt, FiT, out);
}

// TLMPM rate of deformation (solid.cpp:1218-1234): L = Fdot F^-1,
// D = (R^T (L + L^T) R) / 2, returned as the symmetric xx,yy,zz,xy,xz,yz.
__host__ __device__ inline void tl_rate(const double *Fdot, const double *Finv, const double *R,
                                        double *Dsym) {
  double L[9], Lt[9], S[9], Rt[9], T1[9], T2[9];
  m3_mul(Fdot, Finv, L);
  m3_transpose(L, Lt);
  for (int e = 0; e < 9; e++) S[e] = L[e] + Lt[e];
  m3_transpose(R, Rt);
  m3_mul(Rt, S, T1);
  m3_mul(T1, R, T2);
  Dsym[0] = 0.5 * T2[0];
  Dsym[1] = 0.5 * T2[4];
  Dsym[2] = 0.5 * T2[8];
  Dsym[3] = 0.5 * T2[1];
  Dsym[4] = 0.5 * T2[2];
  Dsym[5] = 0.5 * T2[5];
}

}  // namespace kmpm
