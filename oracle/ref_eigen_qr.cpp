// ref_eigen_qr.cpp -- ORACLE helper (test infrastructure only).
// Exposes Eigen 3.4.0's FullPivHouseholderQR, the decomposition the reference
// uses for its Pulay Gram inverse (mpDPFT.cpp:5089-5107), so the tests can
// check this repository's own full-pivoting QR against it.  Compiled against
// the Eigen headers vendored in the reference tree by oracle/build_ref.py;
// output goes to oracle/_ref/ only.
#include <Eigen/Dense>

extern "C" int ref_pulay_coefficients(int n, const double *A, double *coeff, double *testval) {
    Eigen::MatrixXd M(n, n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) M(i, j) = A[i * n + j];
    Eigen::FullPivHouseholderQR<Eigen::MatrixXd> qr(M);
    *testval = -1.;
    if (!qr.isInvertible()) return 0;
    Eigen::MatrixXd inv = qr.inverse();
    Eigen::MatrixXd T = inv * M - Eigen::MatrixXd::Identity(n, n);
    double tv = T.squaredNorm();
    *testval = tv;
    if (tv > 1.0e-8) return 0;
    Eigen::VectorXd x = inv * Eigen::VectorXd::Constant(n, 1.0);
    double s = 0.;
    for (int i = 0; i < n; i++) s += x(i);
    for (int i = 0; i < n; i++) coeff[i] = x(i) / s;
    return 1;
}
