// Drop-in replacement of lib/bem2d_mathroutines.h:17-73 (dense matrices and the solve).
// Matrix keeps the reference's layout: column-major complex128 in a shared vector, dim[0] x dim[1].
// SolveSystem runs on the GPU (rocSOLVER zgetrf/zgetrs behind the C-ABI) and throws LapackError
// for an exactly singular matrix, as the reference does (lib/bem2d_mathroutines.cpp:323-351).
#ifndef BEM2D_DROPIN_MATHROUTINES_H
#define BEM2D_DROPIN_MATHROUTINES_H

#include <string>

#include "bem2d_defs.h"
#include "bem2d_exceptions.h"
#include "bem2d_kernel.h"

namespace bem2d {

struct Matrix {
    explicit Matrix(int n) : Matrix(n, n) {}
    explicit Matrix(int m, int n)
    {
        dim[0] = m;
        dim[1] = n;
        msize = m;
        nsize = n;
        data = pcvector(new cvector((std::size_t)m * n));
    }
    Matrix(const Matrix& m) : dim{m.dim[0], m.dim[1]}, msize(m.msize), nsize(m.nsize)
    {
        data = pcvector(new cvector(*m.data));
    }
    Matrix& operator=(const Matrix& m)
    {
        dim[0] = m.dim[0];
        dim[1] = m.dim[1];
        msize = m.msize;
        nsize = m.nsize;
        data = pcvector(new cvector(*m.data));
        return *this;
    }
    pcvector data;
    int dim[2];
    int msize;
    int nsize;
};

typedef std::shared_ptr<Matrix> pMatrix;

Matrix operator+(const Matrix& lhs, const Matrix& rhs);
Matrix operator-(const Matrix& lhs, const Matrix& rhs);
Matrix operator*(const Matrix& lhs, const complex& alpha);
Matrix operator*(const complex& alpha, const Matrix& rhs);
Matrix operator*(const Matrix& lhs, const double& alpha);
Matrix operator*(const double& alpha, const Matrix& rhs);
Matrix operator*(const Matrix& lhs, const Matrix& rhs);

// inpoly[j] = number of the polygons (closed vertex lists Geometry::points(), the first points of
// the elements) that contain testpoints[j] (lib/bem2d_mathroutines.cpp:732-758).
class Geometry;
void InPolygon(const std::vector<std::shared_ptr<Geometry>> polygons,
               const std::vector<Point>& testpoints, std::vector<int>& inpoly);
void InPolygon(const std::shared_ptr<Geometry> polygon, const std::vector<Point>& testpoints,
               std::vector<int>& inpoly);

// X = m^{-1} rhs (FP64 LU with partial pivoting on the GPU); m is not modified.
pMatrix SolveSystem(Matrix& m, Matrix& rhs);

// Spectral drivers (lib/bem2d_mathroutines.h:48-85), on the GPU (rocSOLVER / rocBLAS and the
// library's Hessenberg + multishift QR): largest singular value and condition number; all singular
// values (descending); eigenvalues of a Hermitian matrix (ascending); eigenvalues of a general matrix
// (zgeev's set, in the order of the computed triangular form); k in the orthonormal basis of the
// mass matrix m, U^{-H} k U^{-1} with m = U^H U; the boundary of the numerical range at n angles
// written to filename as "re im" lines; the conjugate transpose.  The eigenvector overload
// Eigenvalues(k, evalues, evectors) is not provided.
void L2NormCond(const Matrix& stiff, double& norm, double& cond);
void SingularValues(const Matrix& stiff, pdvector& singvals);
void HermitianEigenvalues(const Matrix& k, pdvector& evalues);
void Eigenvalues(const Matrix& k, pcvector& evalues);
Matrix ChangeBasis(const Matrix& k, const Matrix& m);
void NumRange(const Matrix& m, int n, std::string filename);
Matrix ConjTranspose(const Matrix& m);

} // namespace bem2d

#endif
