// Example driver of the C++ drop-in API in the shape of the reference's numrangecomps/disk.cpp:59-140:
// an AnalyticCurve mesh with PolBasis::AddBasis(2), the single-layer and conjugate double-layer
// operators and the mass matrix assembled separately, the combined operator
// Id + 2 CDL - 2 i eta SL moved to the mass matrix's orthonormal basis (ChangeBasis), then its
// eigenvalues (Eigenvalues), 2-norm and condition number (L2NormCond) and numerical range (NumRange).
//
//   numrange_driver <circle|kite> <n> <k> <nsteps> <numrange file>
//
// Prints one JSON object: the eigenvalues, the norm and condition number; NumRange writes its
// boundary points to the file.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/bem2d/bem2d.h"

using namespace bem2d;

int main(int argc, char** argv)
{
    std::string cname = argc > 1 ? argv[1] : "circle";
    int n = argc > 2 ? std::atoi(argv[2]) : 200;
    freqtype k = {argc > 3 ? std::atof(argv[3]) : 10.0, 0.0};
    int nsteps = argc > 4 ? std::atoi(argv[4]) : 100;
    std::string file = argc > 5 ? argv[5] : "numrange.txt";
    double eta1 = k.re;
    pCurve cobj = cname == "kite" ? pCurve(new Kite) : pCurve(new Circle);
    AnalyticCurve curve(n, cobj);
    pGeometry pgeom = curve.GetGeometry();
    PolBasis::AddBasis(2, pgeom);

    SingleLayer sl(k);
    ConjDoubleLayer cdl(k);
    QuadOption quadopts;
    quadopts.L = 3;
    quadopts.N = 5;
    quadopts.sigma = 0.15;
    try {
        Matrix dsl = *(DiscreteKernel(*pgeom, quadopts, sl));
        Matrix dcdl = *(DiscreteKernel(*pgeom, quadopts, cdl));
        Matrix Id = *(EvalIdent(*pgeom, quadopts));
        Matrix combined1 = Id + 2.0 * dcdl - complex(0, 2.0) * eta1 * dsl;
        combined1 = ChangeBasis(combined1, Id);
        pcvector eigvals;
        Eigenvalues(combined1, eigvals);
        double norm, cond;
        L2NormCond(combined1, norm, cond);
        NumRange(combined1, nsteps, file);
        std::printf("{\n\"size\": %zu,\n\"norm\": %.17g,\n\"cond\": %.17g,\n\"eigenvalues\": [",
                    eigvals->size(), norm, cond);
        for (std::size_t i = 0; i < eigvals->size(); i++)
            std::printf("%s[%.17g, %.17g]", i ? ", " : "", (*eigvals)[i].real(), (*eigvals)[i].imag());
        std::printf("]\n}\n");
    } catch (LapackError&) {
        std::fprintf(stderr, "lapack error\n");
        return 2;
    } catch (DeviceError&) {
        std::fprintf(stderr, "device error\n");
        return 3;
    } catch (ParameterException&) {
        std::fprintf(stderr, "parameter exception\n");
        return 4;
    }
    return 0;
}
