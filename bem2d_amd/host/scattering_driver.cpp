// Example driver of the C++ drop-in API, written the way the reference's drivers use bem2d
// (examples/disk.cpp:57-111): build a geometry from AnalyticCurveElements, add the P0 basis, run
// SoundSoftScattering<PlaneWave, CombinedPlaneWave>, then evaluate the field and the far field.
//
//   scattering_driver <circle|kite|polygon> <n> <k> [N]
//
// Prints one JSON object: the solution, the far field at 360 angles, the total field at (-1.5, 0)
// and the diagonal and first column of the system matrix.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/bem2d/bem2d.h"

using namespace bem2d;

static void print_cvec(const char* name, const complex* v, std::size_t n, bool last = false)
{
    std::printf("\"%s\": [", name);
    for (std::size_t i = 0; i < n; i++)
        std::printf("%s[%.17g, %.17g]", i ? ", " : "", v[i].real(), v[i].imag());
    std::printf("]%s\n", last ? "" : ",");
}

int main(int argc, char** argv)
{
    std::string shape = argc > 1 ? argv[1] : "circle";
    int n = argc > 2 ? std::atoi(argv[2]) : 200;
    freqtype k = {argc > 3 ? std::atof(argv[3]) : 10.0, argc > 5 ? std::atof(argv[5]) : 0.0};
    int N = argc > 4 ? std::atoi(argv[4]) : 16;

    pGeometry pgeom;
    if (shape == "polygon") {
        std::vector<Point> pts = {Point(0, 0), Point(1.5, 0.2), Point(0.4, 1.1)};
        pgeom = Polygon(pts, n / 3).GetGeometry();
    } else {
        pCurve curve = shape == "kite" ? pCurve(new Kite) : pCurve(new Circle);
        std::vector<pElement> elements(n);
        for (int i = 0; i < n; i++) {
            elements[i] = pElement(new AnalyticCurveElement((1.0 * i) / n, (1.0 * (i + 1)) / n, curve, i));
            elements[i]->set_next((i + 1) % n);
            elements[i]->set_prev((i - 1 + n) % n);
        }
        pgeom = pGeometry(new Geometry(elements));
    }
    PolBasis::AddBasis(0, pgeom);

    Point direction = normalize(Point(1, 0));
    PlaneWave pw(direction, k);
    CombinedPlaneWave cpw(direction, k);
    SoundSoftScattering<PlaneWave, CombinedPlaneWave> soundsoft(pgeom, k, pw, cpw);
    soundsoft.SetQuadOption(3, N, 0.15);
    try {
        soundsoft.Discretize();
        soundsoft.Solve();
    } catch (LapackError&) {
        std::fprintf(stderr, "singular system\n");
        return 2;
    } catch (DeviceError&) {
        std::fprintf(stderr, "device error\n");
        return 3;
    }
    // the far field is defined for real wavenumbers only (an empty list for complex k)
    dvector theta(k.im == 0 ? 360 : 0);
    for (std::size_t a = 0; a < theta.size(); a++) theta[a] = 2 * PI * a / 360;
    pMatrix ff = k.im == 0 ? soundsoft.FarField(theta) : pMatrix(new Matrix(0, 1));
    std::vector<Point> pts{Point(-1.5, 0)};
    pcvector u = soundsoft.EvalSol(pts);
    pMatrix A = soundsoft.GetMatrix();
    const std::size_t nn = pgeom->size();
    cvector diag(nn);
    for (std::size_t i = 0; i < nn; i++) diag[i] = (*A->data)[i + nn * i];

    std::printf("{\n\"n\": %zu,\n", nn);
    print_cvec("solution", soundsoft.GetSolution()->data->data(), nn);
    print_cvec("farfield", ff->data->data(), theta.size());
    print_cvec("u_at_point", u->data(), 1);
    print_cvec("diag", diag.data(), nn);
    print_cvec("col0", A->data->data(), nn, true);
    std::printf("}\n");
    return 0;
}
