// Example driver of the C++ drop-in API in the shape of the reference's examples/disk.cpp:57-111:
// an AnalyticCurve mesh of the unit circle (n pieces of equal arc length, n from points per
// wavelength), a basis list added with the AddBasis helpers, SoundSoftScattering with
// SetQuadOption(L, N, sigma), set_polygons + set_plotInterior, then EvalSol and the far field.
//
//   disk_driver <lin|p0|pol2|wave|wavepol1> <ppw> <k> [L] [N] [curve] [graded]
//
// curve: circle (default), kite or ellipse (Ellipse(2)); graded = 1 builds the mesh with
// AnalyticCurve(ppw, k, curve, 1, L, sigma) instead of AnalyticCurve(n, curve).
// Prints one JSON object: the element parameter intervals, the solution, the far field at 360
// angles, the total field at (-1.5, 0) and at the interior point (0.1, 0.2), and the diagonal and
// first column of the system matrix.
#include <cstdio>
#include <cstdlib>
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
    std::string basis = argc > 1 ? argv[1] : "lin";
    int ppw = argc > 2 ? std::atoi(argv[2]) : 10;
    freqtype k = {argc > 3 ? std::atof(argv[3]) : 5.0, 0.0};
    int L = argc > 4 ? std::atoi(argv[4]) : 7;
    int N = argc > 5 ? std::atoi(argv[5]) : 30;
    std::string cname = argc > 6 ? argv[6] : "circle";
    bool graded = argc > 7 && std::atoi(argv[7]) != 0;

    pCurve cobj = cname == "kite" ? pCurve(new Kite)
                  : cname == "ellipse" ? pCurve(new Ellipse(2.0))
                                       : pCurve(new Circle);
    pGeometry pgeom;
    if (graded) {
        AnalyticCurve curve(ppw, k, cobj, 1, 2, 0.15);
        pgeom = curve.GetGeometry();
    } else {
        int n = (int)(cobj->Length() * k.re * ppw / 2.0 / PI);
        AnalyticCurve curve(n, cobj);
        pgeom = curve.GetGeometry();
    }
    if (basis == "lin")
        LinBasis::AddBasis(pgeom);
    else if (basis == "pol2")
        PolBasis::AddBasis(2, pgeom);
    else if (basis == "wave")
        WaveBasis::AddBasis(k, pgeom);
    else if (basis == "wavepol1")
        WavePolBasis::AddBasis(1, k, pgeom);
    else
        PolBasis::AddBasis(0, pgeom);

    Point direction = normalize(Point(1, 0));
    PlaneWave pw(direction, k);
    CombinedPlaneWave cpw(direction, k);
    SoundSoftScattering<PlaneWave, CombinedPlaneWave> soundsoft(pgeom, k, pw, cpw);
    soundsoft.SetQuadOption(L, N, 0.15);
    soundsoft.set_polygons(pgeom);
    soundsoft.set_plotInterior();
    try {
        soundsoft.Discretize();
        soundsoft.Solve();
    } catch (LapackError&) {
        std::fprintf(stderr, "singular system\n");
        return 2;
    } catch (DeviceError&) {
        std::fprintf(stderr, "device error\n");
        return 3;
    } catch (ParameterException&) {
        std::fprintf(stderr, "parameter exception\n");
        return 4;
    }
    dvector theta(360);
    for (std::size_t a = 0; a < theta.size(); a++) theta[a] = 2 * PI * a / 360;
    pMatrix ff = soundsoft.FarField(theta);
    std::vector<Point> pts{Point(-1.5, 0), Point(0.1, 0.2)};
    pcvector u = soundsoft.EvalSol(pts);
    pMatrix A = soundsoft.GetMatrix();
    const std::size_t nn = pgeom->size();
    cvector diag(nn);
    for (std::size_t i = 0; i < nn; i++) diag[i] = (*A->data)[i + nn * i];

    const auto& elems = pgeom->elements();
    std::printf("{\n\"n\": %zu,\n\"nn\": %zu,\n\"t\": [", elems.size(), nn);
    for (std::size_t e = 0; e < elems.size(); e++) {
        int type;
        double p[8];
        elems[e]->device_desc(type, p);
        std::printf("%s[%.17g, %.17g]", e ? ", " : "", p[0], p[1]);
    }
    std::printf("],\n");
    print_cvec("solution", soundsoft.GetSolution()->data->data(), nn);
    print_cvec("farfield", ff->data->data(), theta.size());
    print_cvec("u_at_points", u->data(), 2);
    print_cvec("diag", diag.data(), nn);
    print_cvec("col0", A->data->data(), nn, true);
    std::printf("}\n");
    return 0;
}
