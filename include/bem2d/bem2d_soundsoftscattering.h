// Drop-in replacement of lib/bem2d_soundsoftscattering.h:17-168 (the sound-soft scattering facade).
// DiscretizeMatrix assembles A = M + 2 K_combined on the GPU in one pass (the reference's
// DiscreteKernel + EvalIdent + (*pA_)=(*pId_)+2.0*(*pA_), lib/bem2d_soundsoftscattering.h:95-103);
// Solve uses the FP64 GPU LU.
// The interior mask of EvalIncident (set_polygons, set_plotInterior) is kept; the plotting/output
// members (SetOutput, Write*) and NormCond are not part of this build.
#ifndef BEM2D_DROPIN_SOUNDSOFTSCATTERING_H
#define BEM2D_DROPIN_SOUNDSOFTSCATTERING_H

#include <limits>
#include <vector>

#include "bem2d_fun.h"
#include "bem2d_geometry.h"
#include "bem2d_mathroutines.h"
#include "bem2d_quadrature.h"

namespace bem2d {

namespace device {
// A = mass_scale * M + kernel_scale * K_combined(k, eta) in one GPU pass.
pMatrix AssembleCombined(const Geometry& geom, QuadOption opts, freqtype k, double eta,
                         double mass_scale, double kernel_scale);
} // namespace device

template <typename T1, typename T2>
class SoundSoftScattering {
public:
    SoundSoftScattering(pGeometry geom, freqtype kvalue, T1 inc, T2 rhs)
        : pgeom_(geom), k_(kvalue), incident_(inc), frhs_(rhs)
    {
        quadopts_.L = 3;
        quadopts_.N = 5;
        quadopts_.sigma = 0.15;
    }
    virtual ~SoundSoftScattering() {}

    void SetQuadOption(int L, int N, double sigma)
    {
        quadopts_.L = L;
        quadopts_.N = N;
        quadopts_.sigma = sigma;
    }
    // The mass matrix the reference keeps beside A (pId_) is folded into A's assembly; GetIdent
    // assembles it on first use.
    virtual void DiscretizeMatrix()
    {
        pA_ = device::AssembleCombined(*pgeom_, quadopts_, k_, frhs_.eta(), 1.0, 2.0);
        pId_.reset();
    }
    virtual void DiscretizeRhs()
    {
        prhs_ = DotProdBasFuns(*pgeom_, quadopts_, frhs_);
        (*prhs_) = 2.0 * (*prhs_);
    }
    void Discretize()
    {
        DiscretizeMatrix();
        DiscretizeRhs();
    }
    void Solve() { psol_ = SolveSystem(*pA_, *prhs_); }

    inline pMatrix GetMatrix() const { return pA_; }
    inline pMatrix GetIdent()
    {
        if (!pId_ && pA_) pId_ = EvalIdent(*pgeom_, quadopts_);
        return pId_;
    }
    inline pMatrix GetSolution() const { return psol_; }

    // Points inside the polygons given by set_polygons get NaN once set_plotInterior was called
    // (lib/bem2d_soundsoftscattering.h:128-145).
    pcvector EvalIncident(const std::vector<Point>& points)
    {
        pcvector pz(new cvector(points.size()));
        std::vector<int> inpoly(points.size(), 0);
        if (plotInterior_ && !polygons_.empty()) InPolygon(polygons_, points, inpoly);
        for (std::size_t i = 0; i < points.size(); i++)
            (*pz)[i] = inpoly[i] == 0 ? incident_(points[i])
                                      : complex(std::numeric_limits<double>::quiet_NaN(), 0.0);
        return pz;
    }
    void set_polygons(std::vector<pGeometry>& polygons) { polygons_ = polygons; }
    void set_polygons(pGeometry polygon) { polygons_ = std::vector<pGeometry>{polygon}; }
    void set_plotInterior() { plotInterior_ = true; }

    // u = u_inc - S phi at the points (lib/bem2d_soundsoftscattering.h:149-168).
    pcvector EvalSol(const std::vector<Point>& points)
    {
        std::shared_ptr<Geometry::flat_basis_map> bfuns = pgeom_->FlatMap();
        SingleLayer sl(k_);
        pcvector kerneleval(new cvector(points.size()));
        Gauss1D g1d(quadopts_.N);
        EvalKernel(*bfuns, points, *kerneleval, *psol_, sl, g1d);
        pcvector result = EvalIncident(points);
        for (std::size_t i = 0; i < points.size(); i++) (*result)[i] -= (*kerneleval)[i];
        return result;
    }

    // New: far field of the scattered field at the angles theta (radians), one column per RHS.
    pMatrix FarField(const dvector& theta) { return bem2d::FarField(*pgeom_, quadopts_, k_, *psol_, theta); }

private:
    pGeometry pgeom_;
    freqtype k_;
    T1 incident_;
    T2 frhs_;
    QuadOption quadopts_;
    pMatrix pA_, pId_, prhs_, psol_;
    std::vector<pGeometry> polygons_;
    bool plotInterior_ = false;
};

} // namespace bem2d

#endif
