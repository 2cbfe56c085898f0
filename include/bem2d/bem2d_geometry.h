// Drop-in replacement of lib/bem2d_geometry.h:17-75 and lib/bem2d_geometry.cpp:9-64.
// AddBasis does not print "Elems: N" (lib/bem2d_geometry.h:36).
#ifndef BEM2D_DROPIN_GEOMETRY_H
#define BEM2D_DROPIN_GEOMETRY_H

#include <map>
#include <vector>

#include "bem2d_basis.h"
#include "bem2d_element.h"

namespace bem2d {

class Geometry;
typedef std::shared_ptr<Geometry> pGeometry;

class Geometry {
public:
    typedef std::vector<std::map<std::size_t, pBasis>> basis_map;
    typedef std::vector<std::pair<pElement, pBasis>> flat_basis_map;

    Geometry(const std::vector<pElement>& elems) : elements_(elems), size_(0)
    {
        elements_bases_.resize(elems.size());
        for (std::size_t i = 0; i < elements_.size(); i++) points_.push_back(elements_[i]->First());
    }

    // Concatenation of several geometries; re-indexes the elements and closes the ring
    // (lib/bem2d_geometry.cpp:37-54, including for open pieces).
    Geometry(std::vector<pGeometry>& geoms) : size_(0)
    {
        for (auto& g : geoms)
            for (auto& e : g->elements()) elements_.push_back(e);
        int nelems = (int)elements_.size();
        elements_bases_.resize(nelems);
        for (int i = 1; i < nelems - 1; i++) {
            elements_[i]->set_index(i);
            elements_[i]->set_next(i + 1);
            elements_[i]->set_prev(i - 1);
        }
        elements_.front()->set_index(0);
        elements_.front()->set_next(1);
        elements_.front()->set_prev(nelems - 1);
        elements_.back()->set_index(nelems - 1);
        elements_.back()->set_prev(nelems - 2);
        elements_.back()->set_next(0);
        for (std::size_t i = 0; i < elements_.size(); i++) points_.push_back(elements_[i]->First());
    }

    inline std::vector<pElement>& elements() { return elements_; }
    inline const std::vector<pElement>& elements() const { return elements_; }
    inline const basis_map elements_bases() { return elements_bases_; }

    inline void AddBasis(pBasis b)
    {
        for (std::size_t i = 0; i < elements_.size(); i++) {
            elements_bases_[i].insert(std::pair<std::size_t, pBasis>(size_, b));
            size_++;
        }
    }
    inline void AddBasis(std::size_t i, pBasis b)
    {
        elements_bases_[i].insert(std::pair<std::size_t, pBasis>(size_, b));
        size_++;
    }

    // Element-major list of (element, basis) = the matrix row/column order.
    inline std::shared_ptr<flat_basis_map> FlatMap() const
    {
        std::shared_ptr<flat_basis_map> p(new flat_basis_map);
        for (std::size_t i = 0; i < elements_bases_.size(); i++)
            for (auto jt = elements_bases_[i].begin(); jt != elements_bases_[i].end(); jt++)
                p->push_back(std::pair<pElement, pBasis>(elements_[i], jt->second));
        return p;
    }

    inline std::size_t size() const { return size_; }
    inline const PointVector& points() const { return points_; }

private:
    std::vector<pElement> elements_;
    basis_map elements_bases_;
    PointVector points_;
    int size_;
};

} // namespace bem2d

#endif
