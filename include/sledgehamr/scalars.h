// Scalar-field registration.
// Reference interface: sledgehamr::ScalarField (source/scalars.h:10-44).
#pragma once

#include <string>
#include <vector>

namespace sledgehamr {

/** A registered field component.  Construction appends the field to `sfv`
 *  and records its component id (its position in that vector). */
class ScalarField {
  public:
    ScalarField(std::string str, std::vector<ScalarField *> &sfv,
                bool is_momentum)
        : name(std::move(str)), id(static_cast<int>(sfv.size())),
          is_conjugate_momentum(is_momentum) {
        sfv.push_back(this);
    }
    operator int() const { return id; }

    std::string name;
    const int id;
    bool is_conjugate_momentum;
};

}  // namespace sledgehamr
