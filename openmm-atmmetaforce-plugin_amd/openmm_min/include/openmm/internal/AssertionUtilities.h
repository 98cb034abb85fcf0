// Minimal OpenMM-compatible core: assertion helpers used by the OpenMM API and
// its tests (ASSERT, ASSERT_EQUAL, ASSERT_VALID_INDEX).
#pragma once
#include <cmath>
#include <sstream>
#include <string>
#include "openmm/OpenMMException.h"

namespace OpenMM {

inline void throwException(const char* file, int line, const std::string& details) {
    std::stringstream m;
    m << "Assertion failure at " << file << ":" << line;
    if (!details.empty()) m << ".  " << details;
    throw OpenMMException(m.str());
}

}  // namespace OpenMM

#define ASSERT(cond) { if (!(cond)) OpenMM::throwException(__FILE__, __LINE__, ""); }
#define ASSERT_EQUAL(expected, found) \
    { if (!((expected) == (found))) { std::stringstream d_; d_ << "Expected " << (expected) << ", found " << (found); \
      OpenMM::throwException(__FILE__, __LINE__, d_.str()); } }
#define ASSERT_EQUAL_TOL(expected, found, tol) \
    { double e_ = (expected), f_ = (found); double s_ = std::fabs(e_) > 1.0 ? std::fabs(e_) : 1.0; \
      if (!(std::fabs(e_ - f_) / s_ <= (tol))) { std::stringstream d_; d_ << "Expected " << e_ << ", found " << f_; \
      OpenMM::throwException(__FILE__, __LINE__, d_.str()); } }
#define ASSERT_VALID_INDEX(index, vector) \
    { if ((index) < 0 || (index) >= (int)(vector).size()) OpenMM::throwException(__FILE__, __LINE__, "Index out of range"); }
