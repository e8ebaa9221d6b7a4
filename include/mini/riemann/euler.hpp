// mini::riemann solver selection for the C++ host surface.
//
// In the reference the Riemann solver is a template argument of the spatial
// scheme (rotated::Euler<euler::Exact<Gas, 3>> etc., riemann/rotated/euler.hpp,
// riemann/euler/{exact,hllc}.hpp).  Here the solvers run inside the device
// kernels, selected at creation; these tag types keep the reference's names
// and carry the C-ABI identifier.
#ifndef MINI_RIEMANN_EULER_HPP_
#define MINI_RIEMANN_EULER_HPP_

#include <cstdint>

#include "minicfd.h"

namespace mini {
namespace riemann {
namespace euler {

// riemann/euler/exact.hpp:16 (iterative star state, FindRoot to 1e-8)
struct Exact { static constexpr int32_t kId = MCFD_RIEMANN_EXACT; };
// riemann/euler/hllc.hpp:17
struct HartenLaxLeerContact { static constexpr int32_t kId = MCFD_RIEMANN_HLLC; };
// python/riemann.py:405-451, extended to 3-D in the rotated frame
struct Roe { static constexpr int32_t kId = MCFD_RIEMANN_ROE; };

}  // namespace euler
}  // namespace riemann
}  // namespace mini

#endif  // MINI_RIEMANN_EULER_HPP_
