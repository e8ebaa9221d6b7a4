// Boundary values that vary in space and time: the reference registers a
// boundary condition as a function f(x, t) of the global point and the
// system's time (spatial/fem.hpp:196-222) and evaluates it at every boundary
// quadrature point whenever a residual is formed (dg/general.hpp:92-202,
// fr/general.hpp:435-640).  A BoundaryTable holds, per boundary face point,
// the given state the stage kernels read; tags with a function are
// re-evaluated on the host at each new stage time and uploaded, the other
// tags keep their constant states.
#ifndef MINICFD_HOST_BC_TABLE_HPP_
#define MINICFD_HOST_BC_TABLE_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../../include/minicfd.h"
#include "common.hpp"

namespace mcfd {

struct BoundaryTable {
  int nfp = 0;                       // points per boundary face
  std::vector<int32_t> face_tag;     // [boundary face]
  std::vector<double> xyz;           // [boundary face][point][3], in the kernels' point order
  std::vector<double> const_state;   // [tag][5]
  struct Function {
    mcfd_bc_function f = nullptr;
    void *user = nullptr;
    std::vector<int64_t> faces;      // boundary faces of the tag
    std::vector<double> xyz, values;  // their points and the function's values there
  };
  std::vector<Function> fns;         // [tag]
  double *host = nullptr;            // pinned [boundary face][point][5]
  DeviceArray<double> dev;
  hipEvent_t copied = nullptr;       // the last upload has read `host`
  bool have_time = false;
  double time = 0;
  bool dirty = false;                // `host` holds values the device table lacks

  BoundaryTable() = default;
  BoundaryTable(const BoundaryTable &) = delete;
  BoundaryTable &operator=(const BoundaryTable &) = delete;
  ~BoundaryTable();

  // Geometry and constant states of the mesh's boundary faces.
  void Init(int points_per_face, std::vector<int32_t> tags, std::vector<double> points,
            const std::vector<double> &bc_states);
  // Points of the faces with this tag ([n][3], boundary-face order).
  std::vector<double> Points(int tag) const;
  // Installs (or with f == null removes) the function of a tag.
  void SetFunction(int tag, mcfd_bc_function f, void *user);
  // Fixed values at every point of a tag ([n][5] in Points(tag) order): a
  // boundary function of the point only, evaluated once; removes the tag's
  // function.
  void SetValues(int tag, const double *states);
  // Device table to read at time t (tags with a function re-evaluated and the
  // table uploaded on `stream` if t is new; per-point values uploaded once),
  // or null while no tag has a function or per-point values.
  const double *At(double t, hipStream_t stream);
  bool Active() const { return dev.ptr != nullptr; }

 private:
  // Pinned host table, device table and event, filled with the tags' constants.
  void EnsureStorage();
};

}  // namespace mcfd

#endif  // MINICFD_HOST_BC_TABLE_HPP_
