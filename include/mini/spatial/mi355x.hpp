// mini::spatial::mi355x::FiniteElement: a mini::temporal::System<double>
// whose residual and stages run on one MI355X through libminicfd.
//
// It takes the place of the reference's FiniteElement family behind the
// System contract (spatial/fem.hpp:48,224-258): fr::General
// (spatial/fr/general.hpp:218-231), dg::General (spatial/dg/general.hpp:38-40)
// and WithLimiter (spatial/with_limiter.hpp:40-66).  Columns keep the
// reference layout column[(cell * N + i) * 5 + k].  Boundary conditions are
// given per boundary tag at creation (constant states, the SetSmartBoundary /
// SetInviscidWall / SetSupersonic* registrations of fem.hpp:196-222), and a
// tag's f(xyz, t) after it (SetBoundaryFunction).  Errors of the C ABI are
// thrown as std::runtime_error carrying its message.
#ifndef MINI_SPATIAL_MI355X_HPP_
#define MINI_SPATIAL_MI355X_HPP_

#include <algorithm>
#include <array>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "mini/temporal/ode.hpp"
#include "minicfd.h"

namespace mini {
namespace spatial {
namespace mi355x {

inline void Check(int rc) {
  if (rc != MCFD_OK) throw std::runtime_error(std::string("libminicfd: ") + mcfd_last_error());
}

// fem.hpp:196-222 as constant states per boundary tag
inline mcfd_bc_desc SmartBoundary(double const (&state)[5]) {
  mcfd_bc_desc bc{ MCFD_BC_SMART, 0, { state[0], state[1], state[2], state[3], state[4] } };
  return bc;
}
inline mcfd_bc_desc SupersonicInlet(double const (&state)[5]) {
  mcfd_bc_desc bc{ MCFD_BC_SUPERSONIC_INLET, 0, { state[0], state[1], state[2], state[3], state[4] } };
  return bc;
}
inline mcfd_bc_desc InviscidWall() { return mcfd_bc_desc{ MCFD_BC_INVISCID_WALL, 0, { 0, 0, 0, 0, 0 } }; }
inline mcfd_bc_desc SupersonicOutlet() { return mcfd_bc_desc{ MCFD_BC_SUPERSONIC_OUTLET, 0, { 0, 0, 0, 0, 0 } }; }

class FiniteElement : public temporal::System<double> {
 public:
  using Column = temporal::System<double>::Column;

  FiniteElement(mcfd_mesh_desc const &mesh, mcfd_scheme_desc const &scheme, std::vector<mcfd_bc_desc> const &bcs) {
    Check(mcfd_create(&mesh, &scheme, bcs.data(), (int32_t)bcs.size(), &h_));
    Check(mcfd_sizes(h_, &n_cells_, &n_per_cell_, &size_));
    degree_ = scheme.degree;
    cell_type_ = mesh.cell_type;
  }
  FiniteElement(FiniteElement const &) = delete;
  FiniteElement &operator=(FiniteElement const &) = delete;
  FiniteElement(FiniteElement &&that) noexcept
      : h_(std::exchange(that.h_, nullptr)), n_cells_(that.n_cells_), n_per_cell_(that.n_per_cell_),
        size_(that.size_), degree_(that.degree_), cell_type_(that.cell_type_), lumped_lobatto_(that.lumped_lobatto_),
        bc_functions_(std::move(that.bc_functions_)),
        field_names_(std::move(that.field_names_)) {}
  ~FiniteElement() override {
    if (h_) mcfd_destroy(h_);
  }
  using Global = std::array<double, 3>;  // a point, as the reference's Global

  // fr::General<Part, rotated::Euler<Riemann>> with the Vincent correction
  // c_next (1 = DG, p/(2p+1) = Huynh g2) on Gauss-Legendre or Lobatto nodes
  template <class Riemann>
  static FiniteElement FluxReconstruction(mcfd_mesh_desc const &mesh, int degree,
                                          std::vector<mcfd_bc_desc> const &bcs, int nodes = MCFD_NODES_LEGENDRE,
                                          double c_next = 1.0, int precision = MCFD_PRECISION_FP64) {
    mcfd_scheme_desc s = Scheme(MCFD_SCHEME_FR_GENERAL, degree, Riemann::kId);
    s.nodes = nodes;
    s.c_next = c_next;
    s.precision = precision;
    return FiniteElement(mesh, s, bcs);
  }
  // fr::Lobatto<Part, rotated::Euler<Riemann>> (spatial/fr/lobatto.hpp): FR
  // on Lobatto nodes with the lumped correction c_next = p / (2p + 1), and
  // its doubled time step
  template <class Riemann>
  static FiniteElement FluxReconstructionLobatto(mcfd_mesh_desc const &mesh, int degree,
                                                 std::vector<mcfd_bc_desc> const &bcs) {
    FiniteElement fe = FluxReconstruction<Riemann>(mesh, degree, bcs, MCFD_NODES_LOBATTO,
                                                   degree / (2.0 * degree + 1.0));
    fe.lumped_lobatto_ = true;
    return fe;
  }
  // dg::General<Part<..., Projection<double, 3, p, 5>>, rotated::Euler<Riemann>>
  template <class Riemann>
  static FiniteElement DiscontinuousGalerkin(mcfd_mesh_desc const &mesh, int degree,
                                             std::vector<mcfd_bc_desc> const &bcs) {
    return FiniteElement(mesh, Scheme(MCFD_SCHEME_DG_GENERAL, degree, Riemann::kId), bcs);
  }
  // WithLimiter<dg::General<...>, Limiter>: the limiter is applied after every
  // SetSolutionColumn, Approximate and fused stage (with_limiter.hpp:56-59)
  template <class Riemann, class Limiter>
  static FiniteElement WithLimiter(mcfd_mesh_desc const &mesh, int degree, std::vector<mcfd_bc_desc> const &bcs,
                                   bool all_cells = false, double w0 = 0.001, double eps = 1e-6) {
    mcfd_scheme_desc s = Scheme(MCFD_SCHEME_DG_GENERAL, degree, Riemann::kId);
    s.limiter = all_cells ? Limiter::kAll : Limiter::kFlagged;
    s.weno_w0 = w0;
    s.weno_eps = eps;
    return FiniteElement(mesh, s, bcs);
  }
  // WithLimiter(limiter_ptr, ...) of with_limiter.hpp:40-43: the limiter
  // object's w0 and eps and its own troubled-cell rule (Limiter::kDefault)
  template <class Riemann, class Limiter>
  static FiniteElement WithLimiter(Limiter const *limiter_ptr, mcfd_mesh_desc const &mesh, int degree,
                                   std::vector<mcfd_bc_desc> const &bcs) {
    mcfd_scheme_desc s = Scheme(MCFD_SCHEME_DG_GENERAL, degree, Riemann::kId);
    s.limiter = limiter_ptr ? Limiter::kDefault : MCFD_LIMITER_NONE;
    if (limiter_ptr) {
      s.weno_w0 = limiter_ptr->w0();
      s.weno_eps = limiter_ptr->eps();
    }
    return FiniteElement(mesh, s, bcs);
  }

  // ---- temporal::System (ode.hpp:17-48, fem.hpp:224-258)
  void SetTime(double t_curr) override { Check(mcfd_set_time(h_, t_curr)); }
  void SetSolutionColumn(Column const &column) override {
    if ((int64_t)column.size() != size_) throw std::runtime_error("column size mismatch");
    Check(mcfd_set_solution_column(h_, column.data()));
  }
  Column GetSolutionColumn() const override {
    Column column(size_);
    Check(mcfd_get_solution_column(h_, column.data()));
    return column;
  }
  Column GetResidualColumn() const override {
    Column column(size_);
    Check(mcfd_get_residual_column(h_, column.data()));
    return column;
  }

  // ---- the rest of the FiniteElement surface used by the drivers
  // FiniteElement::GetTimeStep(dt_guess, rk_order) of fem.hpp:412-435:
  // min(dt_guess, CFL * min over cells of length / max |u| + c), with the CFL
  // number of the reference's table for this degree; fr::Lobatto (a handle
  // from FluxReconstructionLobatto) returns min(dt_guess, 2 dt)
  // (fr/lobatto.hpp:220-222).  Entries the table leaves empty give 0, as
  // they do in the reference.
  double GetTimeStep(double dt_guess, int rk_order) const {
    const double dt = GetTimeStepForCfl(CflNumber(rk_order, degree_), dt_guess);
    return lumped_lobatto_ ? std::min(dt_guess, 2.0 * dt) : dt;
  }
  // The same reduction with an explicit CFL number.
  double GetTimeStepForCfl(double cfl, double dt_guess = 1e300) const {
    double dt = 0;
    Check(mcfd_get_time_step(h_, cfl, dt_guess, &dt));
    return dt;
  }
  // cfl_[rk_order][degree] of fem.hpp:26-37
  static double CflNumber(int rk_order, int degree) {
    static const double kCfl[5][9] = {
      { 1.000 },
      { 1.000, 0.333 },
      { 1.000, 0.333, 0.209, 0.130, 0.089, 0.066, 0.051, 0.040, 0.033 },
      { 1.256, 0.409, 0.209, 0.130, 0.089, 0.066, 0.051, 0.040, 0.033 },
      { 1.392, 0.464, 0.235, 0.145, 0.100, 0.073, 0.056, 0.045, 0.037 },
    };
    if (rk_order < 0 || rk_order > 4 || degree < 0 || degree > 8)
      throw std::invalid_argument("no CFL number for this RK order and degree");
    return kCfl[rk_order][degree];
  }
  // Approximate(f) of fem.hpp:147-154: values of f at SamplePoints()
  std::vector<double> SamplePoints() const {
    int32_t n = 0;
    Check(mcfd_sample_points(h_, &n, nullptr));
    std::vector<double> xyz((size_t)n_cells_ * n * 3);
    Check(mcfd_sample_points(h_, &n, xyz.data()));
    return xyz;
  }
  void Approximate(std::vector<double> const &values) { Check(mcfd_approximate(h_, values.data())); }
  // Approximate(func) of fem.hpp:147-154 with the reference drivers' initial
  // condition, a callable Value(Global const &xyz) (demo/euler/riemann2d.cpp:20-25):
  // evaluated here at SamplePoints(); its result is read as r[0..4]
  template <class Callable, class = std::enable_if_t<std::is_invocable_v<Callable &, Global const &>>>
  void Approximate(Callable &&func) {
    const std::vector<double> xyz = SamplePoints();
    const size_t n = xyz.size() / 3;
    std::vector<double> values(n * 5);
    for (size_t p = 0; p < n; ++p) {
      const Global x = { xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2] };
      const auto v = func(x);
      for (int k = 0; k < 5; ++k) values[5 * p + k] = v[k];
    }
    Approximate(values);
  }

  // ---- boundary values that vary in space and time
  // The f(xyz, t) of SetSmartBoundary / SetSupersonicInlet / SetSubsonicInlet /
  // SetSubsonicOutlet (fem.hpp:196-222; demo/euler/riemann2d.cpp:28-33) for
  // the boundary tag `tag`, whose kind was given at creation: a callable
  // Value(Global const &xyz, double t), read as r[0..4], evaluated at every
  // point of the tag at each residual's time.  It is called from this
  // object's methods on the calling thread and must not throw.
  template <class Callable>
  void SetBoundaryFunction(int32_t tag, Callable &&func) {
    auto fn = std::make_unique<BoundaryFunction>(
        [f = std::forward<Callable>(func)](int64_t n, double const *xyz, double t, double *states) {
          for (int64_t i = 0; i < n; ++i) {
            const Global x = { xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2] };
            const auto v = f(x, t);
            for (int k = 0; k < 5; ++k) states[5 * i + k] = v[k];
          }
        });
    Check(mcfd_set_boundary_function(h_, tag, &CallBoundaryFunction, fn.get()));
    bc_functions_[tag] = std::move(fn);
  }
  // A boundary function of the point only, given as its values [n][5] at
  // BoundaryPoints(tag): evaluated once instead of at every stage time.
  void SetBoundaryValues(int32_t tag, std::vector<double> const &states) {
    Check(mcfd_set_boundary_values(h_, tag, states.data()));
    bc_functions_.erase(tag);
  }
  // The points of boundary tag `tag`, [n][3], in the order the functions see them.
  std::vector<double> BoundaryPoints(int32_t tag) const {
    int64_t n = 0;
    Check(mcfd_boundary_points(h_, tag, &n, nullptr));
    std::vector<double> xyz((size_t)n * 3);
    Check(mcfd_boundary_points(h_, tag, &n, xyz.data()));
    return xyz;
  }
  // polynomial().LocalToGlobalAndValue of every cell at n local points
  // (mesh/vtk.hpp:431-460): xyz [cell][n][3] and values [cell][n][5] of the
  // current solution, or of `column` (which a host-only handle needs)
  void Evaluate(std::vector<double> const &local, std::vector<double> *xyz, std::vector<double> *values,
                double const *column = nullptr) const {
    const int32_t n = (int32_t)(local.size() / 3);
    xyz->resize((size_t)n_cells_ * n * 3);
    values->resize((size_t)n_cells_ * n * 5);
    Check(mcfd_evaluate(h_, column, n, local.data(), xyz->data(), values->data()));
  }
  // Part::SetFieldNames / GetFieldName (mesh/part.hpp:477-482), the names a frame gives the components
  void SetFieldNames(std::array<std::string, 5> const &names) { field_names_ = names; }
  std::string const &GetFieldName(int k) const { return field_names_.at(k); }
  int32_t CellType() const { return cell_type_; }
  // RungeKutta<3>::Update (rk.hpp:31-49) fused on the device; the state stays in HBM
  void Rk3Update(double t_curr, double dt) { Check(mcfd_rk3_update(h_, t_curr, dt)); }

  int64_t CountCells() const { return n_cells_; }
  int Degree() const { return degree_; }
  int32_t UnknownsPerCell() const { return n_per_cell_; }
  int64_t ColumnSize() const { return size_; }
  mcfd_solver *handle() const { return h_; }

 private:
  static mcfd_scheme_desc Scheme(int32_t scheme, int degree, int32_t riemann) {
    mcfd_scheme_desc s{};
    s.scheme = scheme;
    s.degree = degree;
    s.nodes = MCFD_NODES_LEGENDRE;
    s.riemann = riemann;
    s.precision = MCFD_PRECISION_FP64;
    s.c_next = 1.0;
    s.limiter = MCFD_LIMITER_NONE;
    s.weno_w0 = 0.001;
    s.weno_eps = 1e-6;
    return s;
  }

  mcfd_solver *h_ = nullptr;
  int64_t n_cells_ = 0;
  int32_t n_per_cell_ = 0;
  int64_t size_ = 0;
  int degree_ = 0;
  int32_t cell_type_ = MCFD_HEXA_8;
  bool lumped_lobatto_ = false;  // fr::Lobatto
  // the boundary functions the library calls back, owned per tag (their
  // addresses are the callbacks' user pointers)
  using BoundaryFunction = std::function<void(int64_t, double const *, double, double *)>;
  static void CallBoundaryFunction(int64_t n, double const *xyz, double t, double *states, void *user) noexcept {
    (*static_cast<BoundaryFunction *>(user))(n, xyz, t, states);
  }
  std::map<int32_t, std::unique_ptr<BoundaryFunction>> bc_functions_;
  // demo/euler/sourceless.cpp:109-110
  std::array<std::string, 5> field_names_{ "Density", "MomentumX", "MomentumY", "MomentumZ",
                                           "EnergyStagnationDensity" };
};

}  // namespace mi355x
}  // namespace spatial
}  // namespace mini

#endif  // MINI_SPATIAL_MI355X_HPP_
