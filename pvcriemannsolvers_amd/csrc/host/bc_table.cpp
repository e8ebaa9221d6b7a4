// BoundaryTable: per-point boundary values evaluated on the host.
#include "bc_table.hpp"

#include <algorithm>
#include <stdexcept>

namespace mcfd {

BoundaryTable::~BoundaryTable() {
  if (copied) {
    (void)hipEventSynchronize(copied);
    (void)hipEventDestroy(copied);
  }
  if (host) (void)hipHostFree(host);
}

void BoundaryTable::Init(int points_per_face, std::vector<int32_t> tags, std::vector<double> points,
                         const std::vector<double> &bc_states) {
  nfp = points_per_face;
  face_tag = std::move(tags);
  xyz = std::move(points);
  const_state = bc_states;
  fns.assign(const_state.size() / 5, Function{});
}

std::vector<double> BoundaryTable::Points(int tag) const {
  std::vector<double> out;
  for (size_t b = 0; b < face_tag.size(); ++b)
    if (face_tag[b] == tag) out.insert(out.end(), xyz.begin() + b * nfp * 3, xyz.begin() + (b + 1) * nfp * 3);
  return out;
}

void BoundaryTable::SetFunction(int tag, mcfd_bc_function f, void *user) {
  if (tag < 0 || tag >= (int)fns.size()) throw std::invalid_argument("boundary tag out of range");
  Function &fn = fns[tag];
  fn.f = f;
  fn.user = user;
  fn.faces.clear();
  for (size_t b = 0; b < face_tag.size(); ++b)
    if (face_tag[b] == tag) fn.faces.push_back((int64_t)b);
  fn.xyz = Points(tag);
  fn.values.assign(fn.faces.size() * nfp * 5, 0.0);
  EnsureStorage();
  if (!host) return;  // a mesh without boundary faces
  // the tag's constant state until the function writes it at a stage time
  Check(hipEventSynchronize(copied), "event sync");
  for (int64_t b : fn.faces)
    for (int p = 0; p < nfp; ++p)
      std::copy(const_state.begin() + tag * 5, const_state.begin() + tag * 5 + 5, host + (b * nfp + p) * 5);
  dirty = true;
  have_time = false;
}

void BoundaryTable::SetValues(int tag, const double *states) {
  if (tag < 0 || tag >= (int)fns.size()) throw std::invalid_argument("boundary tag out of range");
  Function &fn = fns[tag];
  fn.f = nullptr;
  fn.user = nullptr;
  EnsureStorage();
  if (!host) return;  // a mesh without boundary faces
  Check(hipEventSynchronize(copied), "event sync");
  int64_t i = 0;
  for (size_t b = 0; b < face_tag.size(); ++b)
    if (face_tag[b] == tag)
      for (int p = 0; p < nfp; ++p, ++i) std::copy(states + i * 5, states + i * 5 + 5, host + (b * nfp + p) * 5);
  dirty = true;
  have_time = false;
}

void BoundaryTable::EnsureStorage() {
  const size_t n = face_tag.size() * nfp * 5;
  if (host || n == 0) return;
  Check(hipHostMalloc((void **)&host, n * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
  Check(hipEventCreateWithFlags(&copied, hipEventDisableTiming), "hipEventCreate");
  Check(hipEventRecord(copied, 0), "hipEventRecord");
  dev.Allocate(n);
  // constant states everywhere; function and per-point tags overwrite theirs
  for (size_t b = 0; b < face_tag.size(); ++b)
    for (int p = 0; p < nfp; ++p)
      std::copy(const_state.begin() + face_tag[b] * 5, const_state.begin() + face_tag[b] * 5 + 5,
                host + (b * nfp + p) * 5);
}

const double *BoundaryTable::At(double t, hipStream_t stream) {
  if (!dev.ptr) return nullptr;
  bool timed = false;
  for (auto &fn : fns) timed |= fn.f && !fn.faces.empty();
  if (!timed && !dirty) return dev.ptr;                 // per-point values already on the device
  if (timed && have_time && t == time && !dirty) return dev.ptr;
  Check(hipEventSynchronize(copied), "event sync");  // the previous upload has read `host`
  for (auto &fn : fns) {
    if (!fn.f || fn.faces.empty()) continue;
    fn.f((int64_t)(fn.faces.size() * nfp), fn.xyz.data(), t, fn.values.data(), fn.user);
    for (size_t i = 0; i < fn.faces.size(); ++i)
      std::copy(fn.values.begin() + i * nfp * 5, fn.values.begin() + (i + 1) * nfp * 5,
                host + fn.faces[i] * nfp * 5);
  }
  Check(hipMemcpyAsync(dev.ptr, host, face_tag.size() * nfp * 5 * sizeof(double), hipMemcpyHostToDevice, stream),
        "boundary values upload");
  Check(hipEventRecord(copied, stream), "hipEventRecord");
  dirty = false;
  have_time = timed;
  time = t;
  return dev.ptr;
}

}  // namespace mcfd
