// cfd_core_hip_backend.cpp -- the reference-side wiring of `--backend hip`, as a maintainer adds it to
// AeroCFD's cfd_core (built with -DCFD_HAS_HIP=1 and linked against libaero_hip.so).
//
// It is written against the reference's own headers (cfd_core/mesh.hpp, solvers/euler/residual_assembly.hpp,
// backend.hpp) and include/aero_hip.h, in the reference's error style (bool + std::string* for the backend
// calls, std::invalid_argument / std::runtime_error where the driver throws):
//   * hip_available()               beside cuda_available()        (numerics/residual_stub.cpp:300-306)
//   * normalize_backend_hip()       normalize_backend with "hip"   (numerics/residual_stub.cpp:308-327)
//   * EulerHipMeshGuard             beside EulerCudaBuffersGuard   (solvers/euler_solver.cpp:452-460)
//   * init_euler_hip_mesh()         init_euler_device_buffers      (cuda_backend.hpp:44-46)
//   * make_hip_residual_config()    EulerResidualAssemblyConfig -> aero_residual_config
//                                   (solvers/euler/residual_assembly.hpp:12-31)
//   * euler_residual_hip()          euler_residual_cuda            (cuda_backend.hpp:49-52), the per-stage call
//                                   of the driver dispatch (solvers/euler_solver.cpp:870-891)
// The build's tests compile this file against the reference headers; nothing in the HIP library uses it.
#include <algorithm>
#include <cctype>
#include <stdexcept>
#include <string>
#include <vector>

#include "cfd_core/backend.hpp"
#include "cfd_core/mesh.hpp"
#include "cfd_core/solvers/euler/mesh_geometry.hpp"
#include "cfd_core/solvers/euler/residual_assembly.hpp"

#ifndef CFD_HAS_HIP
#define CFD_HAS_HIP 1
#endif

#if CFD_HAS_HIP
#include "aero_hip.h"
#endif

namespace cfd::core {

bool hip_available() {
#if CFD_HAS_HIP
  char err[256];
  return aero_hip_available(err, sizeof err) == 1;
#else
  return false;
#endif
}

// normalize_backend with the "hip" choice: the same checks, exception types and message style.
std::string normalize_backend_hip(std::string requested_backend) {
  std::transform(requested_backend.begin(), requested_backend.end(), requested_backend.begin(),
                 [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
  if (requested_backend.empty()) requested_backend = "cpu";
  if (requested_backend != "cpu" && requested_backend != "cuda" && requested_backend != "hip") {
    throw std::invalid_argument("Unsupported backend. Use 'cpu', 'cuda' or 'hip'.");
  }
  if (requested_backend == "hip") {
    if (!hip_available()) {
      throw std::runtime_error(
        "HIP backend requested but unavailable. Build with -DCFD_ENABLE_HIP=ON and ensure a gfx950 device is "
        "visible.");
    }
    return requested_backend;
  }
  return normalize_backend(requested_backend);
}

#if CFD_HAS_HIP
struct EulerHipMeshGuard {
  aero_hip_mesh* handle = nullptr;
  EulerHipMeshGuard() = default;
  EulerHipMeshGuard(const EulerHipMeshGuard&) = delete;
  EulerHipMeshGuard& operator=(const EulerHipMeshGuard&) = delete;
  ~EulerHipMeshGuard() { aero_hip_mesh_destroy(handle); }
};

// The device mesh of a driver run: the reference mesh widened to FP64, its boundary types
// (build_euler_boundary_types, whose values are AERO_BC_*) and the wall patch range.
bool init_euler_hip_mesh(const UnstructuredMesh& mesh, const std::vector<EulerBoundaryConditionType>& boundary_type,
                         int wall_start, int wall_count, EulerHipMeshGuard* guard, std::string* error_message) {
  static_assert(static_cast<int>(EulerBoundaryConditionType::kInterior) == AERO_BC_INTERIOR &&
                static_cast<int>(EulerBoundaryConditionType::kFarfield) == AERO_BC_FARFIELD &&
                static_cast<int>(EulerBoundaryConditionType::kSlipWall) == AERO_BC_SLIP_WALL,
                "boundary codes");
  std::vector<int> bc(boundary_type.size());
  for (std::size_t f = 0; f < bc.size(); ++f) bc[f] = static_cast<int>(boundary_type[f]);
  const std::vector<double> n(mesh.face_normal.begin(), mesh.face_normal.end());
  const std::vector<double> fc(mesh.face_center.begin(), mesh.face_center.end());
  const std::vector<double> area(mesh.face_area.begin(), mesh.face_area.end());
  const std::vector<double> vol(mesh.cell_volume.begin(), mesh.cell_volume.end());
  const std::vector<double> cc(mesh.cell_center.begin(), mesh.cell_center.end());
  aero_mesh_desc d{mesh.num_cells, mesh.num_faces, mesh.face_owner.data(), mesh.face_neighbor.data(),
                   bc.data(), n.data(), fc.data(), area.data(), vol.data(), cc.data(), wall_start, wall_count};
  const aero_hip_options opts{1, 0, AERO_GRAD_GREEN_GAUSS, AERO_ACCUM_AUTO};  // block size and accumulation: the layout probe
  char err[512];
  if (aero_hip_mesh_create(&d, &opts, &guard->handle, err, sizeof err) != AERO_OK) {
    if (error_message) *error_message = err;
    return false;
  }
  return true;
}

// The per-evaluation scalars of EulerResidualAssemblyConfig, one to one.
aero_residual_config make_hip_residual_config(const EulerResidualAssemblyConfig& c) {
  aero_residual_config r{};
  r.farfield[0] = c.farfield_state.rho;
  r.farfield[1] = c.farfield_state.rhou;
  r.farfield[2] = c.farfield_state.rhov;
  r.farfield[3] = c.farfield_state.rhow;
  r.farfield[4] = c.farfield_state.rhoE;
  r.gamma = c.gamma;
  r.flux_scheme = c.flux_scheme == EulerFluxScheme::kHllc ? AERO_FLUX_HLLC : AERO_FLUX_RUSANOV;
  r.limiter = c.limiter == LimiterType::kVenkat ? AERO_LIMITER_VENKAT : AERO_LIMITER_MINMOD;
  r.second_order = c.use_second_order ? 1 : 0;
  r.low_mach_fix = c.low_mach_dissipation_fix ? 1 : 0;
  r.low_mach_cutoff = c.low_mach_dissipation_cutoff;
  r.stab_mach_floor = c.stabilization_mach_floor;
  r.precond_farfield_bc = c.preconditioned_farfield_characteristic ? 1 : 0;
  r.precond_enabled = c.precondition.enabled ? 1 : 0;
  r.precond_mach_ref = c.precondition.mach_ref;
  r.precond_mach_min = c.precondition.mach_min;
  r.precond_beta_min = c.precondition.beta_min;
  r.precond_beta_max = c.precondition.beta_max;
  r.all_speed_fix = c.all_speed_flux_fix ? 1 : 0;
  r.all_speed_cutoff = c.all_speed_mach_cutoff;
  r.all_speed_f_min = c.all_speed_f_min;
  r.all_speed_ramp = c.all_speed_ramp_weight;
  r.limiter_frozen = 0;
  return r;
}

// euler_residual_cuda's role: one residual evaluation of a stage state ([N][5] float, the reference
// numbering) into residual [N][5] and spectral radius [N]; the device computes the gradients.
bool euler_residual_hip(const EulerHipMeshGuard& guard, const EulerResidualAssemblyConfig& config,
                        const std::vector<float>& state, std::vector<float>* residual,
                        std::vector<float>* spectral_radius, EulerResidualAssemblyDiagnostics* diagnostics,
                        std::string* error_message) {
  if (!guard.handle || !residual || !spectral_radius || !config.mesh) {
    if (error_message) *error_message = "euler_residual_hip: missing mesh handle or output";
    return false;
  }
  const int n = config.mesh->num_cells;
  if (static_cast<int>(state.size()) != 5 * n) {
    if (error_message) *error_message = "euler_residual_hip: state size does not match the mesh";
    return false;
  }
  const aero_residual_config rc = make_hip_residual_config(config);
  const std::vector<double> u(state.begin(), state.end());
  std::vector<double> r(static_cast<std::size_t>(5) * n), lam(n);
  aero_residual_diag dg{};
  char err[512];
  if (aero_hip_residual(guard.handle, &rc, u.data(), r.data(), lam.data(), &dg, err, sizeof err) != AERO_OK) {
    if (error_message) *error_message = err;
    return false;
  }
  residual->assign(r.begin(), r.end());
  spectral_radius->assign(lam.begin(), lam.end());
  if (diagnostics) {
    diagnostics->num_faces_first_order_fallback = dg.first_order_fallback;
    diagnostics->num_faces_diffusive_fallback = dg.diffusive_fallback;
    diagnostics->first_failure_face = dg.first_failure_face;
  }
  return true;
}

// The driver's per-stage dispatch (euler_solver.cpp:870-891) for "hip": throws like the CUDA branch.
void dispatch_stage_residual_hip(const EulerHipMeshGuard& guard, const EulerResidualAssemblyConfig& config,
                                 const std::vector<float>& state, std::vector<float>* residual,
                                 std::vector<float>* spectral_radius,
                                 EulerResidualAssemblyDiagnostics* diagnostics) {
  std::string error;
  if (!euler_residual_hip(guard, config, state, residual, spectral_radius, diagnostics, &error)) {
    throw std::runtime_error("HIP Euler residual failed: " + error);
  }
}
#endif  // CFD_HAS_HIP

}  // namespace cfd::core
