// Coupled elastic-acoustic 2D meshes: one assembly per medium plus the
// interface terms (coupled2d.cpp) and the driver that advances both media
// in the reference's order (coupled2d.hip).
#pragma once

#include <memory>
#include <vector>

#include "solver2d.hpp"

namespace smx {

// True when the mesh has both elastic and acoustic elements (an element is
// acoustic when mu == 0 at its 25 points).
bool mesh_is_coupled2d(const smx_mesh2d &mesh);

struct CoupledAssembly2D {
  std::unique_ptr<HostAssembly2D> el, ac; // one assembly per medium
  // station r of the run is receiver rec_sub[r] of medium rec_med[r]
  // (0 elastic, 1 acoustic)
  std::vector<int> rec_med, rec_sub;
  // a summary the C ABI reports (sizes summed over both media, medium 2)
  HostAssembly2D summary;
};

// Splits the mesh by medium (sources and receivers go to the medium of the
// element that contains them; on the interface itself, to the solid),
// assembles each part with the run's start time, and builds the interface
// entries of both sides (coupling_terms/{acoustic_elastic,elastic_acoustic}
// .hpp with the weights and normals of conforming_interfaces/dim2/impl/
// interface_container.tpp:20-80).
void assemble_coupled2d(const smx_mesh2d &mesh, const smx_sources2d &sources,
                        const smx_receivers2d &receivers, double dt, int nstep,
                        const AssemblyOptions &opt, CoupledAssembly2D &out);

// The driver: each step runs the acoustic update and corrector, then the
// elastic ones (core/specfem/solver/time_marching.tpp:33-77), as one fused
// pass per medium; FP32 only, forward runs only.
std::unique_ptr<Solver2DBase>
make_coupled_solver2d(std::unique_ptr<CoupledAssembly2D> a, int precision,
                      int graph_steps);

} // namespace smx
