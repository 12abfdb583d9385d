// bm_topology.h -- the rest of analyze on the device (bone-morphometry
// 0.7.0-rc.3): the cubical complex of the analysis bone, its Betti numbers,
// BoneJ's stack edge correction, the bone-marrow interface faces and the
// ROI-boundary bone fraction.
//   analyze_topology             src/topology.rs:29-51
//   cubical_measures             src/topology.rs:53-171
//   bonej_connectivity           src/topology.rs:175-427
//   bone_marrow_interface_faces  src/topology.rs:442-490
//   count_components / count_enclosed_background_components_6 / flood
//                                src/topology.rs:504-701
//   roi_boundary_bone_fraction   src/metrics.rs:166-202
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bm_engine.h"

namespace bm {

// Integer counts of one volume (the analysis bone: data != 0 inside the ROI;
// without a ROI every voxel is inside).
struct TopologyCounts {
  // cubical_measures: cells of the closed cubical complex
  int64_t vertices = 0, edges = 0, faces = 0, cubes = 0;
  // bonej_connectivity: stack_vertices, stack_edges, stack_faces,
  // face_vertices, face_edges, edge_vertices
  int64_t stack_vertices = 0, stack_edges = 0, stack_faces = 0, face_vertices = 0,
          face_edges = 0, edge_vertices = 0;
  // bone_marrow_interface_faces (analysis bone against the ROI)
  int64_t interface_faces = 0;
  // roi_boundary_bone_fraction: bone voxels, and those with a 6-neighbour
  // outside the ROI or the image
  uint64_t bone_voxels = 0, boundary_bone_voxels = 0;
  // Betti numbers: 26-connected bone components, 6-connected cavities
  uint64_t beta0 = 0, beta2 = 0;
};

enum TopologyParts : unsigned {
  kTopoCubical = 1,   // cells, interface faces, ROI-boundary counts
  kTopoBonej = 2,     // the stack counts
  kTopoBetti = 4,     // beta0, beta2
};

// Computes the requested parts on stream s and waits for them.  d_roi may be
// null.  Uses the engine's scratch slots 2 (labels) and 3 (counters).
void topology_device(Engine& e, const uint8_t* d_bone, const uint8_t* d_roi, const Dims& dims,
                     unsigned parts, TopologyCounts* out, hipStream_t s);

}  // namespace bm
