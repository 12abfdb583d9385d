"""MI355X (gfx950) exact squared EDT and local thickness of bone-morphometry.

Python face of the C ABI in include/bm.h, mirroring the reference crate's
interface for this path (bone-morphometry 0.7.0-rc.3):

  BinaryVolume.new / phase_inside    src/volume.rs:12-41, 84-105
  Phase                               src/edt.rs:8-32
  euclidean_distance_transform_squared  src/edt.rs:57-156
  local_thickness                     src/local_thickness.rs:16-44
  analyze_thickness_spacing           src/metrics.rs:64-132 (thickness/spacing part)
  analyze                             src/metrics.rs:64-164 (every Metrics field)
  cubical_measures, analyze_topology, bonej_connectivity,
  bone_marrow_interface_faces         src/topology.rs:29-490
  FeatureMode                         src/metrics.rs:12-24
  read_binary_tiff, read_roi_tiff,
  read_isotropic_spacing_um           src/tiff_io.rs:18-317 (the input step)
  require_isotropic_spacing           src/sdf_curvature.rs:147-160
  signed_distance_field_um            src/sdf_curvature.rs:162-180
  gaussian_kernel                     src/sdf_curvature.rs:182-198
  gaussian_smooth_nearest_in_place    src/sdf_curvature.rs:200-269
  smoothed_signed_distance_field_um   the two above in sequence, as
                                      src/sdf_curvature.rs:98-105 and
                                      src/minkowski.rs:64-71 call them

and two forms with no counterpart in the reference:

  Batch                               many device-resident analyze calls as one
                                      plan, replayed from a hipGraph
  Server                              request loop with HBM-budget admission and
                                      continuous batching
  autotune / tuning_load              per-shape-class kernel tunables (JSON cache)
  set_stats_order / stats_order       exact device sums (default) or the
                                      reference's sequential f64 order

Every call runs on the GPU through libbm.so; errors raise BmError carrying the
reference's message.
"""
from .api import (Batch, BinaryVolume, BmError, DistributionStats, FeatureMode,  # noqa: F401
                  Metrics, Phase, Server, ThicknessSpacing, analyze, analyze_device,
                  analyze_thickness_spacing, analyze_topology, bone_marrow_interface_faces,
                  bonej_connectivity, cubical_measures,
                  autotune, euclidean_distance_transform_squared, gaussian_kernel,
                  gaussian_smooth_nearest_in_place, gaussian_smooth_nearest_in_place_device,
                  local_thickness, local_thickness_stages, read_binary_tiff,
                  read_binary_tiff_device, read_isotropic_spacing_um, read_roi_tiff,
                  read_roi_tiff_device, require_isotropic_spacing,
                  signed_distance_field_um, signed_distance_field_um_device,
                  smoothed_signed_distance_field_um, smoothed_signed_distance_field_um_device,
                  StatsOrder, set_stats_order, stats_order, tuning_load)
from . import synth  # noqa: F401

__version__ = "0.1.0"
