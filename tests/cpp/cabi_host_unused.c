/* The C ABI entry points the C++ layer references but the CPU stand-in
 * (cabi_host_standin.cpp) does not provide: the FD, RTM, health and terrain
 * families.  Each reports STP_ERR_STATE (-4), so a driver command that
 * reaches one fails with a status instead of running.  Test infrastructure
 * only; declared without prototypes on purpose (C linkage, return value only). */
#define UNUSED(name) int name() { return -4; }
UNUSED(stp_accumulate_f32_f64)
UNUSED(stp_clamp_field_f64)
UNUSED(stp_count_nonfinite_f64)
UNUSED(stp_fd_adjoint)
UNUSED(stp_fd_coefficients)
UNUSED(stp_fd_plan_create)
UNUSED(stp_fd_plan_destroy)
UNUSED(stp_fd_propagate)
UNUSED(stp_fd_set_receivers)
UNUSED(stp_fd_set_sources)
UNUSED(stp_fd_sponge_factors)
UNUSED(stp_fd_verify_cfl)
UNUSED(stp_field_energy_f32)
UNUSED(stp_field_interior_f64)
UNUSED(stp_gaussian_filter_3d_f64)
UNUSED(stp_repair_field_nan_f64)
UNUSED(stp_rtm_finalize_f64)
UNUSED(stp_rtm_image_condition_f64)
UNUSED(stp_terrain_absorbing_f64)
UNUSED(stp_terrain_bathymetry_synthetic_f64)
UNUSED(stp_terrain_boundary_mask)
UNUSED(stp_terrain_free_surface_f64)
UNUSED(stp_terrain_seafloor_f64)
UNUSED(stp_terrain_velocity_mask_f64)
