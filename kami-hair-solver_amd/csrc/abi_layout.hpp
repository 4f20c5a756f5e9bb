// Compile-time checks that every ABI struct keeps the x86-64 size and field
// offsets of the reference ABI v5 (include/kami_hair_solver.h:25-204).
#pragma once

#include "kami_hair_solver.h"
#include "kami_hair_solver_ext.h"

#include <cstddef>

static_assert(sizeof(KhsVec3) == 24, "KhsVec3 size");
static_assert(sizeof(KhsTriangle) == 12, "KhsTriangle size");
static_assert(sizeof(KhsSolverDesc) == 112, "KhsSolverDesc size");
static_assert(offsetof(KhsSolverDesc, gravity) == 8, "KhsSolverDesc.gravity");
static_assert(offsetof(KhsSolverDesc, substeps) == 32, "KhsSolverDesc.substeps");
static_assert(offsetof(KhsSolverDesc, absolute_tolerance) == 48, "KhsSolverDesc.absolute_tolerance");
static_assert(offsetof(KhsSolverDesc, fixed_root_nodes) == 104, "KhsSolverDesc.fixed_root_nodes");
static_assert(offsetof(KhsSolverDesc, thread_count) == 108, "KhsSolverDesc.thread_count");
static_assert(sizeof(KhsHairMaterial) == 96, "KhsHairMaterial size");
static_assert(offsetof(KhsHairMaterial, density) == 8, "KhsHairMaterial.density");
static_assert(offsetof(KhsHairMaterial, collider_offset) == 88, "KhsHairMaterial.collider_offset");
static_assert(sizeof(KhsBuildStats) == 96, "KhsBuildStats size");
static_assert(offsetof(KhsBuildStats, virtual_extension_rest_length) == 64, "KhsBuildStats.extension length");
static_assert(offsetof(KhsBuildStats, degree_of_freedom_count) == 72, "KhsBuildStats.dof");
static_assert(offsetof(KhsBuildStats, initial_minimum_gap) == 88, "KhsBuildStats.initial_minimum_gap");
static_assert(sizeof(KhsStepStats) == 144, "KhsStepStats size");
static_assert(offsetof(KhsStepStats, contact_candidate_count) == 24, "KhsStepStats.candidates");
static_assert(offsetof(KhsStepStats, phase) == 136, "KhsStepStats.phase");
static_assert(sizeof(KhsGpuInfo) == 160, "KhsGpuInfo size");
static_assert(offsetof(KhsGpuInfo, total_vram_bytes) == 24, "KhsGpuInfo.total_vram_bytes");
static_assert(offsetof(KhsGpuInfo, device_name) == 32, "KhsGpuInfo.device_name");
static_assert(sizeof(KhsGpuStats) == 56, "KhsGpuStats size");
static_assert(sizeof(KhsProgress) == 48, "KhsProgress size");
static_assert(offsetof(KhsProgress, frame_elapsed_seconds) == 40, "KhsProgress.elapsed");
static_assert(sizeof(KhsFailureDiagnostics) == 160, "KhsFailureDiagnostics size");
static_assert(offsetof(KhsFailureDiagnostics, distance) == 48, "KhsFailureDiagnostics.distance");
static_assert(offsetof(KhsFailureDiagnostics, collider_point) == 136, "KhsFailureDiagnostics.collider_point");
