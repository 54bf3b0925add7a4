/*
 * nbx_types.h - plain C types shared with the n_body_sim_pro engine ABI.
 *
 * Every struct and enum here is layout-compatible with the reference engine's
 * public headers, so a view, gravity record, error record or statistics record
 * built by the reference's C17 code can be handed to this library unchanged:
 *
 *   NBodySimProStatus / NBodySimProError ..... include/n_body_sim_pro/core/status.h:18-35
 *   NBodySimProParticleSystemView ............ include/n_body_sim_pro/core/particle_system.h:38-50
 *   NBodySimProGravity ....................... include/n_body_sim_pro/physics/gravity.h:25-29
 *   NBodySimProIntegratorType ................ include/n_body_sim_pro/physics/integrator.h:31-35
 *   NBodySimProForceFunction ................. include/n_body_sim_pro/physics/integrator.h:47-49
 *   NBodySimProBarnesHutStats ................ include/n_body_sim_pro/barnes_hut/barnes_hut.h:31-40
 *   NBodySimProDiagnosticsQuantities ......... include/n_body_sim_pro/diagnostics/numerics.h:22-34
 *
 * When the reference's own headers are included first, their include guards
 * are already defined and the definitions below are skipped, so both headers
 * can live in one translation unit (reference headers first).
 */
#ifndef NBX_TYPES_H
#define NBX_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef N_BODY_SIM_PRO_CORE_STATUS_H
typedef enum NBodySimProStatus {
    N_BODY_SIM_PRO_STATUS_OK = 0,
    N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
    N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY,
    N_BODY_SIM_PRO_STATUS_INVALID_STATE,
    N_BODY_SIM_PRO_STATUS_NOT_IMPLEMENTED,
    N_BODY_SIM_PRO_STATUS_UNSUPPORTED_PLATFORM,
    N_BODY_SIM_PRO_STATUS_OVERFLOW,
    N_BODY_SIM_PRO_STATUS_IO_ERROR,
    N_BODY_SIM_PRO_STATUS_INSUFFICIENT_MEMORY
} NBodySimProStatus;

typedef struct NBodySimProError {
    NBodySimProStatus status;
    const char* source_file;
    int source_line;
    char message[256];
} NBodySimProError;
#endif

#ifndef N_BODY_SIM_PRO_CORE_PARTICLE_SYSTEM_H
/* Non-owning SoA snapshot: N plus ten double arrays. */
typedef struct NBodySimProParticleSystemView {
    size_t particle_count;
    double* positions_x;
    double* positions_y;
    double* positions_z;
    double* velocities_x;
    double* velocities_y;
    double* velocities_z;
    double* accelerations_x;
    double* accelerations_y;
    double* accelerations_z;
    double* masses;
} NBodySimProParticleSystemView;
#endif

#ifndef N_BODY_SIM_PRO_PHYSICS_GRAVITY_H
typedef struct NBodySimProGravity {
    double gravitational_constant;
    double softening_length;
    double softening_squared;
} NBodySimProGravity;
#endif

#ifndef N_BODY_SIM_PRO_PHYSICS_INTEGRATOR_H
typedef enum NBodySimProIntegratorType {
    N_BODY_SIM_PRO_INTEGRATOR_EULER,
    N_BODY_SIM_PRO_INTEGRATOR_LEAPFROG,
    N_BODY_SIM_PRO_INTEGRATOR_VELOCITY_VERLET
} NBodySimProIntegratorType;

typedef NBodySimProStatus (*NBodySimProForceFunction)(const NBodySimProParticleSystemView* view,
                                                      const NBodySimProGravity* gravity,
                                                      void* context, NBodySimProError* error);
#endif

#ifndef N_BODY_SIM_PRO_BARNES_HUT_BARNES_HUT_H
typedef struct NBodySimProBarnesHutStats {
    size_t node_count;
    size_t leaf_count;
    size_t internal_node_count;
    size_t accepted_approximations;
    size_t exact_interactions;
    size_t maximum_depth;
    double build_time_seconds;
    double evaluation_time_seconds;
} NBodySimProBarnesHutStats;
#endif

#ifndef N_BODY_SIM_PRO_DIAGNOSTICS_NUMERICS_H
typedef struct NBodySimProDiagnosticsQuantities {
    double kinetic_energy;
    double total_mass;
    double total_momentum_x;
    double total_momentum_y;
    double total_momentum_z;
    double total_angular_momentum_x;
    double total_angular_momentum_y;
    double total_angular_momentum_z;
    double center_of_mass_x;
    double center_of_mass_y;
    double center_of_mass_z;
} NBodySimProDiagnosticsQuantities;
#endif

#ifdef __cplusplus
}
#endif

#endif /* NBX_TYPES_H */
