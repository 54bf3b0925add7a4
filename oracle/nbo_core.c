/*
 * nbo_core.c - oracle: all-pairs sum, leapfrog KDK, diagnostics.
 *
 * TEST INFRASTRUCTURE ONLY (see nbo.h). Built with -ffp-contract=off so that
 * every expression rounds exactly as the reference's ISO C17 build does.
 */
#include "nbo.h"

#include <math.h>

/*
 * Follows src/c/physics/gravity.c:37-67: per target i, ascending j with j == i
 * skipped; d = p_j - p_i; r2 = ((dx*dx + dy*dy) + dz*dz) + eps2;
 * inv = 1/sqrt(r2); inv3 = (inv*inv)*inv; s = (G*m_j)*inv3; a += s*d.
 */
static void direct_one_target(const NBodySimProParticleSystemView* view, double g, double eps2,
                              size_t i, double* ax_out, double* ay_out, double* az_out) {
    const size_t n = view->particle_count;
    const double* px = view->positions_x;
    const double* py = view->positions_y;
    const double* pz = view->positions_z;
    const double* m = view->masses;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (size_t j = 0; j < n; ++j) {
        if (j == i) {
            continue;
        }
        const double dx = px[j] - px[i];
        const double dy = py[j] - py[i];
        const double dz = pz[j] - pz[i];
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double inv = 1.0 / sqrt(r2);
        const double inv3 = inv * inv * inv;
        const double s = g * m[j] * inv3;
        ax += s * dx;
        ay += s * dy;
        az += s * dz;
    }
    *ax_out = ax;
    *ay_out = ay;
    *az_out = az;
}

int nbo_direct_fp64(const NBodySimProParticleSystemView* view, const NBodySimProGravity* gravity) {
    if (view == NULL || gravity == NULL) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    for (size_t i = 0; i < view->particle_count; ++i) {
        direct_one_target(view, gravity->gravitational_constant, gravity->softening_squared, i,
                          &view->accelerations_x[i], &view->accelerations_y[i],
                          &view->accelerations_z[i]);
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

/* Follows src/c/physics/gravity_openmp.c:52-83 (static schedule over i; same per-i order). */
int nbo_direct_openmp(const NBodySimProParticleSystemView* view, const NBodySimProGravity* gravity) {
    if (view == NULL || gravity == NULL) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    const long long n = (long long)view->particle_count;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        direct_one_target(view, gravity->gravitational_constant, gravity->softening_squared,
                          (size_t)i, &view->accelerations_x[i], &view->accelerations_y[i],
                          &view->accelerations_z[i]);
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

int nbo_direct_openmp_subset(const NBodySimProParticleSystemView* view,
                             const NBodySimProGravity* gravity, const int64_t* targets,
                             size_t target_count, double* out_xyz) {
    if (view == NULL || gravity == NULL || targets == NULL || out_xyz == NULL) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    const long long t = (long long)target_count;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long k = 0; k < t; ++k) {
        const int64_t i = targets[k];
        if (i < 0 || (size_t)i >= view->particle_count) {
            out_xyz[3 * k + 0] = out_xyz[3 * k + 1] = out_xyz[3 * k + 2] = NAN;
            continue;
        }
        direct_one_target(view, gravity->gravitational_constant, gravity->softening_squared,
                          (size_t)i, &out_xyz[3 * k + 0], &out_xyz[3 * k + 1],
                          &out_xyz[3 * k + 2]);
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

/*
 * Follows src/c/physics/integrator.c:47-85 (leapfrog KDK):
 * v += a*(dt/2); x += v*dt (one fused loop); a = F(x); v += a*(dt/2).
 */
int nbo_leapfrog_advance(NBodySimProParticleSystemView* view, const NBodySimProGravity* gravity,
                         double timestep, nbo_force_fn force) {
    if (view == NULL || gravity == NULL || force == NULL) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    if (timestep <= 0.0) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    const size_t n = view->particle_count;
    const double half = 0.5 * timestep;
    for (size_t i = 0; i < n; ++i) {
        view->velocities_x[i] += view->accelerations_x[i] * half;
        view->velocities_y[i] += view->accelerations_y[i] * half;
        view->velocities_z[i] += view->accelerations_z[i] * half;
        view->positions_x[i] += view->velocities_x[i] * timestep;
        view->positions_y[i] += view->velocities_y[i] * timestep;
        view->positions_z[i] += view->velocities_z[i] * timestep;
    }
    const int status = force(view, gravity);
    if (status != N_BODY_SIM_PRO_STATUS_OK) {
        return status;
    }
    for (size_t i = 0; i < n; ++i) {
        view->velocities_x[i] += view->accelerations_x[i] * half;
        view->velocities_y[i] += view->accelerations_y[i] * half;
        view->velocities_z[i] += view->accelerations_z[i] * half;
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

/* Follows src/c/diagnostics/numerics.c:5-81 (sequential sums in index order). */
int nbo_diagnostics_global(const NBodySimProParticleSystemView* view,
                           NBodySimProDiagnosticsQuantities* q) {
    if (view == NULL || q == NULL) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    double ke = 0.0, mt = 0.0, px = 0.0, py = 0.0, pz = 0.0, lx = 0.0, ly = 0.0, lz = 0.0;
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (size_t i = 0; i < view->particle_count; ++i) {
        const double m = view->masses[i];
        const double vx = view->velocities_x[i];
        const double vy = view->velocities_y[i];
        const double vz = view->velocities_z[i];
        ke += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        mt += m;
        px += m * vx;
        py += m * vy;
        pz += m * vz;
        const double x = view->positions_x[i];
        const double y = view->positions_y[i];
        const double z = view->positions_z[i];
        lx += m * (y * vz - z * vy);
        ly += m * (z * vx - x * vz);
        lz += m * (x * vy - y * vx);
        cx += m * x;
        cy += m * y;
        cz += m * z;
    }
    q->kinetic_energy = ke;
    q->total_mass = mt;
    q->total_momentum_x = px;
    q->total_momentum_y = py;
    q->total_momentum_z = pz;
    q->total_angular_momentum_x = lx;
    q->total_angular_momentum_y = ly;
    q->total_angular_momentum_z = lz;
    if (mt != 0.0) {
        q->center_of_mass_x = cx / mt;
        q->center_of_mass_y = cy / mt;
        q->center_of_mass_z = cz / mt;
    } else {
        q->center_of_mass_x = q->center_of_mass_y = q->center_of_mass_z = 0.0;
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

/* Follows src/c/diagnostics/numerics.c:83-116: U = -G * sum_{i<j} m_i m_j / sqrt(r2 + eps2). */
int nbo_potential_energy(const NBodySimProParticleSystemView* view,
                         const NBodySimProGravity* gravity, double* out) {
    if (view == NULL || gravity == NULL || out == NULL) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    const size_t n = view->particle_count;
    const double eps2 = gravity->softening_squared;
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = i + 1; j < n; ++j) {
            const double dx = view->positions_x[j] - view->positions_x[i];
            const double dy = view->positions_y[j] - view->positions_y[i];
            const double dz = view->positions_z[j] - view->positions_z[i];
            const double d = sqrt(dx * dx + dy * dy + dz * dz + eps2);
            sum += view->masses[i] * view->masses[j] / d;
        }
    }
    *out = -gravity->gravitational_constant * sum;
    return N_BODY_SIM_PRO_STATUS_OK;
}
