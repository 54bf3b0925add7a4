"""A translation unit of the reference that includes its own headers first and
this library's headers second compiles warning-free as C17 and C++20 (the
reference's languages): the shared types come from the reference's headers,
the drop-ins convert to its NBodySimProForceFunction without a cast, and the
integrator accepts them with an NBXEngine* as the force context. Needs the
reference's headers (skipped where they are absent); compiles only."""
import pathlib
import subprocess

import pytest

from conftest import ROOT

REF_INCLUDE = pathlib.Path("/root/reference/include")

TU = r"""
#include "n_body_sim_pro/core/status.h"
#include "n_body_sim_pro/core/particle_system.h"
#include "n_body_sim_pro/physics/gravity.h"
#include "n_body_sim_pro/physics/integrator.h"
#include "n_body_sim_pro/barnes_hut/barnes_hut.h"
#include "n_body_sim_pro/diagnostics/numerics.h"
#include "n_body_sim_pro/checkpoint/checkpoint.h"

#include "nbx_engine.h"
#include "nbx_learned.h"
#include "nbx_ensemble.h"
#include "nbx_presets.h"
#include "nbx_checkpoint.h"

static const NBodySimProForceFunction kForces[] = {
    n_body_sim_pro_gravity_compute_acceleration_reference,
    n_body_sim_pro_barnes_hut_compute_acceleration,
    nbx_hip_direct_fp64, nbx_hip_direct_fp32, nbx_hip_barnes_hut,
    nbx_hip_barnes_hut_fp64, nbx_hip_bh_learned,
};

NBodySimProStatus nbx_tu_step(NBodySimProParticleSystemView* view,
                              const NBodySimProGravity* gravity, NBXEngine* engine,
                              int which, NBodySimProError* error) {
    return n_body_sim_pro_integrator_advance(view, gravity, N_BODY_SIM_PRO_INTEGRATOR_LEAPFROG,
                                             1.0e-3, kForces[which], engine, error);
}

NBodySimProStatus nbx_tu_stats(NBXEngine* engine, NBodySimProBarnesHutStats* stats) {
    return nbx_engine_stats(engine, stats) ? N_BODY_SIM_PRO_STATUS_INVALID_STATE
                                           : N_BODY_SIM_PRO_STATUS_OK;
}
"""


@pytest.mark.skipif(not REF_INCLUDE.is_dir(), reason="the reference's headers are not here")
@pytest.mark.parametrize("lang", ["c17", "c++20"])
def test_reference_headers_first_then_nbx_headers_compile(tmp_path, lang):
    if lang == "c17":
        src, cmd = tmp_path / "tu.c", ["gcc", "-std=c17"]
    else:
        src, cmd = tmp_path / "tu.cpp", ["g++", "-std=c++20"]
    src.write_text(TU)
    r = subprocess.run([*cmd, "-Wall", "-Wextra", "-Werror", "-pedantic",
                        f"-I{REF_INCLUDE}", f"-I{ROOT / 'include'}", "-c", str(src),
                        "-o", str(tmp_path / "tu.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
