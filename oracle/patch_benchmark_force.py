"""Writes the reference's force benchmark (benchmarks/benchmark_force.c) with
the libnbx.so dispatch that INTEGRATION.md section 2 describes added to its
measure_force_evaluation (benchmark_force.c:84-147), so that the reference's
own harness accepts --algorithm hip_direct_fp64 / hip_direct_fp32 /
hip_barnes_hut / hip_barnes_hut_fp64 / hip_bh_learned next to its CPU names.

TEST INFRASTRUCTURE: the output goes to oracle/_ref/ only (git-ignored); the
reference file is read where it lies and never copied into the repository.
Each insertion is anchored on a line of the reference that must occur exactly
once, so a changed reference fails loudly instead of being patched wrongly.

    python3 oracle/patch_benchmark_force.py SRC OUT
"""
import sys

INCLUDE_ANCHOR = '#include "n_body_sim_pro/threading/threading.h"\n'
INCLUDE_ADD = '#include "nbx_engine.h"\n'

TREE_ANCHOR = "    NBodySimProBarnesHutTree* tree = NULL;\n"
ENGINE_ADD = """    NBXEngine* nbx = NULL;
    const int nbx_algo = nbx_algorithm_from_name(algorithm);
    if (nbx_algo >= 0) {
        NBXConfig nbx_config = {view->particle_count, 0, theta, 0u};
        nbx = nbx_engine_create(&nbx_config, &error);
        if (nbx == NULL) {
            fprintf(stderr, "failed to create HIP engine: %s\\n", error.message);
            return -1.0;
        }
    }
"""

LOOP_ANCHOR = '        if (strcmp(algorithm, "reference") == 0) {\n'
LOOP_REPLACE = """        if (nbx_algo == NBX_ALGORITHM_DIRECT_FP64) {
            status = nbx_hip_direct_fp64(view, gravity, nbx, &error);
        } else if (nbx_algo == NBX_ALGORITHM_DIRECT_FP32) {
            status = nbx_hip_direct_fp32(view, gravity, nbx, &error);
        } else if (nbx_algo == NBX_ALGORITHM_BARNES_HUT) {
            status = nbx_hip_barnes_hut(view, gravity, nbx, &error);
        } else if (nbx_algo == NBX_ALGORITHM_BARNES_HUT_FP64) {
            status = nbx_hip_barnes_hut_fp64(view, gravity, nbx, &error);
        } else if (nbx_algo == NBX_ALGORITHM_BH_LEARNED) {
            status = nbx_hip_bh_learned(view, gravity, nbx, &error);
        } else if (strcmp(algorithm, "reference") == 0) {
"""

FAIL_ANCHOR = ('            fprintf(stderr, "force evaluation failed: %s\\n", '
               'n_body_sim_pro_status_string(status));\n')
FAIL_ADD = "            nbx_engine_destroy(nbx);\n"

END_ANCHOR = "    const double elapsed = wall_time_seconds() - start;\n"
END_ADD = "    nbx_engine_destroy(nbx);\n"


def insert_after(text, anchor, addition):
    if text.count(anchor) != 1:
        raise SystemExit(f"anchor not found exactly once: {anchor.strip()!r}")
    return text.replace(anchor, anchor + addition)


def replace_once(text, anchor, replacement):
    if text.count(anchor) != 1:
        raise SystemExit(f"anchor not found exactly once: {anchor.strip()!r}")
    return text.replace(anchor, replacement)


def main(src, out):
    with open(src, encoding="utf-8") as f:
        text = f.read()
    text = insert_after(text, INCLUDE_ANCHOR, INCLUDE_ADD)
    text = insert_after(text, TREE_ANCHOR, ENGINE_ADD)
    text = replace_once(text, LOOP_ANCHOR, LOOP_REPLACE)
    text = insert_after(text, FAIL_ANCHOR, FAIL_ADD)
    text = insert_after(text, END_ANCHOR, END_ADD)
    with open(out, "w", encoding="utf-8") as f:
        f.write(text)


if __name__ == "__main__":
    if len(sys.argv) != 3:
        raise SystemExit(__doc__)
    main(sys.argv[1], sys.argv[2])
