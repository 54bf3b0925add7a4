import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnbx.so on cuda:0)")


@pytest.fixture(scope="session")
def oracle():
    from nbx_testlib import ORACLE_SO, load_oracle
    if not ORACLE_SO.exists():
        import subprocess
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    return load_oracle()


@pytest.fixture(scope="session")
def ref():
    from nbx_testlib import have_ref, load_ref
    if not have_ref():
        pytest.skip("oracle/_ref was not built (needs the reference sources)")
    return load_ref()


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    from nbx_testlib import GOLDEN
    with np.load(GOLDEN / "reference_vectors.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def nbx():
    """The product library. GPU tests fail (not skip) when it is missing."""
    import n_body_sim_pro_amd as nbx_mod
    nbx_mod.load()
    return nbx_mod
