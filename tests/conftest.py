import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    # build what is missing (never rebuild what exists: the GPU box runs the
    # libraries built in this tree)
    lib = os.path.join(ROOT, "simulador_quantico_amd", "libhgm.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "simulador_quantico_amd", "csrc"), "-j8", "-s"], check=True)
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "oracle"], check=True)


@pytest.fixture(scope="session")
def golden():
    import json

    import numpy as np

    def load(name):
        with open(os.path.join(GOLDEN, name)) as f:
            return json.load(f)

    return {
        "strings": load("strings.json"),
        "parse": load("parse.json"),
        "circuits": load("circuits.json"),
        "measure": load("measure.json"),
        "measure_edges": load("measure_edges.json"),
        "states": dict(np.load(os.path.join(GOLDEN, "states.npz"))),
    }


@pytest.fixture(scope="session")
def golden_grover23():
    import json
    with open(os.path.join(GOLDEN, "grover23.json")) as f:
        return json.load(f)
