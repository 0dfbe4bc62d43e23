import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    """Build the native libraries only if they are absent (the GPU box runs
    the prebuilt files that travel with the tree)."""
    from mdet_amd import LIB_PATH
    if not os.path.exists(LIB_PATH):
        from mdet_amd import build
        build.build()
    from oracle import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def ctx():
    import mdet_amd as M
    c = M.Ctx(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def golden():
    import json
    import numpy as np
    import mdet_amd as M
    here = os.path.join(ROOT, "tests", "golden")
    fa = open(os.path.join(here, "tnf_fixture.fa"), "rb").read()
    meta = json.load(open(os.path.join(here, "tnf_fixture.json")))
    tnf = np.fromfile(os.path.join(here, "tnf_fixture.f64"), np.float64).reshape(-1, 136)
    recs = M.fasta_read(fa)
    return {"fasta": fa, "meta": meta, "tnf": tnf, "records": recs}
