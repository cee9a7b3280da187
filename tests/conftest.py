"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs
on the CPU (oracle against golden vectors, host logic, ABI surface)."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    return lambda name: np.load(os.path.join(GOLDEN, name))


@pytest.fixture(scope="session")
def small_state():
    from atlas12_amd.state import SyntheState
    return SyntheState.load(os.path.join(GOLDEN, "synthe_small_state.npz"))


@pytest.fixture(scope="session")
def small_ref():
    from atlas12_amd.workloads import LTE_DTYPE
    z = np.load(os.path.join(GOLDEN, "synthe_small_ref.npz"))
    return {"lines": z["lines"].view(LTE_DTYPE), "spec": z["spec"], "gfall": z["gfall"],
            "taunu": z["linform_taunu"]}


@pytest.fixture(scope="session")
def coef():
    from atlas12_amd.state import load_coef
    return load_coef(os.path.join(GOLDEN, "josh_coef.npz"))


@pytest.fixture(scope="session")
def atl_small():
    """tests/golden/atlas_small.npz and its C-ABI state."""
    from atlas12_amd import native
    z = np.load(os.path.join(GOLDEN, "atlas_small.npz"))
    return z, native.atlas_state_from(z)


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def select_case():
    """tests/golden/atlas_select.npz with its raw line lists regenerated
    from the recorded seed: (state, predicted records, MOL2 records)."""
    import hashlib

    from atlas12_amd import workloads as wl
    z = np.load(os.path.join(GOLDEN, "atlas_select.npz"))
    lo, hi = (float(x) for x in z["wl_range"])
    pred, tio = wl.atlas_iteration_lists(int(z["n_raw"]), lo, hi, seed=int(z["seed"]),
                                         molecules=tuple(int(m) for m in z["mol_nelion"]),
                                         tio_loggf=tuple(float(x) for x in z["tio_loggf"]))
    digest = hashlib.sha256(pred.tobytes() + tio.tobytes()).hexdigest()
    assert digest == str(z["raw_sha256"]), "line-list generator drifted from the fixture"
    return z, pred, tio


@pytest.fixture(scope="session")
def select_all_case():
    """tests/golden/atlas_select_all.npz with its seven raw lists
    regenerated from the recorded configuration: (state, lists for
    atlas_select_all with the diatomics read by the ASCII reader, the
    diatomic list texts)."""
    import hashlib
    import json

    from atlas12_amd import records
    from atlas12_amd import workloads as wl
    z = dict(np.load(os.path.join(GOLDEN, "atlas_select_all.npz")))
    cfg = json.loads(str(z["config"]))
    mols = tuple(int(m) for m in z["mol_nelion"])
    pred, tio, lowobs, hilines, h2o, h3p, dia = wl.atlas_all_lists(mols, **cfg)
    files = {"lowobsat12.bin": lowobs.tobytes(), "hilines.bin": hilines.tobytes(),
             "h2opokazatel.bin": h2o.tobytes(), "h3plus.dat": h3p.tobytes()}
    files.update({k: v.encode() for k, v in dia.items()})
    digest = hashlib.sha256(b"".join([pred.tobytes(), tio.tobytes()] +
                                     [files[k] for k in sorted(files)])).hexdigest()
    assert digest == str(z["raw_sha256"]), "line-list generator drifted from the fixture"
    wt = z["wavetab"]
    drec = np.concatenate([records.read_diatomic_ascii(t.encode().split(b"\n"), wt[0], wt[-1])
                           for t in dia.values()])
    lists = dict(pred=pred, lowobs=lowobs, hilines=hilines, dia=drec, mol2=tio, h2o=h2o, h3p=h3p)
    return z, lists, dia


@pytest.fixture(scope="session")
def gpu_ctx():
    """A device context on cuda:0 -- fails (does not skip) when the HIP
    library is missing, so a GPU run can never pass on a fallback."""
    from atlas12_amd import native
    ctx = native.Context(0)
    yield ctx
    ctx.close()
