"""What the trained cost model gains, on phases of graphs it never saw.

picker/heldout_phases.npz holds the SWEEP timings (every candidate's count
time, per phase) of whole graphs held out of training (every 7th graph of the
seeded R-MAT / Chung-Lu list), with the phase features the engine computed.
The committed model runs on the GPU (kt_mlp_forward, the engine's own
kernels) and picks one candidate per phase; its summed cost is compared with
the per-phase best (the exhaustive sweep, the judge of its regret), with the
fixed default variant's count (candidate 0, the reference's method) and with
every candidate used alone for all phases."""
import numpy as np
import pytest

import ktruss_amd as K
import learn

pytestmark = pytest.mark.gpu


def test_model_gain_on_heldout_graphs():
    K.init(0)
    model, configs = learn.load_model()
    with np.load(learn.HELDOUT_PATH, allow_pickle=False) as z:
        feats, costs, meta = z["feats"], z["costs"], z["meta"]
    assert costs.shape[1] == len(configs) and len(costs) > 500
    assert np.all(meta[:, 0].astype(np.int64) % learn.HOLDOUT_MOD == learn.HOLDOUT_REM)
    names = [c.describe() for c in configs]
    delta = names.index(next(n for n in names if n.startswith("delta")))
    choice = model.forward(feats).argmin(1)
    picked = float(costs[np.arange(len(costs)), choice].sum())
    best = float(costs.min(1).sum())
    default = float(costs[:, 0].sum())
    singles = costs.sum(0)
    print(f"held-out phases {len(costs)} of {len(np.unique(meta[:, 0]))} graphs: model {picked:.1f} ms, "
          f"per-phase best {best:.1f} ms, fixed DELTA {singles[delta]:.1f} ms, default {default:.1f} ms, "
          f"best single {singles.min():.1f} ms ({names[int(singles.argmin())]})")
    assert picked <= 1.10 * best            # regret against the exhaustive sweep
    assert picked < singles.min()           # below every candidate used alone (fixed DELTA included)
    assert picked < default / 5             # far below the default variant's recounts
