"""Colony-level statistics with the learned JKR radius against the exact
FP64 radius: 16 colonies grown from single founder cells (seeds 0-15,
Bogacki-Shampine at tolerance 1e-8, gamma = 100 Pa um, the growth-run
parameters of the 3-D driver) to 3 simulated hours (about 28 cells each),
batched (bfm_ensemble_simulate), once with each radius.

Individual colonies diverge between the two runs (a 1e-6 change of a force
moves a division by a step and the colonies then differ), so the comparison
is between ensembles: the mean cell count, the distribution of contact
overlaps (quartiles, pooled over colonies; the JKR force sets the
equilibrium overlap of adhering cells) and the in-plane orientation order.
A third run with gamma raised by half shows that the overlap quartiles do
move by far more than the bound when the adhesion force is wrong."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_COL = 16
T_END = 3.0          # h
R = 0.8


def _grow(jkr, gamma=100.0):
    import bfm
    from bfm import sim
    from bfm import surrogate as sur
    from bfm.core import make_params, JKR_SURROGATE, JKR_EXACT
    reps = [sim.founder_cell(gamma=gamma) for _ in range(N_COL)]
    cfg = sim.sim_config(max_time=T_END, tableau="bogacki_shampine", seed=0)
    cap = 8192
    p = make_params(jkr_mode=JKR_SURROGATE if jkr == "surrogate" else JKR_EXACT)
    c = bfm.Context(cap, max_pairs=32 * cap, params=p)
    if jkr == "surrogate":
        sur.install(c, sur.load())
    print(f"growing {N_COL} colonies, jkr={jkr}, gamma={gamma}", flush=True)
    out, t_out, it_out, st = sim.ensemble_simulate(c, cfg, reps, list(range(N_COL)), autotune=False, cap=cap)
    assert st["failed_replicates"] == [] and np.all(t_out >= T_END)
    return out


def _stats(orc, colonies):
    from gpu_common import LDIV, THR
    overlaps, order = [], []
    for cells in colonies:
        nb = orc.get_cell_neighbors(cells, THR, R, LDIV)
        d = np.linalg.norm(nb[:, 2:5], axis=1)
        overlaps.append((2 * R - d)[d < 2 * R])
        th = np.arctan2(cells[:, 5], cells[:, 4])
        order.append(np.abs(np.mean(np.exp(2j * th))))
    ov = np.concatenate(overlaps)
    return dict(cells=float(np.mean([c.shape[0] for c in colonies])),
                q=np.percentile(ov, [25, 50, 75]), order=float(np.mean(order)), contacts=int(ov.size))


@pytest.fixture(scope="module")
def grown(orc):
    return {k: _stats(orc, _grow(*k)) for k in (("exact", 100.0), ("surrogate", 100.0), ("exact", 150.0))}


def test_learned_radius_keeps_colony_statistics(grown):
    ex, su, wrong = grown[("exact", 100.0)], grown[("surrogate", 100.0)], grown[("exact", 150.0)]
    for name, s in (("exact", ex), ("learned", su), ("gamma x1.5", wrong)):
        print(f"{name}: mean cells {s['cells']:.1f}, contacts {s['contacts']}, overlap quartiles "
              f"{np.round(s['q'], 5).tolist()}, in-plane order {s['order']:.3f}", flush=True)
    assert abs(su["cells"] - ex["cells"]) <= 0.10 * ex["cells"]
    assert np.all(np.abs(su["q"] - ex["q"]) <= 0.05 * ex["q"])
    assert abs(su["order"] - ex["order"]) <= 0.15
    # the statistic is sensitive: a 50 % error of the adhesion moves the overlaps well past the bound
    assert np.any(np.abs(wrong["q"] - ex["q"]) > 0.15 * ex["q"])
