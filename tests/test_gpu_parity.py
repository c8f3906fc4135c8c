"""HIP objective against the reference's numbers and the oracle (1e-10 relative,
FP64; summation order is the only permitted difference)."""
import numpy as np
import pytest

from conftest import golden_cases, load_case, rel_err

import pyoracle
from meamzilla_amd import gen
from meamzilla_amd import formats
from meamzilla_amd._lib import MeamzError, Objective

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("case", golden_cases())
def test_gpu_matches_reference_golden(case):
    _say("golden", case)
    pot, db, exp = load_case(case)
    obj = Objective(pot, db)
    r = obj.compute()
    assert abs(r["F"] - exp["F"]) <= TOL * abs(exp["F"])
    assert rel_err(r["errvec"], exp["errvec"]) < TOL
    assert rel_err(r["cell_energy"], exp["cell_energy"]) < TOL
    assert rel_err(r["cell_stress"], exp["cell_stress"]) < TOL
    assert rel_err(r["forces"], exp["forces"]) < TOL
    if formats.is_pair_style(pot.style):
        with pytest.raises(MeamzError):
            obj.densities()  # pair styles have no densities
    else:
        assert rel_err(obj.densities(), exp["densities"]) < TOL


def test_gpu_gmeam_cut_g_counts_match_oracle():
    """gmeam/spline with one g spline (type A, types A B) ending at cos = 0.5:
    triplets beyond it drop out, counted per triplet from the neighbours'
    types; counts and objective against the oracle."""
    _say("gmeam cut g")
    w = gen.workload(2, scale=16 / 512)
    rng = np.random.Generator(np.random.PCG64(23))
    nk = dict(phi=10, rho=10, U=10, f=10, g=10)
    w.pot = gen.make_potential(rng, "gmeam/spline", 2, 5.0, nk, (-1.0, 3.0))
    w.truth = gen.make_potential(rng, "gmeam/spline", 2, 5.0, nk, (-1.0, 3.0))
    g0 = sum(formats.function_counts("gmeam/spline", 2)[:4])
    w.pot.splines[g0 + 1].xn = 0.5
    gen.add_targets(w, pyoracle.evaluator)
    o = pyoracle.Oracle(w.pot, w.db)
    obj = Objective(w.pot, w.db)
    assert (obj.info.npairs, obj.info.ntriplets) == o.counts()
    e, r = o.eval(), obj.compute()
    assert abs(r["F"] - e["F"]) <= TOL * abs(e["F"])
    assert rel_err(r["errvec"], e["errvec"]) < TOL


@pytest.mark.parametrize("style", ["jmeam/spline", "jmeam", "jmeam2/spline", "jmeam2"])
def test_gpu_jmeam_styles_match_oracle(style):
    """jmeam styles on the binary workload slice (a few thousand atoms): the
    p_ij p_ik q terms ride the forward walk beside g; plain and batched calls
    against the oracle."""
    _say("jmeam style", style)
    w = gen.workload(2, scale=24 / 512)
    rng = np.random.Generator(np.random.PCG64(91))
    nk = dict(phi=12, rho=12, U=10, f=12, g=10, p=11, q=9)
    w.pot = gen.make_potential(rng, style, 2, 5.0, nk, (-1.0, 3.0))
    w.truth = gen.make_potential(rng, style, 2, 5.0, nk, (-1.0, 3.0))
    if not style.endswith("spline"):
        w.pot.splines[13].xn = 4.3  # p_AB ends early: triplets with the g term only
    gen.add_targets(w, pyoracle.evaluator)
    o = pyoracle.Oracle(w.pot, w.db)
    obj = Objective(w.pot, w.db)
    assert (obj.info.npairs, obj.info.ntriplets) == o.counts()
    e = o.eval()
    r = obj.compute()
    assert abs(r["F"] - e["F"]) <= TOL * abs(e["F"])
    for k in ("errvec", "cell_energy", "cell_stress", "forces"):
        assert rel_err(r[k], e[k]) < TOL, k
    assert rel_err(obj.densities(), o.densities()) < TOL
    kv = w.pot.knot_vector()
    ks = np.stack([kv, kv * 1.01])
    Fb, Eb = obj.residuals_batch(ks)
    for n in range(2):
        en = o.eval(ks[n])
        assert abs(Fb[n] - en["F"]) <= TOL * abs(en["F"])
        assert rel_err(Eb[n], en["errvec"]) < TOL


@pytest.mark.parametrize("style", ["pair/spline", "pair"])
def test_gpu_pair_styles_match_oracle(style):
    """Pair styles on the binary workload slice: the device walks the full pair
    list (each pair from both ends, half phi each), the oracle the reference's
    half list with self-image pairs at half weight."""
    _say("pair style", style)
    w = gen.workload(2, scale=40 / 512)
    rng = np.random.Generator(np.random.PCG64(77))
    w.pot = gen.make_potential(rng, style, 2, 5.0, dict(phi=12), (-1.0, 3.0))
    w.truth = gen.make_potential(rng, style, 2, 5.0, dict(phi=12), (-1.0, 3.0))
    gen.add_targets(w, pyoracle.evaluator)
    o = pyoracle.Oracle(w.pot, w.db)
    obj = Objective(w.pot, w.db)
    assert obj.info.ntriplets == 0 and obj.info.nerr == 7 * w.db.ncells + 3 * w.db.total_atoms
    e = o.eval()
    r = obj.compute()
    assert abs(r["F"] - e["F"]) <= TOL * abs(e["F"])
    for k in ("errvec", "cell_energy", "cell_stress", "forces"):
        assert rel_err(r[k], e[k]) < TOL, k
    # batched candidates (cand tile 4 path) against the plain call and the oracle
    kv = w.pot.knot_vector()
    ks = np.stack([kv, kv * 1.01, kv * 0.99])
    Fb, Eb = obj.residuals_batch(ks)
    for n in range(3):
        en = o.eval(ks[n])
        assert abs(Fb[n] - en["F"]) <= TOL * abs(en["F"])
        assert rel_err(Eb[n], en["errvec"]) < TOL


@pytest.mark.parametrize("cfg,scale", [(0, 1.0), (1, 24 / 256), (2, 40 / 512), (3, 30 / 1000), (4, 3 / 2000)])
def test_gpu_matches_oracle_workloads(cfg, scale):
    _say("workload", cfg, scale)
    w = gen.workload(cfg, scale=scale)
    gen.add_targets(w, pyoracle.evaluator)
    o = pyoracle.Oracle(w.pot, w.db)
    obj = Objective(w.pot, w.db)
    npairs, ntrip = o.counts()
    assert obj.info.npairs == npairs and obj.info.ntriplets == ntrip
    e = o.eval()
    r = obj.compute()
    assert abs(r["F"] - e["F"]) <= TOL * abs(e["F"])
    for k in ("errvec", "cell_energy", "cell_stress", "forces"):
        assert rel_err(r[k], e[k]) < TOL, k


def test_gpu_batch_matches_plain():
    _say("batch vs plain")
    w = gen.workload(2, scale=20 / 512)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    K = w.candidate_knots(5)
    Fb = obj.batch(K)
    Fp = np.array([obj.objective(k) for k in K])
    assert np.max(np.abs(Fb - Fp) / np.abs(Fp)) < 1e-12
    o = pyoracle.Oracle(w.pot, w.db)
    Fo = np.array([o.objective(k) for k in K])
    assert np.max(np.abs(Fb - Fo) / np.abs(Fo)) < TOL


@pytest.mark.parametrize("tile", [(16, 256, 256), (64, 512, 256), (32, 256, 128), (8, 512, 512), (64, 512, 128)])
def test_gpu_tiles_agree(tile):
    _say("tile", tile)
    w = gen.workload(1, scale=6 / 256)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    obj.set_tile(*tile)
    r = obj.compute()
    e = pyoracle.Oracle(w.pot, w.db).eval()
    assert abs(r["F"] - e["F"]) <= TOL * abs(e["F"])
    assert rel_err(r["forces"], e["forces"]) < TOL


def test_gpu_graph_replays_match_eager():
    """The captured forward graph, replayed with eager evaluations and host
    synchronisation in between, agrees with the eager launch sequence."""
    _say("graph replay")
    w = gen.workload(1, scale=8 / 256)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    K = w.candidate_knots(3)
    for rep in range(4):
        for k in K:
            eager = obj.compute(k)["F"]          # outputs requested: eager path
            graph = obj.objective(k)             # no outputs: graph replay
            assert abs(graph - eager) <= 1e-12 * abs(eager), (rep, graph, eager)
        Fb = obj.batch(K)                        # batched graph of 3 vectors
        Fe = np.array([obj.compute(k)["F"] for k in K])
        assert np.max(np.abs(Fb - Fe) / np.abs(Fe)) < 1e-12


@pytest.mark.parametrize("cand_tile,ncand", [(2, 5), (4, 7), (4, 4), (2, 1)])
def test_gpu_candidate_tiles_agree(cand_tile, ncand):
    """Workgroups scoring several candidates of a batch (tails included) agree
    with plain evaluations and with the oracle; so do their gradients."""
    _say("candidate tile", cand_tile, ncand)
    w = gen.workload(3, scale=16 / 1000)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    K = w.candidate_knots(ncand)
    Fp = np.array([obj.objective(k) for k in K])
    Gp = np.array([obj.objective_grad(k)[1] for k in K])
    t = obj.get_tile()
    obj.set_tile(t.max_atoms, t.max_pairs, t.block, cand_tile)
    Fb, Gb = obj.batch(K, grad=True)
    assert np.max(np.abs(Fb - Fp) / np.abs(Fp)) < 1e-12
    assert np.max(np.abs(Fb - obj.batch(K)) / np.abs(Fp)) < 1e-12
    for q in range(ncand):
        assert rel_err(Gb[q], Gp[q]) < 1e-10, q
    o = pyoracle.Oracle(w.pot, w.db)
    Fo = np.array([o.objective(k) for k in K])
    assert np.max(np.abs(Fb - Fo) / np.abs(Fo)) < TOL


def test_gpu_autotune_keeps_results():
    """The autotuner picks a supported tile and the objective of every
    candidate is unchanged under it."""
    _say("autotune")
    w = gen.workload(1, scale=8 / 256)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    K = w.candidate_knots(4)
    F0 = obj.batch(K)
    t = obj.autotune(ncand=4, reps=1)
    assert t.cand_tile in (1, 2, 4) and t.block in (64, 128, 256, 512)
    g = obj.get_tile()
    assert (g.max_atoms, g.max_pairs, g.block, g.cand_tile) == (t.max_atoms, t.max_pairs, t.block, t.cand_tile)
    F1 = obj.batch(K)
    assert np.max(np.abs(F1 - F0) / np.abs(F0)) < 1e-12


def test_gpu_autotune_grad_keeps_results():
    """The autotuner timing the caller's vectors in gradient calls picks a
    supported tile and the batched
    objective and gradient of every candidate are unchanged under it (to the
    round-off of the adjoint's LDS accumulation order)."""
    _say("autotune grad")
    w = gen.workload(1, scale=8 / 256)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    K = w.candidate_knots(4)
    F0, G0 = obj.batch(K, grad=True)
    t = obj.autotune(ncand=4, reps=1, grad=True, knots=K)
    assert t.cand_tile in (1, 2, 4) and t.block in (64, 128, 256, 512)
    g = obj.get_tile()
    assert (g.max_atoms, g.max_pairs, g.block, g.cand_tile) == (t.max_atoms, t.max_pairs, t.block, t.cand_tile)
    F1, G1 = obj.batch(K, grad=True)
    assert np.max(np.abs(F1 - F0) / np.abs(F0)) < 1e-12
    for q in range(4):
        assert rel_err(G1[q], G0[q]) < 1e-12, q


def test_gpu_residuals_batch_matches_plain():
    """Batched residual vectors (Powell's Jacobian columns) equal the plain
    evaluation's residual vector for every candidate."""
    _say("residuals batch")
    w = gen.workload(3, scale=12 / 1000)
    gen.add_targets(w, pyoracle.evaluator)
    obj = Objective(w.pot, w.db)
    K = w.candidate_knots(5)
    F, E = obj.residuals_batch(K)
    for q in range(len(K)):
        r = obj.compute(K[q])
        assert abs(F[q] - r["F"]) <= 1e-12 * abs(r["F"])
        assert rel_err(E[q], r["errvec"]) < 1e-12, q
        assert abs(np.dot(E[q], E[q]) - F[q]) <= 1e-10 * F[q]


def _cells_db(cells):
    """Database from (positions (n,3) cartesian, box (3,3), types) triples."""
    from meamzilla_amd.formats import Database
    natoms = np.array([len(c[0]) for c in cells], np.int32)
    box = np.array([c[1] for c in cells], float)
    pos = np.concatenate([np.asarray(c[0], float) for c in cells])
    types = np.concatenate([np.asarray(c[2], np.int32) for c in cells])
    nc, na = len(cells), len(pos)
    return Database(natoms, box, types, pos, np.zeros((na, 3)), np.zeros(nc), np.zeros((nc, 6)),
                    np.ones(nc), np.ones(nc), np.ones(nc), 1.0, 1.0)


@pytest.mark.parametrize("cfg", [0, 1, 2])
def test_gpu_edge_geometries(cfg):
    """Isolated atom (no neighbours), dimer, trimer, and one atom whose only
    neighbours are its own periodic images, in one database; objective,
    residuals and forces against the oracle, and the adjoint against central
    differences."""
    _say("edge geometries", cfg)
    w = gen.workload(cfg, scale=0.01)
    nt = w.pot.ntypes
    L = 20.0
    big = np.diag([L, L, L])
    cells = [
        (np.array([[5.0, 5.0, 5.0]]), big, [0]),
        (np.array([[5.0, 5.0, 5.0], [7.9, 5.1, 5.0]]), big, [0, nt - 1]),
        (np.array([[5.0, 5.0, 5.0], [7.7, 5.0, 5.0], [6.3, 7.3, 5.2]]), big, [0, nt - 1, 0]),
        (np.array([[0.3, 0.2, 0.1]]), np.array([[2.9, 0.0, 0.0], [0.2, 3.0, 0.0], [0.1, -0.3, 3.1]]), [nt - 1]),
    ]
    db = _cells_db(cells)
    gen.add_targets(gen.Workload("edge", w.pot, w.truth, db, w.cutoff, 1, 5), pyoracle.evaluator)
    o = pyoracle.Oracle(w.pot, db)
    obj = Objective(w.pot, db)
    assert (obj.info.npairs, obj.info.ntriplets) == o.counts()
    e = o.eval()
    r = obj.compute()
    assert abs(r["F"] - e["F"]) <= TOL * abs(e["F"])
    for k in ("errvec", "cell_energy", "cell_stress", "forces"):
        assert rel_err(r[k], e[k]) < TOL, k
    kv = w.pot.knot_vector()
    F, g = obj.objective_grad(kv)
    h = 1e-6 * np.maximum(np.abs(kv), 1.0)
    idx = np.arange(len(kv))
    gcd = np.array([(obj.objective(kv + h[i] * (idx == i)) - obj.objective(kv - h[i] * (idx == i))) / (2 * h[i])
                    for i in idx])
    assert rel_err(g, gcd) < 1e-6
    # component-wise: own size plus the rounding floor of each central difference
    floor = 1e3 * np.finfo(np.float64).eps * abs(F) / h
    worst = float(np.max(np.abs(g - gcd) / (1e-6 * np.abs(gcd) + floor)))
    _say("edge geometries worst component error / bound", worst)
    assert worst <= 1.0
