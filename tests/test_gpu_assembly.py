"""Device constraint set, barrier and friction assembly against the oracle
(through the C ABI: rigc_construct_constraint_set, rigc_compute_barrier_term,
rigc_compute_minimum_distance, rigc_update_friction_constraints,
rigc_compute_friction_term).

Bars: the constraint set (kinds, elements, multiplicities, bodies, eps_x),
the minimum distance and the friction frames are bit-identical (same
operation order on both sides); energies and gradients agree within 1e-10
relative; Hessian entries within 1e-9 relative to the largest entry of the
assembled matrix (each constraint's 12x12 block is projected to PSD by an
eigen-decomposition on both sides, in different rotation orders)."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import context, rigidipc_amd
from rigidipc_amd import scenes as S
from test_oracle_barrier import DHAT, contact_scene

pytestmark = pytest.mark.gpu

E_TOL = 1e-10
G_TOL = 1e-10
H_TOL = 1e-9


def perturbed(sc, seed, rot=0.02):
    rng = np.random.default_rng(seed + 100)
    poses = sc.poses.copy()
    dyn = ~sc.is_static.astype(bool)
    poses[dyn, 3:] += rng.normal(size=(int(dyn.sum()), 3)) * rot
    return poses


def config2_contact(seed=0):
    """The 1,000-body scene with its bottom layer lowered to 4 mm above the
    plane, so that vertices, edges and faces of ~100 bodies are inside the
    barrier's activation band."""
    sc = S.config2(seed=seed)
    poses = sc.poses.copy()
    O.set_scene(sc)
    W = O.world_vertices(poses, len(sc.vertices))
    vs = sc.body_vertex_start
    for b in range(len(sc.group_ids)):
        if sc.is_static[b]:
            continue
        miny = W[vs[b]:vs[b + 1], 1].min()
        if miny < 0.5:
            poses[b, 1] -= miny - 0.004
    return sc, poses


def key_of(rec):
    return [tuple(int(x) for x in r[:3]) for r in rec]


def match(gpu_rec, ora_rec):
    """Index of each oracle record in the device list (by type, a, b)."""
    pos = {k: i for i, k in enumerate(key_of(gpu_rec))}
    return np.array([pos[k] for k in key_of(ora_rec)])


def rel_err(a, b):
    scale = np.abs(b).max() if np.size(b) else 0.0
    return np.abs(a - b).max() / scale if scale > 0 else np.abs(a - b).max()


def oracle_blocks(rec, lh):
    blocks = {}
    for r, h in zip(rec, lh):
        h = h.reshape(12, 12)
        b = (int(r[4]), int(r[5]))
        for i in range(2):
            for j in range(2):
                if b[i] > b[j]:
                    continue
                blk = h[6 * i:6 * i + 6, 6 * j:6 * j + 6]
                key = (b[i], b[j])
                blocks[key] = blocks.get(key, 0) + blk
    return blocks


def gpu_blocks(hess):
    diag, pairs, off = hess
    blocks = {(b, b): diag[b] for b in range(len(diag)) if np.any(diag[b])}
    for (i, j), B in zip(pairs, off):
        blocks[(int(i), int(j))] = B
    return blocks


def compare_blocks(gb, ob):
    keys = set(ob) | set(gb)
    scale = max(np.abs(B).max() for B in ob.values())
    worst = 0.0
    for k in keys:
        a = gb.get(k, np.zeros((6, 6)))
        b = ob.get(k, np.zeros((6, 6)))
        worst = max(worst, np.abs(a - b).max() / scale)
    return worst


SCENES = {
    "contact": lambda: (contact_scene(0), None, DHAT),
    "contact_rot": lambda: (contact_scene(1), 1, DHAT),
    "config2": lambda: (*config2_contact(0), 0.05),
}


def load(name):
    if name == "config2":
        sc, poses, dhat = SCENES[name]()
    else:
        sc, seed, dhat = SCENES[name]()
        poses = perturbed(sc, seed or 0)
    return sc, poses, dhat


@pytest.fixture(scope="module")
def ctx():
    c = context()
    yield c
    c.close()


@pytest.mark.parametrize("name", list(SCENES))
def test_constraint_set_matches_oracle(ctx, name):
    sc, poses, dhat = load(name)
    print(f"constraint set: scene {name}, {len(sc.group_ids)} bodies, dhat {dhat}", flush=True)
    O.set_scene(sc)
    orec, oeps = O.constraint_set(poses, dhat)
    ctx.set_scene(sc)
    counts = ctx.construct_constraint_set(poses, dhat)
    grec, geps = ctx.constraint_set()
    assert len(orec) > 0
    assert counts.sum() == len(grec) == len(orec)
    assert np.array_equal(np.bincount(orec[:, 0], minlength=4), counts)
    idx = match(grec, orec)
    assert np.array_equal(grec[idx].astype(np.int64), orec)
    assert np.array_equal(geps[idx], oeps)
    if name != "config2":
        assert len(set(orec[:, 0].tolist())) >= 3


@pytest.mark.parametrize("name", list(SCENES))
def test_barrier_matches_oracle(ctx, name):
    sc, poses, dhat = load(name)
    nb = len(sc.group_ids)
    print(f"barrier: scene {name}, {nb} bodies", flush=True)
    O.set_scene(sc)
    orec, _ = O.constraint_set(poses, dhat)
    O.set_psd_projection(True)
    oe, og, olh = O.barrier(poses, dhat, want_hess=True, n_constraints=len(orec))
    ctx.set_scene(sc)
    ge, gg, gh = ctx.barrier_term(poses, dhat, grad=True, hess=True, rebuild=True)
    assert ge == pytest.approx(oe, rel=E_TOL, abs=0)
    assert rel_err(gg, og) <= G_TOL
    assert compare_blocks(gpu_blocks(gh), oracle_blocks(orec, olh)) <= H_TOL
    # energy-only and gradient-only paths give the same numbers
    e1, _, _ = ctx.barrier_term(poses, dhat)
    e2, g2, _ = ctx.barrier_term(poses, dhat, grad=True)
    assert e1 == ge and e2 == ge
    assert rel_err(g2, gg) <= 1e-13


def test_barrier_hessian_is_psd_and_symmetric(ctx):
    sc, poses, dhat = load("contact")
    nb = len(sc.group_ids)
    ctx.set_scene(sc)
    _, _, gh = ctx.barrier_term(poses, dhat, grad=True, hess=True, rebuild=True)
    H = rigidipc_amd.dense_hessian(gh, nb)
    assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
    lam = np.linalg.eigvalsh(0.5 * (H + H.T))
    assert lam.min() >= -1e-10 * lam.max()


@pytest.mark.parametrize("name", list(SCENES))
def test_minimum_distance_matches_oracle(ctx, name):
    sc, poses, dhat = load(name)
    O.set_scene(sc)
    O.constraint_set(poses, dhat)
    om = O.min_distance(poses)
    ctx.set_scene(sc)
    gm = ctx.minimum_distance(poses, dhat)
    assert gm == np.sqrt(om)


def test_minimum_distance_without_constraints_is_inf(ctx):
    sc = S.config1(seed=0)
    ctx.set_scene(sc)
    assert ctx.minimum_distance(sc.poses, 1e-6) == np.inf
    e, g, h = ctx.barrier_term(sc.poses, 1e-6, grad=True, hess=True, rebuild=True)
    assert e == 0 and not np.any(g) and not np.any(h[0]) and len(h[1]) == 0


@pytest.mark.parametrize("name", ["contact", "contact_rot", "config2"])
def test_friction_matches_oracle(ctx, name):
    sc, poses0, dhat = load(name)
    nb = len(sc.group_ids)
    print(f"friction: scene {name}, {nb} bodies", flush=True)
    kappa, mu, epsv_h = 3.7e3, 0.4, 1e-3 * 0.01
    rng = np.random.default_rng(9)
    poses = poses0.copy()
    dyn = ~sc.is_static.astype(bool)
    poses[dyn, :3] += rng.normal(size=(int(dyn.sum()), 3)) * 2e-4
    poses[dyn, 3:] += rng.normal(size=(int(dyn.sum()), 3)) * 2e-4
    O.set_scene(sc)
    orec, _ = O.constraint_set(poses0, dhat)
    ofr = O.friction_update(poses0, dhat, 0.0, kappa, mu)
    O.set_psd_projection(True)
    oe, og, olh = O.friction(poses, poses0, epsv_h, want_hess=True, n_constraints=len(orec))
    ctx.set_scene(sc)
    n = ctx.update_friction_constraints(poses0, dhat, 0.0, kappa, mu)
    assert n == len(orec)
    grec, _ = ctx.constraint_set()
    gfr = ctx.friction_constraints()
    idx = match(grec, orec)
    assert np.array_equal(gfr[idx][:, 1:], ofr[:, 1:])
    assert rel_err(gfr[idx][:, 0], ofr[:, 0]) <= 1e-14
    ge, gg, gh = ctx.friction_term(poses, poses0, epsv_h, grad=True, hess=True)
    assert ge == pytest.approx(oe, rel=E_TOL, abs=0)
    assert rel_err(gg, og) <= G_TOL
    assert compare_blocks(gpu_blocks(gh), oracle_blocks(orec, olh)) <= H_TOL
