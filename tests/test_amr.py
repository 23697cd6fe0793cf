"""Remeshing of the time loop (src/cipher.c:1607-1700) on the host forest:
adapt by label (DMAdaptLabel with the level limits of
src/cipher.c:1153-1154) and the state transfer (DMForestTransferVec).

PETSc/p4est are not in the image, so the adapt is pinned by its invariants:
2:1 balance holds after every adapt (balancing again changes nothing), a
leaf labelled for refinement below the maximum level is split, no leaf ends
more than one level coarser than before, a coarsened leaf is exactly one
family of unlabelled leaves above the minimum level, uniform forests refine
and coarsen to uniform forests and stop at the level limits; the transfer
copies records into refined leaves, averages them into coarsened ones, and
so conserves the volume integral of every field.
"""
import numpy as np
import pytest

from cipher_amd import mesh as M
from cipher_amd import timeloop as TL
from tests import helpers as H


def _balanced(f):
    g = M.balance(f)
    return g.ncells == f.ncells and np.array_equal(g.level, f.level) and np.array_equal(g.anchor, f.anchor)


def _covering(old, new):
    """For every new leaf, the old leaf covering its anchor voxel."""
    return old.id_map()[tuple(new.anchor.T)]


def _volume(f):
    return np.prod(f.size / f.res) * f.cell_extent().astype(np.float64) ** f.dim


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_forests_refine_and_coarsen_uniformly(dim):
    res = (16,) * dim
    f = M.uniform_forest(res, [1] * dim, [1.0] * dim, [True] * dim, level=2)
    g = M.adapt(f, np.ones(f.ncells, bool), 0, 4)
    assert g.ncells == f.ncells * 2 ** dim and np.all(g.level == 3)
    h = M.adapt(g, np.zeros(g.ncells, bool), 2, 4)
    assert h.ncells == f.ncells and np.all(h.level == 2)
    # limits: nothing refines at maxlevel, nothing coarsens at minlevel
    assert M.adapt(g, np.ones(g.ncells, bool), 0, 3).ncells == g.ncells
    assert M.adapt(g, np.zeros(g.ncells, bool), 3, 4).ncells == g.ncells


@pytest.mark.parametrize("name", ["gg2d_amr", "gg3d_amr"])
def test_adapt_follows_the_labels_and_stays_balanced(name):
    wl = H.gg3d_amr() if name == "gg3d_amr" else H.gg2d_amr()
    f = wl.forest
    cfg = wl.cfg
    rng = np.random.default_rng(5)
    for trial in range(3):
        flags = rng.random(f.ncells) < 0.15
        g = M.adapt(f, flags, cfg.minnrefine, cfg.maxnrefine)
        assert _balanced(g)
        cov = _covering(f, g)
        # every new leaf is at most one level coarser than what it covers
        assert np.all(g.level >= f.level[cov] - 1)
        # labelled leaves below the maximum are split (exactly once, or more by balance)
        old_of = f.level[cov]
        split = flags & (f.level < cfg.maxnrefine)
        for i in np.flatnonzero(split)[:50]:
            sel = cov == i
            assert sel.sum() >= 2 ** f.dim and np.all(g.level[sel] >= f.level[i] + 1)
        # a coarsened leaf covers a whole family, none of it labelled, above minlevel
        coarse = g.level < old_of
        if coarse.any():
            assert np.all(g.level[coarse] >= cfg.minnrefine)
            idm = f.id_map()
            ext = g.cell_extent()
            for j in np.flatnonzero(coarse)[:50]:
                s = int(ext[j])
                region = idm[tuple(slice(int(a), int(a) + s) for a in g.anchor[j])]
                ids = np.unique(region)
                assert ids.size == 2 ** f.dim and not flags[ids].any()
                assert np.all(f.level[ids] == g.level[j] + 1)
        f = g


def test_transfer_copies_averages_and_conserves():
    wl = H.gg2d_amr()
    f = wl.forest
    rng = np.random.default_rng(11)
    flags = rng.random(f.ncells) < 0.2
    g = M.adapt(f, flags, wl.cfg.minnrefine, wl.cfg.maxnrefine)
    X = rng.standard_normal((f.ncells, 5))
    Y = M.transfer(f, g, X)
    cov = _covering(f, g)
    same = f.level[cov] <= g.level
    assert np.array_equal(Y[same], X[cov[same]])
    vf, vg = _volume(f), _volume(g)
    np.testing.assert_allclose((vg[:, None] * Y).sum(0), (vf[:, None] * X).sum(0), rtol=1e-12, atol=1e-12)
    # a coarsened leaf holds the mean of the records it covers
    idm = f.id_map()
    for j in np.flatnonzero(~same)[:20]:
        s = int(g.cell_extent()[j])
        ids = np.unique(idm[tuple(slice(int(a), int(a) + s) for a in g.anchor[j])])
        np.testing.assert_allclose(Y[j], X[ids].mean(0), rtol=1e-14, atol=1e-15)
    with pytest.raises(ValueError):
        M.transfer(f, g, X[:-1])


def test_remesh_of_a_poststep_state_keeps_valid_active_sets():
    """After PostStep a single-phase cell's face neighbours hold only its
    phase, so a coarsened family is one phase and its averaged record is an
    exact copy: every transferred active set is a valid sorted list with
    fractions summing to 1."""
    import oracle
    wl = H.gg2d_amr()
    L = wl.layout
    X = H.evolve(wl, H.initial_state(wl), steps=3)
    X, rc = oracle.poststep(wl.cfg, wl.mesh, X)
    assert rc == 0
    f2, X2 = TL.adapt_state(wl.forest, X, L, wl.cfg.minnrefine, wl.cfg.maxnrefine)
    assert f2 is not wl.forest
    D = L["ndof"]
    V = X2.reshape(-1, D)
    n = V[:, L["as_offset"]]
    assert np.all(n == np.rint(n)) and np.all(n >= 1)
    for c in range(f2.ncells):
        k = int(n[c])
        ids = V[c, L["as_offset"] + 1:L["as_offset"] + 1 + k]
        assert np.all(ids == np.rint(ids)) and np.all(np.diff(ids) > 0)
        assert V[c, L["pf_offset"]:L["pf_offset"] + k].sum() == pytest.approx(1.0, abs=1e-12)
    # the labels: every multi-phase leaf below the maximum level was split
    flags = TL.refine_flags(X, L)
    assert np.array_equal(flags, np.rint(X.reshape(-1, D)[:, 0]) > 1)
    cov = _covering(wl.forest, f2)
    split = flags & (wl.forest.level < wl.cfg.maxnrefine)
    assert np.all(f2.level[np.isin(cov, np.flatnonzero(split))] > wl.forest.level[cov][np.isin(cov, np.flatnonzero(split))])
