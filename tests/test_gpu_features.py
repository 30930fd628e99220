"""GPU checks of the rows around the core path: tail correction, batched
replicas, position refresh and skin test, hipGraph replay, adjoint, autotuner
and the reference-style float accumulation output.  Everything goes through
the C ABI; the FP64 oracle is the checker."""
import numpy as np
import pytest

from oracle import oracle as O
from sre import Engine, SkinExceeded, configs

pytestmark = pytest.mark.gpu

TOL_F_MIXED = 1e-5


def rel_force_err(F, ref):
    rms = np.sqrt((ref ** 2).sum(1).mean())
    return np.sqrt(((F - ref) ** 2).sum(1)).max() / rms


def oracle_system(w, r=0, xyz=None, tail=False, exact=False, rc_xyz=None):
    """rc_xyz: positions at which rc was assigned (the engine assigns rc at
    rebuild time, like AssignRCut::assign); defaults to the evaluated ones."""
    sl = w.replica_slice(r)
    x = w.xyz[sl] if xyz is None else xyz
    box, types = w.boxes[r], w.types[sl]
    rc, _ = O.assign_rc(x if rc_xyz is None else rc_xyz, box, w.rc_tables[r])
    off, j = O.build_pairs(x, box, rc, w.skin)
    tl = (O.tail_atoms(x, box, types, w.eps, w.sigma, w.rc_tables[r].ravel(), w.rho[r],
                       exact=exact) if tail else None)
    F, E, W, nin = O.forces(x, box, types, w.eps, w.sigma, rc, off, j, tail=tl)
    return dict(rc=rc, off=off, j=j, F=F, E=E, W=W, tail=tl)


@pytest.fixture(scope="module")
def slab():
    return configs.slab(seed=31, n=12000, box=(18.0, 18.0, 64.0), nz=96, name="slab_12k")


@pytest.fixture(scope="module")
def binary():
    return configs.slab(seed=32, n=12000, box=(18.0, 18.0, 64.0), nz=96, binary=True,
                        name="binary_12k")


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("exact", [False, True])
def test_tail_correction_matches_oracle(torch_cuda, slab, prec, exact):
    e = Engine(skin=slab.skin, precision=prec)
    slab.setup(e, tail=True)
    e.set_tail_mode(exact=exact)
    e.rebuild()
    e.compute(stats=True, tail=True)
    ref = oracle_system(slab, tail=True, exact=exact)
    F = e.forces()
    tol = TOL_F_MIXED if prec == "mixed" else 1e-11
    assert rel_force_err(F, ref["F"]) <= tol
    a = e.atom_stats()
    np.testing.assert_allclose(a[:, 0], ref["E"], rtol=1e-4 if prec == "mixed" else 1e-10,
                               atol=1e-5 if prec == "mixed" else 1e-11)
    # the tail term is really there: it moves the z forces by the oracle's amount
    e.compute(stats=True, tail=False)
    F0 = e.forces()
    dz = F[:, 2] - F0[:, 2]
    np.testing.assert_allclose(dz, ref["tail"][:, 0], rtol=1e-6, atol=1e-9)


def test_binary_mixture_matches_oracle(torch_cuda, binary):
    for prec, tol in (("mixed", TOL_F_MIXED), ("fp64", 1e-11)):
        e = Engine(n_types=2, skin=binary.skin, precision=prec)
        binary.setup(e, tail=True)
        e.rebuild()
        e.compute(stats=True, tail=True)
        ref = oracle_system(binary, tail=True)
        goff, gj = e.export_pairs()
        assert np.array_equal(goff, ref["off"]) and np.array_equal(gj, ref["j"])
        assert rel_force_err(e.forces(), ref["F"]) <= tol
        st = e.stats()[0]
        assert abs(st[0] - ref["E"].sum()) / np.abs(ref["E"]).sum() <= 1e-6


def test_batched_replicas_match_oracle_per_replica(torch_cuda):
    w = configs.replicas(seed=9, count=12, n_lo=2000, n_hi=6000)
    e = Engine(skin=w.skin)
    w.setup(e, tail=True)
    e.rebuild()
    e.compute(stats=True, tail=True)
    F = e.forces()
    st = e.stats()
    goff, gj = e.export_pairs()
    for r in range(len(w.n_atoms)):
        sl = w.replica_slice(r)
        ref = oracle_system(w, r, tail=True)
        o = goff[sl.start:sl.stop + 1] - goff[sl.start]
        jj = gj[goff[sl.start]:goff[sl.stop]] - sl.start
        assert np.array_equal(o, ref["off"]) and np.array_equal(jj, ref["j"]), r
        assert rel_force_err(F[sl], ref["F"]) <= TOL_F_MIXED, r
        assert abs(st[r, 0] - ref["E"].sum()) / np.abs(ref["E"]).sum() <= 1e-6, r
        np.testing.assert_allclose(st[r, 1:], ref["W"].sum(0), rtol=1e-6)
    # the persistent launch (wavefronts taking chunks of tiles across replica
    # boundaries) gives the same fixed-point forces and per-replica statistics
    f0 = e.forces_fixed()
    for pw in (1, 4):
        e.set_tuning(persistent=pw)
        e.compute(stats=True, tail=True)
        assert np.array_equal(e.forces_fixed(), f0), pw
        np.testing.assert_allclose(e.stats(), st, rtol=1e-12)


def test_update_positions_and_skin(torch_cuda, slab):
    import torch
    e = Engine(skin=slab.skin)
    slab.setup(e, tail=False)
    pos = torch.from_numpy(slab.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    # small moves: no rebuild needed, forces at the new positions with the old list
    rng = np.random.default_rng(5)
    moved = slab.xyz.copy()
    moved[:, :3] += rng.normal(scale=0.02, size=(slab.n, 3)).astype(np.float32)
    pos.copy_(torch.from_numpy(moved))
    torch.cuda.synchronize()
    assert e.update_positions() is False
    e.compute(stats=True)
    ref = oracle_system(slab, xyz=moved, rc_xyz=slab.xyz)
    # the list and rc were assigned at the old positions; with the skin the list
    # still holds every pair inside rc at the new ones, so the force sets agree
    assert rel_force_err(e.forces(), ref["F"]) <= TOL_F_MIXED
    # a large move trips the skin test, and step() then rebuilds
    moved[7, :3] += np.float32(0.2)
    pos.copy_(torch.from_numpy(moved))
    torch.cuda.synchronize()
    assert e.update_positions() is True
    assert e.step(stats=True) is True
    ref = oracle_system(slab, xyz=moved)
    goff, gj = e.export_pairs()
    assert np.array_equal(goff, ref["off"]) and np.array_equal(gj, ref["j"])
    assert rel_force_err(e.forces(), ref["F"]) <= TOL_F_MIXED


def test_graph_replay_equals_eager_and_skin_flag(torch_cuda, slab):
    import torch
    e = Engine(skin=slab.skin)
    slab.setup(e, tail=True)
    pos = torch.from_numpy(slab.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    e.set_perturbation(0.01, 123, 0, anchored=False)
    results = []
    for rep in range(3):
        e.graph_run(1, stats=True, tail=True, perturb=True)
        f_graph = e.forces_fixed()
        s_graph = e.stats()
        e.update_positions(check_skin=False)       # eager, same positions
        e.compute(stats=True, tail=True)
        assert np.array_equal(e.forces_fixed(), f_graph)
        np.testing.assert_array_equal(e.stats(), s_graph)
        results.append(f_graph)
    assert not np.array_equal(results[0], results[1])   # the positions moved
    e.check_skin_flag()
    # 200 more perturbation steps without a rebuild must trip the flag
    e.graph_run(200, stats=False, tail=False, perturb=True)
    with pytest.raises(SkinExceeded):
        e.check_skin_flag()


@pytest.mark.parametrize("cell_z", [0.25, 1.0])
def test_graph_replay_across_rebuilds(torch_cuda, cell_z):
    """A captured graph keeps reading the sorted permutation after rebuilds.
    The 64k slab's cell key has 17 bits at cell_z 0.25 (3 radix passes) and
    15 bits at 1.0 (2 passes): the permutation must end in the same buffer
    either way, or the replays refresh the wrong atoms."""
    import torch
    w = configs.slab()
    e = Engine(skin=w.skin)
    e.set_tuning(cell_xy=0.5, cell_z=cell_z)
    w.setup(e, tail=True)
    pos = torch.from_numpy(w.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    e.set_perturbation(0.01, 11, 0, anchored=True)
    for seg in range(3):
        e.graph_run(4, stats=True, tail=True, perturb=True)
        f_graph = e.forces_fixed()
        e.update_positions(check_skin=False)       # eager, same positions
        e.compute(stats=True, tail=True)
        assert np.array_equal(e.forces_fixed(), f_graph), seg
        e.rebuild()
        assert e.last_rebuild_displacement() <= 0.5 * w.skin, seg


def test_perturbation_is_reproducible(torch_cuda, slab):
    import torch
    outs = []
    for _ in range(2):
        e = Engine(skin=slab.skin)
        slab.setup(e, tail=False)
        pos = torch.from_numpy(slab.xyz.copy()).cuda()
        torch.cuda.synchronize()
        e.bind_positions(pos)
        e.rebuild()
        e.set_perturbation(0.01, 99, 0)
        e.graph_run(5, perturb=True)
        e.synchronize()
        outs.append(pos.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    d = outs[0][:, :3] - slab.xyz[:, :3]
    assert abs(d.std() / np.sqrt(5) - 0.01) < 5e-4


def test_anchored_perturbation_stays_bounded(torch_cuda, slab):
    """Anchored vibration: x = x0 + sigma N(0,1) at every replay, so 300 steps
    neither trip the skin test nor bring atoms closer than the set-up allows."""
    import torch
    e = Engine(skin=slab.skin)
    slab.setup(e, tail=True)
    pos = torch.from_numpy(slab.xyz.copy()).cuda()
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    e.set_perturbation(0.01, 5, 0, anchored=True)
    e.graph_run(1, perturb=True)
    e.synchronize()
    first = pos.cpu().numpy()
    for _ in range(50):
        e.graph_run(6, stats=True, tail=True, perturb=True)
        e.rebuild()
        assert e.last_rebuild_displacement() <= 0.5 * slab.skin
    e.check_skin_flag()
    e.compute(stats=True, tail=True)
    e.stats()                      # raises on a force-range overflow
    last = pos.cpu().numpy()
    d = last[:, :3] - slab.xyz[:, :3]
    assert abs(d.std() - 0.01) < 5e-4         # one draw from x0, not a walk
    assert not np.array_equal(first, last)
    # the same step counter gives the same positions
    pos.copy_(torch.from_numpy(slab.xyz.copy()))
    torch.cuda.synchronize()
    e.set_perturbation(0.01, 5, 0, anchored=True)
    e.graph_run(1, perturb=True)
    e.synchronize()
    np.testing.assert_array_equal(pos.cpu().numpy(), first)


def test_add_forces_f32_soa_accumulates(torch_cuda, slab):
    import torch
    e = Engine(skin=slab.skin)
    slab.setup(e, tail=False)
    e.rebuild()
    e.compute()
    F = e.forces()
    fx = torch.ones(slab.n, device="cuda")
    fy = torch.zeros(slab.n, device="cuda")
    fz = torch.zeros(slab.n, device="cuda")
    e.add_forces_f32_soa(fx, fy, fz)
    e.synchronize()
    np.testing.assert_allclose(fx.cpu().numpy(), (F[:, 0] + 1).astype(np.float32), rtol=1e-6,
                               atol=1e-6)
    np.testing.assert_allclose(fz.cpu().numpy(), F[:, 2].astype(np.float32), rtol=1e-6)


def test_autotune_keeps_forces_bitwise(torch_cuda, tmp_path, slab):
    e = Engine(skin=slab.skin)
    slab.setup(e, tail=True)
    e.rebuild()
    e.compute(stats=True, tail=True)
    f0 = e.forces_fixed()
    cache = str(tmp_path / "tune.txt")
    t = e.autotune(stats=True, tail=True, cache_path=cache)
    e.compute(stats=True, tail=True)
    assert np.array_equal(e.forces_fixed(), f0)
    # a second engine reads the cache and lands on the same tuning
    e2 = Engine(skin=slab.skin)
    slab.setup(e2, tail=True)
    t2 = e2.autotune(stats=True, tail=True, cache_path=cache)
    assert t2 == t


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_adjoint_matches_oracle_gradient(torch_cuda, binary, prec):
    import torch
    w = binary
    e = Engine(n_types=2, skin=w.skin, precision=prec)
    w.setup(e, tail=True)
    e.rebuild()
    rng = np.random.default_rng(17)
    g = rng.normal(size=(w.n, 3))
    dLdE = 0.37
    ge, gs, grc = e.adjoint(dLdE=[dLdE], dLdF=torch.from_numpy(g).cuda(), tail=True,
                            n_rc=w.rc_tables[0].size)
    box = w.boxes[0]
    rc, _ = O.assign_rc(w.xyz, box, w.rc_tables[0])
    off, j = O.build_pairs(w.xyz, box, rc, w.skin)
    pe, ps = O.grad_pairs(w.xyz, box, w.types, w.eps, w.sigma, rc, off, j, dLdE, g)
    te, ts, trc = O.grad_tail(w.xyz, box, w.types, w.eps, w.sigma, w.rc_tables[0].ravel(),
                              w.rho[0], dLdE, g, h=1e-4)
    tol = 2e-5 if prec == "mixed" else 1e-7
    np.testing.assert_allclose(ge, pe + te, rtol=tol, atol=tol * np.abs(pe).max())
    np.testing.assert_allclose(gs, ps + ts, rtol=tol, atol=tol * np.abs(ps).max())
    np.testing.assert_allclose(grc, trc, rtol=1e-6, atol=1e-6 * np.abs(trc).max())


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_adjoint_bitwise_independent_of_tuning(torch_cuda, binary, prec):
    """The adjoint's per-pair terms are summed as integers, so its gradients
    are the same bits whatever the lanes per atom or block size."""
    import torch
    w = binary
    g = torch.from_numpy(np.random.default_rng(3).normal(size=(w.n, 3))).cuda()
    res = []
    for lanes, block in ((4, 256), (8, 128), (2, 64)):
        e = Engine(n_types=2, skin=w.skin, precision=prec)
        w.setup(e, tail=True)
        e.set_tuning(force_lanes=lanes, build_lanes=lanes, block=block)
        e.rebuild()
        res.append(e.adjoint(dLdE=[0.3], dLdF=g, tail=True, n_rc=w.rc_tables[0].size))
        e.close()
    for r in res[1:]:
        for a, b in zip(res[0], r):
            assert np.array_equal(a, b)


def test_overlapping_atoms_raise_force_range(torch_cuda):
    """Two atoms 0.3 sigma apart: their pair force leaves the range where the
    fixed-point sums are exact; the compute finishes and the read raises."""
    from sre.engine import ForceRange
    from sre import gen_fcc
    xyz = gen_fcc(5, 1.6796, 0.02, 3)
    xyz[1, :3] = xyz[0, :3] + np.array([0.3, 0.0, 0.0], np.float32)
    L = np.float32(5 * 1.6796)
    e = Engine(n_types=1, skin=0.3)
    e.set_system([len(xyz)], np.array([[L, L, L]], np.float32))
    e.set_types(np.zeros(len(xyz), np.int32))
    e.set_pair_params(np.ones((1, 1)), np.ones((1, 1)))
    e.set_rc_table(np.full((1, 1, 1), 2.5, np.float32))
    e.upload_positions(xyz)
    e.rebuild()
    e.compute(stats=True)
    with pytest.raises(ForceRange):
        e.stats()


def _segments_vs_manual(w, n_seg, steps, sort_mode=0, prec="mixed"):
    import torch
    res = []
    for queued in (False, True):
        e = Engine(n_types=w.n_types, skin=w.skin, precision=prec)
        e.set_tuning(sort_mode=sort_mode)
        w.setup(e, tail=True)
        pos = torch.from_numpy(w.xyz.copy()).cuda()
        torch.cuda.synchronize()
        e.bind_positions(pos)
        e.rebuild()
        e.set_perturbation(0.01, 21, 0, anchored=True)
        if queued:
            nls, disp = e.run_segments(n_seg, steps, stats=True, tail=True, perturb=True)
            st = e.stats().copy()     # the last step's, read after the final rebuild
        else:
            nls, disp = [], 0.0
            for _ in range(n_seg):
                e.graph_run(steps, stats=True, tail=True, perturb=True)
                st = e.stats().copy()
                e.rebuild()
                nls.append(e.neighbor_count(in_rc=False)[0])
                disp = max(disp, e.last_rebuild_displacement())
        e.graph_run(5, stats=True, tail=True, perturb=True)
        F = e.forces_fixed() if prec == "mixed" else e.forces()
        res.append((list(nls), disp, F, e.export_pairs(), pos.cpu().numpy(), st))
        e.close()
    (n0, d0, f0, p0, x0, s0), (n1, d1, f1, p1, x1, s1) = res
    assert n0 == n1 and d0 == d1 and 0.0 < d1 <= 0.5 * w.skin
    np.testing.assert_array_equal(s0, s1)
    assert np.array_equal(x0, x1) and np.array_equal(f0, f1)
    assert np.array_equal(p0[0], p1[0]) and np.array_equal(p0[1], p1[1])


@pytest.mark.parametrize("n_seg,steps,sort_mode", [(3, 5, 0), (19, 2, 0), (4, 3, 1)])
def test_run_segments_equals_graph_and_rebuild(torch_cuda, slab, n_seg, steps, sort_mode):
    """sre_run_segments (segments and rebuilds queued as graphs of up to 16
    segments, the cell histogram of each rebuild made by the segment's last
    force launch, checks at the end) does what graph_run + rebuild per
    segment does: the same positions, lists and forces bit for bit, the same
    list lengths, largest displacement and last-step statistics.  sort_mode
    1 (radix sort): no histogram in the force launch."""
    _segments_vs_manual(slab, n_seg, steps, sort_mode)


def test_run_segments_equals_graph_and_rebuild_species_and_replicas(torch_cuda, binary, slab):
    """The same with per-type parameters, with several replicas (the
    histogram blocks compute every replica's cell keys) and in the FP64 form
    (its force launches carry the same refresh and histogram blocks)."""
    _segments_vs_manual(binary, 5, 4)
    _segments_vs_manual(slab, 3, 4, prec="fp64")
    _segments_vs_manual(configs.replicas(seed=15, count=5, n_lo=1500, n_hi=5000, nz=32),
                        5, 4)
