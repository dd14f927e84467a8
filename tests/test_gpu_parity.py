"""HIP path vs the oracle, through the C ABI (libgodar_amd.so).

Bars:
  * candidate and contact pair sets: identical to the oracle's, every step;
  * FP64 state: bitwise equal to the oracle's in the lockstep runs (config 0,
    stiff bonds, switches off, gridded wind) and to the reference-made
    10^4-step fixture.  exp, log, acos and asin are correctly rounded on the
    device; glibc, which the oracle and the reference call, misrounds about
    one call in a thousand, so a force column can differ by an ulp
    (sheltering minima) without the state differing; FP64_REL_TOL bounds the
    force columns and the dense ridging fixture;
  * batch invariance: a member's trajectory does not depend on its batch
    mates (bitwise), in FP64 and in mixed precision.
"""
import numpy as np
import pytest

from fixtures import load, load_long, pairs_at
from godar_amd import generators as g
from workloads import config0_params, kick, packed_box, squeeze

pytestmark = pytest.mark.gpu

STATE = ["x", "y", "u", "v", "omega", "r", "h"]
STRESS = ["tsigxx", "tsigyy", "tsigxy", "tsigyx"]
FP64_REL_TOL = 1e-8


def _ctx(n, members=1, precision="fp64", **kw):
    from godar_amd.sim import Context
    return Context(n, members, precision, record_forces=True, **kw)


def _rel(a, b):
    scale = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / scale))


def _pairs(ctx, m, kind):
    i, j = ctx.pairs(m, kind)
    return list(zip(i.tolist(), j.tolist()))


def _oracle_pairs(O):
    li, lj, lf = O.log()
    cand = sorted(zip(li.tolist(), lj.tolist()))
    c = (lf & 1).astype(bool)
    cont = sorted(zip(li[c].tolist(), lj[c].tolist()))
    return cand, cont


def _lockstep(field, params, nsteps, precision="fp64", check_every=1):
    from oracle_py import OracleSim
    ctx = _ctx(field.n, 1, precision)
    m = ctx.add_field(params, field)
    O = OracleSim(params, field)
    worst = 0.0
    seeds = repaired = 0
    for t in range(1, nsteps + 1):
        ctx.step(1)
        O.step()
        st = ctx.stats(m)
        seeds += st["seeds"]
        repaired += st["repaired"]
        if t % check_every and t != nsteps:
            continue
        cand, cont = _oracle_pairs(O)
        assert _pairs(ctx, m, "candidate") == cand, f"candidate set differs at step {t}"
        assert _pairs(ctx, m, "contact") == cont, f"contact set differs at step {t}"
        a, b = ctx.floes(m), O.state()
        for k in STATE:
            worst = max(worst, _rel(a[k], b[k]) if k in ("x", "y", "r", "h") else
                        float(np.max(np.abs(a[k] - b[k]))) / (float(np.max(np.abs(b[k]))) + 1e-300))
    ctx.close()
    return worst, seeds, repaired


def test_one_step_identical_inputs():
    """Step 1 from identical inputs: pair sets exact, state equal to the last bits."""
    from oracle_py import OracleSim
    for field, params in ((g.config0(0), None), (kick(squeeze(packed_box(64, 3)), 1), "dense")):
        p = config0_params(field) if params is None else config0_params(field, sigmanc_crit=1e5)
        ctx = _ctx(field.n)
        m = ctx.add_field(p, field)
        O = OracleSim(p, field)
        ctx.step(1)
        O.step()
        cand, cont = _oracle_pairs(O)
        assert _pairs(ctx, m, "candidate") == cand
        assert _pairs(ctx, m, "contact") == cont
        a, b = ctx.floes(m), O.state()
        fa, fb = ctx.forces(m), O.forces()
        for k in STATE:
            assert np.array_equal(a[k], b[k]), k
        for k in ("fcx", "fcy", "mc", "fbx", "fby", "mb", "fx_bc", "fy_bc", "m_bc", "tac", "pc",
                  "tab", "pb", "ta_bc", "p_bc", "tp", "hsfa_min", "hsfw_min", *STRESS):
            scale = float(np.max(np.abs(fb[k]))) + 1e-300
            assert float(np.max(np.abs(fa[k] - fb[k]))) / scale <= 1e-13, k
        ctx.close()


def test_stress_diagnostics_match_oracle():
    """Stress and pressure diagnostics (stepper.f90:131-146,184-199,238-271,
    328-331; boundaries.f90:141-144) in a colliding, ridging box: candidate
    pairs that separated keep contributing their last contact forces; FP64
    vs the oracle every step, within the FP64 tolerance."""
    from oracle_py import OracleSim
    f = kick(squeeze(packed_box(64, 3)), 1)
    p = config0_params(f, sigmanc_crit=1e5)
    ctx = _ctx(f.n)
    m = ctx.add_field(p, f)
    O = OracleSim(p, f)
    ever, stale_seen, worst = set(), 0, 0.0
    for t in range(1, 81):
        ctx.step(1)
        O.step()
        cand, cont = _oracle_pairs(O)
        ever |= set(cont)
        stale_seen += len((set(cand) - set(cont)) & ever)
        fa, fb = ctx.forces(m), O.forces()
        for k in (*STRESS, "tp", "tac", "pc", "tab", "pb", "ta_bc", "p_bc"):
            scale = float(np.max(np.abs(fb[k]))) + 1e-300
            worst = max(worst, float(np.max(np.abs(fa[k] - fb[k]))) / scale)
    ctx.close()
    assert stale_seen > 0, "no separated pair exercised the last-contact forces"
    assert worst < FP64_REL_TOL, worst


def test_config0_trajectory_fp64():
    f, p, d = load("config0_seed0.npz")
    worst, _, _ = _lockstep(f, p, 300, check_every=20)
    assert worst == 0.0, worst


def test_config0_golden_fp64():
    """1,000 steps against the reference-made fixture."""
    f, p, d = load("config0_seed0.npz")
    ctx = _ctx(f.n)
    m = ctx.add_field(p, f)
    for t in (1, 100, 1000):
        ctx.step(t - ctx.tstep(m) + 1)
        if t in (1, 100, 1000):
            assert _pairs(ctx, m, "candidate") == pairs_at(d, "cand", t)
            assert _pairs(ctx, m, "contact") == pairs_at(d, "cont", t)
            a = ctx.floes(m)
            for k in ("x", "y", "r", "h"):
                assert np.array_equal(a[k], d[f"{k}_{t}"]), (k, t)
    ctx.close()


def test_dense_ridging_repair_fp64():
    """Heavy in-loop ridging: the ordered repair must reproduce the sequential semantics."""
    f, p, d = load("dense64.npz")
    ctx = _ctx(f.n)
    m = ctx.add_field(p, f)
    seeds = 0
    for t in range(1, 51):
        ctx.step(1)
        seeds += ctx.stats(m)["seeds"]
        assert _pairs(ctx, m, "candidate") == pairs_at(d, "cand", t), t
        assert _pairs(ctx, m, "contact") == pairs_at(d, "cont", t), t
        if t in (1, 10, 50):
            a = ctx.floes(m)
            for k in ("x", "y", "r", "h"):
                assert _rel(a[k], d[f"{k}_{t}"]) < FP64_REL_TOL, (k, t)
    assert seeds > 300
    ctx.close()


def test_repair_modes_bitwise():
    """The ridging repair gives the same trajectory, bitwise, whether each
    member evaluates 64, 5 or 1 pending events at once, and when its pending
    events outgrow on-chip memory (8 or 3 kept) and it continues on the heap
    in HBM; heavy in-loop ridging, FP64 and mixed precision."""
    fd, pd, _ = load("dense64.npz")
    fb = kick(squeeze(packed_box(64, 3)), 1)
    pb = config0_params(fb, sigmanc_crit=1e5)
    for prec in ("fp64", "mixed"):
        runs = {}
        for mode in ((64, 1024), (1, 1024), (5, 1024), (64, 8), (64, 3)):
            ctx = _ctx(fd.n + fb.n, 2, prec)
            ctx.set_repair(*mode)
            ms = [ctx.add_field(pd, fd), ctx.add_field(pb, fb)]
            repaired = 0
            snaps = []
            for t in range(1, 31):
                ctx.step(1)
                repaired += sum(ctx.stats(m)["repaired"] for m in ms)
                if t in (1, 5, 30):
                    snaps.append([(ctx.floes(m), _pairs(ctx, m, "contact"), ctx.forces(m)["tfx"]) for m in ms])
            runs[mode] = (snaps, repaired)
            ctx.close()
        ref, rep0 = runs[(64, 1024)]
        assert rep0 > 300, rep0
        for mode, (snaps, rep) in runs.items():
            assert rep == rep0, (prec, mode, rep, rep0)
            for sa, sb in zip(snaps, ref):
                for (fa, ca, ta), (fb_, cb, tb) in zip(sa, sb):
                    for key in STATE:
                        assert np.array_equal(fa[key], fb_[key]), (prec, mode, key)
                    assert ca == cb, (prec, mode)
                    assert np.array_equal(ta, tb), (prec, mode)


def test_repair_components_bitwise():
    """Nine separate ridging clusters in one member: a round of the repair
    commits the events of different contact components together.  The
    trajectory, contact sets, total forces and the number of events
    re-evaluated equal those of the one-at-a-time repair (the reference's
    key order) bitwise, in FP64 and mixed precision, with the pending events
    on chip and spilling to the heap; FP64 also against the oracle."""
    from workloads import clusters
    f = clusters()
    p = config0_params(f, sigmanc_crit=1e5)
    for prec in ("fp64", "mixed"):
        runs = {}
        for mode in ((64, 1024), (1, 1024), (7, 1024), (64, 12)):
            ctx = _ctx(f.n, 1, prec)
            ctx.set_repair(*mode)
            m = ctx.add_field(p, f)
            repaired = seeds = 0
            snaps = []
            for t in range(1, 41):
                ctx.step(1)
                st = ctx.stats(m)
                repaired += st["repaired"]
                seeds += st["seeds"]
                if t in (1, 10, 40):
                    snaps.append((ctx.floes(m), _pairs(ctx, m, "contact"), ctx.forces(m)["tfx"]))
            runs[mode] = (snaps, repaired, seeds)
            ctx.close()
        ref, rep0, seeds0 = runs[(1, 1024)]
        print(f"[components] {prec}: {seeds0} seeds, {rep0} events re-evaluated", flush=True)
        assert seeds0 > 50 and rep0 > 100, (seeds0, rep0)
        for mode, (snaps, rep, _) in runs.items():
            assert rep == rep0, (prec, mode, rep, rep0)
            for sa, sb in zip(snaps, ref):
                assert sa[1] == sb[1], (prec, mode)
                assert np.array_equal(sa[2], sb[2]), (prec, mode)
                for k in STATE:
                    assert np.array_equal(sa[0][k], sb[0][k]), (prec, mode, k)
    worst, seeds, _ = _lockstep(f, p, 40, check_every=5)
    print(f"[components] fp64 vs the oracle: worst {worst:.3e}", flush=True)
    assert worst < FP64_REL_TOL and seeds > 50, (worst, seeds)


def test_superset_skin_bitwise():
    """Candidate lists taken from skinned supersets (rebuilt only after a floe
    has moved far enough) give the same trajectory, bitwise, as a full cell
    search every step (skin 0): sparse and dense (wavefront) searches, moving
    fields with ridging, FP64 and mixed."""
    fa = kick(g.config0_member(41, 600), 3, speed=4.0)   # rebuilds every few steps at small skins
    fd, pd, _ = load("dense64.npz")
    pa = config0_params(fa)
    for prec in ("fp64", "mixed"):
        runs = {}
        for skin in (0.0, 0.02, 0.1, 0.5):
            ctx = _ctx(fa.n + 2 * fd.n, 3, prec)
            ctx.set_skin(skin)
            ms = [ctx.add_field(pa, fa), ctx.add_field(pd, fd)]
            ctx.set_dense_search(0)          # this member: wavefront search
            ms.append(ctx.add_field(pd, fd))
            snaps = []
            for t in range(1, 41):
                ctx.step(1)
                if t in (1, 7, 40):
                    snaps.append([(ctx.floes(m), _pairs(ctx, m, "candidate"), ctx.forces(m)["tfx"])
                                  for m in ms])
            runs[skin] = snaps
            ctx.close()
        ref = runs[0.0]
        for skin, snaps in runs.items():
            for sa, sb in zip(snaps, ref):
                for (fa_, ca, ta), (fb_, cb, tb) in zip(sa, sb):
                    for key in STATE:
                        assert np.array_equal(fa_[key], fb_[key]), (prec, skin, key)
                    assert ca == cb, (prec, skin)
                    assert np.array_equal(ta, tb), (prec, skin)


def test_stiff_bond_breaking_fp64():
    f = kick(squeeze(packed_box(64, 5), 0.995), 2, speed=0.5)
    p = config0_params(f, sigmanc_crit=1e5, e_modul=6e9, eb=6e9, bond_lim=50.0, bond_frac=1.0)
    worst, seeds, _ = _lockstep(f, p, 150, check_every=5)
    assert worst == 0.0 and seeds > 0, (worst, seeds)


def test_switches_off_fp64():
    f = kick(squeeze(packed_box(64, 4)), 3)
    p = config0_params(f, shelter=0, ridging=0, slipping=0, cohesion=0)
    worst, seeds, _ = _lockstep(f, p, 120, check_every=10)
    assert worst == 0.0 and seeds == 0, (worst, seeds)


def test_wind_grid_fp64():
    f = kick(squeeze(packed_box(64, 6)), 4)
    f.wind_grid = g.gaussian_wind_grid(f.box, 8.0, 3.0, 5e3, seed=9, spacing=1e3)
    p = config0_params(f, sigmanc_crit=1e5)
    worst, _, _ = _lockstep(f, p, 120, check_every=10)
    assert worst == 0.0, worst


def test_bonds_and_bond_state_match_oracle():
    from oracle_py import OracleSim
    f, p, _ = load("config0_seed0.npz")
    ctx = _ctx(f.n)
    m = ctx.add_field(p, f)
    O = OracleSim(p, f)
    ctx.step(50)
    O.step(50)
    bi, bj = ctx.pairs(m, "bond")
    oi, oj = O.bonds()
    assert np.array_equal(bi, oi) and np.array_equal(bj, oj) and bi.size > 200
    st = ctx.bond_state(m)
    ost = O.pairs(oi, oj)
    for k, ko in (("damage", "damage"), ("deltanb", "deltanb"), ("deltatb", "deltatb"),
                  ("rb", "rb"), ("hb", "hb"), ("lb", "lb"), ("sb", "sb"), ("ib", "ib")):
        scale = float(np.max(np.abs(ost[ko]))) + 1e-300
        assert float(np.max(np.abs(st[k] - ost[ko]))) / scale < 1e-9, k
    ctx.close()


def test_batch_invariance_bitwise():
    """Three fields of different sizes in one context == each alone, bitwise."""
    fields = [g.config0_member(11, 600), kick(squeeze(packed_box(64, 3)), 1), g.config0_member(12, 900)]
    params = [config0_params(fields[0]), config0_params(fields[1], sigmanc_crit=1e5),
              config0_params(fields[2])]
    for prec in ("fp64", "mixed"):
        tot = sum(f.n for f in fields)
        ctx = _ctx(tot, 3, prec)
        ms = [ctx.add_field(p, f) for f, p in zip(fields, params)]
        ctx.step(60)
        for f, p, m in zip(fields, params, ms):
            alone = _ctx(f.n, 1, prec)
            ma = alone.add_field(p, f)
            alone.step(60)
            a, b = ctx.floes(m), alone.floes(ma)
            for k in STATE:
                assert np.array_equal(a[k], b[k]), (prec, k)
            assert _pairs(ctx, m, "contact") == _pairs(alone, ma, "contact")
            alone.close()
        ctx.close()


def test_graph_replay_matches_eager():
    """Graph-replayed steps == kernel-by-kernel steps, bitwise, on every
    replay: replays alternate with eager work (pair and force read-backs) and
    synchronisations, and a member added mid-run forces a recapture."""
    fa = g.config0_member(21, 700)
    fb = kick(squeeze(packed_box(64, 3)), 1)
    pa, pb = config0_params(fa), config0_params(fb, sigmanc_crit=1e5)
    for prec in ("fp64", "mixed"):
        eager, graph = _ctx(fa.n + fb.n, 2, prec), _ctx(fa.n + fb.n, 2, prec)
        eager.use_graphs(False)
        ms_e, ms_g = [eager.add_field(pa, fa)], [graph.add_field(pa, fa)]
        for rnd in range(8):
            if rnd == 3:
                ms_e.append(eager.add_field(pb, fb))
                ms_g.append(graph.add_field(pb, fb))
            k = 1 + rnd % 3
            eager.step(k)
            graph.step(k)
            for me, mg in zip(ms_e, ms_g):
                a, b = graph.floes(mg), eager.floes(me)
                for key in STATE:
                    assert np.array_equal(a[key], b[key]), (prec, rnd, key)
                assert _pairs(graph, mg, "contact") == _pairs(eager, me, "contact")
                assert np.array_equal(graph.forces(mg)["tfx"], eager.forces(me)["tfx"])
        eager.close()
        graph.close()


def test_side_streams_match_serial_launches():
    """The contact and wall kernels run on a side stream beside the bond
    kernel; steps timed stage by stage launch the three one after another.
    Both give the same trajectory, forces and pair sets, bitwise, on a packed
    box where pairs are bonded and in contact at once (both kernels set bits
    of such a pair's outer flag word, with atomicOr)."""
    fb = kick(squeeze(packed_box(64, 3)), 1)
    p = config0_params(fb, sigmanc_crit=1e5, bond_lim=50.0, bond_frac=1.0)
    for prec in ("fp64", "mixed"):
        serial, forked = _ctx(fb.n, 1, prec), _ctx(fb.n, 1, prec)
        serial.profile(True)
        forked.use_graphs(False)
        ms, mf = serial.add_field(p, fb), forked.add_field(p, fb)
        both = 0
        for rnd in range(6):
            serial.step(10)
            forked.step(10)
            a, b = serial.floes(ms), forked.floes(mf)
            for key in STATE:
                assert np.array_equal(a[key], b[key]), (prec, rnd, key)
            fs, ff = serial.forces(ms), forked.forces(mf)
            for key in ("tfx", "tfy", "m", "fcx", "fbx", "mb", "fx_bc"):
                assert np.array_equal(fs[key], ff[key]), (prec, rnd, key)
            cs, cf = _pairs(serial, ms, "contact"), _pairs(forked, mf, "contact")
            bs, bf = _pairs(serial, ms, "bond"), _pairs(forked, mf, "bond")
            assert cs == cf and bs == bf, (prec, rnd)
            both += len(set(cs) & set(bs))
        assert both > 0, prec     # pairs bonded and in contact at once
        serial.close()
        forked.close()


def test_block_sizes_and_autotune_bitwise():
    """Results do not depend on the kernels' block sizes: 64-thread blocks,
    the autotuner's choice (which itself steps the members) and the default
    all give the same trajectory, bitwise."""
    f = g.config0_member(31, 800)
    fb = kick(squeeze(packed_box(64, 3)), 1)
    pa, pb = config0_params(f), config0_params(fb, sigmanc_crit=1e5)
    runs = []
    for mode in ("default", "b64", "auto"):
        ctx = _ctx(f.n + fb.n, 2, "mixed")
        ms = [ctx.add_field(pa, f), ctx.add_field(pb, fb)]
        if mode == "b64":
            ctx.set_block(64)
        if mode == "auto":
            blocks = ctx.autotune(1)
            assert set(blocks.values()) <= {64, 128, 256}
            done = ctx.tstep(ms[0]) - 1
        else:
            done = 0
        assert done < 250
        ctx.step(250 - done)
        assert ctx.tstep(ms[0]) == 251
        if mode == "auto":
            sh = ctx.search()
            assert sh["skin_frac"] in (0.05, 0.1, 0.2) and sh["thread_span"] in (1, 2)
            assert sh["dense_scan"] in (48, 96, 192)
        runs.append([ctx.floes(m) for m in ms])
        ctx.close()
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            for k in STATE:
                assert np.array_equal(a[k], b[k]), k


def test_search_shape_bitwise():
    """Results do not depend on the search shape: superset skin, the thread
    search's cell span and the dense-search threshold, set before the run or
    changed in the middle of it (the next step builds the supersets anew),
    give the same trajectory as the defaults, bitwise."""
    from godar_amd import workloads as W
    fa, pa = W.config4_field((3000, 78))
    fb = g.config0_member(43, 700)
    pb = config0_params(fb)
    shapes = [None, (0.05, 2, 96), (0.2, 1, 0), (0.1, 2, 10 ** 9)]
    for prec in ("fp64", "mixed"):
        runs = []
        for shape in shapes + ["switch"]:
            ctx = _ctx(fa.n + fb.n, 2, prec)
            ms = [ctx.add_field(pa, fa), ctx.add_field(pb, fb)]
            if shape == "switch":
                ctx.step(20)
                ctx.set_search(0.05, 2, 0)
                ctx.step(20)
                ctx.set_search(0.2, 1, 10 ** 9)
                ctx.step(20)
            else:
                if shape is not None:
                    ctx.set_search(*shape)
                    sh = ctx.search()
                    assert (sh["skin_frac"], sh["thread_span"], sh["dense_scan"]) == shape
                ctx.step(60)
            runs.append([ctx.floes(m) for m in ms])
            ctx.close()
        for other in runs[1:]:
            for a, b in zip(runs[0], other):
                for k in STATE:
                    assert np.array_equal(a[k], b[k]), (prec, k)


def test_dense_search_matches_thread_search():
    """The wavefront-per-floe neighbour search (dense members) gives the same
    lists, sheltering and trajectories as the thread-per-floe search,
    bitwise: a power-law field (many small floes per cell) and a sparse one,
    each forced through both searches."""
    from godar_amd import workloads as W
    fa, pa = W.config4_field((3000, 77))
    fb = g.config0_member(41, 900)
    pb = config0_params(fb)
    for prec in ("fp64", "mixed"):
        runs = []
        for thr in (0, 1 << 30):
            ctx = _ctx(fa.n + fb.n, 2, prec)
            ctx.set_dense_search(thr)
            ms = [ctx.add_field(pa, fa), ctx.add_field(pb, fb)]
            ctx.step(25)
            runs.append(([ctx.floes(m) for m in ms], [_pairs(ctx, m, "candidate") for m in ms],
                         [ctx.forces(m)["hsfa_min"] for m in ms]))
            ctx.close()
        (sa, ca, ha), (sb, cb, hb) = runs
        for a, b in zip(sa, sb):
            for k in STATE:
                assert np.array_equal(a[k], b[k]), (prec, k)
        assert ca == cb
        for a, b in zip(ha, hb):
            assert np.array_equal(a, b)


def test_dense_search_long_lists_and_unordered_supersets():
    """Dense floes with long candidate lists give the same lists, sheltering
    and trajectories through the dense search as through the thread search,
    bitwise, in FP64 and mixed precision: a member with a search radius of
    5 km (many floes with more than 48 candidates on a side, so that entries
    often go below the register part of the outer list, in supersets short
    enough for the build to order), and one with a radius of 12 km whose
    supersets exceed 1,024 entries (left in scan order at the build and
    handed to the thread search)."""
    cases = [(packed_box(600, 5, r_lo=200.0, r_hi=400.0), 5.0e3, 48, 6),
             (packed_box(1500, 8, box=1.4e4, r_lo=100.0, r_hi=200.0), 1.2e4, 1024, 3)]
    for f, rtree, big, steps in cases:
        p = config0_params(f, rtree=rtree)
        for prec in ("fp64", "mixed"):
            runs = []
            for thr in (0, 1 << 30):
                ctx = _ctx(f.n, 1, prec)
                ctx.set_dense_search(thr)
                m = ctx.add_field(p, f)
                ctx.step(steps)
                runs.append((ctx.floes(m), _pairs(ctx, m, "candidate"), ctx.forces(m)["hsfa_min"]))
                ctx.close()
            (sa, ca, ha), (sb, cb, hb) = runs
            outer = np.bincount([i for i, _ in ca], minlength=f.n)
            inner = np.bincount([j for _, j in ca], minlength=f.n)
            many = ((outer > big) | (inner > big)) if big < 1024 else (outer + inner > big)
            assert 0.02 < many.mean(), (rtree, many.mean())
            if big < 1024:      # most supersets short enough to be ordered (256 entries)
                assert np.mean(outer + inner < 200) > 0.9, (rtree, np.mean(outer + inner))
            for k in STATE:
                assert np.array_equal(sa[k], sb[k]), (prec, rtree, k)
            assert ca == cb
            assert np.array_equal(ha, hb)


def test_checkpoint_restore_continues_exactly():
    """A member restored from a checkpoint (floes, walls, contact histories
    with last contact forces, bond strain state, step number) continues
    bitwise as the uninterrupted run, alone or next to another member, in FP64
    and mixed precision; forces and stresses of the next steps match too."""
    fa = kick(squeeze(packed_box(64, 3)), 1)
    pa = config0_params(fa, sigmanc_crit=1e5)
    fb = g.config0_member(51, 700)
    pb = config0_params(fb)
    for prec in ("fp64", "mixed"):
        ref = _ctx(fa.n, 1, prec)
        m0 = ref.add_field(pa, fa)
        ref.step(30)
        block = ref.checkpoint(m0, pa)
        ref.step(25)
        want, wantf = ref.floes(m0), ref.forces(m0)
        ref.close()
        ctx = _ctx(fa.n + fb.n, 2, prec)
        ctx.add_field(pb, fb)
        m1 = ctx.restore(block)
        assert ctx.tstep(m1) == 31
        ctx.step(25)
        got, gotf = ctx.floes(m1), ctx.forces(m1)
        ctx.close()
        for k in STATE + ["theta"]:
            assert np.array_equal(got[k], want[k]), (prec, k)
        for k in ("tfx", "tfy", "m", *STRESS):
            assert np.array_equal(gotf[k], wantf[k]), (prec, k)


def test_forcing_change_takes_effect_next_step():
    """A wind or current change between steps (godar_set_forcing) acts from
    the next step on: the run continues bitwise like a member restored from a
    checkpoint that carries the new forcing."""
    import copy
    f = kick(g.config0_member(61, 500), 5)
    p = config0_params(f)
    p2 = copy.copy(p)
    p2.ua, p2.va, p2.uw, p2.vw = -6.0, 4.0, 0.05, -0.2
    for prec in ("fp64", "mixed"):
        a = _ctx(f.n, 1, prec)
        ma = a.add_field(p, f)
        a.step(12)
        a.set_forcing(ma, p2.ua, p2.va, p2.uw, p2.vw)
        block = a.checkpoint(ma, p2)
        a.step(9)
        b = _ctx(f.n, 1, prec)
        mb = b.restore(block)
        b.step(9)
        fa, fb = a.floes(ma), b.floes(mb)
        for k in STATE:
            assert np.array_equal(fa[k], fb[k]), (prec, k)
        assert np.array_equal(a.forces(ma)["tfx"], b.forces(mb)["tfx"]), prec
        a.close()
        b.close()


def test_capacity_overflow_rolls_back():
    """A too-small candidate list aborts the step without advancing the state;
    after reserving more room the run continues exactly as an unhindered one.
    godar_step does the same by itself (automatic growth)."""
    from godar_amd._lib import GodarError
    f = kick(squeeze(packed_box(64, 3)), 1)
    p = config0_params(f, sigmanc_crit=1e5, rtree=6000.0)
    ref = _ctx(f.n)
    mr = ref.add_field(p, f)
    ref.step(20)
    ctx = _ctx(f.n, outer=4, inner=4)
    m = ctx.add_field(p, f)
    ctx.step_async(20)
    with pytest.raises(GodarError):
        ctx.sync()
    assert ctx.tstep(m) == 1
    x0 = ctx.floes(m)["x"]
    assert np.array_equal(x0, f.x)
    ctx.reserve(outer=64, inner=64)
    ctx.step(20)
    a, b = ctx.floes(m), ref.floes(mr)
    for k in STATE:
        assert np.array_equal(a[k], b[k]), k
    auto = _ctx(f.n, outer=2, inner=2, hist=1, bonds=1)
    ma = auto.add_field(p, f)
    auto.step(7)
    auto.step(13)
    a = auto.floes(ma)
    for k in STATE:
        assert np.array_equal(a[k], b[k]), k
    auto.close()
    ctx.close()
    ref.close()


def test_restart_with_loaded_bonds():
    """Bonds handed over through godar_set_bonds (restart) continue exactly."""
    f, p, _ = load("config0_seed0.npz")
    ctx = _ctx(f.n)
    m = ctx.add_field(p, f)
    ctx.step(30)
    st = ctx.floes(m)
    bi, bj = ctx.pairs(m, "bond")
    bs = ctx.bond_state(m)
    cols = np.stack([bs[k] for k in ("bond", "damage", "deltanb", "deltatb", "thetarelb_ji",
                                      "thetarelb_ij", "rb", "hb", "lb", "sb", "ib")], axis=1)
    ctx.step(1)
    a = ctx.floes(m)
    # fresh member at step 31 with the same floes and bonds (contact histories
    # are not part of the reference's restart files; pick a contact-free state)
    f2 = g.FloeField(st["x"], st["y"], st["r"], st["h"], st["theta"], st["omega"], st["u"], st["v"],
                     f.box, 0, f.wind, f.current)
    p2 = p.copy(restart=1)
    if _pairs(ctx, m, "contact"):
        pytest.skip("state has contacts; restart does not carry contact history")
    c2 = _ctx(f.n)
    m2 = c2.add_field(p2, f2, tstep0=31)
    c2.set_bonds(m2, bi, bj, cols)
    c2.step(1)
    b = c2.floes(m2)
    for k in STATE:
        assert np.array_equal(a[k], b[k]), k
    ctx.close()
    c2.close()


def test_config1_scale_pair_sets_exact():
    """262,144 floes (config 1 generator): one step from identical inputs gives the
    oracle's candidate and contact sets, and ridging conserves every r^2 h."""
    from oracle_py import OracleSim
    f = g.config1(1, n=262144)
    p = config0_params(f, dt=0.5, bond_frac=0.3, bond_seed=77)
    ctx = _ctx(f.n)
    m = ctx.add_field(p, f)
    O = OracleSim(p, f)
    ctx.step(1)
    O.step()
    cand, cont = _oracle_pairs(O)
    assert _pairs(ctx, m, "candidate") == cand
    assert _pairs(ctx, m, "contact") == cont
    a, b = ctx.floes(m), O.state()
    for k in ("x", "y", "r", "h"):
        assert _rel(a[k], b[k]) < 1e-12, k
    v0 = f.r ** 2 * f.h
    ctx.step(20)
    a = ctx.floes(m)
    assert np.max(np.abs(a["r"] ** 2 * a["h"] / v0 - 1)) < 1e-12
    ctx.close()


def test_long_run_error_budget():
    """10,000 steps of the 1,000-floe field against the reference-made fixture.

    The reference's dynamics are chaotic: moving every initial centre by one
    ulp changes final positions by up to ~4-6e-2 (relative) after 10^4 steps
    (envelope stored in the fixture).  FP64 results that are not bit-identical
    are judged against that envelope; mixed precision is reported beside it.
    Total ice volume (sum r^2 h, ridging conserves it per floe) must hold."""
    f, p, _ = load("config0_seed0.npz")
    L = load_long()
    env_max = max(float(L["ulp_env_max_10000_0"]), float(L["ulp_env_max_10000_1"]))
    env_med = max(float(L["ulp_env_median_10000_0"]), float(L["ulp_env_median_10000_1"]))
    errs = {}
    for prec in ("fp64", "mixed"):
        ctx = _ctx(f.n, 1, prec)
        m = ctx.add_field(p, f)
        ctx.step(2000)
        a = ctx.floes(m)
        xy2 = np.concatenate([a["x"], a["y"]])
        ctx.step(8000)
        a = ctx.floes(m)
        xy = np.concatenate([a["x"], a["y"]])
        ref = np.concatenate([L["x_10000"], L["y_10000"]])
        ref2 = np.concatenate([L["x_2000"], L["y_2000"]])
        rel = np.abs(xy - ref) / np.abs(ref)
        rel2 = np.abs(xy2 - ref2) / np.abs(ref2)
        errs[prec] = dict(max_2000=float(rel2.max()), max=float(rel.max()),
                          median=float(np.median(rel)))
        vol = np.sum(a["r"] ** 2 * a["h"])
        assert abs(vol / np.sum(L["r_10000"] ** 2 * L["h_10000"]) - 1) < 1e-9
        ctx.close()
    print("relative position error after 1e4 steps:", errs,
          "reference 1-ulp envelope: max", env_max, "median", env_med)
    assert errs["fp64"]["max"] <= env_max
    assert errs["fp64"]["median"] <= env_med
