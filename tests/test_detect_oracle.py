"""CPU tests of the detectors: the crossing oracle pinned against the
continuous definition and hand-fixed boundary cases, and the host logic of
pyffa_amd.FFAG_dump (trigger rules, file names and contents, and that the
steps for which ``needs_step`` is False may be skipped) driven by a host
stand-in for FFAG_Bunch."""
import os

import numpy as np
import pytest

from oracle import detect as D

TWO_PI = 2 * np.pi


def brute_force_cross(fi_pre, fi_post, target):
    """Forward motion by less than a turn: some target + 2 pi k lies in
    [fi_pre, fi_post]."""
    out = np.zeros(fi_pre.size, dtype=bool)
    for i in range(fi_pre.size):
        a, b = fi_pre[i], fi_post[i]
        k0 = np.floor((a - target) / TWO_PI) - 1
        out[i] = any(a <= target + TWO_PI * k <= b for k in np.arange(k0, k0 + 4))
    return out


def test_oracle_matches_the_continuous_definition():
    rng = np.random.default_rng(3)
    n = 20000
    pre = rng.uniform(-30.0, 60.0, n)
    post = pre + rng.uniform(0.0, 6.0, n)
    alive = np.ones(n, bool)
    for target in (0.0, 0.3, np.pi, 5.9, TWO_PI - 1e-3):
        m = D.crossing_mask(pre, post, alive, target)
        want = brute_force_cross(pre, post, target)
        # reduction rounding can move an azimuth within ~1e-14 of the target
        # across it; exclude those
        a, b = D.reduce_azimuth(pre), D.reduce_azimuth(post)
        far = (np.abs(a - target) > 1e-9) & (np.abs(b - target) > 1e-9)
        assert far.sum() > 0.99 * n
        assert np.array_equal(m[far], want[far]), target
        assert 0.01 < m.mean() < 0.99


def test_oracle_boundary_cases():
    t = 1.25
    pre = np.array([t, 1.0, t - TWO_PI, 6.0, 6.0, 6.0, 6.0, np.nan, 1.0, 1.0, 2.0])
    post = np.array([2.0, t, 2.0, TWO_PI + 0.0, TWO_PI + 1.0, TWO_PI + t, TWO_PI + 1.3, 2.0, np.nan, 2.0, 1.0])
    alive = np.array([1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 1], bool)
    want = np.array([True,   # target equal to the pre azimuth
                     True,   # target equal to the post azimuth
                     True,   # pre one turn below the target value
                     False,  # the step ends exactly at 2 pi == 0: no wrap, [6, 2 pi) misses 1.25
                     False,  # wrapped, target past neither end
                     True,   # wrapped, target equal to post
                     True,   # wrapped, target inside [0, post]
                     True,   # NaN pre: "post >= pre" is false, so the wrap branch tests target <= post
                     True,   # NaN post: the wrap branch tests target >= pre
                     False,  # dead before the step
                     False])  # backward step read as a wrap: 1.25 is neither >= 2.0 nor <= 1.0
    m = D.crossing_mask(pre, post, alive, t)
    assert np.array_equal(m, want), np.flatnonzero(m != want)
    # backward motion is read as a wrap (the reference has no direction test)
    assert D.crossing_mask(np.array([2.0]), np.array([1.0]), np.array([True]), 0.5)[0]
    assert D.crossing_mask(np.array([2.0]), np.array([1.0]), np.array([True]), 3.0)[0]
    # target 0 with a wrap to exactly 0
    assert D.crossing_mask(np.array([6.0]), np.array([TWO_PI]), np.array([True]), 0.0)[0]


def test_planted_mistake_is_caught():
    """A strict comparison at the ends (a plausible slip) fails the boundary
    check: the pins above discriminate it."""
    def strict(fi_pre, fi_post, alive, target):
        a, b = D.reduce_azimuth(fi_pre), D.reduce_azimuth(fi_post)
        return np.where(b >= a, (a < target) & (target < b), (target > a) | (target < b)) & alive
    pre, post, alive = np.array([1.25]), np.array([2.0]), np.array([True])
    assert D.crossing_mask(pre, post, alive, 1.25)[0] != strict(pre, post, alive, 1.25)[0]


# ---------------------------------------------------------------- detectors
class HostBunch:
    """Host stand-in exposing the FFAG_Bunch methods the detectors call."""

    def __init__(self, M):
        from pyffa_amd.FFAG_Bunch import FFAG_ManageBunchAttribute
        self.BunchAttribute = FFAG_ManageBunchAttribute()
        self.M = M.copy()
        self.pre_fi = None
        self.pre_alive = None

    @property
    def LocalBunch(self):
        return self.M

    def global_ids(self):
        return self.M[:, 15].astype(np.int64)

    def rows(self, r):
        return self.M[np.asarray(r, dtype=np.int64)]

    def _assemble(self, pos):
        return self.M[pos]

    def UpdatePreSteps(self):
        self.pre_fi = self.M[:, 4].copy()
        self.pre_alive = self.M[:, 8] == 1

    @property
    def Step_SurviveFlag(self):
        return self.pre_alive

    def detect_crossing(self, target, input_rows=None):
        if input_rows is None:
            rows = np.flatnonzero(D.crossing_mask(self.pre_fi, self.M[:, 4], self.pre_alive, target))
            return rows, rows
        return D.crossing_mask_portion(self.pre_fi, self.M[:, 4], self.pre_alive, target,
                                       np.asarray(input_rows, dtype=np.int64))


def synthetic_bunch(n, seed):
    rng = np.random.default_rng(seed)
    M = np.zeros((n, 16))
    M[:, 0] = rng.uniform(8.4, 8.8, n)
    M[:, 1] = rng.normal(0, 1e6, n)
    M[:, 2] = rng.normal(0, 1e-3, n)
    M[:, 3] = rng.normal(0, 1e5, n)
    M[:, 4] = rng.uniform(0, TWO_PI, n)
    M[:, 5] = 1.9e8 / M[:, 0]
    M[:, 7] = 1
    M[:, 8] = 1
    M[:, 9] = rng.uniform(0, 1, n)
    M[:, 13] = 0
    M[:, 14] = np.arange(n)
    M[:, 15] = 1000 + rng.permutation(n)
    return M, rng


def advance(B, rng, dt, k):
    M = B.M
    alive = M[:, 8] == 1
    M[alive, 4] += M[alive, 5] * dt * (1 + 0.5 * np.sin(k + M[alive, 0]))
    M[alive, 9] += 2.5e9 * dt
    M[alive, 0] += 1e-5 * np.cos(M[alive, 4])
    die = alive & (rng.random(M.shape[0]) < 0.002)
    M[die, 8] = 0


def make_dumps(folder, gids):
    from pyffa_amd import FFAG_dump as FD
    d = FD.Dumps()
    d.add_dump(FD.StepDump(0.5e-9, 10e-9, 1e-9, gids[:6], save_folder=os.path.join(folder, "step")))
    d.add_dump(FD.StepDumpBunch(2e-9, 8e-9, 2e-9, save_folder=os.path.join(folder, "stepbunch")))
    d.add_dump(FD.PositionDump(30.0, 0.0, 13e-9, gids[::3], save_folder=os.path.join(folder, "pos"),
                               max_records=5))
    d.add_dump(FD.PositionDumpBunch(200.0, 1e-9, 12e-9, save_folder=os.path.join(folder, "posbunch")))
    return d


def run(folder, every_step, n=400, steps=320, dt=0.05e-9):
    M, _ = synthetic_bunch(n, 5)
    rng = np.random.default_rng(11)
    B = HostBunch(M)
    d = make_dumps(folder, M[:, 15].astype(np.int64))
    t = 0.0
    examined = 0
    for k in range(steps):
        tn = t + dt
        act = every_step or d.needs_step(t, tn)
        if act:
            B.UpdatePreSteps()
            examined += 1
        advance(B, rng, dt, k)
        if act:
            d.check_and_dump(t, tn, B, data_type="Boris")
        t = tn
    return examined


def load_all(folder):
    out = {}
    for root, _, files in os.walk(folder):
        for f in files:
            with np.load(os.path.join(root, f)) as z:
                out[os.path.relpath(os.path.join(root, f), folder)] = z["particles"]
    return out


def test_skipping_the_steps_no_detector_needs_changes_nothing(tmp_path):
    n_all = run(str(tmp_path / "a"), True)
    n_some = run(str(tmp_path / "b"), False)
    assert n_some < n_all
    a, b = load_all(str(tmp_path / "a")), load_all(str(tmp_path / "b"))
    assert sorted(a) == sorted(b)
    assert len(a) > 10
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_detector_outputs(tmp_path):
    """File names and contents against a direct computation."""
    from pyffa_amd.FFAG_ParasAndConversion import FFAG_ConversionTools
    conv = FFAG_ConversionTools()
    n, steps, dt = 400, 320, 0.05e-9
    M, _ = synthetic_bunch(n, 5)
    rng = np.random.default_rng(11)
    B = HostBunch(M)
    gids = M[:, 15].astype(np.int64)
    d = make_dumps(str(tmp_path), gids)
    target = np.deg2rad(200.0) % TWO_PI
    want_pb, want_step, times_sb = [], {}, []
    t = 0.0
    for k in range(steps):
        tn = t + dt
        B.UpdatePreSteps()
        pre = B.M.copy()
        advance(B, rng, dt, k)
        d.check_and_dump(t, tn, B, data_type="Boris")
        if 1e-9 < tn < 12e-9:
            hit = np.flatnonzero(brute_force_cross(pre[:, 4], B.M[:, 4], target) & (pre[:, 8] == 1))
            if hit.size:
                c = conv.ConvertVrzek2Przek_boris(B.M[hit])
                c[:, 9] = np.mod(c[:, 9], TWO_PI)
                want_pb.append(np.hstack((c, np.full((hit.size, 1), tn))))
        if 0.5e-9 < t < 10e-9 and 0.5e-9 < tn < 10e-9 and \
                np.floor((t - 0.5e-9) / 1e-9 + 1e-12) != np.floor((tn - 0.5e-9) / 1e-9 + 1e-12):
            for row in range(6):
                if pre[row, 8] == 1:
                    c = conv.ConvertVrzek2Przek_boris(B.M[row:row + 1])
                    want_step.setdefault(int(gids[row]), []).append(np.hstack((c, [[tn]]))[0])
        if 2e-9 < tn < 8e-9 and (t - 2e-9) // 2e-9 != (tn - 2e-9) // 2e-9:
            times_sb.append(tn)
        t = tn
    got = load_all(str(tmp_path))
    pb = got.pop("posbunch/crossing_angle_200.0_start_1.000_end_12.000_rank_0.npz")
    assert pb.shape[1] == 17
    assert np.array_equal(pb, np.vstack(want_pb))
    for gid, rows in want_step.items():
        assert np.array_equal(got.pop(f"step/particle_{gid}_rank_0.npz"), np.vstack(rows))
    assert len(want_step) >= 4        # a tracked particle may be lost before the first record
    assert sorted(k for k in got if k.startswith("stepbunch")) == \
        sorted(f"stepbunch/time_{tt * 1e9:.2f}_rank_0.npz" for tt in times_sb)
    for tt in times_sb:
        a = got.pop(f"stepbunch/time_{tt * 1e9:.2f}_rank_0.npz")
        assert np.all(a[:, 16] == tt) and np.all(a[:, 8] == 1)
    pos = [k for k in got if k.startswith("pos/")]
    assert pos and all(k.endswith("_angle_30.0_rank_0.npz") for k in pos)
    for k in pos:
        a = got.pop(k)
        assert 1 <= a.shape[0] <= 5 and a.shape[1] == 17
        assert np.all(a[:, 9] >= 0) and np.all(a[:, 9] < TWO_PI)
    assert not got


def test_non_empty_folder_is_refused(tmp_path):
    from pyffa_amd import FFAG_dump as FD
    (tmp_path / "x").write_text("")
    with pytest.raises(FileExistsError):
        FD.PositionDumpBunch(10.0, 0.0, 1.0, save_folder=str(tmp_path))
