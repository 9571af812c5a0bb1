"""Detectors of the tracking loop (src/FFAG_dump.py).

Same classes, constructor signatures, trigger rules and output files as the
reference:

* ``StepDump``: tracked particles (by Global_ID) each time the step crosses a
  point of the time grid start + k*interval; one ``particle_{gid}_rank_0.npz``
  per particle (src/FFAG_dump.py:143-299).
* ``StepDumpBunch``: all surviving particles at each grid crossing, appended
  to ``time_{t[ns]:.2f}_rank_0.npz`` (src/FFAG_dump.py:302-395).
* ``PositionDump``: tracked particles each time they cross an azimuth, one
  ``particle_{gid}_angle_{deg:.1f}_rank_0.npz`` per particle
  (src/FFAG_dump.py:534-786).
* ``PositionDumpBunch``: every particle crossing an azimuth, one
  ``crossing_angle_{deg:.1f}_start_{ns:.3f}_end_{ns:.3f}_rank_0.npz``
  (src/FFAG_dump.py:789-1131).
* ``Dumps``: the collection the tracker calls (src/FFAG_dump.py:1255-1284).

Every file holds one array ``particles`` of 17 columns: the row converted by
``ConvertVrzek2Przek_boris`` (Boris) or ``ConvertVrzek2Przek`` (RK4), then
the time in seconds.  The position detectors reduce column 9 (RF phase) mod
2 pi.  The rows are the state after the step (no interpolation to the grid
point or to the azimuth).

What runs where: the crossing test of the position detectors runs on the
GPU over the pre/post-step azimuths (``FFAG_Bunch.detect_crossing``) and only
the rows that are written leave the device (``FFAG_Bunch.rows``).  Each
detector also answers ``needs_step(t_pre, t_post)``: whether its
``check_and_dump`` would do anything for that step, so the tracker issues
multi-step launches between detector events and one-step launches (with the
pre-step state recorded) only for the steps a detector examines.

This build runs one process per GPU: the reference's MPI rank is 0, and
where the reference aborts every rank on a non-empty output folder
(``comm.Abort(1)``) this raises ``FileExistsError``.
"""
from __future__ import annotations

import os

import numpy as np

from .FFAG_ParasAndConversion import FFAG_ConversionTools

_RANK = 0


def _prepare_folder(save_folder):
    """Create the output folder, or refuse one that is not empty
    (src/FFAG_dump.py:175-187)."""
    if not os.path.exists(save_folder):
        os.makedirs(save_folder)
    elif os.listdir(save_folder):
        raise FileExistsError(f"data folder {save_folder} exists and is not empty")


def _convert(states, data_type):
    conv = FFAG_ConversionTools()
    if data_type == "RK4":
        return conv.ConvertVrzek2Przek(states)
    if data_type == "Boris":
        return conv.ConvertVrzek2Przek_boris(states)
    raise ValueError("data_type must be 'RK4' or 'Boris'")


def _local_rows(bunch_obj, tracked_global_ids):
    """Rows of the tracked Global_IDs, in the order of the tracked list;
    ids absent from the bunch are skipped and a repeated Global_ID maps to
    its last row (the gid -> row dict of src/FFAG_dump.py:201-209)."""
    g = bunch_obj.global_ids()
    order = np.argsort(g, kind="stable")
    gs = g[order]
    tracked = np.asarray(tracked_global_ids, dtype=np.int64)
    pos = np.searchsorted(gs, tracked, side="right") - 1
    ok = (pos >= 0) & (gs[np.clip(pos, 0, None)] == tracked)
    return order[pos[ok]].astype(np.int64)


def append_to_npz(filepath, new_data):
    """Append rows to the ``particles`` array of an .npz file
    (src/FFAG_dump.py:1428-1448)."""
    if os.path.exists(filepath):
        with np.load(filepath) as f:
            combined = np.vstack((f["particles"], new_data))
    else:
        combined = new_data
    np.savez_compressed(filepath, particles=combined)


class StepDump:
    def __init__(self, start_time, end_time, time_interval, tracked_particle_id_global, save_folder="dump_data"):
        self.start_time = float(start_time)
        self.end_time = float(end_time)
        self.time_interval = float(time_interval)
        assert self.time_interval > 0.0, "time_interval must be > 0"
        self.save_folder = save_folder
        self.rank = _RANK
        _prepare_folder(save_folder)
        self.tracked_global_ids = np.asarray(tracked_particle_id_global, dtype=np.int64)
        assert self.tracked_global_ids.ndim == 1, "tracked_particle_id_global must be 1-D"
        self.save_flag = False
        self.local_tracked_ids = np.zeros((0,), dtype=np.int64)
        self.particle_buffer_dict: dict[int, list] = {}

    def determine_local_ids(self, bunch_obj):
        self.local_tracked_ids = _local_rows(bunch_obj, self.tracked_global_ids)
        return self.local_tracked_ids

    def _grid_crossed(self, time_Pre, time_Post):
        # floor with a 1e-12 margin against rounding at the grid points
        # (src/FFAG_dump.py:266-273)
        eps = 1e-12
        pre_idx = int(np.floor((time_Pre - self.start_time) / self.time_interval + eps))
        post_idx = int(np.floor((time_Post - self.start_time) / self.time_interval + eps))
        return pre_idx != post_idx

    def needs_step(self, time_Pre, time_Post):
        if self.save_flag:
            return False
        if not (self.start_time < time_Pre < self.end_time):
            return time_Pre > self.end_time
        return self.start_time < time_Post < self.end_time and self._grid_crossed(time_Pre, time_Post)

    def save_all_particles(self, bunch_obj, data_type, time_stamp):
        """Record the tracked particles that were alive before the step
        (src/FFAG_dump.py:211-240)."""
        if self.local_tracked_ids.size == 0:
            return
        alive = bunch_obj.Step_SurviveFlag
        rows = np.array([i for i in self.local_tracked_ids if alive[i]], dtype=np.int64)
        if rows.size == 0:
            return
        states = bunch_obj.rows(rows)
        converted = _convert(states, data_type)
        tcol = np.full((rows.size, 1), time_stamp, dtype=np.float64)
        full = np.hstack((converted, tcol)).astype(np.float64)
        for k in range(rows.size):
            gid = int(states[k, -1])
            self.particle_buffer_dict.setdefault(gid, []).append(full[k].copy())

    def check_and_dump(self, time_Pre, time_Post, bunch_obj, data_type):
        """src/FFAG_dump.py:242-279: lazily map the ids; past end_time write
        the files once; inside (start, end) record the step's end state when
        the step crosses a grid point (time written is time_Post)."""
        if self.save_flag:
            return
        if self.local_tracked_ids.size == 0:
            self.determine_local_ids(bunch_obj)
            if self.local_tracked_ids.size == 0:
                return
        if not (self.start_time < time_Pre < self.end_time):
            if time_Pre > self.end_time and not self.save_flag:
                self.dump_all_particles()
                self.save_flag = True
            return
        if self.start_time < time_Post < self.end_time and self._grid_crossed(time_Pre, time_Post):
            self.save_all_particles(bunch_obj, data_type, time_Post)

    def dump_all_particles(self):
        total = 0
        for gid, rows in self.particle_buffer_dict.items():
            if not rows:
                continue
            data = np.vstack(rows)
            path = os.path.join(self.save_folder, f"particle_{gid}_rank_{self.rank}.npz")
            np.savez_compressed(path, particles=data)
            total += data.shape[0]
        return total


class StepDumpBunch:
    def __init__(self, start_time, end_time, time_interval, filter_survive_flag=True, save_folder="dump_data"):
        self.start_time = start_time
        self.end_time = end_time
        self.time_interval = time_interval
        self.save_folder = save_folder
        self.filter_survive_flag = filter_survive_flag
        self.rank = _RANK
        _prepare_folder(save_folder)

    def _grid_crossed(self, time_Pre, time_Post):
        # floor division without a margin (src/FFAG_dump.py:346-349)
        return (time_Pre - self.start_time) // self.time_interval != \
            (time_Post - self.start_time) // self.time_interval

    def needs_step(self, time_Pre, time_Post):
        return self.start_time < time_Post < self.end_time and self._grid_crossed(time_Pre, time_Post)

    def check_and_dump(self, time_Pre, time_Post, bunch_obj, data_type):
        if self.needs_step(time_Pre, time_Post):
            self.dump_bunch(bunch_obj, time_Post, data_type)

    def dump_bunch(self, bunch_obj, time_Post, data_type):
        """Append the (surviving) particles after the step, in input order
        (src/FFAG_dump.py:358-395)."""
        M = bunch_obj.LocalBunch
        if self.filter_survive_flag:
            M = M[M[:, bunch_obj.BunchAttribute.Attribute["Survive"]] == 1, :]
        if M.shape[0] == 0:
            return
        path = os.path.join(self.save_folder, f"time_{time_Post * 1e9:.2f}_rank_{self.rank}.npz")
        converted = _convert(M, data_type)
        append_to_npz(path, np.hstack((converted, np.full((converted.shape[0], 1), time_Post))))


class PositionDump:
    def __init__(self, target_azimuth_angle_deg, start_time, end_time, tracked_particle_id_global,
                 save_folder="dump_data", max_records=10000):
        self.target_azimuth_angle = np.deg2rad(target_azimuth_angle_deg) % (2 * np.pi)
        self.start_time = start_time
        self.end_time = end_time
        self.tracked_global_ids = np.asarray(tracked_particle_id_global, dtype=np.int64)
        assert self.tracked_global_ids.ndim == 1, "tracked_particle_id_global must be 1-D"
        self.save_folder = save_folder
        self.max_records = max_records
        self.save_flag = False
        self.rank = _RANK
        _prepare_folder(save_folder)
        self.particle_buffer_dict = {}
        self.cursor_dict = {}
        self.write_done_dict = {}
        self.local_tracked_ids = np.zeros((0,), dtype=np.int64)

    def determine_local_ids(self, bunch_obj):
        self.local_tracked_ids = _local_rows(bunch_obj, self.tracked_global_ids)
        return self.local_tracked_ids

    def needs_step(self, step_i, time_i):
        if self.save_flag:
            return False
        if not (self.start_time < time_i < self.end_time):
            return time_i > self.end_time
        return True

    def check_azimuth_crossing(self, bunch_obj):
        """Crossing mask over the tracked rows (detect_azimuth_crossing_portion,
        src/FFAG_MathTools.py:1697-1757), evaluated on the GPU."""
        return bunch_obj.detect_crossing(self.target_azimuth_angle, self.local_tracked_ids)

    def _init_gid_buffer_if_needed(self, gid, dim):
        if gid not in self.particle_buffer_dict:
            self.particle_buffer_dict[gid] = np.empty((self.max_records, dim), dtype=np.float64)
            self.cursor_dict[gid] = 0
            self.write_done_dict[gid] = False

    def check_and_dump(self, step_i, time_i, bunch_obj, data_type):
        """src/FFAG_dump.py:677-770 (``step_i`` receives the pre-step time
        from ``Dumps``, as in the reference, and is unused)."""
        if self.save_flag:
            return
        if not (self.start_time < time_i < self.end_time):
            if time_i > self.end_time and not self.save_flag:
                self.save_all_particles()
                self.save_flag = True
            return
        if self.local_tracked_ids.size < 1:
            self.determine_local_ids(bunch_obj)
            if self.local_tracked_ids.size < 1:
                return
        hits_mask = self.check_azimuth_crossing(bunch_obj)
        if not np.any(hits_mask):
            return
        hits = self.local_tracked_ids[hits_mask]
        states = bunch_obj.rows(hits)
        converted = _convert(states, data_type)
        converted[:, 9] = np.mod(converted[:, 9], 2 * np.pi)
        full = np.concatenate((converted, np.full((converted.shape[0], 1), time_i, dtype=np.float64)), axis=1)
        gids = states[:, -1].astype(np.int64)
        order = np.argsort(gids, kind="stable")
        gids_sorted, full_sorted = gids[order], full[order]
        starts = np.r_[0, np.flatnonzero(gids_sorted[1:] != gids_sorted[:-1]) + 1]
        ends = np.r_[starts[1:], gids_sorted.size]
        for s, e in zip(starts, ends):
            gid = int(gids_sorted[s])
            block = full_sorted[s:e, :]
            self._init_gid_buffer_if_needed(gid, block.shape[1])
            cursor = self.cursor_dict[gid]
            room = self.max_records - cursor
            if room <= 0:
                self.write_done_dict[gid] = True
                continue
            n = min(block.shape[0], room)
            self.particle_buffer_dict[gid][cursor:cursor + n, :] = block[:n, :]
            self.cursor_dict[gid] = cursor + n
            if self.cursor_dict[gid] >= self.max_records:
                self.write_done_dict[gid] = True

    def save_all_particles(self):
        angle_deg = np.rad2deg(self.target_azimuth_angle)
        for gid, buffer in self.particle_buffer_dict.items():
            count = self.cursor_dict.get(gid, 0)
            if count == 0:
                continue
            path = os.path.join(self.save_folder, f"particle_{gid}_angle_{angle_deg:.1f}_rank_{self.rank}.npz")
            np.savez(path, particles=buffer[:count, :])


class PositionDumpBunch:
    def __init__(self, target_azimuth_angle_deg, start_time, end_time, save_folder="dump_data"):
        self.target_azimuth_angle = np.deg2rad(target_azimuth_angle_deg) % (2 * np.pi)
        self.save_folder = save_folder
        self.start_time = start_time
        self.end_time = end_time
        self.save_flag = False
        self.saved_data = np.empty((50000, 17), dtype=np.float64)
        self.write_cursor = 0
        self.rank = _RANK
        _prepare_folder(save_folder)

    def needs_step(self, step_i, time_i):
        return (self.start_time < time_i < self.end_time) or (time_i > self.end_time and self.save_flag is False)

    def check_azimuth_crossing(self, bunch_obj):
        """Crossing particles of this step (detect_azimuth_crossing,
        src/FFAG_MathTools.py:1633-1695) on the GPU: their input rows in
        increasing order, as the reference's boolean mask selects them, and
        their device positions."""
        return bunch_obj.detect_crossing(self.target_azimuth_angle)

    def add_particles(self, particle_data, time_i):
        n_new = particle_data.shape[0]
        rows = np.hstack((particle_data, np.full((n_new, 1), time_i, dtype=np.float64)))
        need = self.write_cursor + n_new
        if need > self.saved_data.shape[0]:
            grown = np.empty((max(need, self.saved_data.shape[0] * 2), 17), dtype=np.float64)
            grown[:self.write_cursor] = self.saved_data[:self.write_cursor]
            self.saved_data = grown
        self.saved_data[self.write_cursor:need] = rows
        self.write_cursor = need

    def save_all_particles(self, data_type):
        """One file with every recorded crossing (src/FFAG_dump.py:1021-1049);
        written even when no particle crossed, as in the reference."""
        data = self.saved_data[:self.write_cursor, :]
        name = (f"crossing_angle_{np.rad2deg(self.target_azimuth_angle):.1f}_start_{self.start_time * 1e9:.3f}"
                f"_end_{self.end_time * 1e9:.3f}_rank_{self.rank}.npz")
        converted = _convert(data[:, :-1], data_type)
        converted[:, 9] = np.mod(converted[:, 9], 2 * np.pi)
        np.savez(os.path.join(self.save_folder, name), particles=np.hstack((converted, data[:, [-1]])))

    def check_and_dump(self, step_i, time_i, bunch_obj, data_type):
        """src/FFAG_dump.py:1098-1131."""
        if self.start_time < time_i < self.end_time:
            rows, pos = self.check_azimuth_crossing(bunch_obj)
        elif time_i > self.end_time and self.save_flag is False:
            self.save_all_particles(data_type)
            self.save_flag = True
            return
        else:
            return
        if rows.size == 0:
            return
        self.add_particles(bunch_obj._assemble(pos), time_i)


class Dumps:
    def __init__(self):
        self.dumps = []

    def add_dump(self, dump_obj):
        self.dumps.append(dump_obj)

    def needs_step(self, time_Pre, time_Post):
        """Whether any detector acts on the step (time_Pre, time_Post); a
        detector without ``needs_step`` is asked on every step."""
        for d in self.dumps:
            f = getattr(d, "needs_step", None)
            if f is None or f(time_Pre, time_Post):
                return True
        return False

    def check_and_dump(self, time_Pre, time_Post, bunch_obj, data_type="RK4"):
        for d in self.dumps:
            d.check_and_dump(time_Pre, time_Post, bunch_obj, data_type)

    def list_dumps(self):
        for i, d in enumerate(self.dumps):
            print(f"Dump {i + 1}: {d.__class__.__name__}")
