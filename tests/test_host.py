"""Host-side logic of the product, CPU only: the C ABI library exports every
entry point include/gaffa_hip.h declares, gaffa's plan/delay arithmetic in
the C++ host layer equals the oracle's, and the Python surface raises gaffa's
errors before touching a device."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gaffa_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|uint32_t|uint64_t|const char\*)\s+(gaffa_hip_\w+)\s*\(", text, re.M)
    assert len(names) > 20
    return names


def test_c_abi_exports_every_declared_symbol():
    import gaffa_amd

    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []
    lib.gaffa_hip_abi_version.restype = ctypes.c_int
    assert lib.gaffa_hip_abi_version() == 3


def test_c_abi_rejects_bad_arguments_without_a_device():
    import gaffa_amd

    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    lib.gaffa_hip_last_error.restype = ctypes.c_char_p
    assert lib.gaffa_hip_dedisperse_multi_u8(None, None) == 1
    assert b"must not be null" in lib.gaffa_hip_last_error()
    assert lib.gaffa_hip_ffa_search_batch(None, None) == 1
    assert lib.gaffa_hip_detrend_normalise_batch(None, None) == 1
    lib.gaffa_hip_preprocess_workspace_bytes.restype = ctypes.c_uint64
    assert lib.gaffa_hip_preprocess_workspace_bytes(4, 1000, 10) > 0


def test_product_plans_equal_oracle(core, port):
    g = np.load(os.path.join(GOLDEN, "plans.npz"))
    for name in ("small", "lband", "large", "riptide_default"):
        n, ts, pmin, pmax, duty, spacing = g[f"{name}_args"]
        plan = core._make_riptide_ffa_plan(nsamples=int(n), tsamp=ts, period_min=pmin,
                                           period_max=pmax, duty_cycle_max=duty,
                                           width_trial_spacing=spacing)
        got = plan.tasks
        want = g[f"{name}_tasks"]
        for field in want.dtype.names:
            np.testing.assert_array_equal(got[field], want[field])
        assert list(plan.width_trials) == list(g[f"{name}_widths"])


def test_product_delays_equal_oracle(core, port):
    from gaffa_amd.pipeline import channel_frequencies

    f = channel_frequencies(1024, 1520.0, -0.293)
    for dm in (0.0, 0.1643, 30.0, 348.6, 498.8):
        for c in (0, 1, 511, 1023):
            assert core.dedispersion_delay_bins(dm, f[c], f.max(), 64e-6) == \
                port.delay_bins(dm, f[c], f.max(), 64e-6)


def test_python_surface_errors(core):
    """gaffa's own Python assertions (tests/test_ffa.py:45-80), restated: the
    same exception types and message substrings for the same calls, with
    gaffa's default backend="cpu", which gaffa_amd refuses after the input
    checks."""
    import inspect

    from gaffa_amd import ffa

    plan = ffa.make_riptide_plan(nsamples=8, tsamp=1.0, period_min=2.0, period_max=4.0,
                                 bins_min=2, bins_max=2)
    assert plan.task_count > 0 and plan.width_trials == (1,)
    with pytest.raises(TypeError, match="numpy.ndarray"):
        ffa.ffa_search([0.0] * 8, plan)
    with pytest.raises(TypeError, match="dtype float32"):
        ffa.ffa_search(np.zeros(8, dtype=np.float64), plan)
    with pytest.raises(ValueError, match="1D"):
        ffa.ffa_search(np.zeros((1, 8), dtype=np.float32), plan)
    with pytest.raises(ValueError, match="C-contiguous"):
        ffa.ffa_search(np.zeros(16, dtype=np.float32)[::2], plan)
    plan9 = ffa.make_riptide_plan(nsamples=9, tsamp=1.0, period_min=2.0, period_max=4.0,
                                  bins_min=2, bins_max=2)
    with pytest.raises(ValueError, match="input_nsamples"):
        ffa.ffa_search(np.zeros(8, np.float32), plan9)
    with pytest.raises(ValueError, match="positive or None"):
        ffa.ffa_search(np.zeros(8, np.float32), plan, max_peaks=0)
    with pytest.raises(ValueError, match="'cpu' or 'cuda'"):
        ffa.ffa_search(np.zeros(8, np.float32), plan, backend="opencl")
    with pytest.raises(ValueError, match="only valid"):
        ffa.ffa_search(np.zeros(8, np.float32), plan, device_id=1)
    # gaffa's default backend is its CPU path, which is not part of gaffa_amd
    with pytest.raises(ValueError, match="backend='cuda'"):
        ffa.ffa_search(np.zeros(8, np.float32), plan)
    with pytest.raises(ValueError, match="non-negative"):
        ffa.ffa_search(np.zeros(8, np.float32), plan, backend="cuda", device_id=-1)
    assert inspect.signature(ffa.ffa_search).parameters["backend"].default == "cpu"


def test_dedispersion_surface_errors(core):
    """gaffa's dedispersion backend names, defaults and messages
    (dedispersion_bindings.cpp:56-64, 648-672; tests/test_dedispersion.py:109-113)."""
    from gaffa_amd import dedispersion
    from gaffa_amd.io import Filterbank

    fb = Filterbank(os.path.join(GOLDEN, "basetest.fil"))
    with pytest.raises(ValueError, match="unknown dedispersion backend"):
        dedispersion.dedisperse_single_dm(fb, dm=0.0, backend="bad")
    calls = (
        lambda **k: dedispersion.dedisperse_multi_dm(fb, dm_low=0.0, dm_step=1.0, ndm=2, **k),
        lambda **k: dedispersion.dedisperse_single_dm(fb, dm=0.0, **k),
        lambda **k: dedispersion.dedisperse_spectrum(fb, dm=0.0, **k),
    )
    for call in calls:
        # gaffa's default for these three is backend="cpu"
        with pytest.raises(ValueError, match="backend 'cpu' is gaffa's CPU path"):
            call()
        with pytest.raises(ValueError, match="backend 'cpu'"):
            call(backend="cpu")
    doc = dedispersion.dedisperse_subband.__doc__
    assert "backend: str = 'cuda'" in doc and "ndm_per_nominal: " in doc


def test_filterbank_from_array_roundtrip():
    from gaffa_amd.io import Filterbank
    from gaffa_amd.pipeline import channel_frequencies

    x = np.arange(4 * 3, dtype=np.uint8).reshape(4, 3)
    fb = Filterbank.from_array(x, tsamp=1e-3, frequency_mhz=channel_frequencies(3, 1500.0, -4.0))
    assert fb.data.shape == (4, 1, 3) and fb.header.nbits == 8
    assert fb.header.frequency_table == [1492.0, 1496.0, 1500.0]


def test_workload_shapes():
    from gaffa_amd.pipeline import lband_workload, small_workload

    s = small_workload()
    assert (s.nchans, s.ndm, s.window_samples()) == (64, 64, 3907)
    w = lband_workload()
    assert (w.nchans, w.ndm, w.window_samples()) == (1024, 1024, 31251)
    dms = w.dms
    assert dms[0] == 0.0 and 495.0 < dms[-1] <= 500.0
    assert np.all(np.diff(dms) > 0)
    steps = [s.dm_step for s in w.dm_segments]
    assert steps == sorted(steps) and all(s.ndm % 32 == 0 for s in w.dm_segments)


@pytest.mark.parametrize("ranked,parts", [(True, 1), (False, 1), (True, 7), (False, 3)])
def test_dm_peak_key_sort_matches_comparator(core, ranked, parts):
    """expand_sorted_dm_peaks sorts (key, index) pairs and fixes equal-key runs
    with the full comparator: the order must equal a plain sort by
    is_better_dm_peak (finite first, S/N desc, period, frequency, width, phase,
    shift, bins asc, then dm_index, dm), ties and +-0 included, with and
    without the period ranks in the key.  With parts > 1 the keys are sorted
    and the peaks expanded part by part and then merged, as the search
    pipeline does per FFA tile."""
    plan = core._make_riptide_ffa_plan(nsamples=130870, tsamp=256e-6, period_min=0.2,
                                       period_max=2.0, duty_cycle_max=0.09,
                                       width_trial_spacing=2.0)
    tasks = plan.tasks
    nw = len(plan.width_trials)
    rng = np.random.default_rng(11)
    n = 20000
    task = rng.integers(0, len(tasks), n).astype(np.uint32)
    shift = (rng.random(n) * tasks["rows_eval"][task]).astype(np.uint32)
    phase = (rng.random(n) * tasks["bins"][task]).astype(np.uint32)
    widx = rng.integers(0, nw, n).astype(np.uint32)
    series = rng.integers(0, 64, n).astype(np.uint32)
    # few distinct S/N values -> long equal-S/N runs; both signed zeros
    snr = rng.choice(np.array([7.5, 7.25, 6.0, 0.0, -0.0, 12.125], np.float32), n)
    dms = list(np.arange(64) * 0.5)
    got = core._expand_sorted_dm_peaks(plan, series, task, shift, phase, widx, snr, dms,
                                       ranked=ranked, parts=parts)
    assert len(got) == n

    def key(p):
        q = p.peak
        finite = np.isfinite(q.snr) and np.isfinite(q.period) and np.isfinite(q.frequency)
        return (not finite, -q.snr, q.period, q.frequency, q.width, q.phase, q.shift, q.bins,
                p.dm_index, p.dm)

    want = sorted(got, key=key)
    assert [key(p) for p in got] == [key(p) for p in want]


def test_native_filterbank_reader_and_writer(core, tmp_path):
    """The header-only read agrees with the full read on basetest.fil, and a
    written file reads back (descending files come back ascending)."""
    from gaffa_amd.io import Filterbank, write_filterbank

    path = os.path.join(GOLDEN, "basetest.fil")
    h, _ = core._read_filterbank(path, header_only=True)
    fb = Filterbank(path)
    assert (h["nsamples"], h["nchans"], h["nbits"], h["tsamp"]) == \
        (fb.header.nsamples, fb.header.nchans, fb.header.nbits, fb.header.tsamp)
    assert h["frequency_table"] == list(fb.header.frequency_table)
    assert h["header_size"] == fb.header.header_size
    rng = np.random.default_rng(0)
    x = rng.integers(0, 256, size=(1000, 16), dtype=np.uint8)
    for foff in (-2.0, 2.0):
        out = tmp_path / f"w{foff}.fil"
        write_filterbank(out, x, tsamp=1e-3, fch1=1400.0, foff=foff, source_name="synthetic")
        hw, _ = core._read_filterbank(str(out), header_only=True)
        back = Filterbank(str(out))
        assert hw["reversed"] == (foff < 0) and hw["nsamples"] == 1000
        assert hw["frequency_table"] == list(back.header.frequency_table)
        want = x[:, ::-1] if foff < 0 else x
        np.testing.assert_array_equal(back.data[:, 0, :], want)


def _sigproc_string(s: str) -> bytes:
    import struct

    return struct.pack("<i", len(s)) + s.encode()


def _write_raw_header(path, fields, payload: bytes):
    """A SIGPROC file from (key, kind, value) fields, kind in i/d/s/- (a bare
    keyword such as FREQUENCY_START)."""
    import struct

    out = _sigproc_string("HEADER_START")
    for key, kind, value in fields:
        out += _sigproc_string(key)
        if kind == "i":
            out += struct.pack("<i", value)
        elif kind == "d":
            out += struct.pack("<d", value)
        elif kind == "s":
            out += _sigproc_string(value)
    out += _sigproc_string("HEADER_END") + payload
    with open(path, "wb") as fh:
        fh.write(out)


def test_native_filterbank_reader_equals_gaffa(core, ref, tmp_path):
    """The native reader against gaffa's own read_filterbank (io/filterbank.cpp,
    compiled from the reference sources): every header field, the frequency
    table and the sample bytes, in both channel orders, for basetest.fil, for
    written ascending / descending 8-bit files, and for 16- and 32-bit files
    with a descending FREQUENCY table; and the same errors for a partial
    trailing row and an fchannel outside a FREQUENCY block."""
    from gaffa_amd.io import Filterbank, write_filterbank

    rng = np.random.default_rng(0)
    x = rng.integers(0, 256, size=(1000, 16), dtype=np.uint8)
    paths = [os.path.join(GOLDEN, "basetest.fil")]
    for foff in (-2.0, 2.0):
        out = tmp_path / f"w{foff}.fil"
        write_filterbank(out, x, tsamp=1e-3, fch1=1400.0, foff=foff, source_name="synthetic")
        paths.append(str(out))
    table = [1500.0 - 3.0 * c for c in range(12)]
    for nbits, dtype in ((16, np.uint16), (32, np.float32)):
        y = rng.integers(0, 4000, size=(300, 2, 12)).astype(dtype)
        fields = [("rawdatafile", "s", "raw.dat"), ("source_name", "s", "J0000+0000"),
                  ("telescope_id", "i", 6), ("machine_id", "i", 10), ("data_type", "i", 1),
                  ("nbits", "i", nbits), ("nifs", "i", 2), ("tsamp", "d", 6.4e-5),
                  ("tstart", "d", 60000.5), ("refdm", "d", 3.5), ("FREQUENCY_START", "-", 0)]
        fields += [("fchannel", "d", f) for f in table] + [("FREQUENCY_END", "-", 0),
                                                          ("nsamples", "i", 999)]
        out = tmp_path / f"table{nbits}.fil"
        _write_raw_header(out, fields, y.tobytes())
        paths.append(str(out))
    for path in paths:
        for order in ("frequency_ascending", "preserve_file_order"):
            want_h, want_f, want_d = ref.read_filterbank(path, order == "preserve_file_order")
            fb = Filterbank(path, channel_order=order)
            got = fb.header
            for key, value in want_h.items():
                assert getattr(got, key) == value, (path, order, key)
            assert list(got.frequency_table) == list(want_f), (path, order)
            np.testing.assert_array_equal(np.frombuffer(fb.data.tobytes(), np.uint8), want_d)
            assert fb.data.shape == (want_h["nsamples"], want_h["nifs"], want_h["nchans"])
            # the cpu_scalar and cpu_openmp reversals give the same bytes
            for backend in ("cpu_scalar", "cpu_openmp"):
                alt = Filterbank(path, channel_order=order, reverse_backend=backend,
                                 io_buffer_bytes=4096)
                np.testing.assert_array_equal(alt.data, fb.data)

    partial = tmp_path / "partial.fil"
    _write_raw_header(partial, [("nbits", "i", 8), ("nchans", "i", 4), ("tsamp", "d", 1e-3),
                                ("fch1", "d", 1400.0), ("foff", "d", -1.0)], bytes(4 * 10 + 3))
    stray = tmp_path / "stray.fil"
    _write_raw_header(stray, [("nbits", "i", 8), ("fchannel", "d", 1400.0)], bytes(8))
    for bad, msg in ((partial, "not a whole number of rows"),
                     (stray, "fchannel outside FREQUENCY block")):
        with pytest.raises(RuntimeError, match=msg):
            ref.read_filterbank(bad)
        with pytest.raises(RuntimeError, match=msg):
            Filterbank(bad)


def test_filterbank_options_and_messages(core):
    """gaffa's Filterbank keyword names and messages
    (tests/test_filterbank_io.py:12-60)."""
    from gaffa_amd.io import Filterbank

    path = os.path.join(GOLDEN, "basetest.fil")
    fb = Filterbank(path)
    assert fb.header.nbits == 8 and fb.data.dtype == np.uint8
    assert fb.shape == fb.data.shape == (fb.header.nsamples, fb.header.nifs, fb.header.nchans)
    assert fb.dtype == fb.data.dtype and fb.nbytes == fb.data.nbytes
    assert Filterbank(path, channel_order="preserve_file_order").header.foff < 0
    with pytest.raises(ValueError, match="unknown channel_order"):
        Filterbank(path, channel_order="bad")
    with pytest.raises(ValueError, match="unknown reverse_backend"):
        Filterbank(path, reverse_backend="bad")

    def load_data():
        return Filterbank(path).data

    data = load_data()
    assert data.shape[0] > 0 and data[0, 0, 0] >= 0


def test_large_workload_shape(core):
    """The 4096-channel configuration: 4096 trials 0-1000 in 32 smearing
    segments, 64-channel subbands, 5 RFI channels, 32 pulsars."""
    from gaffa_amd.pipeline import large_workload

    w = large_workload()
    assert (w.nchans, w.ndm, w.subband_channels, len(w.pulsars), len(w.rfi_channels)) == \
        (4096, 4096, 64, 32, 5)
    assert w.dms[0] == 0.0 and 990.0 < w.dms[-1] <= 1000.0 and np.all(np.diff(w.dms) > 0)
    f = w.frequencies
    dmax = core.dedispersion_delay_bins(float(w.dms[-1]), float(f[0]), float(f.max()), w.tsamp)
    assert 57000 < dmax < 57400  # n_out ~ 4.137 M of 2^22


class _Merge(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint32) for name in
                ("head", "tail", "wrap", "out", "head2", "wrap2", "tail2", "wrap3")]


@pytest.mark.parametrize("rows,bins", [(2, 7), (3, 5), (5, 3), (7, 9), (16, 4), (37, 11),
                                       (84, 180), (113, 181), (512, 2)])
def test_subtree_merge_tables_replay_the_oracle_transform(port, rows, bins):
    """The host-built subtree merge tables (gaffa_hip_ffa_subtree_merges),
    replayed pass by pass in float32 exactly as the subtree kernel applies
    them (two levels per pass as (A + B) + (C + D), the bottom pass single
    when the level count is odd, each rolled operand at its split-at-wrap
    index, -0.0 for a copied operand, padding entries skipped), give the
    oracle's FFA transform of the node bit for bit, every row written exactly
    once per pass."""
    import gaffa_amd

    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    lib.gaffa_hip_ffa_subtree_merges.restype = ctypes.c_uint32
    lib.gaffa_hip_ffa_subtree_merges.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    n = lib.gaffa_hip_ffa_subtree_merges(rows, bins, None)
    levels = int(np.ceil(np.log2(rows))) if rows > 1 else 0
    npasses = (levels + 1) // 2
    assert n % max(npasses, 1) == 0 and (n == 0) == (levels == 0)
    table = (_Merge * max(n, 1))()
    assert lib.gaffa_hip_ffa_subtree_merges(rows, bins, ctypes.cast(table, ctypes.c_void_p)) == n
    field = {name: np.array([getattr(m, name) for m in table[:n]], np.int64)
             for name, _ in _Merge._fields_}
    rng = np.random.default_rng(rows * 1000 + bins)
    x = rng.normal(size=rows * bins).astype(np.float32)
    cur = x.copy()
    per_pass = n // npasses if npasses else 0
    p = np.arange(bins)
    copy = 0xFFFFFFFF

    def rolled(src, base, wrap):
        if base == copy:
            return np.full(bins, -0.0, np.float32)
        return src[np.where(p < wrap, base + p, base + p - bins)]

    for k in range(npasses - 1, -1, -1):
        two = 2 * k + 1 < levels
        nxt = np.full(rows * bins, np.nan, np.float32)
        sl = slice(k * per_pass, (k + 1) * per_pass)
        for e in range(sl.start, sl.stop):
            f = {name: int(v[e]) for name, v in field.items()}
            if f["out"] == copy:  # padding of a wave list: skipped
                continue
            v = cur[f["head"] + p] + rolled(cur, f["tail"], f["wrap"])
            if two:
                v = v + (rolled(cur, f["head2"], f["wrap2"]) + rolled(cur, f["tail2"], f["wrap3"]))
            assert np.isnan(nxt[f["out"]])  # each row once per pass
            nxt[f["out"] + p] = v
        assert not np.isnan(nxt).any()  # every row written at every pass
        cur = nxt
    want = port.ffa_transform(x, bins).reshape(-1)
    assert cur.tobytes() == want.tobytes()


def _pick_row(child_rows, rows, s):
    # the FFA's FP32 row selection (gaffa's ffa.cpp:133-143)
    scale = np.float32(child_rows - 1) / np.float32(rows - 1)
    return int(scale * np.float32(s) + np.float32(0.5))


def _merge_desc(start, size, r, bins):
    s = r - start
    hn = size // 2
    h = _pick_row(hn, size, s)
    t = _pick_row(size - hn, size, s)
    return start + h, start + hn + t, (s - t) % bins


def _node(rows, depth, r):
    start, size = 0, rows
    for _ in range(depth):
        if size <= 1:
            break
        h = size // 2
        if r < start + h:
            size = h
        else:
            start, size = start + h, size - h
    return start, size


def _leaves(rows, depth, levels, r, bins, with_rolls=False):
    """Depth-(depth + levels) rows an output row r of a depth-`depth` node
    sums, in the order of its one-level merges (and each one's roll: the
    tail rolls along its path, mod bins)."""
    out, rolls = [r], [0]
    nodes = [_node(rows, depth, r)]
    for _ in range(levels):
        nxt_rows, nxt_rolls, nxt_nodes = [], [], []
        for row, acc, (start, size) in zip(out, rolls, nodes):
            h, t, roll = _merge_desc(start, size, row, bins)
            hn = size // 2
            nxt_rows += [h, t]
            nxt_rolls += [acc, (acc + roll) % bins]
            nxt_nodes += [(start, hn), (start + hn, size - hn)]
        out, rolls, nodes = nxt_rows, nxt_rolls, nxt_nodes
    return (out, rolls) if with_rolls else out


class _LevelRow(ctypes.Structure):
    _fields_ = [("row", ctypes.c_uint32 * 8), ("roll", ctypes.c_uint32 * 8)]


class _LevelSpan(ctypes.Structure):
    _fields_ = [("row_offset", ctypes.c_uint32), ("staged", ctypes.c_uint32),
                ("first", ctypes.c_uint32 * 8), ("count", ctypes.c_uint32 * 8),
                ("reserved", ctypes.c_uint32 * 2)]


@pytest.mark.parametrize("rows,bins,depth,levels,first,count",
                         [(5000, 200, 5, 3, 160, 32), (4100, 181, 3, 2, 530, 17),
                          (777, 255, 2, 1, 0, 32), (96, 7, 1, 3, 48, 20)])
def test_level_item_tables_match_the_tree(rows, bins, depth, levels, first, count):
    """gaffa_hip_ffa_level_item writes each row's leaves (rows and combined
    rolls, zero past 2^L) and the item's runs as the tree restated above
    gives them."""
    import gaffa_amd

    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    f = lib.gaffa_hip_ffa_level_item
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_uint32] * 6 + [ctypes.c_void_p, ctypes.c_void_p]
    span = _LevelSpan()
    out = (_LevelRow * count)()
    need = f(rows, bins, depth, levels, first, count, ctypes.addressof(span), ctypes.addressof(out))
    nleaf = 1 << levels
    got_rows = np.array([list(r.row) for r in out])
    got_rolls = np.array([list(r.roll) for r in out])
    want = [_leaves(rows, depth, levels, first + i, bins, True) for i in range(count)]
    want_rows = np.array([w[0] for w in want])
    want_rolls = np.array([w[1] for w in want])
    assert np.array_equal(got_rows[:, :nleaf], want_rows)
    assert np.array_equal(got_rolls[:, :nleaf], want_rolls)
    assert not got_rows[:, nleaf:].any() and not got_rolls[:, nleaf:].any()
    lo, hi = want_rows.min(axis=0), want_rows.max(axis=0)
    assert list(span.first)[:nleaf] == lo.tolist()
    assert list(span.count)[:nleaf] == (hi - lo + 1).tolist()
    assert span.staged == int((hi - lo + 1).sum())
    assert need == (span.staged * bins + 256) * 4


@pytest.mark.parametrize("rows,bins,depth,levels", [(5000, 200, 5, 3), (4100, 181, 3, 2),
                                                    (777, 255, 2, 1), (96, 7, 1, 3)])
def test_level_item_lds_need_matches_leaf_runs(rows, bins, depth, levels):
    """gaffa_hip_ffa_level_lds_bytes gives a level item's leaf runs (from
    the smallest to the largest row of each leaf over the item's rows, one
    run per leaf) plus a 256-float guard, the LDS the level kernel stages.
    Inside one node every leaf is monotone in the output row, so a run holds
    only rows the item reads; an item across two nodes needs more."""
    import gaffa_amd

    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    f = lib.gaffa_hip_ffa_level_lds_bytes
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_uint32] * 6
    starts = sorted({_node(rows, depth, r)[0] for r in range(rows)})
    bounds = list(zip(starts, starts[1:] + [rows]))
    checked = 0
    for a, b in bounds[:3] + bounds[-2:]:
        for first in range(a, b, 32):
            count = min(32, b - first)
            leaves = np.array([_leaves(rows, depth, levels, r, bins)
                               for r in range(first, first + count)])
            assert np.all(np.diff(leaves, axis=0) >= 0)  # monotone inside the node
            runs = int((leaves.max(axis=0) - leaves.min(axis=0) + 1).sum())
            assert f(rows, bins, depth, levels, first, count) == (runs * bins + 256) * 4
            checked += 1
    assert checked > 0
    if len(bounds) > 1:  # an item across the first node boundary
        first = bounds[0][1] - 4
        inside = f(rows, bins, depth, levels, first, 4)
        across = f(rows, bins, depth, levels, first, 8)
        assert across > inside
    bad = (1 << 64) - 1
    assert f(rows, bins, depth, 0, 0, 1) == bad and f(rows, bins, depth, 4, 0, 1) == bad
    assert f(rows, bins, depth, levels, rows - 1, 2) == bad  # past the last row
    assert f(rows, bins, depth, levels, 0, 33) == bad        # more rows than an item holds


@pytest.mark.parametrize("nsamples,level_rows", [(1 << 20, 0), (1 << 21, 7)])
def test_level_layout_items_lie_in_nodes_and_carry_their_tables(nsamples, level_rows):
    """The plan builder's level work items each lie inside one node of their
    output depth and cover every row of every pass exactly once; each item's
    span and leaf rows (kept parallel to the items after they are ordered by
    depth) are what gaffa_hip_ffa_level_item gives for it; and the launch LDS
    is the largest item need within 64 KB."""
    import collections

    import gaffa_amd
    from gaffa_amd import _core as core
    from gaffa_amd.ffa import make_riptide_plan

    plan = make_riptide_plan(nsamples=nsamples, tsamp=6.4e-5, period_min=0.1, period_max=10.0,
                             bins_min=180, bins_max=256, duty_cycle_max=0.18,
                             width_trial_spacing=2.0)
    d = core._ffa_level_layout(plan, level_rows)
    items, spans, leaf = d["items"], d["spans"], d["leaf_rows"]
    assert len(items) == len(spans) > 0
    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    f = lib.gaffa_hip_ffa_level_item
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_uint32] * 6 + [ctypes.c_void_p, ctypes.c_void_p]
    tasks = plan.tasks
    cover = collections.defaultdict(list)
    max_need = 0
    for k, (task, a, first, count) in enumerate(items.tolist()):
        depth, levels = a & 0xFFFF, a >> 16
        rows, bins = int(tasks["rows"][task]), int(tasks["bins"][task])
        assert 1 <= levels <= 3 and 1 <= count <= (level_rows or 32)
        assert _node(rows, depth, first) == _node(rows, depth, first + count - 1)
        cover[(task, depth)].append((first, count))
        span = _LevelSpan()
        out = (_LevelRow * count)()
        need = f(rows, bins, depth, levels, first, count, ctypes.addressof(span),
                 ctypes.addressof(out))
        sp = spans[k]
        assert sp[1] == span.staged
        assert sp[2:10].tolist() == list(span.first) and sp[10:18].tolist() == list(span.count)
        want = np.array([list(r.row) + list(r.roll) for r in out], np.uint32)
        assert np.array_equal(leaf[sp[0]: sp[0] + count], want)
        if need <= 64 * 1024:
            max_need = max(max_need, need)
    assert d["level_lds_bytes"] == max_need
    for (task, depth), runs in cover.items():
        pos = 0
        for first, count in sorted(runs):
            assert first == pos
            pos += count
        assert pos == int(tasks["rows"][task])


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    so before it touches a device."""
    import subprocess
    import sys as _sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([_sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and r.stdout == ""
