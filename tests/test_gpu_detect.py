"""The step after the path: the device boxcar peak search (dd_detect_peaks)
against the oracle's straight-loop restatement, and on single-step
dedispersed output, where an injected pulse must be found at its DM and
time."""
import numpy as np
import pytest

import oracle
from dedispersion_amd import PEAK_DTYPE, Plan, detect_peaks, detect_peaks_batch, workloads
from tests.conftest import host_threads
from tests.gpu_helpers import stream_of

pytestmark = pytest.mark.gpu


def run(torch, rows: np.ndarray, S: int, nw: int, offset: int = 0):
    d_rows = torch.from_numpy(np.ascontiguousarray(rows).ravel()).cuda()
    d_peaks = torch.zeros(rows.shape[0] * PEAK_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    detect_peaks(d_rows.data_ptr() + 4 * offset, rows.shape[0], S, rows.shape[1], nw, d_peaks.data_ptr(),
                 stream_of(torch))
    torch.cuda.synchronize()
    return d_peaks.cpu().numpy().view(PEAK_DTYPE)


def compare(got, want, what):
    for k in ("sample", "width"):
        bad = np.flatnonzero(got[k] != want[k])
        assert bad.size == 0, f"{what}: {k} differs at rows {bad[:5]}: {got[k][bad[:5]]} vs {want[k][bad[:5]]}"
    for k in ("snr", "mean", "sigma"):
        np.testing.assert_allclose(got[k], want[k], rtol=1e-6, atol=1e-30, err_msg=f"{what} {k}")


@pytest.mark.parametrize("S,pitch", [(1000, 1024), (20000, 20032), (8192 + 30, 8192 + 30), (32768, 32768), (5, 5)],
                         ids=["1000", "3-segments", "unaligned-pitch", "32768", "tiny"])
@pytest.mark.parametrize("nw", [1, 4, 7])
def test_random_rows_against_oracle(torch_cuda, S, pitch, nw):
    rng = np.random.default_rng(S + nw)
    rows = rng.integers(0, 60000, size=(24, pitch)).astype(np.float32)
    for r in range(0, 24, 3):  # pulses of widths 1..64 across segment borders
        w = 1 << (r % 7)
        t0 = (8192 - w // 2 + 37 * r) % max(1, S - w) if S > w else 0
        rows[r, t0:t0 + w] += 40000
    rows[23, :] = 123.0  # a constant row
    got = run(torch_cuda, rows, S, nw)
    want = oracle.detect_peaks(rows, S, nw)
    compare(got, want, f"S={S} widths={nw}")


# The search walks each row in chunks of 63 runs of 68 starts (4284 starts),
# reading 4352 samples: sizes around those borders, pulses straddling runs
# and chunks, equal maxima in different runs and chunks (the first must win),
# maxima at the row's last window, zero and constant rows.
@pytest.mark.parametrize("S", [1, 67, 68, 131, 4284, 4285, 4352, 8568, 8569, 12859])
@pytest.mark.parametrize("aligned", [True, False], ids=["aligned", "unaligned"])
def test_chunk_and_run_borders(torch_cuda, S, aligned):
    pitch = (S + 3) // 4 * 4 if aligned else S + (1 if S % 4 == 0 else 0)
    rng = np.random.default_rng(S * 2 + aligned)
    rows = rng.integers(0, 200, size=(40, pitch)).astype(np.float32)
    borders = [b for b in (67, 68, 135, 136, 4283, 4284, 4351, 4352, 8567, 8568, 12851, 12852) if b < S]
    for r in range(24):
        w = 1 << (r % 7)
        if borders:
            t0 = max(0, min(S - w, borders[r % len(borders)] - w // 2))
            rows[r, t0:t0 + w] += 5000
    # equal maxima: an impulse in a late run and the same in a later chunk
    if S > 10:
        rows[24, :] = 3
        rows[24, S // 3] = rows[24, S - 2] = 900
    # the row's last window is the best of its width
    rows[25, :] = 1
    rows[25, max(0, S - 64):S] = 400
    rows[26, :] = 0                       # zero row
    rows[27, :] = 9                       # constant row
    rows[28, :] = 0
    rows[28, S - 1] = 50                  # a spike in the last sample only
    got = run(torch_cuda, rows, S, 7)
    want = oracle.detect_peaks(rows, S, 7)
    compare(got, want, f"S={S} pitch={pitch}")


def test_rows_beyond_one_pass_of_the_grid(torch_cuda):
    """More rows than the launch has wavefronts (each takes several rows)."""
    rng = np.random.default_rng(11)
    S, pitch, n = 300, 304, 2 * 256 * 8 + 37
    rows = rng.integers(0, 1000, size=(n, pitch)).astype(np.float32)
    for r in range(0, n, 7):
        w = 1 << (r % 7)
        t0 = (r * 13) % (S - w)
        rows[r, t0:t0 + w] += 3000
    got = run(torch_cuda, rows, S, 7)
    want = oracle.detect_peaks(rows, S, 7)
    compare(got, want, f"{n} rows")


def test_rows_shared_by_wavefronts_over_several_rows(torch_cuda):
    """Few enough rows that each is split over three wavefronts (three chunks
    of 4284 starts), and more rows than workgroups, so a workgroup merges
    several rows in turn."""
    rng = np.random.default_rng(13)
    S, n = 9000, 700
    rows = rng.integers(0, 3000, size=(n, S)).astype(np.float32)
    for r in range(n):
        w = 1 << (r % 7)
        t0 = (r * 4283 + 17) % (S - w)
        rows[r, t0:t0 + w] += 9000
    got = run(torch_cuda, rows, S, 7)
    want = oracle.detect_peaks(rows, S, 7)
    compare(got, want, f"{n} rows of {S}")


def test_long_rows_against_oracle(torch_cuda):
    """Rows of the 40-beam outputs' longest length, values of 1536-channel sums."""
    rng = np.random.default_rng(12)
    S = 32768
    rows = rng.integers(180000, 210000, size=(48, S)).astype(np.float32)
    for r in range(48):
        w = 1 << (r % 7)
        t0 = int(rng.integers(0, S - w))
        rows[r, t0:t0 + w] += 20000
    got = run(torch_cuda, rows, S, 7)
    want = oracle.detect_peaks(rows, S, 7)
    compare(got, want, "32768-sample rows")


@pytest.mark.parametrize("dm_index,t0,width", [(40, 1000, 1), (17, 2500, 6), (55, 3000, 16)])
def test_pulse_found_in_dedispersed_output(torch_cuda, dm_index, t0, width):
    """Single step, 256 channels, 64 DMs of 0.5, 4096 samples of Gaussian
    noise with one dispersed boxcar pulse: its DM row's best window overlaps
    the pulse; the search agrees with the oracle on the rows around it."""
    from dedispersion_amd import DD_SINGLE_STEP, synth_gaussian_u8

    torch = torch_cuda
    w = workloads.tiny(channels=256, samples=4096, dms=64, step=0.5, min_freq=1300.0, bandwidth=0.5, seed=5)
    plan = Plan(w.obs, 128)
    z = plan.sizes
    s = stream_of(torch)
    d_in = torch.zeros(z.input_elements, dtype=torch.uint8, device="cuda")
    synth_gaussian_u8(d_in.data_ptr(), 256, z.nr_dispersed_samples, z.in_pitch, 128.0, 16.0, 77, s)
    dm_value = float(np.float32(dm_index) * np.float32(0.5))
    plan.inject_pulse(d_in.data_ptr(), z.in_pitch, z.nr_dispersed_samples, dm_value, t0, width, 12, s)
    d_out = torch.empty(z.output_elements, dtype=torch.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=s)
    rows = w.obs.nr_dms
    d_peaks = torch.zeros(rows * PEAK_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    detect_peaks(d_out.data_ptr(), rows, z.nr_samples, z.out_pitch, 7, d_peaks.data_ptr(), s)
    torch.cuda.synchronize()
    peaks = d_peaks.cpu().numpy().view(PEAK_DTYPE)
    out = d_out.cpu().numpy().reshape(rows, z.out_pitch)
    near = list(range(max(0, dm_index - 2), min(rows, dm_index + 3)))
    compare(peaks[near], oracle.detect_peaks(out[near], z.nr_samples, 7), "dedispersed rows")
    pk = peaks[dm_index]
    start, found = int(pk["sample"]), int(pk["width"])
    assert pk["snr"] > 10, pk
    assert start < t0 + width and t0 < start + found, pk
    assert np.argmax(peaks["snr"]) in near, peaks["snr"][near]
    plan.close()


def _batch_case(torch, shapes, seed):
    """Device arrays of random rows with pulses, one per (rows, S, pitch)."""
    rng = np.random.default_rng(seed)
    host, dev, peaks = [], [], []
    for n, S, pitch in shapes:
        rows = rng.integers(0, 5000, size=(max(n, 1), pitch)).astype(np.float32)
        for r in range(n):
            w = 1 << (r % 7)
            if S > w:
                t0 = int(rng.integers(0, S - w))
                rows[r, t0:t0 + w] += 20000
        host.append(rows)
        dev.append(torch.from_numpy(rows.ravel()).cuda())
        peaks.append(torch.zeros(max(n, 1) * PEAK_DTYPE.itemsize, dtype=torch.uint8, device="cuda"))
    jobs = [(d.data_ptr(), n, S, pitch, pk.data_ptr()) for d, pk, (n, S, pitch) in zip(dev, peaks, shapes)]
    return host, dev, peaks, jobs


def test_batch_matches_single_sets_and_oracle(torch_cuda):
    """One batch over sets of different lengths, pitches (aligned and not)
    and row counts (one empty): each set's peaks equal one dd_detect_peaks
    call's, record for record, and the oracle's."""
    torch = torch_cuda
    shapes = [(40, 8192, 8192), (7, 30000, 30001), (0, 100, 100), (300, 4285, 4288), (3, 1, 1), (25, 12859, 12860),
              (64, 700, 703)]
    host, dev, peaks, jobs = _batch_case(torch, shapes, 21)
    detect_peaks_batch(jobs, 7, stream_of(torch))
    torch.cuda.synchronize()
    for (ptr, n, S, pitch, pk), rows, d in zip(jobs, host, peaks):
        if n == 0:
            assert not d.any().item()
            continue
        got = d.cpu().numpy().view(PEAK_DTYPE)
        single = torch.zeros_like(d)
        detect_peaks(ptr, n, S, pitch, 7, single.data_ptr(), stream_of(torch))
        torch.cuda.synchronize()
        assert np.array_equal(got.view(np.uint8), single.cpu().numpy()), f"S={S}: batch differs from a single call"
        compare(got, oracle.detect_peaks(rows[:n], S, 7), f"batch set S={S} pitch={pitch}")


def test_batch_of_more_sets_than_one_launch_holds(torch_cuda):
    """150 sets (three launches of at most 64), widths 1-16."""
    torch = torch_cuda
    rng = np.random.default_rng(22)
    shapes = [(int(rng.integers(1, 6)), int(S), int(S) + int(rng.integers(0, 4)))
              for S in rng.integers(50, 9000, size=150)]
    host, dev, peaks, jobs = _batch_case(torch, shapes, 23)
    detect_peaks_batch(jobs, 5, stream_of(torch))
    torch.cuda.synchronize()
    for (ptr, n, S, pitch, pk), rows, d in zip(jobs, host, peaks):
        compare(d.cpu().numpy().view(PEAK_DTYPE), oracle.detect_peaks(rows[:n], S, 5), f"S={S} pitch={pitch}")


def test_random_sets_against_oracle(torch_cuda):
    """25 seeded random batches of one to four sets: lengths 1 to 40000,
    pitches up to four floats past the row, 0-24 rows, value ranges from
    0-1 to 0-400000, and rows with a boxcar pulse, a constant value, two
    equal impulses or a spike in the last sample; random width counts.
    Batch, single calls and the oracle agree."""
    torch = torch_cuda
    rng = np.random.default_rng(31)
    for case in range(25):
        shapes, hosts = [], []
        for _ in range(int(rng.integers(1, 5))):
            S = int(rng.choice([rng.integers(1, 200), rng.integers(200, 9000), rng.integers(9000, 40000)]))
            pitch = S + int(rng.integers(0, 5))
            n = int(rng.integers(0, 25))
            hi = int(rng.choice([2, 50, 5000, 400000]))
            rows = rng.integers(0, hi, size=(max(n, 1), pitch)).astype(np.float32)
            for r in range(n):
                kind = int(rng.integers(0, 4))
                if kind == 0 and S > 1:
                    w = 1 << int(rng.integers(0, 7))
                    t0 = int(rng.integers(0, max(1, S - w)))
                    rows[r, t0:t0 + w] += 3 * hi
                elif kind == 1:
                    rows[r, :] = float(rng.integers(0, hi))
                elif kind == 2 and S > 2:
                    a, b = rng.choice(S, 2, replace=False)
                    rows[r, :] = 0
                    rows[r, a] = rows[r, b] = hi
                else:
                    rows[r, S - 1] += 5 * hi
            shapes.append((n, S, pitch))
            hosts.append(rows)
        nw = int(rng.integers(1, 8))
        dev = [torch.from_numpy(h.ravel()).cuda() for h in hosts]
        peaks = [torch.zeros(max(n, 1) * PEAK_DTYPE.itemsize, dtype=torch.uint8, device="cuda") for n, _, _ in shapes]
        detect_peaks_batch([(d.data_ptr(), n, S, pitch, pk.data_ptr()) for d, pk, (n, S, pitch) in
                            zip(dev, peaks, shapes)], nw, stream_of(torch))
        torch.cuda.synchronize()
        for d, pk, h, (n, S, pitch) in zip(dev, peaks, hosts, shapes):
            if n == 0:
                continue
            single = torch.zeros_like(pk)
            detect_peaks(d.data_ptr(), n, S, pitch, nw, single.data_ptr(), stream_of(torch))
            torch.cuda.synchronize()
            got = pk.cpu().numpy().view(PEAK_DTYPE)
            assert np.array_equal(got.view(np.uint8), single.cpu().numpy()), f"case {case} S={S}: batch vs single"
            compare(got, oracle.detect_peaks(h[:n], S, nw), f"case {case} S={S} pitch={pitch} widths={nw}")


def test_batch_rejects_bad_sets(torch_cuda):
    from dedispersion_amd import DDError

    x = torch_cuda.zeros(256, dtype=torch_cuda.float32, device="cuda")
    good = (x.data_ptr(), 1, 64, 64, x.data_ptr())
    with pytest.raises(DDError, match="nr_widths"):
        detect_peaks_batch([good], 8)
    with pytest.raises(DDError, match="job 1: need 0 < nr_samples <= pitch"):
        detect_peaks_batch([good, (x.data_ptr(), 1, 65, 64, x.data_ptr())], 3)
    with pytest.raises(DDError, match="job 0: null argument"):
        detect_peaks_batch([(0, 1, 64, 64, x.data_ptr())], 3)


@pytest.mark.parametrize("nbytes", [16, 1 << 20, (1 << 26) + 48])
def test_read_probe_reads_every_byte_once(torch_cuda, nbytes):
    """The read probe behind the bench's read ceiling: its per-thread partial
    sums add up to the buffer's sum (small integers: every sum exact), so it
    reads every 16-byte piece exactly once."""
    from dedispersion_amd import check, load

    torch = torch_cuda
    n = nbytes // 4
    x = (torch.arange(n, device="cuda", dtype=torch.int64) % 7).to(torch.float32)
    sink = torch.full((2097152,), -1.0, dtype=torch.float32, device="cuda")
    check(load().dd_read_probe(x.data_ptr(), nbytes, sink.data_ptr(), stream_of(torch)))
    torch.cuda.synchronize()
    got = sink.double().sum().item()
    want = float((torch.arange(n, dtype=torch.int64) % 7).sum().item())
    assert got == want, (got, want)


def test_rejects_bad_arguments(torch_cuda):
    from dedispersion_amd import DDError

    x = torch_cuda.zeros(64, dtype=torch_cuda.float32, device="cuda")
    with pytest.raises(DDError, match="nr_widths"):
        detect_peaks(x.data_ptr(), 1, 64, 64, 8, x.data_ptr())
    with pytest.raises(DDError, match="nr_samples"):
        detect_peaks(x.data_ptr(), 1, 65, 64, 3, x.data_ptr())
